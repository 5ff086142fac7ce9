"""Batched fragments: many AFQMC problems of uneven size advanced together.

An LNO-style calculation (lno_afqmc/lno_afqmc.py:442-605) runs one phaseless AFQMC per
fragment, each a small problem (norb, nocc, nchol, walkers all different).  Here every
fragment keeps its own device context and resident state, and one grouped step
(``DeviceGroup.propagate``, afqmc_group_propagate) advances all of them: each phase of
the step is a single launch over every fragment's walkers, so small fragments fill the
GPU together.

A fragment's block follows sampler.propagate_phaseless (sampling.py:319-341) with
n_ene_blocks = 1: fields for n_prop_steps steps are drawn on the device from the
fragment's key, the steps run grouped, then QR, overlaps, local energies and the block
statistics run per fragment, and stochastic reconfiguration closes the block.  Each
fragment's result depends only on its own inputs and seed: a fragment gives bit-identical
results alone and in any mix (no split-K, fixed summation order in every kernel).
"""

import numpy as np
import torch

from . import hamiltonian, propagation, rng, wavefunctions
from .device import DeviceGroup


class Fragment:
    """One AFQMC problem with resident device state.

    ham_data / wave_data: the reference's dicts (h0, h1 (2,N,N), chol (G,N*N); mo_coeff,
    rdm1); walker_type "rhf" or "uhf"; nelec (n_up, n_dn)."""

    def __init__(self, ham_data, wave_data, walker_type, nelec, n_walkers, dt=0.005, seed=0,
                 precision="bf16x3", n_prop_steps=50, n_exp_terms=6, tag=None):
        norb = int(np.asarray(ham_data["h1"]).shape[-1])
        self.tag = tag
        self.norb, self.nelec, self.n_walkers = norb, tuple(nelec), int(n_walkers)
        self.n_prop_steps = int(n_prop_steps)
        if walker_type == "rhf":
            self.trial = wavefunctions.rhf(norb, tuple(nelec))
            self.prop = propagation.propagator_restricted(dt, n_walkers, n_exp_terms,
                                                          precision=precision)
        else:
            self.trial = wavefunctions.uhf(norb, tuple(nelec))
            self.prop = propagation.propagator_unrestricted(dt, n_walkers, n_exp_terms,
                                                            precision=precision)
        self.wave_data = dict(wave_data)
        ham = hamiltonian.hamiltonian(norb)
        hd = ham.build_measurement_intermediates(dict(ham_data), self.trial, self.wave_data)
        hd = ham.build_propagation_intermediates(hd, self.prop, self.trial, self.wave_data)
        self.ham_data = hd
        self.dev = hd["_afqmc_device"]
        pd = self.prop.init_prop_data(self.trial, self.wave_data, hd)
        pd["key"] = rng.PRNGKey(seed)
        self.prop_data = pd
        ddev = self.dev.device
        e0 = pd["e_estimate"].reshape(()).to(torch.float64)
        self.scalars = torch.stack([e0, e0]).contiguous()
        pd["_scalars"] = self.scalars
        self.fields = torch.empty((self.n_prop_steps, self.n_walkers, self.dev.nchol),
                                  dtype=torch.float64, device=ddev)
        self.energies = torch.empty(self.n_walkers, dtype=torch.complex128, device=ddev)
        self.stats = []  # one (3,) device tensor per finished block
        self.blocks_done = 0
        self.dev.overlap(pd["walkers"], out=pd["overlaps"])

    @property
    def walker_steps(self):
        return self.n_walkers * self.n_prop_steps * self.blocks_done

    def device_bytes(self):
        fields = self.fields.numel() * 8
        walkers = sum(w.numel() * 16 for w in (self.prop_data["walkers"]
                      if isinstance(self.prop_data["walkers"], list)
                      else [self.prop_data["walkers"]]))
        return int(self.dev.device_bytes()) + fields + walkers

    def begin_block(self):
        """sampling.py:60-69: split the key and draw this block's fields."""
        pd = self.prop_data
        pd["key"], subkey = rng.split(pd["key"])
        self.dev.fields_normal(subkey[0], subkey[1], self.fields.shape, out=self.fields)

    def end_block(self):
        """sampling.py:74-91 and :142: QR, overlaps, energies, block statistics, SR."""
        pd = self.prop_data
        self.dev.orthonormalize(pd["walkers"])
        self.dev.overlap(pd["walkers"], out=pd["overlaps"])
        en = self.dev.energy(pd["walkers"], out=self.energies)
        st = torch.zeros(3, dtype=torch.float64, device=self.dev.device)
        self.dev.block_stats(en, pd["weights"], self.scalars, st)
        self.stats.append(st)
        self.prop.stochastic_reconfiguration_local(pd)
        self.dev.overlap(pd["walkers"], out=pd["overlaps"])
        self.blocks_done += 1

    def block_energies(self):
        """(block energies, block weights) as numpy arrays (synchronises)."""
        if not self.stats:
            return np.zeros(0), np.zeros(0)
        st = torch.stack(self.stats).cpu().numpy()
        return st[:, 0], st[:, 1]

    def energy(self):
        be, bw = self.block_energies()
        return float(np.sum(be * bw) / np.sum(bw)) if be.size else float("nan")


class FragmentBatch:
    """A set of fragments advanced by grouped steps (one group, one stream).

    With use_graph, the n_prop_steps grouped steps of a block are captured once into a
    HIP graph (torch.cuda.CUDAGraph over the batch's stream, which is then a stream of its
    own unless one is given) and replayed every block while the member set is unchanged;
    the first block after a change runs eagerly (it rebuilds the group's tables), the
    next one captures.  The caller's current stream waits for each block."""

    def __init__(self, fragments=(), max_fragments=1024, tune=False, use_graph=False,
                 stream=None):
        # graph capture needs a stream other than the default one (default priority)
        if stream is None:
            stream = torch.cuda.Stream() if use_graph else torch.cuda.current_stream()
        self.stream = stream
        with torch.cuda.stream(self.stream):
            self.group = DeviceGroup(max_fragments)
        self.fragments = []
        self.use_graph = use_graph
        self._version = 0        # bumped on every membership change
        self._warm_version = -1  # version whose tables an eager block has built
        self._graph = None
        self._graph_version = -1
        for f in fragments:
            self.add(f)
        self.tuned = None
        if tune and self.fragments:
            with torch.cuda.stream(self.stream):
                self.tuned = self.group.tune()

    def add(self, frag):
        if self.fragments and frag.n_prop_steps != self.fragments[0].n_prop_steps:
            raise ValueError("fragments of a batch share n_prop_steps")
        pd = frag.prop_data
        # work that set the fragment up on the caller's stream comes first
        self.stream.wait_stream(torch.cuda.current_stream())
        self.group.add(frag.dev, pd["walkers"], pd["weights"], pd["overlaps"], frag.fields,
                       frag.scalars)
        self.fragments.append(frag)
        self._version += 1

    def remove(self, frags):
        """Drop finished fragments (the group's tables are rebuilt on the next step)."""
        gone = {id(f) for f in frags}
        keep = [f for f in self.fragments if id(f) not in gone]
        self.group.clear()
        self.fragments = []
        for f in keep:
            self.add(f)

    def _steps(self):
        for s in range(self.fragments[0].n_prop_steps):
            self.group.propagate(s, reuse_green=s > 0)

    def run_block(self):
        """One block of every fragment: fields, n_prop_steps grouped steps, block end."""
        if not self.fragments:
            return
        with torch.cuda.stream(self.stream):
            for f in self.fragments:
                f.begin_block()
            if not self.use_graph or self._warm_version != self._version:
                self._steps()
                self._warm_version = self._version
            else:
                if self._graph_version != self._version:
                    self._graph = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(self._graph, stream=self.stream):
                        self._steps()
                    self._graph_version = self._version
                self._graph.replay()
            for f in self.fragments:
                f.end_block()
        # later work on the caller's stream (reading results) follows the block
        torch.cuda.current_stream().wait_stream(self.stream)
