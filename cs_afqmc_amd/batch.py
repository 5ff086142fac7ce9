"""Batched fragments: many AFQMC problems of uneven size advanced together.

An LNO-style calculation (lno_afqmc/lno_afqmc.py:442-605) runs one phaseless AFQMC per
fragment, each a small problem (norb, nocc, nchol, walkers all different).  Here every
fragment keeps its own device context and resident state, and one grouped step
(``DeviceGroup.propagate``, afqmc_group_propagate) advances all of them: each phase of
the step is a single launch over every fragment's walkers, so small fragments fill the
GPU together.  At the block end the QR and the overlaps are grouped launches too; each
fragment's energy, statistics and resampling run on side streams, where they overlap
one another (faster in the captured block than the grouped energy chain).

A fragment's block is one sampler call of the fragment driver
(lno_afqmc/ccsd_pt2/run_afqmc_nompi.py:83-90 around sampling.py:319-341 with one
energy block): fields for n_prop_steps steps are drawn on the device from the
fragment's device-resident key, the steps run grouped, then QR, overlaps, local
energies and the block statistics, stochastic reconfiguration with the key's next
uniform draw, and the driver's e_estimate <- 0.9 e_estimate + 0.1 E_block.  Nothing in
a block reads back to the host, so the whole block of every fragment -- draws, grouped
steps, block ends -- is captured once into a HIP graph and replayed while the member
set is unchanged.  Each fragment's result depends only on its own inputs and seed: a
fragment gives bit-identical results alone and in any mix (no split-K, fixed summation
order in every kernel).
"""

import numpy as np
import torch

from . import _lib, hamiltonian, propagation, rng, wavefunctions
from .device import DeviceGroup
from .graphs import Replay
from .prop_unrestricted import propagation as pu_prop
from .prop_unrestricted import wavefunctions_unrestricted as pu_wf


class Fragment:
    """One AFQMC problem with resident device state.

    ham_data / wave_data: the reference's dicts (h0, h1 (2,N,N), chol (G,N*N); mo_coeff,
    rdm1); walker_type "rhf" or "uhf"; nelec (n_up, n_dn).  Unrestricted fragments whose
    chol is (2, G, N*N) -- one set of vectors per spin, the layout of the reference's
    prop_unrestricted package -- step with that package's trial and propagator."""

    def __init__(self, ham_data, wave_data, walker_type, nelec, n_walkers, dt=0.005, seed=0,
                 precision="bf16x3", n_prop_steps=50, n_exp_terms=6, tag=None):
        norb = int(np.asarray(ham_data["h1"]).shape[-1])
        self.tag = tag
        self.walker_type = walker_type
        self.norb, self.nelec, self.n_walkers = norb, tuple(nelec), int(n_walkers)
        self.n_prop_steps = int(n_prop_steps)
        if walker_type == "rhf":
            self.trial = wavefunctions.rhf(norb, tuple(nelec))
            self.prop = propagation.propagator_restricted(dt, n_walkers, n_exp_terms,
                                                          precision=precision)
        elif len(getattr(ham_data["chol"], "shape", np.shape(ham_data["chol"]))) == 3:
            self.trial = pu_wf.uhf(norb, tuple(nelec))
            self.prop = pu_prop.propagator_unrestricted(dt, n_walkers, n_exp_terms,
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
        self.key = rng.device_key(pd["key"], ddev)
        self.fields = torch.empty((self.n_prop_steps, self.n_walkers, self.dev.nchol),
                                  dtype=torch.float64, device=ddev)
        self.energies = torch.empty(self.n_walkers, dtype=torch.complex128, device=ddev)
        self.slot = torch.zeros(_lib.BLOCK_STATS_LEN, dtype=torch.float64, device=ddev)
        self.idx = torch.zeros(self.n_walkers, dtype=torch.int32, device=ddev)
        self.stats = []  # one (BLOCK_STATS_LEN,) device tensor per finished block
        self.sr_indices = []  # the resampling indices of each finished block (device)
        self.blocks_done = 0
        self._single = None  # captured single-context block (the plain form)
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
        """sampling.py:60-69: split the (device) key and draw this block's fields."""
        self.dev.fields_normal_key(self.key, self.fields)

    def end_block(self):
        """sampling.py:74-91 and :142-143, then the driver's estimate update
        (run_afqmc_nompi.py:89): QR, overlaps, energies, block statistics into
        self.slot, SR, overlaps, e_estimate mix.  Host bookkeeping is in record()."""
        pd = self.prop_data
        self.dev.orthonormalize(pd["walkers"])
        self.dev.overlap(pd["walkers"], out=pd["overlaps"])
        self.dev.energy(pd["walkers"], out=self.energies)
        self.end_block_tail()
        self.dev.overlap(pd["walkers"], out=pd["overlaps"])

    def end_block_energy_tail(self):
        self.dev.energy(self.prop_data["walkers"], out=self.energies)
        self.end_block_tail()

    def end_block_tail(self):
        """The part of end_block between the energies and the last overlaps: block
        statistics, SR, e_estimate mix (the batch runs the rest as grouped launches)."""
        pd = self.prop_data
        self.dev.block_stats(self.energies, pd["weights"], self.scalars, self.slot)
        self.dev.sr_key(pd["walkers"], pd["weights"], self.key, idx_out=self.idx)
        self.dev.estimate_mix(self.scalars, self.slot)

    def record(self, stats_row, sr_idx=None):
        """Host side of a finished block: keep its statistics row and resampling indices
        (device tensors), advance the host copy of the key by the block's two draws."""
        self.stats.append(stats_row)
        if sr_idx is not None:
            self.sr_indices.append(sr_idx)
        self.prop_data["key"] = rng.advance(self.prop_data["key"], 2)
        self.blocks_done += 1

    def _steps_single(self):
        pd = self.prop_data
        for s in range(self.n_prop_steps):
            self.dev.propagate(pd["walkers"], pd["weights"], pd["overlaps"], self.fields[s],
                               self.scalars, reuse_green=s > 0)

    def _block_single(self):
        self.begin_block()
        self._steps_single()
        self.end_block()

    def run_block_single(self, use_graph=True):
        """The plain (ungrouped) form: this fragment's block through its own context
        (afqmc_propagate per step), captured once and replayed."""
        if self._single is None or self._single.use_graph != use_graph:
            if self._single is not None:
                self._single.release()
            self._single = Replay(self._block_single, use_graph)
        self._single()
        self.record(self.slot.clone(), self.idx.clone())

    def block_energies(self):
        """(block energies, block weights) as numpy arrays (synchronises)."""
        st = self.block_stats()
        return st[:, 0], st[:, 1]

    def block_stats(self):
        """(blocks, BLOCK_STATS_LEN) numpy array: energy, weight, zero weights, clamped."""
        if not self.stats:
            return np.zeros((0, _lib.BLOCK_STATS_LEN))
        return torch.stack(self.stats).cpu().numpy()

    def energy(self):
        be, bw = self.block_energies()
        return float(np.sum(be * bw) / np.sum(bw)) if be.size else float("nan")


class FragmentBatch:
    """A set of fragments advanced by grouped steps (one group, one stream).

    With use_graph, a block of every fragment (fields, n_prop_steps grouped steps, block
    ends) is captured once into a HIP graph and replayed every block while the member
    set is unchanged; the first block after a change runs eagerly (it rebuilds the
    group's tables), the next one captures.  The caller's current stream waits for
    each block."""

    def __init__(self, fragments=(), max_fragments=1024, tune=False, use_graph=False,
                 stream=None, keep_sr_indices=False, side_streams=4):
        # graph capture needs a stream other than the default one (default priority)
        if stream is None:
            stream = torch.cuda.Stream() if use_graph else torch.cuda.current_stream()
        self.stream = stream
        # the per-fragment draws and block ends are independent of one another: they are
        # spread over side streams (default priority) that fork from and join back into
        # the batch stream, so the small launches of different fragments can overlap
        self._side = [torch.cuda.Stream() for _ in range(max(0, int(side_streams)))]
        with torch.cuda.stream(self.stream):
            self.group = DeviceGroup(max_fragments)
        self.fragments = []
        self.use_graph = use_graph
        self.keep_sr_indices = keep_sr_indices  # each block's SR indices, per fragment
        self._replay = None
        self.tuned = None
        for f in fragments:
            self.add(f)
        if tune and self.fragments:
            self.tune()

    @property
    def _graph(self):
        return None if self._replay is None else self._replay.graph

    def tune(self):
        """Pick GEMM tiles for the current members by timing (afqmc_group_tune)."""
        self._drop_replay()  # re-capture with the chosen tiles
        with torch.cuda.stream(self.stream):
            self.tuned = self.group.tune()
        return self.tuned

    def _drop_replay(self):
        if self._replay is not None:
            self._replay.release()
        self._replay = None

    def add(self, frag):
        if self.fragments and frag.n_prop_steps != self.fragments[0].n_prop_steps:
            raise ValueError("fragments of a batch share n_prop_steps")
        pd = frag.prop_data
        # work that set the fragment up on the caller's stream comes first
        self.stream.wait_stream(torch.cuda.current_stream())
        self._drop_replay()
        self.group.add(frag.dev, pd["walkers"], pd["weights"], pd["overlaps"], frag.fields,
                       frag.scalars)
        self.fragments.append(frag)

    def remove(self, frags):
        """Drop finished fragments (the group's tables are rebuilt on the next step)."""
        gone = {id(f) for f in frags}
        keep = [f for f in self.fragments if id(f) not in gone]
        self._drop_replay()
        self.group.clear()
        self.fragments = []
        for f in keep:
            self.add(f)

    def _per_fragment(self, fn):
        """fn(fragment) for every fragment, round robin over the side streams, which
        wait for the work queued so far and are joined back before returning."""
        if not self._side:
            for f in self.fragments:
                fn(f)
            return
        main = torch.cuda.current_stream()
        fork = torch.cuda.Event()
        fork.record(main)
        used = self._side[:len(self.fragments)]
        for st in used:
            st.wait_event(fork)
        for i, f in enumerate(self.fragments):
            with torch.cuda.stream(used[i % len(used)]):
                fn(f)
        for st in used:
            join = torch.cuda.Event()
            join.record(st)
            main.wait_event(join)

    def _block(self):
        self.group.fields_normal_key([f.key for f in self.fragments])
        for s in range(self.fragments[0].n_prop_steps):
            self.group.propagate(s, reuse_green=s > 0)
        # block ends: QR and overlaps of every fragment as grouped launches, then each
        # fragment's energy, statistics, SR and estimate on the side streams, and the
        # overlaps of the resampled walkers grouped again
        self.group.orthonormalize()
        self.group.overlap()
        self._per_fragment(lambda f: f.end_block_energy_tail())
        self.group.overlap()

    def run_block(self):
        """One block of every fragment: fields, n_prop_steps grouped steps, block ends."""
        if not self.fragments:
            return
        with torch.cuda.stream(self.stream):
            if self._replay is None:
                self._replay = Replay(self._block, self.use_graph)
            self._replay()
            rows = torch.stack([f.slot for f in self.fragments])
            idx = [f.idx.clone() for f in self.fragments] if self.keep_sr_indices else None
            for i, f in enumerate(self.fragments):
                f.record(rows[i], None if idx is None else idx[i])
        # later work on the caller's stream (reading results) follows the block
        torch.cuda.current_stream().wait_stream(self.stream)
