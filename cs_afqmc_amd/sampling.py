"""Block sampler with the reference's API, device resident.

Mirrors ad_afqmc/sampling.py ``sampler`` (:15-20) and ``propagate_phaseless``
(:319-341) -> ``_sr_block_scan`` (:126-144) -> ``_block_scan`` (:50-92) ->
``_step_scan`` (:22-34).  Walkers, weights, overlaps and the energy scalars stay on the
GPU for the whole call.  The key lives on the device as well: each block draws its
fields (n_prop_steps, W, nchol) from it (Philox4x32-10 + Box-Muller in place of
jax.random.normal) and each stochastic reconfiguration its uniform draw, so one block
-- fields, n_prop_steps steps, QR, overlaps, local energies, block statistics -- is a
fixed launch sequence that is captured once into a HIP graph and replayed (the
reference jit-compiles the same scan into one XLA program).  The host synchronises
once per call, to return the block-averaged energy; prop_data["key"] (host) advances
by one per block and one per resampling, as random.split does in the reference.
"""

from dataclasses import dataclass

import torch

from . import _lib, rng
from .graphs import Replay


class _BlockState:
    """Fixed device buffers and captured graphs of one (context, walker buffers,
    n_prop_steps) combination."""

    def __init__(self, dev, prop_data, n_prop_steps, W, use_graph):
        ddev = dev.device
        self.dev = dev
        self.n_prop_steps = n_prop_steps
        self.scalars = torch.zeros(2, dtype=torch.float64, device=ddev)
        self.energies = torch.empty(W, dtype=torch.complex128, device=ddev)
        self.fields = torch.empty((n_prop_steps, W, dev.nchol), dtype=torch.float64, device=ddev)
        self.key = torch.zeros(2, dtype=torch.int64, device=ddev)
        self.slot = torch.zeros(_lib.BLOCK_STATS_LEN, dtype=torch.float64, device=ddev)
        self.idx = torch.zeros(W, dtype=torch.int32, device=ddev)  # the last SR's indices
        self.walkers = prop_data["walkers"]
        self.weights = prop_data["weights"]
        self.overlaps = prop_data["overlaps"]
        self.block = Replay(self._block, use_graph)
        self.sr = Replay(self._sr, use_graph)
        self.sig = self.signature(dev, prop_data, n_prop_steps)

    @staticmethod
    def signature(dev, prop_data, n_prop_steps):
        w = prop_data["walkers"]
        ws = w if isinstance(w, (list, tuple)) else [w]
        return (id(dev), n_prop_steps, tuple(x.data_ptr() for x in ws),
                prop_data["weights"].data_ptr(), prop_data["overlaps"].data_ptr())

    def _block(self):
        """sampling.py:60-91: split the key and draw the fields, n_prop_steps phaseless
        steps, QR, overlaps, local energies, block statistics (the energy clamp, block
        energy / weight and the population-control mix) into self.slot."""
        dev = self.dev
        dev.fields_normal_key(self.key, self.fields)
        for s in range(self.n_prop_steps):
            dev.propagate(self.walkers, self.weights, self.overlaps, self.fields[s], self.scalars,
                          reuse_green=s > 0)
        dev.orthonormalize(self.walkers)
        dev.overlap(self.walkers, out=self.overlaps)
        en = dev.energy(self.walkers, out=self.energies)
        dev.block_stats(en, self.weights, self.scalars, self.slot)

    def _sr(self):
        """sampling.py:142-143 (propagation.py:234-241): local SR with the key's next
        uniform draw, then the overlaps of the resampled walkers."""
        self.dev.sr_key(self.walkers, self.weights, self.key, idx_out=self.idx)
        self.dev.overlap(self.walkers, out=self.overlaps)


@dataclass
class sampler:
    n_prop_steps: int = 50
    n_ene_blocks: int = 50
    n_sr_blocks: int = 1
    n_blocks: int = 50
    use_graph: bool = True

    def _state(self, dev, prop, prop_data):
        st = prop_data.get("_block_state")
        sig = _BlockState.signature(dev, prop_data, self.n_prop_steps)
        if st is None or st.sig != sig or st.block.use_graph != self.use_graph:
            if st is not None:
                st.block.release()
                st.sr.release()
            st = _BlockState(dev, prop_data, self.n_prop_steps, prop.n_walkers, self.use_graph)
            prop_data["_block_state"] = st
        return st

    def propagate_phaseless(self, ham, ham_data, prop, prop_data, trial, wave_data):
        """sampling.py:319-341.  Returns (block-averaged energy as a float, prop_data)."""
        dev = prop._device(ham_data)
        prop_data["_afqmc_device"] = dev
        W = prop.n_walkers
        st = self._state(dev, prop, prop_data)
        e_est = prop_data["e_estimate"]
        e_est = e_est if isinstance(e_est, torch.Tensor) else torch.tensor(float(e_est))
        e_est = e_est.to(dev.device, torch.float64).reshape(())
        # pop_control_ene_shift starts each call at e_estimate (sampling.py:327-329)
        st.scalars[0].copy_(e_est)
        st.scalars[1].copy_(e_est)
        prop_data["_scalars"] = st.scalars
        prop_data["_energies"] = st.energies
        rng.device_key(prop_data["key"], dev.device, out=st.key)
        dev.overlap(prop_data["walkers"], out=prop_data["overlaps"])
        nblk = self.n_sr_blocks * self.n_ene_blocks
        stats = torch.zeros((nblk, _lib.BLOCK_STATS_LEN), dtype=torch.float64, device=dev.device)
        b = 0
        for _ in range(self.n_sr_blocks):
            for _ in range(self.n_ene_blocks):
                st.block()
                stats[b].copy_(st.slot)
                b += 1
            st.sr()
        prop_data["key"] = rng.advance(prop_data["key"], nblk + self.n_sr_blocks)
        prop_data["pop_control_ene_shift"] = st.scalars[1].clone()
        be, bw = stats[:, 0], stats[:, 1]
        prop_data["n_killed_walkers"] = stats[:, 2].sum() / (nblk * W)
        prop_data["block_energies"] = be
        prop_data["block_weights"] = bw
        prop_data["block_stats"] = stats
        energy = float((torch.sum(be * bw) / torch.sum(bw)).item())
        return energy, prop_data
