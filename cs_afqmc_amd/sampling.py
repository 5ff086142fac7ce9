"""Block sampler with the reference's API, device resident.

Mirrors ad_afqmc/sampling.py ``sampler`` (:15-20) and ``propagate_phaseless``
(:319-341) -> ``_sr_block_scan`` (:126-144) -> ``_block_scan`` (:50-92) ->
``_step_scan`` (:22-34).  Walkers, weights, overlaps and the energy scalars stay on the
GPU for the whole call; auxiliary fields are drawn on the device per block (shape
(n_prop_steps, W, nchol), Philox4x32-10 + Box-Muller keyed by prop_data["key"]) in
place of jax.random.normal.  The host synchronises once per call, to return the
block-averaged energy.
"""

from dataclasses import dataclass

import torch

from . import rng


@dataclass
class sampler:
    n_prop_steps: int = 50
    n_ene_blocks: int = 50
    n_sr_blocks: int = 1
    n_blocks: int = 50

    def _fields(self, dev, prop, key, out=None):
        seed, counter = key
        return dev.fields_normal(seed, counter, (self.n_prop_steps, prop.n_walkers, dev.nchol),
                                 out=out)

    def _block_scan(self, ham_data, prop, prop_data, trial, wave_data, stats_out, fields_buf):
        """sampling.py:50-92: n_prop_steps phaseless steps, then QR, overlaps, local
        energies, clamped block energy / weight and the population-control mix."""
        dev = prop._device(ham_data)
        prop_data["key"], subkey = rng.split(prop_data["key"])
        fields = self._fields(dev, prop, subkey, out=fields_buf)
        scalars = prop_data["_scalars"]
        walkers = prop_data["walkers"]
        for s in range(self.n_prop_steps):
            dev.propagate(walkers, prop_data["weights"], prop_data["overlaps"], fields[s],
                          scalars, reuse_green=s > 0)
        dev.orthonormalize(walkers)
        dev.overlap(walkers, out=prop_data["overlaps"])
        energies = dev.energy(walkers, out=prop_data["_energies"])
        dev.block_stats(energies, prop_data["weights"], scalars, stats_out)
        return prop_data

    def propagate_phaseless(self, ham, ham_data, prop, prop_data, trial, wave_data):
        """sampling.py:319-341.  Returns (block-averaged energy as a float, prop_data)."""
        dev = prop._device(ham_data)
        prop_data["_afqmc_device"] = dev
        W = prop.n_walkers
        ddev = dev.device
        e_est = prop_data["e_estimate"]
        e_est = e_est if isinstance(e_est, torch.Tensor) else torch.tensor(float(e_est))
        scalars = torch.stack([e_est.to(ddev, torch.float64).reshape(()),
                               e_est.to(ddev, torch.float64).reshape(())]).contiguous()
        prop_data["_scalars"] = scalars
        if "_energies" not in prop_data or prop_data["_energies"].shape[0] != W:
            prop_data["_energies"] = torch.empty(W, dtype=torch.complex128, device=ddev)
        dev.overlap(prop_data["walkers"], out=prop_data["overlaps"])
        nblk = self.n_sr_blocks * self.n_ene_blocks
        stats = torch.zeros((nblk, 3), dtype=torch.float64, device=ddev)
        fields_buf = torch.empty((self.n_prop_steps, W, dev.nchol), dtype=torch.float64,
                                 device=ddev)
        b = 0
        for _ in range(self.n_sr_blocks):
            for _ in range(self.n_ene_blocks):
                self._block_scan(ham_data, prop, prop_data, trial, wave_data, stats[b],
                                 fields_buf)
                b += 1
            prop.stochastic_reconfiguration_local(prop_data)
            dev.overlap(prop_data["walkers"], out=prop_data["overlaps"])
        prop_data["pop_control_ene_shift"] = scalars[1]
        be, bw = stats[:, 0], stats[:, 1]
        prop_data["n_killed_walkers"] = stats[:, 2].sum() / (nblk * W)
        prop_data["block_energies"] = be
        prop_data["block_weights"] = bw
        energy = float((torch.sum(be * bw) / torch.sum(bw)).item())
        return energy, prop_data
