"""Correlated sampling: two systems propagated with the same auxiliary fields.

Mirrors ad_afqmc/corr_sample/corr_sample.py: ``field_block_scan`` (:513-548),
``corr_otler_rm`` (:551-585), ``cs_block_scan`` (:588-618) and ``cs_steps_scan``
(:667-694).  Both systems have the same walker count and number of Cholesky vectors,
and each block draws one field tensor (n_prop_steps, W, nchol) from system 1's key;
after the block, walkers whose weight is a z-score outlier in either system get weight
0 in both, and the block contributes sum(E_L w) and sum(w) per system, whose ratio
difference estimates the energy difference with correlated noise.

The two systems' steps run as one grouped step (afqmc_group_propagate with both
contexts bound to the same fields block), or through their own contexts in turn.
"""

import numpy as np
import torch

from . import _lib, rng
from .device import DeviceGroup


class _System:
    def __init__(self, ham_data, prop, trial, wave_data, prop_data):
        self.ham_data, self.prop, self.trial, self.wave_data = ham_data, prop, trial, wave_data
        self.pd = prop_data
        self.dev = prop._device(ham_data)
        e = prop_data["e_estimate"]
        e = e if isinstance(e, torch.Tensor) else torch.tensor(float(e))
        e = e.to(self.dev.device, torch.float64).reshape(())
        pcs = prop_data.get("pop_control_ene_shift", e)
        pcs = pcs if isinstance(pcs, torch.Tensor) else torch.tensor(float(pcs))
        pcs = pcs.to(self.dev.device, torch.float64).reshape(())
        self.scalars = torch.stack([e, pcs]).contiguous()
        self.energies = torch.empty(prop.n_walkers, dtype=torch.complex128, device=self.dev.device)


def _field_block_end(sy):
    """corr_sample.py:541-548: QR, overlaps, block energy -> population-control shift."""
    pd = sy.pd
    sy.dev.orthonormalize(pd["walkers"])
    sy.dev.overlap(pd["walkers"], out=pd["overlaps"])
    en = sy.dev.energy(pd["walkers"], out=sy.energies)
    st = torch.zeros(_lib.BLOCK_STATS_LEN, dtype=torch.float64, device=sy.dev.device)
    sy.dev.block_stats(en, pd["weights"], sy.scalars, st)  # pcs <- 0.9 pcs + 0.1 E_block
    return en


def _local_sums(sy, en):
    """corr_sample.py:672-684: sum(clamped E_L * w) and sum(w) (device scalars)."""
    scratch = torch.stack([sy.scalars[0], sy.scalars[1]]).contiguous()  # keeps pcs untouched
    st = torch.zeros(_lib.BLOCK_STATS_LEN, dtype=torch.float64, device=sy.dev.device)
    sy.dev.block_stats(en, sy.pd["weights"], scratch, st)
    w = st[1]
    return torch.where(w > 0, st[0] * w, torch.zeros_like(w)), w


class CorrelatedPair:
    """Two systems sharing fields.  prop_data dicts come from each propagator's
    init_prop_data; system 1's prop_data["key"] drives the fields."""

    def __init__(self, sys1, sys2, n_prop_steps=50, z_thresh=5.0, grouped=True):
        self.s = [_System(*sys1), _System(*sys2)]
        a, b = self.s
        if a.prop.n_walkers != b.prop.n_walkers or a.dev.nchol != b.dev.nchol:
            raise ValueError("correlated systems need the same n_walkers and nchol")
        self.n_prop_steps = int(n_prop_steps)
        self.z_thresh = float(z_thresh)
        self.fields = torch.empty((self.n_prop_steps, a.prop.n_walkers, a.dev.nchol),
                                  dtype=torch.float64, device=a.dev.device)
        self.group = None
        if grouped:
            self.group = DeviceGroup(2)
            for sy in self.s:
                self.group.add(sy.dev, sy.pd["walkers"], sy.pd["weights"], sy.pd["overlaps"],
                               self.fields, sy.scalars)
        for sy in self.s:
            sy.dev.overlap(sy.pd["walkers"], out=sy.pd["overlaps"])

    def draw_fields(self):
        """corr_sample.py:592-601: split system 1's key, draw the block's fields."""
        pd = self.s[0].pd
        pd["key"], subkey = rng.split(pd["key"])
        self.s[0].dev.fields_normal(subkey[0], subkey[1], self.fields.shape, out=self.fields)

    def block(self, fields=None):
        """One correlated block; returns device scalars (loc_en1, loc_w1, loc_en2, loc_w2).
        `fields` (n_prop_steps, W, nchol) replaces the drawn fields when given."""
        if fields is None:
            self.draw_fields()
        else:
            self.fields.copy_(fields)
        if self.group is not None:
            for s in range(self.n_prop_steps):
                self.group.propagate(s, reuse_green=s > 0)
        else:
            for sy in self.s:
                for s in range(self.n_prop_steps):
                    sy.dev.propagate(sy.pd["walkers"], sy.pd["weights"], sy.pd["overlaps"],
                                     self.fields[s], sy.scalars, reuse_green=s > 0)
        ens = [_field_block_end(sy) for sy in self.s]
        a, b = self.s
        a.dev.cs_outliers(a.pd["weights"], b.pd["weights"], self.z_thresh)
        out = []
        for sy, en in zip(self.s, ens):
            out += list(_local_sums(sy, en))
        return tuple(out)

    def steps_scan(self, steps):
        """cs_steps_scan (corr_sample.py:667-694): `steps` blocks; numpy arrays
        (loc_en1, loc_w1, loc_en2, loc_w2) of length `steps`."""
        res = [self.block() for _ in range(steps)]
        arr = torch.stack([torch.stack(r) for r in res]).cpu().numpy()
        return arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3]

    def energy_difference(self, loc_en1, loc_w1, loc_en2, loc_w2):
        """E1 - E2 from the accumulated local sums."""
        return float(np.sum(loc_en1) / np.sum(loc_w1) - np.sum(loc_en2) / np.sum(loc_w2))
