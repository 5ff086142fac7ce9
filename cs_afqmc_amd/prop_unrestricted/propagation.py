"""Phaseless propagator over spin-resolved Cholesky vectors, stepping on the GPU.

Mirrors ad_afqmc/prop_unrestricted/propagation.py ``propagator_unrestricted``
(:332-513, phaseless part): the step of :101-165 with the alpha determinant propagated
by VHS_a = i sqrt(dt) x.L^a and the beta determinant by VHS_b = i sqrt(dt) x.L^b
(_apply_trotprop, :361-411), and the intermediates of
_build_propagation_intermediates (:461-510):
  mf_shifts    = i (sum_pq L^a_g[pq] rdm1_a[pq] + sum_pq L^b_g[pq] rdm1_b[pq])
  h0_prop      = -h0 - sum_g mf_g^2 / 2
  v0_s         = 1/2 sum_g L^s_g (L^s_g)^T
  v1_s         = Re(i sum_g mf_g L^s_g)
  exp_h1[s]    = expm(-dt (h1[s] - v0_s - v1_s) / 2).
The Cholesky contractions of the intermediates run on the device in float64, the
matrix exponentials on the host; the step itself runs in an AFQMC_UHF_SPIN_CHOL
context of the HIP library.
"""

from dataclasses import dataclass

import numpy as np
import scipy.linalg
import torch

from .. import propagation as _prop
from ..device import DeviceStep, drop_cached
from ..wavefunctions import _as_np
from .wavefunctions_unrestricted import spin_chol_view


@dataclass
class propagator_unrestricted(_prop.propagator_unrestricted):
    """prop_unrestricted/propagation.py:332-513 (phaseless part)."""

    spin_chol = True

    def _build_propagation_intermediates(self, ham_data, trial, wave_data):
        """prop_unrestricted/propagation.py:461-510 (formulas in the module docstring)."""
        norb = trial.norb
        rdm1 = np.asarray(trial.get_rdm1(wave_data), dtype=np.float64)
        chol = spin_chol_view(ham_data, norb)  # (2, G, N, N)
        G = chol.shape[1]
        flat = chol.reshape(2, G, norb * norb)
        dev = chol.device
        mf_re = sum(flat[s] @ torch.from_numpy(rdm1[s].reshape(-1)).to(dev) for s in range(2))
        h1 = np.asarray(_as_np(ham_data["h1"]), dtype=np.float64)
        exp_h1 = []
        for s in range(2):
            # G independent (N x N x N) products summed, as in the restricted builder
            v0 = 0.5 * torch.bmm(chol[s], chol[s].transpose(1, 2).contiguous()).sum(0)
            v1 = -(mf_re @ flat[s]).view(norb, norb)  # Re(i sum_g mf_g L^s_g), mf = i mf_re
            h1_mod = h1[s] - (v0 + v1).cpu().numpy()
            exp_h1.append(scipy.linalg.expm(-self.dt * h1_mod / 2.0))
        mf_shifts = 1.0j * mf_re.cpu().numpy()
        ham_data["mf_shifts"] = mf_shifts
        h0 = float(np.real(_as_np(ham_data["h0"])))
        ham_data["h0_prop"] = -h0 - np.sum(mf_shifts**2) / 2.0
        ham_data["exp_h1"] = np.array(exp_h1)
        step = DeviceStep(
            "uhf", norb, trial.nelec, self.n_walkers, self.dt, self.n_exp_terms, h0,
            ham_data["h0_prop"], flat.reshape(2 * G, norb * norb), wave_data["mo_coeff"],
            ham_data["rot_h1"], ham_data["rot_chol"], mf_shifts, ham_data["exp_h1"],
            precision=self.precision, spin_chol=True)
        drop_cached(ham_data)  # the context holds its own copies
        ham_data["_afqmc_device"] = step
        wave_data["_afqmc_device"] = step
        return ham_data

    def __hash__(self):
        return hash(tuple(self.__dict__.values()))
