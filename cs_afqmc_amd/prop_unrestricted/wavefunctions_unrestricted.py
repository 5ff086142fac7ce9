"""UHF trial over spin-resolved Cholesky vectors, evaluated on the GPU.

Mirrors ad_afqmc/prop_unrestricted/wavefunctions_unrestricted.py ``uhf`` (:284-476):
overlap det(C_a^T w_up) det(C_b^T w_dn) (:303-313), force bias
sum_ij rot_chol[0][g,i,j] G_up[i,j] + rot_chol[1][g,i,j] G_dn[i,j] (:339-350), local
energy with per-spin rot_h1 / rot_chol (:352-379) and the intermediates
rot_h1[s] = C_s^T h1[s], rot_chol[s][g] = C_s^T L^s_g (:453-475).  Measurements run in
the HIP library (afqmc_overlap / afqmc_force_bias / afqmc_energy of an
AFQMC_UHF_SPIN_CHOL context); ham_data["chol"] is (2, G, N*N).
"""

from dataclasses import dataclass

import numpy as np
import torch

from .. import wavefunctions as _wf
from ..device import device_chol


@dataclass
class uhf(_wf.uhf):
    """wavefunctions_unrestricted.py:284-476.  wave_data["mo_coeff"] is
    [C_up (norb, n_up), C_dn (norb, n_dn)]; ham_data["h1"] (2, N, N) and
    ham_data["chol"] (2, G, N*N) hold one set per spin."""

    spin_chol = True

    def _build_measurement_intermediates(self, ham_data, wave_data):
        """wavefunctions_unrestricted.py:453-475: symmetrised h1 per spin,
        rot_h1[s] = C_s^T h1[s], rot_chol[s][g] = C_s^T L^s_g (one batched float64
        product per spin on the device)."""
        norb = self.norb
        h1 = np.asarray(_wf._as_np(ham_data["h1"]), dtype=np.float64)
        h1 = np.array([(h1[0] + h1[0].T) / 2.0, (h1[1] + h1[1].T) / 2.0])
        ham_data["h1"] = h1
        mo = [np.asarray(_wf._as_np(m), dtype=np.float64) for m in wave_data["mo_coeff"]]
        ham_data["rot_h1"] = [mo[0].T @ h1[0], mo[1].T @ h1[1]]
        _wf._invalidate(ham_data)
        chol = spin_chol_view(ham_data, norb)
        ham_data["rot_chol"] = [
            torch.matmul(torch.from_numpy(np.ascontiguousarray(mo[s].T)).to(chol.device),
                         chol[s])
            for s in range(2)]
        return ham_data


def spin_chol_view(ham_data, norb):
    """ham_data["chol"] on the device as (2, G, N, N) float64 (uploaded once)."""
    chol = device_chol(ham_data, norb)
    if chol.shape[0] % 2:
        raise ValueError("spin-resolved Cholesky vectors must be (2, G, N*N)")
    return chol.view(2, chol.shape[0] // 2, norb, norb)
