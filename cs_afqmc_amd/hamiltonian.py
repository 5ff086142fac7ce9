"""Hamiltonian handler with the reference's API (ad_afqmc/hamiltonian.py:12-84).

Thin wrappers that forward to the trial's and the propagator's intermediate builders;
the propagator's builder also creates the device context of the GPU step.
"""

from dataclasses import dataclass

import numpy as np

from .wavefunctions import _as_np


@dataclass
class hamiltonian:
    norb: int

    def rotate_orbs(self, ham_data, mo_coeff):
        """hamiltonian.py:23-48: h1 -> C^T h1 C, L_g -> C^T L_g C."""
        mo = np.asarray(_as_np(mo_coeff), dtype=np.float64)
        h1 = np.asarray(_as_np(ham_data["h1"]), dtype=np.float64)
        ham_data["h1"] = np.array([mo.T @ h1[0] @ mo, mo.T @ h1[1] @ mo])
        chol = np.asarray(_as_np(ham_data["chol"]), dtype=np.float64).reshape(
            -1, self.norb, self.norb)
        chol = np.einsum("gij,jp->gip", chol, mo, optimize=True)
        chol = np.einsum("qi,gip->gqp", mo.T, chol, optimize=True)
        ham_data["chol"] = chol.reshape(-1, self.norb * self.norb)
        return ham_data

    def build_measurement_intermediates(self, ham_data, trial, wave_data):
        """hamiltonian.py:50-65."""
        return trial._build_measurement_intermediates(ham_data, wave_data)

    def build_propagation_intermediates(self, ham_data, prop, trial, wave_data):
        """hamiltonian.py:67-81."""
        return prop._build_propagation_intermediates(ham_data, trial, wave_data)

    def __hash__(self):
        return hash(self.norb)
