"""Trial wave functions with the reference's measurement API, evaluated on the GPU.

Mirrors ad_afqmc/wavefunctions.py: ``rhf`` (:453-635) and ``uhf`` (:638-844) with
``calc_overlap`` / ``calc_force_bias`` / ``calc_energy`` (:45-241), ``get_rdm1``
(:268-282), ``get_init_walkers`` (:298-355) and ``_build_measurement_intermediates``
(:602-619 / :817-841).  Walkers are cuda complex128 tensors: (W, N, n) for
restricted walkers, [up, dn] for unrestricted ones.  Measurements run in the HIP
library (afqmc_overlap / afqmc_force_bias / afqmc_energy); intermediates are built
once per run in float64, the Cholesky contractions on the device.

Supported pairs (mpi_jax.py:314-352): restricted walkers with an rhf trial, and
unrestricted walkers with a uhf trial.  The trial coefficients must be real.
"""

from dataclasses import dataclass
from typing import Tuple

import numpy as np
import torch

from .device import DeviceStep, device_chol, drop_cached


def _as_np(x):
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def _device_for(trial, ham_data, wave_data, walker_type, n_walkers):
    """Return a DeviceStep for measurements on ``n_walkers`` walkers.

    Reuses the propagator's context stored in ham_data when it fits; otherwise builds
    a measurement-only context (identity exp_h1, zero mf_shifts: unused by the
    measurement kernels)."""
    holders = [ham_data] if ham_data is not None else [wave_data]
    for holder in holders:
        for key in ("_afqmc_device", "_afqmc_measure_device"):
            dev = holder.get(key)
            # overlaps (ham_data None) do not read the Cholesky vectors of either kind
            if (dev is not None and dev.walker_type == walker_type
                    and (ham_data is None
                         or dev.spin_chol == getattr(trial, "spin_chol", False))
                    and dev.n_walkers >= n_walkers):
                return dev
    norb = trial.norb
    eye = np.eye(norb) if walker_type == "rhf" else np.array([np.eye(norb)] * 2)
    if ham_data is not None:
        chol = device_chol(ham_data, norb)
        spin = getattr(trial, "spin_chol", False)
        dev = DeviceStep(
            walker_type, norb, trial.nelec, n_walkers, 0.01, 2, _as_np(ham_data["h0"]), 0.0,
            chol, wave_data["mo_coeff"], ham_data["rot_h1"], ham_data["rot_chol"],
            np.zeros(chol.shape[0] // (2 if spin else 1), dtype=np.complex128), eye,
            spin_chol=spin)
        drop_cached(ham_data)  # the context holds its own copy of the vectors
        ham_data["_afqmc_measure_device"] = dev
        return dev
    # overlap only: the Hamiltonian arrays are placeholders the overlap kernel never reads
    if walker_type == "rhf":
        n = trial.nelec[0]
        rh1, rch = np.zeros((n, norb)), np.zeros((1, n, norb))
    else:
        rh1 = [np.zeros((trial.nelec[s], norb)) for s in range(2)]
        rch = [np.zeros((1, trial.nelec[s], norb)) for s in range(2)]
    dev = DeviceStep(walker_type, norb, trial.nelec, n_walkers, 0.01, 2, 0.0, 0.0,
                     np.zeros((1, norb * norb)), wave_data["mo_coeff"], rh1, rch,
                     np.zeros(1, dtype=np.complex128), eye)
    wave_data["_afqmc_measure_device"] = dev
    return dev


def _nwalkers(walkers):
    return walkers.shape[0] if isinstance(walkers, torch.Tensor) else walkers[0].shape[0]


@dataclass
class rhf:
    """Restricted Hartree-Fock trial (wavefunctions.py:453-635).  wave_data["mo_coeff"]
    is (norb, nocc); ham_data holds rot_h1 (nocc, norb) and rot_chol (nchol, nocc, norb)."""

    norb: int
    nelec: Tuple[int, int]
    n_opt_iter: int = 30
    n_batch: int = 1

    walker_type = "rhf"

    def __post_init__(self):
        assert self.nelec[0] == self.nelec[1], "RHF requires equal up and down electrons."

    def calc_overlap(self, walkers, wave_data):
        dev = _device_for(self, None, wave_data, "rhf", _nwalkers(walkers))
        return dev.overlap(walkers)

    def calc_force_bias(self, walkers, ham_data, wave_data):
        dev = _device_for(self, ham_data, wave_data, "rhf", _nwalkers(walkers))
        return dev.force_bias(walkers)

    def calc_energy(self, walkers, ham_data, wave_data):
        dev = _device_for(self, ham_data, wave_data, "rhf", _nwalkers(walkers))
        return dev.energy(walkers)

    def _calc_rdm1(self, wave_data):
        mo = _as_np(wave_data["mo_coeff"])
        return np.array([mo @ mo.T] * 2)

    def get_rdm1(self, wave_data):
        if "rdm1" in wave_data:
            return _as_np(wave_data["rdm1"])
        return self._calc_rdm1(wave_data)

    def get_init_walkers(self, wave_data, n_walkers, restricted=True):
        return _init_walkers(self, wave_data, n_walkers, restricted)

    def _build_measurement_intermediates(self, ham_data, wave_data):
        """wavefunctions.py:602-619: symmetrised h1, rot_h1 = C^T (h1a + h1b)/2,
        rot_chol[g] = C^T L_g."""
        norb = self.norb
        h1 = np.asarray(_as_np(ham_data["h1"]), dtype=np.float64)
        h1 = np.array([(h1[0] + h1[0].T) / 2.0, (h1[1] + h1[1].T) / 2.0])
        ham_data["h1"] = h1
        mo = np.asarray(_as_np(wave_data["mo_coeff"]), dtype=np.float64)
        ham_data["rot_h1"] = mo.T @ ((h1[0] + h1[1]) / 2.0)
        _invalidate(ham_data)
        _rotate_chol(ham_data, norb, [mo], restricted=True)
        return ham_data


@dataclass
class uhf:
    """Unrestricted Hartree-Fock trial (wavefunctions.py:638-844).  wave_data["mo_coeff"]
    is [C_up (norb, n_up), C_dn (norb, n_dn)]."""

    norb: int
    nelec: Tuple[int, int]
    n_opt_iter: int = 30
    n_batch: int = 1

    walker_type = "uhf"

    def calc_overlap(self, walkers, wave_data):
        dev = _device_for(self, None, wave_data, "uhf", _nwalkers(walkers))
        return dev.overlap(walkers)

    def calc_force_bias(self, walkers, ham_data, wave_data):
        dev = _device_for(self, ham_data, wave_data, "uhf", _nwalkers(walkers))
        return dev.force_bias(walkers)

    def calc_energy(self, walkers, ham_data, wave_data):
        dev = _device_for(self, ham_data, wave_data, "uhf", _nwalkers(walkers))
        return dev.energy(walkers)

    def _calc_rdm1(self, wave_data):
        mo = [_as_np(m) for m in wave_data["mo_coeff"]]
        return np.array([mo[0] @ mo[0].T, mo[1] @ mo[1].T])

    def get_rdm1(self, wave_data):
        if "rdm1" in wave_data:
            return _as_np(wave_data["rdm1"])
        return self._calc_rdm1(wave_data)

    def get_init_walkers(self, wave_data, n_walkers, restricted=False):
        return _init_walkers(self, wave_data, n_walkers, restricted)

    def _build_measurement_intermediates(self, ham_data, wave_data):
        """wavefunctions.py:817-841."""
        norb = self.norb
        h1 = np.asarray(_as_np(ham_data["h1"]), dtype=np.float64)
        h1 = np.array([(h1[0] + h1[0].T) / 2.0, (h1[1] + h1[1].T) / 2.0])
        ham_data["h1"] = h1
        mo = [np.asarray(_as_np(m), dtype=np.float64) for m in wave_data["mo_coeff"]]
        ham_data["rot_h1"] = [mo[0].T @ h1[0], mo[1].T @ h1[1]]
        _invalidate(ham_data)
        _rotate_chol(ham_data, norb, mo, restricted=False)
        return ham_data


def _rotate_chol(ham_data, norb, mos, restricted):
    """ham_data["rot_chol"]: C^T L_g for every Cholesky vector and each trial C in `mos`,
    one batched float64 product on the device over the uploaded vectors.  The result
    stays in device memory (a float64 cuda tensor (G, n, N); a list of two for uhf), as
    the reference's jax arrays do on a GPU."""
    chol = device_chol(ham_data, norb).view(-1, norb, norb)
    rot = [torch.matmul(torch.from_numpy(np.ascontiguousarray(m.T)).to(chol.device), chol)
           for m in mos]
    ham_data["rot_chol"] = rot[0] if restricted else rot


def _invalidate(ham_data):
    ham_data.pop("_afqmc_device", None)
    ham_data.pop("_afqmc_measure_device", None)


def _eigh_vectors(a):
    """Eigenvectors of a symmetric matrix (numpy's LAPACK driver, so the basis chosen in
    degenerate subspaces is the host reference's).  Run on one BLAS thread: next to the
    HIP runtime's threads a threaded decomposition of an orbital-sized matrix was
    measured 50x slower (110 ms against 2 ms at N = 250)."""
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        return np.linalg.eigh(a)[1]
    with threadpool_limits(limits=1, user_api="blas"):
        return np.linalg.eigh(a)[1]


def init_orbitals(rdm1, nelec, restricted):
    """wavefunctions.py:298-355: the starting walker orbitals as numpy arrays -- natural
    orbitals (descending occupation) of each spin rdm1; restricted walkers take the
    spin-up ones when |det(up^T dn)| > 1e-3, else the QR-orthonormalised column sums
    up_i + sign(up_i . dn_i) dn_i, if those overlap both spins' orbitals by more than
    1e-3 in |det(Q^T up) det(Q^T dn)|.  Returns one (N, n) array when restricted, else
    [up (N, n_up), dn (N, n_dn)]."""
    vec_up = _eigh_vectors(rdm1[0])
    # identical spin densities (closed shells) share one decomposition
    vec_dn = vec_up if np.array_equal(rdm1[0], rdm1[1]) else _eigh_vectors(rdm1[1])
    natorbs_up = vec_up[:, ::-1][:, : nelec[0]]
    natorbs_dn = vec_dn[:, ::-1][:, : nelec[1]]
    if not restricted:
        return [natorbs_up, natorbs_dn]
    if nelec[0] != nelec[1]:
        raise NotImplementedError("restricted walkers need nelec[0] == nelec[1]")
    det_overlap = np.linalg.det(natorbs_up.T @ natorbs_dn)
    if np.abs(det_overlap) > 1e-3:
        return natorbs_up
    overlaps = np.array([natorbs_up[:, i].T @ natorbs_dn[:, i] for i in range(nelec[0])])
    new_vecs = natorbs_up + np.einsum("ij,j->ij", natorbs_dn, np.sign(overlaps))
    new_vecs = np.linalg.qr(new_vecs)[0]
    det_overlap = np.linalg.det(new_vecs.T @ natorbs_up) * np.linalg.det(
        new_vecs.T @ natorbs_dn)
    if np.abs(det_overlap) > 1e-3:
        return new_vecs
    raise ValueError("Cannot find a set of RHF orbitals with good trial overlap.")


def _init_walkers(trial, wave_data, n_walkers, restricted):
    """init_orbitals of the trial's rdm1, copied n_walkers times, as cuda complex128
    tensors."""
    orbs = init_orbitals(trial.get_rdm1(wave_data), trial.nelec, restricted)
    dev = torch.device("cuda", torch.cuda.current_device())

    def rep(m):
        t = torch.as_tensor(np.ascontiguousarray(m + 0.0j), device=dev)
        return t.unsqueeze(0).repeat(n_walkers, 1, 1).contiguous()

    return rep(orbs) if restricted else [rep(m) for m in orbs]
