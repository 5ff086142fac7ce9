"""Phaseless AFQMC propagators with the reference's API, stepping on the GPU.

Mirrors ad_afqmc/propagation.py ``propagator_restricted`` (:201-328) and
``propagator_unrestricted`` (:331-512): dataclass fields ``dt``, ``n_walkers``,
``n_exp_terms``, ``n_batch`` and the methods ``init_prop_data``, ``propagate``,
``orthonormalize_walkers``, ``stochastic_reconfiguration_local`` and
``_build_propagation_intermediates``.  One extra field, ``precision``, selects the
arithmetic of the Cholesky contractions ("fp64" is the plain form).

prop_data holds cuda tensors: walkers (complex128; a list [up, dn] for unrestricted
walkers), weights (W,) float64, overlaps (W,) complex128, and 0-d float64 tensors
e_estimate / pop_control_ene_shift; ``key`` is a (seed, counter) pair (see rng.py).
``n_batch`` is accepted for API compatibility: the GPU step always processes the whole
population at once (batching only bounded memory in the reference).
"""

from dataclasses import dataclass

import numpy as np
import scipy.linalg
import torch

from . import rng
from .device import DeviceStep, device_chol, drop_cached
from .wavefunctions import _as_np


def _to_device_walkers(walkers):
    dev = torch.device("cuda", torch.cuda.current_device())

    def one(w):
        if isinstance(w, torch.Tensor):
            return w.to(device=dev, dtype=torch.complex128).contiguous()
        return torch.as_tensor(np.ascontiguousarray(np.asarray(w, dtype=np.complex128)),
                               device=dev)

    if isinstance(walkers, (list, tuple)):
        return [one(walkers[0]), one(walkers[1])]
    return one(walkers)


def _scalar_tensor(x, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float64).reshape(())
    return torch.tensor(float(np.real(x)), dtype=torch.float64, device=dev)


@dataclass
class propagator_restricted:
    """propagation.py:201-328."""

    dt: float = 0.01
    n_walkers: int = 50
    n_exp_terms: int = 6
    n_batch: int = 1
    precision: str = "fp64"

    walker_type = "rhf"

    # ------------------------------------------------------------ intermediates
    def _build_propagation_intermediates(self, ham_data, trial, wave_data):
        """propagation.py:292-325 (restricted) / :464-509 (unrestricted):
        mf_shifts = i sum(L_g * rdm1), h0_prop = -h0 - sum(mf^2)/2,
        v0 = 1/2 sum_g L_g L_g^T, h1_mod = h1 - v0 - Re(i sum_g mf_g L_g),
        exp_h1 = expm(-dt h1_mod / 2).  Float64, once per run: the Cholesky contractions
        on the device (over the vectors uploaded for the context), the rest on the host;
        then the device context (intermediates + workspace) is created."""
        norb = trial.norb
        rdm1 = np.asarray(trial.get_rdm1(wave_data), dtype=np.float64)
        rdm1t = rdm1[0] + rdm1[1]
        chol = device_chol(ham_data, norb)  # (G, N*N)
        G = chol.shape[0]
        mf_re = chol @ torch.from_numpy(rdm1t.reshape(-1)).to(chol.device)
        # v0 = 1/2 sum_g L_g L_g^T: G independent (N x N x N) products summed (one
        # (N x GN x N) product would leave most of the GPU idle on its few output tiles)
        chol3 = chol.view(G, norb, norb)
        v0 = (0.5 * torch.bmm(chol3, chol3.transpose(1, 2).contiguous()).sum(0)).cpu().numpy()
        # Re(i sum_g mf_g L_g) = -sum_g mf_re_g L_g: a real product
        v1 = (-(mf_re @ chol)).view(norb, norb).cpu().numpy()
        mf_re = mf_re.cpu().numpy()
        mf_shifts = 1.0j * mf_re  # purely imaginary
        ham_data["mf_shifts"] = mf_shifts
        h0 = float(np.real(_as_np(ham_data["h0"])))
        ham_data["h0_prop"] = -h0 - np.sum(mf_shifts**2) / 2.0
        h1 = np.asarray(_as_np(ham_data["h1"]), dtype=np.float64)
        if self.walker_type == "rhf":
            h1_mod = (h1[0] + h1[1]) / 2.0 - v0 - v1
            ham_data["exp_h1"] = scipy.linalg.expm(-self.dt * h1_mod / 2.0)
        else:
            h1_mod = h1 - np.array([v0 + v1, v0 + v1])
            ham_data["exp_h1"] = np.array(
                [scipy.linalg.expm(-self.dt * h1_mod[0] / 2.0),
                 scipy.linalg.expm(-self.dt * h1_mod[1] / 2.0)])
        dev = DeviceStep(
            self.walker_type, norb, trial.nelec, self.n_walkers, self.dt, self.n_exp_terms,
            h0, ham_data["h0_prop"], chol, wave_data["mo_coeff"], ham_data["rot_h1"],
            ham_data["rot_chol"], mf_shifts, ham_data["exp_h1"], precision=self.precision)
        drop_cached(ham_data)  # the context holds its own copies
        ham_data["_afqmc_device"] = dev
        wave_data["_afqmc_device"] = dev
        return ham_data

    @staticmethod
    def _device(ham_data):
        dev = ham_data.get("_afqmc_device")
        if dev is None:
            raise RuntimeError(
                "ham_data has no device context: call build_propagation_intermediates first")
        return dev

    # ------------------------------------------------------------ walkers
    def init_prop_data(self, trial, wave_data, ham_data, init_walkers=None):
        """propagation.py:210-232 / :333-357."""
        dev = self._device(ham_data)
        prop_data = {"_afqmc_device": dev}
        prop_data["weights"] = torch.ones(self.n_walkers, dtype=torch.float64, device=dev.device)
        if init_walkers is not None:
            prop_data["walkers"] = _to_device_walkers(init_walkers)
        else:
            prop_data["walkers"] = trial.get_init_walkers(
                wave_data, self.n_walkers, restricted=(self.walker_type == "rhf"))
        energy_samples = torch.real(dev.energy(prop_data["walkers"]))
        e_estimate = torch.sum(energy_samples) / self.n_walkers
        prop_data["e_estimate"] = e_estimate
        prop_data["pop_control_ene_shift"] = e_estimate.clone()
        prop_data["overlaps"] = dev.overlap(prop_data["walkers"])
        if self.walker_type == "uhf":
            prop_data["normed_overlaps"] = prop_data["overlaps"]
            prop_data["norms"] = torch.ones(self.n_walkers, dtype=torch.complex128,
                                            device=dev.device)
        return prop_data

    def propagate(self, trial, ham_data, prop_data, fields, wave_data, force_bias_out=None):
        """propagation.py:100-165: one phaseless step of every walker with the given
        auxiliary fields (W, nchol).  Runs entirely on the GPU (afqmc_propagate)."""
        dev = self._device(ham_data)
        prop_data["_afqmc_device"] = dev
        walkers = prop_data["walkers"]
        if not isinstance(fields, torch.Tensor):
            fields = torch.as_tensor(np.ascontiguousarray(np.asarray(fields, dtype=np.float64)),
                                     device=dev.device)
        fields = fields.to(device=dev.device, dtype=torch.float64).contiguous()
        scalars = torch.stack([
            _scalar_tensor(prop_data["e_estimate"], dev.device),
            _scalar_tensor(prop_data["pop_control_ene_shift"], dev.device),
        ]).contiguous()
        dev.propagate(walkers, prop_data["weights"], prop_data["overlaps"], fields, scalars,
                      force_bias_out)
        prop_data["pop_control_ene_shift"] = scalars[1]
        return prop_data

    def orthonormalize_walkers(self, prop_data):
        """propagation.py:251-253 / :422-424 (linalg_utils.qr_vmap[_uhf])."""
        prop_data["walkers"] = _device_of(prop_data).orthonormalize(prop_data["walkers"])
        return prop_data

    def stochastic_reconfiguration_local(self, prop_data):
        """propagation.py:234-241 / :402-409 -> sr.py:24-48."""
        prop_data["key"], subkey = rng.split(prop_data["key"])
        zeta = rng.uniform(subkey)
        _device_of(prop_data).sr(prop_data["walkers"], prop_data["weights"], zeta)
        return prop_data

    def stochastic_reconfiguration_global(self, prop_data, comm=None):
        """propagation.py:243-249 -> sr.py:66-106 (and :109-156 for two spins): the
        gather / comb / scatter form.  The whole population lives on this device, so
        the gather and scatter are the identity and the comb runs on the device with
        that variant's rounding of the comb points.  `comm` is ignored."""
        prop_data["key"], subkey = rng.split(prop_data["key"])
        zeta = rng.uniform(subkey)
        _device_of(prop_data).sr(prop_data["walkers"], prop_data["weights"], zeta,
                                 global_order=True)
        return prop_data

    def __hash__(self):
        return hash(tuple(self.__dict__.values()))


@dataclass
class propagator_unrestricted(propagator_restricted):
    """propagation.py:331-512 (phaseless part)."""

    walker_type = "uhf"

    def __hash__(self):
        return hash(tuple(self.__dict__.values()))


def _device_of(prop_data):
    dev = prop_data.get("_afqmc_device")
    if dev is None:
        raise RuntimeError("prop_data has no device context (use init_prop_data)")
    return dev
