"""Device step context: one afqmc_ctx of the HIP library plus its torch plumbing.

torch provides device memory and the current HIP stream; every computation of the
step runs in the library's gfx950 kernels (see include/cs_afqmc_amd.h).  There is no
CPU fallback: without a GPU or without the built library the calls raise.
"""

import ctypes

import numpy as np
import torch

from . import _lib

_f64 = np.float64


def _cptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _aptr(a):
    """Pointer to a host numpy array or a device tensor (float64)."""
    if isinstance(a, torch.Tensor):
        return ctypes.cast(ctypes.c_void_p(a.data_ptr()), ctypes.POINTER(ctypes.c_double))
    return _cptr(a)


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


_CACHE = "_afqmc_dev_cache"


def cached_tensor(ham_data, key):
    """Device copy of ham_data[key] made by `cache_tensor`, or None when ham_data[key]
    has been replaced since (the cache is keyed by the host object itself)."""
    hit = ham_data.get(_CACHE, {}).get(key)
    return hit[1] if hit is not None and hit[0] is ham_data.get(key) else None


def cache_tensor(ham_data, key, tensor):
    """Remember `tensor` as the device copy of the current ham_data[key]."""
    ham_data.setdefault(_CACHE, {})[key] = (ham_data.get(key), tensor)


def drop_cached(ham_data):
    ham_data.pop(_CACHE, None)


def device_chol(ham_data, norb):
    """ham_data["chol"] as a contiguous float64 (G, N*N) tensor on the current device,
    uploaded once per host array (the set-up contractions and the context read it)."""
    if not torch.cuda.is_available():
        raise _lib.AfqmcError("cs_afqmc_amd needs a ROCm GPU (torch.cuda is unavailable)")
    t = cached_tensor(ham_data, "chol")
    if t is None or t.device.index != torch.cuda.current_device():
        src = ham_data["chol"]
        if isinstance(src, torch.Tensor):
            t = src.detach().to(device="cuda", dtype=torch.float64)
        else:
            t = torch.from_numpy(np.ascontiguousarray(np.asarray(src, dtype=_f64)))
            t = t.to(device="cuda")
        t = t.reshape(-1, norb * norb).contiguous()
        cache_tensor(ham_data, "chol", t)
    return t


def _dev_f64(t, count, name):
    """A float64 device tensor passed to the context in place of a host array."""
    if t.device.type != "cuda" or t.device.index != torch.cuda.current_device():
        raise ValueError(f"{name}: device tensor on {t.device}, current device "
                         f"{torch.cuda.current_device()}")
    if t.is_complex():
        raise NotImplementedError(f"{name}: complex arrays are not supported")
    t = t.detach().to(torch.float64).contiguous()
    if t.numel() != count:
        raise ValueError(f"{name}: {t.numel()} elements, expected {count}")
    return t


class DeviceStep:
    """Owns the device copies of the intermediates and the step workspace.

    Args mirror the reference's ham_data / wave_data entries (host float64 arrays):
      walker_type: "rhf" (propagator_restricted + rhf trial) or "uhf"
      mo_coeff:    (N, n) for rhf; [(N, n_up), (N, n_dn)] for uhf
      rot_h1:      (n, N) / [(n_up, N), (n_dn, N)]
      rot_chol:    (G, n, N) / [(G, n_up, N), (G, n_dn, N)]
      chol:        (G, N*N)
      mf_shifts:   (G,) complex
      exp_h1:      (N, N) / (2, N, N)
    """

    def __init__(self, walker_type, norb, nelec, n_walkers, dt, n_exp_terms, h0, h0_prop,
                 chol, mo_coeff, rot_h1, rot_chol, mf_shifts, exp_h1, precision="fp64",
                 spin_chol=False):
        if not torch.cuda.is_available():
            raise _lib.AfqmcError("cs_afqmc_amd needs a ROCm GPU (torch.cuda is unavailable)")
        self.walker_type = walker_type
        self.norb = int(norb)
        self.nelec = (int(nelec[0]), int(nelec[1]))
        self.n_walkers = int(n_walkers)
        self.dt = float(dt)
        self.n_exp_terms = int(n_exp_terms)
        self.precision = precision
        # spin_chol: unrestricted walkers with separate alpha / beta Cholesky vectors,
        # chol (2, G, N*N) alpha then beta (prop_unrestricted/propagation.py:361-400)
        self.spin_chol = bool(spin_chol)
        if self.spin_chol and walker_type != "uhf":
            raise ValueError("spin-resolved Cholesky vectors need unrestricted walkers")
        ncs = 2 if self.spin_chol else 1
        if isinstance(chol, torch.Tensor):
            self.nchol = chol.numel() // (ncs * norb * norb)
            chol = _dev_f64(chol, ncs * self.nchol * norb * norb, "chol")
        else:
            chol = np.ascontiguousarray(np.asarray(chol, dtype=_f64).reshape(-1, norb * norb))
            if chol.shape[0] % ncs:
                raise ValueError("spin-resolved chol needs (2, G, N*N)")
            self.nchol = chol.shape[0] // ncs
        if walker_type == "rhf":
            mo = np.ascontiguousarray(np.asarray(mo_coeff, dtype=_f64))
            rh1 = np.ascontiguousarray(np.asarray(rot_h1, dtype=_f64))
            rot_chol = [rot_chol]
            eh1 = np.ascontiguousarray(np.asarray(exp_h1, dtype=_f64).reshape(norb, norb))
            wt = _lib.AFQMC_RHF
        elif walker_type == "uhf":
            mo = np.ascontiguousarray(np.concatenate(
                [np.asarray(m, dtype=_f64).ravel() for m in mo_coeff]))
            rh1 = np.ascontiguousarray(np.concatenate(
                [np.asarray(m, dtype=_f64).ravel() for m in rot_h1]))
            eh1 = np.ascontiguousarray(np.asarray(exp_h1, dtype=_f64).reshape(2, norb, norb))
            wt = _lib.AFQMC_UHF_SPIN_CHOL if self.spin_chol else _lib.AFQMC_UHF
        else:
            raise ValueError(f"walker_type {walker_type!r}")
        if any(isinstance(m, torch.Tensor) for m in rot_chol):
            n_rc = sum(self.nchol * n * norb
                       for n in self.nelec[: 1 if walker_type == "rhf" else 2])
            rch = _dev_f64(torch.cat([torch.as_tensor(m, device="cuda").reshape(-1)
                                      for m in rot_chol]), n_rc, "rot_chol")
        else:
            rch = np.ascontiguousarray(np.concatenate(
                [np.asarray(m, dtype=_f64).ravel() for m in rot_chol]))
        for a in (mo, rh1, rch):
            if (a.is_complex() if isinstance(a, torch.Tensor) else np.iscomplexobj(a)):
                raise NotImplementedError("complex trial coefficients are not supported")
        mf = np.ascontiguousarray(np.asarray(mf_shifts, dtype=np.complex128).reshape(-1))
        if mf.shape[0] != self.nchol:
            raise ValueError("mf_shifts length != nchol")
        mf_ri = np.ascontiguousarray(mf.view(_f64))
        self._keep = (chol, mo, rh1, rch, eh1, mf_ri)
        dims = _lib.Dims(wt, self.norb, self.nelec[0], self.nelec[1], self.nchol,
                         self.n_walkers, self.n_exp_terms, _lib.PRECISIONS[precision], self.dt)
        ham = _lib.HamHost(float(np.real(h0)), float(np.real(h0_prop)), _aptr(chol), _cptr(mo),
                           _cptr(rh1), _aptr(rch), _cptr(mf_ri), _cptr(eh1))
        self.device = torch.device("cuda", torch.cuda.current_device())
        handle = ctypes.c_void_p()
        L = _lib.lib()
        _lib.check(L.afqmc_ctx_create(ctypes.byref(handle), ctypes.byref(dims),
                                      ctypes.byref(ham), _stream()))
        self._h = handle
        self._lib = L

    # ------------------------------------------------------------------ plumbing
    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.afqmc_ctx_destroy(h)
            self._h = None

    def device_bytes(self):
        return int(self._lib.afqmc_ctx_device_bytes(self._h))

    def vhs_entries(self):
        """Entries of one stored VHS plane (N*N, or N(N+1)/2 for the upper triangle)."""
        return int(self._lib.afqmc_ctx_vhs_entries(self._h))

    def _prep(self):
        _lib.check(self._lib.afqmc_set_stream(self._h, _stream()))

    def _walkers(self, walkers):
        if self.walker_type == "rhf":
            up, dn = walkers, None
            self._check_walker_tensor(up, self.nelec[0])
            return up, None, up.shape[0]
        up, dn = walkers
        self._check_walker_tensor(up, self.nelec[0])
        self._check_walker_tensor(dn, self.nelec[1])
        if up.shape[0] != dn.shape[0]:
            raise ValueError("up/dn walker counts differ")
        return up, dn, up.shape[0]

    def _check_walker_tensor(self, t, n):
        if not isinstance(t, torch.Tensor) or t.device.type != "cuda":
            raise TypeError("walkers must be cuda tensors (complex128)")
        if t.dtype != torch.complex128 or not t.is_contiguous():
            raise TypeError("walkers must be contiguous complex128")
        if t.dim() != 3 or t.shape[1] != self.norb or t.shape[2] != n:
            raise ValueError(f"walker shape {tuple(t.shape)} != (W, {self.norb}, {n})")
        if t.shape[0] > self.n_walkers:
            raise ValueError(f"{t.shape[0]} walkers > context capacity {self.n_walkers}")

    def _vec(self, t, W, dtype, name, cols=None):
        shape = (W,) if cols is None else (W, cols)
        if (not isinstance(t, torch.Tensor) or t.device.type != "cuda" or t.dtype != dtype
                or not t.is_contiguous() or tuple(t.shape) != shape):
            raise TypeError(f"{name} must be a contiguous cuda {dtype} tensor of shape {shape}")

    # ------------------------------------------------------------------ entry points
    def propagate(self, walkers, weights, overlaps, fields, scalars, force_bias_out=None,
                  reuse_green=False):
        """reuse_green: the walkers are those the previous propagate on this context
        returned (no QR / SR / edits since), so its Green's function is reused."""
        up, dn, W = self._walkers(walkers)
        self._vec(weights, W, torch.float64, "weights")
        self._vec(overlaps, W, torch.complex128, "overlaps")
        self._vec(fields, W, torch.float64, "fields", self.nchol)
        if (scalars.dtype != torch.float64 or scalars.numel() != 2 or not scalars.is_cuda
                or not scalars.is_contiguous()):
            raise TypeError("scalars must be a contiguous cuda float64 tensor of 2 elements")
        if force_bias_out is not None:
            self._vec(force_bias_out, W, torch.complex128, "force_bias_out", self.nchol)
        self._prep()
        _lib.check(self._lib.afqmc_propagate(
            self._h, W, _dptr(up), _dptr(dn), _dptr(weights), _dptr(overlaps), _dptr(fields),
            _dptr(scalars), _dptr(force_bias_out),
            _lib.AFQMC_STEP_REUSE_GREEN if reuse_green else 0))

    def overlap(self, walkers, out=None):
        up, dn, W = self._walkers(walkers)
        if out is None:
            out = torch.empty(W, dtype=torch.complex128, device=up.device)
        self._prep()
        _lib.check(self._lib.afqmc_overlap(self._h, W, _dptr(up), _dptr(dn), _dptr(out)))
        return out

    def force_bias(self, walkers, out=None):
        up, dn, W = self._walkers(walkers)
        if out is None:
            out = torch.empty((W, self.nchol), dtype=torch.complex128, device=up.device)
        self._prep()
        _lib.check(self._lib.afqmc_force_bias(self._h, W, _dptr(up), _dptr(dn), _dptr(out)))
        return out

    def energy(self, walkers, out=None):
        up, dn, W = self._walkers(walkers)
        if out is None:
            out = torch.empty(W, dtype=torch.complex128, device=up.device)
        self._prep()
        _lib.check(self._lib.afqmc_energy(self._h, W, _dptr(up), _dptr(dn), _dptr(out)))
        return out

    def orthonormalize(self, walkers):
        up, dn, W = self._walkers(walkers)
        self._prep()
        _lib.check(self._lib.afqmc_orthonormalize(self._h, W, _dptr(up), _dptr(dn)))
        return walkers

    def sr(self, walkers, weights, zeta, idx_out=None, global_order=False):
        """Systematic resampling in place; global_order selects the comb-point rounding
        of the gather/scatter variant (sr.py:66-106)."""
        up, dn, W = self._walkers(walkers)
        self._vec(weights, W, torch.float64, "weights")
        if idx_out is not None:
            self._vec(idx_out, W, torch.int32, "idx_out")
        self._prep()
        flags = _lib.AFQMC_SR_GLOBAL_ORDER if global_order else 0
        _lib.check(self._lib.afqmc_sr(self._h, W, _dptr(up), _dptr(dn), _dptr(weights),
                                      float(zeta), _dptr(idx_out), flags))
        return idx_out

    def fields_normal(self, seed, counter, shape, out=None):
        if out is None:
            out = torch.empty(shape, dtype=torch.float64, device=self.device)
        self._prep()
        _lib.check(self._lib.afqmc_fields_normal(
            self._h, int(seed) & (2**64 - 1), int(counter) & (2**64 - 1), _dptr(out),
            out.numel()))
        return out

    def _check_key(self, key):
        if (not isinstance(key, torch.Tensor) or key.device.type != "cuda"
                or key.dtype != torch.int64 or key.numel() != 2 or not key.is_contiguous()):
            raise TypeError("key must be a contiguous cuda int64 tensor {seed, counter}")

    def fields_normal_key(self, key, out):
        """Fields from a device-resident key (int64[2] = {seed, counter} as two's
        complement uint64): draws with (seed, counter), then counter += 1."""
        self._check_key(key)
        if out.dtype != torch.float64 or not out.is_cuda or not out.is_contiguous():
            raise TypeError("out must be a contiguous cuda float64 tensor")
        self._prep()
        _lib.check(self._lib.afqmc_fields_normal_key(self._h, _dptr(key), _dptr(out),
                                                     out.numel()))
        return out

    def sr_key(self, walkers, weights, key, idx_out=None, global_order=False):
        """Systematic resampling with zeta drawn on the device from `key` (see
        fields_normal_key); nothing is read back to the host."""
        up, dn, W = self._walkers(walkers)
        self._vec(weights, W, torch.float64, "weights")
        self._check_key(key)
        if idx_out is not None:
            self._vec(idx_out, W, torch.int32, "idx_out")
        self._prep()
        flags = _lib.AFQMC_SR_GLOBAL_ORDER if global_order else 0
        _lib.check(self._lib.afqmc_sr_key(self._h, W, _dptr(up), _dptr(dn), _dptr(weights),
                                          _dptr(key), _dptr(idx_out), flags))
        return idx_out

    def block_stats(self, energies, weights, scalars, out):
        """out: (BLOCK_STATS_LEN,) float64 = block energy, block weight, zero weights,
        clamped local energies."""
        W = weights.shape[0]
        self._vec(energies, W, torch.complex128, "energies")
        self._vec(weights, W, torch.float64, "weights")
        if (out.dtype != torch.float64 or not out.is_cuda or not out.is_contiguous()
                or out.numel() != _lib.BLOCK_STATS_LEN):
            raise TypeError(f"out must be a contiguous cuda float64 tensor of "
                            f"{_lib.BLOCK_STATS_LEN} elements")
        self._prep()
        _lib.check(self._lib.afqmc_block_stats(self._h, W, _dptr(energies), _dptr(weights),
                                               _dptr(scalars), _dptr(out)))
        return out

    def estimate_mix(self, scalars, stats):
        """scalars[0] <- 0.9 scalars[0] + 0.1 (weighted mean block energy of `stats`,
        rows of BLOCK_STATS_LEN), the driver's update after a sampler call."""
        if (stats.dtype != torch.float64 or not stats.is_cuda or not stats.is_contiguous()
                or stats.numel() % _lib.BLOCK_STATS_LEN or stats.numel() == 0):
            raise TypeError("stats must be contiguous cuda float64 rows of block statistics")
        self._prep()
        _lib.check(self._lib.afqmc_estimate_mix(self._h, _dptr(scalars), _dptr(stats),
                                                stats.numel() // _lib.BLOCK_STATS_LEN))

    def cs_outliers(self, weights1, weights2, z_thresh=5.0):
        """corr_sample.py:551-585: zero the weights of joint z-score outliers."""
        W = weights1.shape[0]
        self._vec(weights1, W, torch.float64, "weights1")
        self._vec(weights2, W, torch.float64, "weights2")
        self._prep()
        _lib.check(self._lib.afqmc_cs_outliers(self._h, W, _dptr(weights1), _dptr(weights2),
                                               float(z_thresh)))

    # ------------------------------------------------------------------ phase timing
    def prof_enable(self, on=True, phases=None):
        """Record event pairs around phases; `phases` (names) restricts them, None = all."""
        mask = 0
        for p in phases or ():
            mask |= 1 << _lib.PHASES.index(p)
        _lib.check(self._lib.afqmc_prof_phases(self._h, mask))
        _lib.check(self._lib.afqmc_prof_enable(self._h, 1 if on else 0))

    def prof_collect(self):
        """{phase: (total ms, launches)} since the last collect (synchronises)."""
        ms = np.zeros(len(_lib.PHASES), dtype=np.float64)
        cnt = np.zeros(len(_lib.PHASES), dtype=np.int64)
        _lib.check(self._lib.afqmc_prof_collect(
            self._h, ms.ctypes.data_as(ctypes.c_void_p), cnt.ctypes.data_as(ctypes.c_void_p)))
        return {p: (float(ms[i]), int(cnt[i])) for i, p in enumerate(_lib.PHASES)}


class DeviceGroup:
    """A fragment group of the HIP library (afqmc_group_*): many DeviceStep contexts
    advanced by one grouped step.  Members bind their walker state and a fields block
    (n_steps, W, nchol); ``propagate(step)`` advances every member by one step."""

    def __init__(self, max_fragments=256):
        if not torch.cuda.is_available():
            raise _lib.AfqmcError("cs_afqmc_amd needs a ROCm GPU (torch.cuda is unavailable)")
        L = _lib.lib()
        handle = ctypes.c_void_p()
        _lib.check(L.afqmc_group_create(ctypes.byref(handle), int(max_fragments), _stream()))
        self._h = handle
        self._lib = L
        self._members = []  # keeps the bound tensors alive

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.afqmc_group_destroy(h)
            self._h = None

    def _prep(self):
        _lib.check(self._lib.afqmc_group_set_stream(self._h, _stream()))

    def clear(self):
        _lib.check(self._lib.afqmc_group_clear(self._h))
        self._members = []

    def add(self, dev, walkers, weights, overlaps, fields_block, scalars):
        up, dn, W = dev._walkers(walkers)
        if W != dev.n_walkers:
            raise ValueError("group members run at their full walker capacity")
        dev._vec(weights, W, torch.float64, "weights")
        dev._vec(overlaps, W, torch.complex128, "overlaps")
        if (fields_block.dtype != torch.float64 or not fields_block.is_contiguous()
                or fields_block.dim() != 3 or tuple(fields_block.shape[1:]) != (W, dev.nchol)):
            raise ValueError("fields_block must be contiguous float64 (n_steps, W, nchol)")
        if scalars.dtype != torch.float64 or scalars.numel() != 2:
            raise ValueError("scalars must be float64 [e_estimate, pop_control_ene_shift]")
        _lib.check(self._lib.afqmc_group_add(
            self._h, dev._h, _dptr(up), _dptr(dn), _dptr(weights), _dptr(overlaps),
            _dptr(fields_block), int(fields_block.shape[0]), _dptr(scalars)))
        self._members.append((dev, walkers, weights, overlaps, fields_block, scalars))

    def __len__(self):
        return int(self._lib.afqmc_group_size(self._h))

    def propagate(self, step, reuse_green=False):
        self._prep()
        flags = _lib.AFQMC_STEP_REUSE_GREEN if reuse_green else 0
        _lib.check(self._lib.afqmc_group_propagate(self._h, int(step), flags))

    def orthonormalize(self):
        """QR of every member's walkers (afqmc_group_orthonormalize)."""
        self._prep()
        _lib.check(self._lib.afqmc_group_orthonormalize(self._h))

    def overlap(self):
        """Every member's overlaps into its bound overlaps tensor (afqmc_group_overlap)."""
        self._prep()
        _lib.check(self._lib.afqmc_group_overlap(self._h))

    def energy(self, energies):
        """Local energies of every member's walkers into energies[i] (complex128, W_i),
        in the order members were added (afqmc_group_energy)."""
        if len(energies) != len(self._members):
            raise ValueError("one energies tensor per member")
        for (dev, *_), e in zip(self._members, energies):
            dev._vec(e, dev.n_walkers, torch.complex128, "energies")
        self._prep()
        self._en_ptrs = (ctypes.c_void_p * len(energies))(*[e.data_ptr() for e in energies])
        _lib.check(self._lib.afqmc_group_energy(self._h, self._en_ptrs))

    def fields_normal_key(self, keys):
        """Every member's bound fields block drawn from its device key (keys[i], int64[2]
        {seed, counter}), then each counter += 1 (afqmc_group_fields_normal_key)."""
        if len(keys) != len(self._members):
            raise ValueError("one key per member")
        for (dev, *_), k in zip(self._members, keys):
            dev._check_key(k)
        self._prep()
        self._key_ptrs = (ctypes.c_void_p * len(keys))(*[k.data_ptr() for k in keys])
        _lib.check(self._lib.afqmc_group_fields_normal_key(self._h, self._key_ptrs))

    def prof_enable(self, on=True, phases=None):
        """Event pairs around each phase of the grouped step (all grouped launches of a
        phase inside one pair); `phases` (names) restricts them, None = all."""
        mask = 0
        for p in phases or ():
            mask |= 1 << _lib.PHASES.index(p)
        _lib.check(self._lib.afqmc_group_prof_phases(self._h, mask))
        _lib.check(self._lib.afqmc_group_prof_enable(self._h, 1 if on else 0))

    def prof_collect(self):
        """{phase: (total ms, launches)} since the last collect (synchronises)."""
        ms = np.zeros(len(_lib.PHASES), dtype=np.float64)
        cnt = np.zeros(len(_lib.PHASES), dtype=np.int64)
        _lib.check(self._lib.afqmc_group_prof_collect(
            self._h, ms.ctypes.data_as(ctypes.c_void_p), cnt.ctypes.data_as(ctypes.c_void_p)))
        return {p: (float(ms[i]), int(cnt[i])) for i, p in enumerate(_lib.PHASES)}

    def tune(self):
        """Pick GEMM tiles by timing; returns {fb, vhs, cg: (ms64, ms128)} (cg: the
        energy's float64 force-bias product)."""
        self._prep()
        ms = np.zeros(6, dtype=np.float64)
        _lib.check(self._lib.afqmc_group_tune(self._h, ms.ctypes.data_as(ctypes.c_void_p)))
        return {"fb": (ms[0], ms[1]), "vhs": (ms[2], ms[3]), "cg": (ms[4], ms[5])}
