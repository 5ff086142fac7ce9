"""ctypes binding of the C ABI in include/cs_afqmc_amd.h.

The library is built in-tree (cs_afqmc_amd/libcs_afqmc_amd.so, see _build.py).  There
is no fallback: if the library is missing or fails to load, every entry point raises.
"""

import ctypes
import os

from . import _build

AFQMC_RHF = 1
AFQMC_UHF = 2
AFQMC_UHF_SPIN_CHOL = 3
AFQMC_STEP_REUSE_GREEN = 1
AFQMC_SR_GLOBAL_ORDER = 1
BLOCK_STATS_LEN = 4  # AFQMC_BLOCK_STATS_LEN: block energy, weight, zero weights, clamped
PRECISIONS = {"fp64": 0, "fp32": 1, "bf16x3": 2}

# every symbol declared in include/cs_afqmc_amd.h
EXPORTS = (
    "afqmc_ctx_create",
    "afqmc_ctx_destroy",
    "afqmc_set_stream",
    "afqmc_ctx_device_bytes",
    "afqmc_ctx_vhs_entries",
    "afqmc_last_error",
    "afqmc_propagate",
    "afqmc_overlap",
    "afqmc_force_bias",
    "afqmc_energy",
    "afqmc_orthonormalize",
    "afqmc_sr",
    "afqmc_sr_key",
    "afqmc_sr_static",
    "afqmc_block_stats",
    "afqmc_estimate_mix",
    "afqmc_prof_enable",
    "afqmc_prof_collect",
    "afqmc_prof_phases",
    "afqmc_fields_normal",
    "afqmc_fields_normal_key",
    "afqmc_cs_outliers",
    "afqmc_group_create",
    "afqmc_group_destroy",
    "afqmc_group_set_stream",
    "afqmc_group_clear",
    "afqmc_group_add",
    "afqmc_group_size",
    "afqmc_group_propagate",
    "afqmc_group_prof_enable",
    "afqmc_group_prof_phases",
    "afqmc_group_prof_collect",
    "afqmc_group_tune",
    "afqmc_group_orthonormalize",
    "afqmc_group_overlap",
    "afqmc_group_energy",
    "afqmc_group_fields_normal_key",
)

# phase indices of afqmc_prof_collect (AFQMC_PH_* in the header)
PHASES = ("green", "fb_gemm", "fields", "vhs_gemm", "taylor", "overlap", "weights", "qr",
          "energy", "sr", "rng")


class AfqmcError(RuntimeError):
    pass


class Dims(ctypes.Structure):
    _fields_ = [
        ("walker_type", ctypes.c_int32),
        ("norb", ctypes.c_int32),
        ("nocc_up", ctypes.c_int32),
        ("nocc_dn", ctypes.c_int32),
        ("nchol", ctypes.c_int32),
        ("n_walkers", ctypes.c_int32),
        ("n_exp_terms", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("dt", ctypes.c_double),
    ]


_dp = ctypes.POINTER(ctypes.c_double)


class HamHost(ctypes.Structure):
    _fields_ = [
        ("h0", ctypes.c_double),
        ("h0_prop", ctypes.c_double),
        ("chol", _dp),
        ("mo_coeff", _dp),
        ("rot_h1", _dp),
        ("rot_chol", _dp),
        ("mf_shifts", _dp),
        ("exp_h1", _dp),
    ]


_LIB = None
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32


def _declare(lib):
    lib.afqmc_ctx_create.argtypes = [
        ctypes.POINTER(_vp), ctypes.POINTER(Dims), ctypes.POINTER(HamHost), _vp]
    lib.afqmc_ctx_destroy.argtypes = [_vp]
    lib.afqmc_set_stream.argtypes = [_vp, _vp]
    lib.afqmc_ctx_device_bytes.argtypes = [_vp]
    lib.afqmc_ctx_device_bytes.restype = ctypes.c_int64
    lib.afqmc_ctx_vhs_entries.argtypes = [_vp]
    lib.afqmc_ctx_vhs_entries.restype = ctypes.c_int64
    lib.afqmc_last_error.argtypes = []
    lib.afqmc_last_error.restype = ctypes.c_char_p
    lib.afqmc_propagate.argtypes = [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32]
    lib.afqmc_overlap.argtypes = [_vp, _i32, _vp, _vp, _vp]
    lib.afqmc_force_bias.argtypes = [_vp, _i32, _vp, _vp, _vp]
    lib.afqmc_energy.argtypes = [_vp, _i32, _vp, _vp, _vp]
    lib.afqmc_orthonormalize.argtypes = [_vp, _i32, _vp, _vp]
    lib.afqmc_sr.argtypes = [_vp, _i32, _vp, _vp, _vp, ctypes.c_double, _vp, _i32]
    lib.afqmc_sr_key.argtypes = [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _i32]
    lib.afqmc_sr_static.argtypes = [_i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp,
                                    ctypes.c_double, _vp, _i32, _vp]
    lib.afqmc_block_stats.argtypes = [_vp, _i32, _vp, _vp, _vp, _vp]
    lib.afqmc_estimate_mix.argtypes = [_vp, _vp, _vp, _i32]
    lib.afqmc_prof_enable.argtypes = [_vp, _i32]
    lib.afqmc_prof_collect.argtypes = [_vp, _vp, _vp]
    lib.afqmc_prof_phases.argtypes = [_vp, ctypes.c_uint32]
    lib.afqmc_fields_normal.argtypes = [
        _vp, ctypes.c_uint64, ctypes.c_uint64, _vp, ctypes.c_int64]
    lib.afqmc_fields_normal_key.argtypes = [_vp, _vp, _vp, ctypes.c_int64]
    lib.afqmc_cs_outliers.argtypes = [_vp, _i32, _vp, _vp, ctypes.c_double]
    lib.afqmc_group_create.argtypes = [ctypes.POINTER(_vp), _i32, _vp]
    lib.afqmc_group_destroy.argtypes = [_vp]
    lib.afqmc_group_set_stream.argtypes = [_vp, _vp]
    lib.afqmc_group_clear.argtypes = [_vp]
    lib.afqmc_group_add.argtypes = [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i32, _vp]
    lib.afqmc_group_size.argtypes = [_vp]
    lib.afqmc_group_propagate.argtypes = [_vp, _i32, _i32]
    lib.afqmc_group_tune.argtypes = [_vp, _vp]
    lib.afqmc_group_orthonormalize.argtypes = [_vp]
    lib.afqmc_group_overlap.argtypes = [_vp]
    lib.afqmc_group_energy.argtypes = [_vp, _vp]
    lib.afqmc_group_fields_normal_key.argtypes = [_vp, _vp]
    lib.afqmc_group_prof_enable.argtypes = [_vp, _i32]
    lib.afqmc_group_prof_phases.argtypes = [_vp, ctypes.c_uint32]
    lib.afqmc_group_prof_collect.argtypes = [_vp, _vp, _vp]
    for name in EXPORTS:
        fn = getattr(lib, name)
        if fn.restype is ctypes.c_int:  # default
            fn.restype = ctypes.c_int


def lib():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        path = _build.LIB_PATH
        if not os.path.exists(path):
            raise AfqmcError(
                f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
            )
        lib_ = ctypes.CDLL(path)
        _declare(lib_)
        _LIB = lib_
    return _LIB


def check(rc):
    if rc != 0:
        msg = lib().afqmc_last_error().decode(errors="replace")
        raise AfqmcError(f"cs_afqmc_amd error {rc}: {msg}")
