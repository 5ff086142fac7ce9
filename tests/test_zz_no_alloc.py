"""afqmc_sr_static (the allocation-free resampling, scratch in a fixed device array)
against afqmc_sr (the context's workspace), through the C ABI: same indices, weights
and walker copies bit for bit, on two calls in a row, the second with fewer walkers;
and its size guard."""

import ctypes

import numpy as np
import pytest
import torch

from cs_afqmc_amd import _lib, hamiltonian, problems, propagation, wavefunctions

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _context(walker_type, norb, nelec, nchol, W):
    if walker_type == "rhf":
        hd, wd, ne = problems.synthetic_rhf(norb, nelec[0], nchol, seed=4)
        trial = wavefunctions.rhf(norb, ne)
        prop = propagation.propagator_restricted(dt=0.01, n_walkers=W)
    else:
        hd, wd, ne = problems.synthetic_uhf(norb, nelec[0], nelec[1], nchol, seed=4)
        trial = wavefunctions.uhf(norb, ne)
        prop = propagation.propagator_unrestricted(dt=0.01, n_walkers=W)
    ham = hamiltonian.hamiltonian(norb)
    wave = dict(wd)
    hdat = ham.build_measurement_intermediates(dict(hd), trial, wave)
    hdat = ham.build_propagation_intermediates(hdat, prop, trial, wave)
    return hdat["_afqmc_device"], ne


@pytest.mark.parametrize("walker_type", ["rhf", "uhf"])
def test_sr_static_matches_context_sr(gpu, walker_type):
    L = _lib.lib()
    norb, nelec, cap = 12, (4, 3), 96
    dev, ne = _context(walker_type, norb, nelec, 30, cap)
    spins = [ne[0]] if walker_type == "rhf" else [ne[0], ne[1]]
    rs = np.random.default_rng(21)
    stream = torch.cuda.current_stream()
    for call, (W, zeta) in enumerate([(96, 0.37), (53, 0.81)]):
        print(f"sr_static {walker_type} call {call} W={W}", flush=True)
        w = [torch.as_tensor(rs.standard_normal((W, norb, n))
                             + 1j * rs.standard_normal((W, norb, n)), device=gpu)
             for n in spins]
        weights = rs.uniform(0.0, 2.0, W) * (rs.uniform(size=W) > 0.2)
        for flags in (0, _lib.AFQMC_SR_GLOBAL_ORDER):
            # existing entry point: context workspace, in place
            ref_w = [x.clone() for x in w]
            ref_wt = torch.as_tensor(weights, device=gpu)
            ref_idx = torch.empty(W, dtype=torch.int32, device=gpu)
            _lib.check(L.afqmc_set_stream(dev._h, ctypes.c_void_p(stream.cuda_stream)))
            _lib.check(L.afqmc_sr(dev._h, W, _ptr(ref_w[0]),
                                  _ptr(ref_w[1]) if len(spins) == 2 else None, _ptr(ref_wt),
                                  zeta, _ptr(ref_idx), flags))
            # allocation-free entry point: no context, out of place
            out = [torch.empty_like(x) for x in w]
            wt = torch.as_tensor(weights, device=gpu)
            idx = torch.full((W,), -1, dtype=torch.int32, device=gpu)
            elems = [norb * n for n in spins] + [0]
            _lib.check(L.afqmc_sr_static(
                W, elems[0], elems[1], _ptr(w[0]), _ptr(w[1]) if len(spins) == 2 else None,
                _ptr(out[0]), _ptr(out[1]) if len(spins) == 2 else None, _ptr(wt), zeta,
                _ptr(idx), flags, ctypes.c_void_p(stream.cuda_stream)))
            torch.cuda.synchronize()
            assert torch.equal(idx, ref_idx)
            assert torch.equal(wt, ref_wt)
            for a, b in zip(out, ref_w):
                assert torch.equal(a, b)
            # the inputs were not touched
            idx_np = idx.cpu().numpy()
            assert np.all(np.diff(idx_np) >= 0) and idx_np.min() >= 0 and idx_np.max() < W


def test_sr_static_rejects_oversize_and_in_place(gpu):
    L = _lib.lib()
    w = torch.zeros((4, 3, 2), dtype=torch.complex128, device=gpu)
    out = torch.zeros_like(w)
    wt = torch.ones(4, dtype=torch.float64, device=gpu)
    rc = L.afqmc_sr_static((1 << 20) + 1, 6, 0, _ptr(w), None, _ptr(out), None, _ptr(wt), 0.5,
                           None, 0, None)
    assert rc == 3 and b"scratch" in L.afqmc_last_error()
    rc = L.afqmc_sr_static(4, 6, 0, _ptr(w), None, _ptr(w), None, _ptr(wt), 0.5, None, 0, None)
    assert rc == 1 and b"out of place" in L.afqmc_last_error()
    rc = L.afqmc_sr_static(4, 6, 0, _ptr(w), None, _ptr(out), None, _ptr(wt), 1.0, None, 0, None)
    assert rc == 1 and b"zeta" in L.afqmc_last_error()
