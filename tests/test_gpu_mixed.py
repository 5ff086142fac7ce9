"""Mixed-precision step (split-bf16 Cholesky contractions and Taylor products in the
register-resident or the streaming kernels, float64 overlaps / determinants / QR / weights /
energies) against the float64 oracle.

Tolerances: one step agrees to 1e-4 relative (split-bf16 carries 16 mantissa bits);
block energies of the device sampler driven by identical fields agree with the plain
float64 GPU run to 1e-4 Eh (the 0.1 mEh target)."""

import numpy as np
import pytest
import torch

from cs_afqmc_amd import hamiltonian, problems, propagation, rng, sampling, wavefunctions
from oracle import afqmc_oracle as o

pytestmark = pytest.mark.gpu

CASES = [
    ("rhf", 40, (8, 8), 100, 32),
    ("rhf", 100, (20, 20), 400, 16),
    # fragment sizes (n = N / 5, nchol = 4 N): every k-block count of the split-bf16
    # Taylor kernel and every measurement-kernel width
    ("rhf", 20, (4, 4), 80, 16),
    ("rhf", 50, (10, 10), 200, 16),
    ("rhf", 70, (14, 14), 280, 8),
    ("rhf", 85, (17, 17), 340, 8),
    ("rhf", 120, (24, 24), 480, 4),
    ("uhf", 36, (7, 5), 90, 24),
    # streaming split-bf16 Taylor (N > 128 or n > 32): one and two row tiles per wave,
    # every column-tile count, N not a multiple of 8 (masked row chunks), N = 256
    ("uhf", 136, (34, 30), 160, 8),
    ("rhf", 96, (40, 40), 200, 8),
    ("uhf", 200, (40, 40), 300, 4),
    ("rhf", 150, (12, 12), 120, 6),
    ("rhf", 250, (50, 50), 200, 4),
    ("uhf", 256, (64, 57), 120, 3),
]


def rel(a, b):
    if isinstance(b, (list, tuple)):
        return max(rel(x, y) for x, y in zip(a, b))
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


def to_np(x):
    if isinstance(x, (list, tuple)):
        return [to_np(y) for y in x]
    return x.detach().cpu().numpy()


def to_dev(x, dev):
    if isinstance(x, (list, tuple)):
        return [to_dev(y, dev) for y in x]
    return torch.as_tensor(np.ascontiguousarray(x), device=dev)


def build(walker_type, norb, nelec, nchol, W, precision, seed=7, dt=0.005, chol_std=None,
          asym=0.0):
    if walker_type == "rhf":
        hd, wd, nelec = problems.synthetic_rhf(norb, nelec[0], nchol, seed=seed, chol_std=chol_std)
        trial = wavefunctions.rhf(norb, nelec)
        prop = propagation.propagator_restricted(dt=dt, n_walkers=W, precision=precision)
    else:
        hd, wd, nelec = problems.synthetic_uhf(norb, nelec[0], nelec[1], nchol, seed=seed,
                                               chol_std=chol_std)
        trial = wavefunctions.uhf(norb, nelec)
        prop = propagation.propagator_unrestricted(dt=dt, n_walkers=W, precision=precision)
    if asym:
        # non-symmetric Cholesky vectors: V is no longer symmetric and is kept whole
        r = np.random.default_rng(seed + 99)
        hd["chol"] = hd["chol"] + asym * r.standard_normal(hd["chol"].shape)
    hd["h0"] = 0.37
    ham = hamiltonian.hamiltonian(norb)
    wave = dict(wd)
    hdat = ham.build_measurement_intermediates(dict(hd), trial, wave)
    hdat = ham.build_propagation_intermediates(hdat, prop, trial, wave)
    return hd, wd, nelec, trial, prop, ham, hdat, wave


# symmetric Cholesky vectors (all synthetic problems, and real molecular orbitals) keep
# V as its upper triangle; these cases break the symmetry so V is kept whole
ASYM_CASES = [("rhf", 40, (8, 8), 100, 32), ("uhf", 36, (7, 5), 90, 24),
              ("uhf", 136, (34, 30), 160, 8), ("rhf", 250, (50, 50), 200, 4)]


@pytest.mark.parametrize("case,asym",
                         [(c, 0.0) for c in CASES] + [(c, 0.01) for c in ASYM_CASES],
                         ids=[f"{c[0]}-N{c[1]}" for c in CASES]
                         + [f"{c[0]}-N{c[1]}-asym" for c in ASYM_CASES])
def test_mixed_step_close_to_oracle(gpu, case, asym):
    walker_type, norb, nelec, nchol, W = case
    print(f"mixed case {case} asym {asym}", flush=True)
    cs = 0.5 / np.sqrt(norb)
    hd, wd, nelec, trial, prop, ham, hdat, wave = build(walker_type, norb, nelec, nchol, W,
                                                        "bf16x3", chol_std=cs, asym=asym * cs)
    dt = prop.dt
    oh = o.build_measurement_intermediates(walker_type, dict(hd), wd, norb)
    oh = o.build_propagation_intermediates(walker_type, oh, wd["rdm1"], norb, nelec, dt)
    r = np.random.default_rng(1)
    walkers = problems.random_walkers(walker_type, norb, nelec, W, r, near=wd["mo_coeff"],
                                      noise=0.15)
    fields = r.standard_normal((W, nchol))
    weights = 0.5 + r.random(W)
    ovl = o.calc_overlap(walker_type, walkers, wd)
    oprop = {"walkers": walkers, "weights": weights.copy(), "overlaps": ovl.copy(),
             "e_estimate": -1.3, "pop_control_ene_shift": -1.25}
    gprop = {"walkers": to_dev(walkers, gpu), "weights": to_dev(weights, gpu),
             "overlaps": to_dev(ovl, gpu),
             "e_estimate": torch.tensor(-1.3, dtype=torch.float64, device=gpu),
             "pop_control_ene_shift": torch.tensor(-1.25, dtype=torch.float64, device=gpu)}
    fb_out = torch.empty((W, nchol), dtype=torch.complex128, device=gpu)
    gprop = prop.propagate(trial, hdat, gprop, to_dev(fields, gpu), wave, force_bias_out=fb_out)
    oprop, fb = o.propagate(walker_type, walker_type, oh, oprop, fields, wd, dt, 6, norb,
                            return_force_bias=True)
    assert rel(to_np(fb_out), fb) < 1e-4
    assert rel(to_np(gprop["walkers"]), oprop["walkers"]) < 1e-4
    assert rel(to_np(gprop["overlaps"]), oprop["overlaps"]) < 1e-4
    assert rel(to_np(gprop["weights"]), oprop["weights"]) < 1e-4
    # and the mixed mode is really not the float64 path
    assert rel(to_np(fb_out), fb) > 1e-13


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}-N{c[1]}" for c in CASES])
def test_mixed_energy_close_to_oracle(gpu, case):
    """Local energies with the split-bf16 exchange contraction (float64 Green's
    function, c_g, one-body term and accumulation): 1e-5 relative."""
    walker_type, norb, nelec, nchol, W = case
    print(f"mixed energy case {case}", flush=True)
    cs = 0.5 / np.sqrt(norb)
    hd, wd, nelec, trial, prop, ham, hdat, wave = build(walker_type, norb, nelec, nchol, W,
                                                        "bf16x3", chol_std=cs)
    oh = o.build_measurement_intermediates(walker_type, dict(hd), wd, norb)
    r = np.random.default_rng(4)
    walkers = problems.random_walkers(walker_type, norb, nelec, W, r, near=wd["mo_coeff"],
                                      noise=0.15)
    got = to_np(trial.calc_energy(to_dev(walkers, gpu), hdat, wave))
    ref = o.calc_energy(walker_type, walkers, oh, wd)
    assert rel(got, ref) < 1e-5
    assert rel(got, ref) > 1e-14  # the exchange really ran in split-bf16


@pytest.mark.parametrize("walker_type", ["rhf", "uhf"])
def test_mixed_energy_matches_fp64_sampler(gpu, walker_type):
    """Same device fields and SR draws: block energies of the mixed and the float64 runs
    differ by less than 0.1 mEh."""
    norb, nchol, W = 60, 200, 128
    nelec = (12, 12) if walker_type == "rhf" else (12, 10)
    energies = {}
    for precision in ("fp64", "bf16x3"):
        hd, wd, ne, trial, prop, ham, hdat, wave = build(walker_type, norb, nelec, nchol, W,
                                                         precision, seed=3, dt=0.005)
        pd = prop.init_prop_data(trial, wave, hdat)
        pd["key"] = rng.PRNGKey(11)
        smp = sampling.sampler(n_prop_steps=25, n_ene_blocks=2, n_sr_blocks=2)
        e, pd = smp.propagate_phaseless(ham, hdat, prop, pd, trial, wave)
        energies[precision] = (e, to_np(pd["block_energies"]))
        print(precision, e, flush=True)
    e64, b64 = energies["fp64"]
    emx, bmx = energies["bf16x3"]
    assert np.all(np.isfinite(b64)) and np.all(np.isfinite(bmx))
    assert np.max(np.abs(bmx - b64)) < 1e-4
    assert abs(emx - e64) < 1e-4
