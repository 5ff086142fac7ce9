"""Spin-resolved Cholesky vectors (the reference's prop_unrestricted package: chol
(2, G, N*N), alpha VHS from L^a, beta VHS from L^b) on the GPU against the oracle.

Tolerances: float64 1e-10 relative (max abs error / max abs value), split-bf16 1e-4,
FP32 2e-5 (the arithmetic of the Cholesky contractions, as for spin-summed vectors).
Weights and overlaps carry the determinants of n_up + n_dn propagated orbitals, whose
relative error can reach (n_up + n_dn) times an element's: for split-bf16 (2^-17 per
product) they are held to max(1e-4, (n_up + n_dn) 2^-17).
With both spins given the same vectors the context reproduces the spin-summed
unrestricted context's block energies to 1e-10 relative in float64 (the mean-field
shifts are summed in another order) and 1e-6 in split-bf16 (a last-bit difference of a
shifted field can move its bf16 rounding); grouped fragments match their
own single contexts to 1e-9 (float64) and 1e-5 (split-bf16) in block energies."""

import numpy as np
import pytest
import torch

from cs_afqmc_amd import batch, hamiltonian, problems, propagation, rng, sampling
from cs_afqmc_amd import wavefunctions
from cs_afqmc_amd.prop_unrestricted import propagation as pu_prop
from cs_afqmc_amd.prop_unrestricted import wavefunctions_unrestricted as pu_wf
from oracle import afqmc_oracle as o

pytestmark = pytest.mark.gpu

DT = 0.005
CASES = [
    # norb, nelec, nchol, W
    (12, (4, 3), 20, 16),
    (36, (7, 5), 90, 48),
    (70, (17, 13), 150, 24),
    (136, (34, 30), 160, 8),  # streaming float64 Taylor kernel (N > 128)
]
IDS = [f"N{c[0]}-n{c[1][0]}.{c[1][1]}-G{c[2]}-W{c[3]}" for c in CASES]
TOLS = {"fp64": 1e-10, "bf16x3": 1e-4, "fp32": 2e-5}


def rel(a, b):
    if isinstance(b, (list, tuple)):
        return max(rel(x, y) for x, y in zip(a, b))
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (s if s > 0 else 1.0)


def to_np(x):
    if isinstance(x, (list, tuple)):
        return [to_np(y) for y in x]
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)


def to_dev(x, dev):
    if isinstance(x, (list, tuple)):
        return [to_dev(y, dev) for y in x]
    return torch.as_tensor(np.ascontiguousarray(x), device=dev)


def problem(norb, nelec, nchol, seed=7):
    cs = 0.5 / np.sqrt(norb)
    hd, wd, ne = problems.synthetic_uhf_spin(norb, nelec[0], nelec[1], nchol, seed=seed,
                                             chol_std=cs)
    hd["h0"] = 0.37
    return hd, wd, ne


def build(hd, wd, norb, nelec, W, precision):
    trial = pu_wf.uhf(norb, nelec)
    prop = pu_prop.propagator_unrestricted(dt=DT, n_walkers=W, precision=precision)
    ham = hamiltonian.hamiltonian(norb)
    wave = dict(wd)
    hdat = ham.build_measurement_intermediates(dict(hd), trial, wave)
    hdat = ham.build_propagation_intermediates(hdat, prop, trial, wave)
    return trial, prop, hdat, wave


def oracle_ham(hd, wd, norb):
    oh = o.build_measurement_intermediates_spin(dict(hd), wd, norb)
    return o.build_propagation_intermediates_spin(oh, wd["rdm1"], norb, DT)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3", "fp32"])
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_spin_chol_step_matches_oracle(gpu, case, precision):
    norb, nelec, nchol, W = case
    print(f"spin-chol {precision} {case}", flush=True)
    tol = TOLS[precision]
    hd, wd, nelec = problem(*case[:3])
    trial, prop, hdat, wave = build(hd, wd, norb, nelec, W, precision)
    oh = oracle_ham(hd, wd, norb)
    assert rel(hdat["mf_shifts"], oh["mf_shifts"]) < 1e-13
    assert rel(hdat["exp_h1"], oh["exp_h1"]) < 1e-12
    rs = np.random.default_rng(3)
    walkers = problems.random_walkers("uhf", norb, nelec, W, rs, near=wd["mo_coeff"],
                                      noise=0.15)
    owd = {"mo_coeff": wd["mo_coeff"], "rdm1": wd["rdm1"]}
    dw = to_dev(walkers, gpu)
    assert rel(to_np(trial.calc_force_bias(dw, hdat, wave)),
               o.calc_force_bias("uhf", walkers, oh, owd)) < tol
    assert rel(to_np(trial.calc_energy(dw, hdat, wave)),
               o.calc_energy("uhf", walkers, oh, owd)) < tol
    weights = 0.5 + rs.random(W)
    ovl = o.calc_overlap("uhf", walkers, owd)
    opd = {"walkers": walkers, "weights": weights.copy(), "overlaps": ovl.copy(),
           "e_estimate": -1.3, "pop_control_ene_shift": -1.25}
    gpd = {"walkers": to_dev(walkers, gpu), "weights": to_dev(weights, gpu),
           "overlaps": to_dev(ovl, gpu),
           "e_estimate": torch.tensor(-1.3, dtype=torch.float64, device=gpu),
           "pop_control_ene_shift": torch.tensor(-1.25, dtype=torch.float64, device=gpu)}
    steps = 5 if precision == "fp64" else 1
    for _ in range(steps):
        fields = rs.standard_normal((W, nchol))
        gpd = prop.propagate(trial, hdat, gpd, to_dev(fields, gpu), wave)
        opd = o.propagate("uhf_spin", "uhf", oh, opd, fields, owd, DT, 6, norb)
    gw = to_np(gpd["walkers"])
    assert rel(gw[0], opd["walkers"][0]) < tol and rel(gw[1], opd["walkers"][1]) < tol
    tol_det = max(tol, sum(nelec) * 2.0**-17) if precision == "bf16x3" else tol
    assert rel(to_np(gpd["weights"]), opd["weights"]) < tol_det
    assert rel(to_np(gpd["overlaps"]), opd["overlaps"]) < tol_det


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("bf16x3", 1e-6)])
def test_shared_vectors_match_spin_summed_context(gpu, precision, tol):
    """Both spins given the same vectors and one-body matrix: the spin-resolved sampler
    and the spin-summed unrestricted sampler draw the same fields and agree."""
    norb, nelec, nchol, W = 40, (8, 6), 120, 64
    hd, wd, ne = problem(norb, nelec, nchol, seed=13)
    hd_same = dict(hd, chol=np.array([hd["chol"][0], hd["chol"][0]]),
                   h1=np.array([hd["h1"][0], hd["h1"][0]]))
    hd_flat = dict(hd, chol=hd["chol"][0], h1=hd_same["h1"])
    energies = []
    for spin, h in ((True, hd_same), (False, hd_flat)):
        if spin:
            trial, prop, hdat, wave = build(h, wd, norb, ne, W, precision)
        else:
            trial = wavefunctions.uhf(norb, ne)
            prop = propagation.propagator_unrestricted(dt=DT, n_walkers=W,
                                                       precision=precision)
            ham = hamiltonian.hamiltonian(norb)
            wave = dict(wd)
            hdat = ham.build_measurement_intermediates(dict(h), trial, wave)
            hdat = ham.build_propagation_intermediates(hdat, prop, trial, wave)
        pd = prop.init_prop_data(trial, wave, hdat)
        pd["key"] = rng.PRNGKey(21)
        smp = sampling.sampler(n_prop_steps=10, n_ene_blocks=2, n_sr_blocks=2)
        _, pd = smp.propagate_phaseless(None, hdat, prop, pd, trial, wave)
        energies.append(to_np(pd["block_energies"]))
        print(precision, "spin" if spin else "flat", energies[-1], flush=True)
    assert np.all(np.isfinite(energies[0]))
    assert np.max(np.abs(energies[0] - energies[1])) <= tol * np.max(np.abs(energies[1]))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-9), ("bf16x3", 1e-5)])
def test_spin_chol_fragments_grouped_match_single(gpu, precision, tol):
    """Spin-resolved fragments grouped with spin-summed ones: each fragment's block
    energies against its own single-context run."""
    specs = [("spin", 30, (6, 5), 100, 32, 1), ("rhf", 24, (5, 5), 80, 16, 2),
             ("spin", 140, (30, 28), 150, 8, 3), ("uhf", 36, (7, 5), 90, 16, 4)]

    def make(spec):
        kind, norb, nelec, nchol, W, seed = spec
        cs = 0.5 / np.sqrt(norb)
        if kind == "spin":
            hd, wd, ne = problems.synthetic_uhf_spin(norb, nelec[0], nelec[1], nchol,
                                                     seed=seed, chol_std=cs)
            wt = "uhf"
        elif kind == "rhf":
            hd, wd, ne = problems.synthetic_rhf(norb, nelec[0], nchol, seed=seed, chol_std=cs)
            wt = "rhf"
        else:
            hd, wd, ne = problems.synthetic_uhf(norb, nelec[0], nelec[1], nchol, seed=seed,
                                                chol_std=cs)
            wt = "uhf"
        return batch.Fragment(hd, wd, wt, ne, W, dt=DT, seed=seed, precision=precision,
                              n_prop_steps=8)

    grouped = [make(s) for s in specs]
    assert grouped[0].dev.spin_chol and not grouped[3].dev.spin_chol
    bt = batch.FragmentBatch(grouped)
    for _ in range(2):
        bt.run_block()
    for spec, fg in zip(specs, grouped):
        fs = make(spec)
        for _ in range(2):
            fs.run_block_single(use_graph=False)
        a, b = fg.block_energies()[0], fs.block_energies()[0]
        print(precision, spec, a, b, flush=True)
        assert np.all(np.isfinite(a))
        assert np.max(np.abs(a - b)) <= tol * np.max(np.abs(b))
