"""FP32 precision mode (AFQMC_PREC_FP32): the Cholesky contractions -- force bias, VHS and
the local-energy exchange -- on float32 MFMA (float64 operands rounded to float32 as they
are staged, float32 accumulation), everything else float64, against the float64 oracle.

Tolerances: force bias, one step and local energies 2e-5 relative (float32 carries 24
mantissa bits; the contractions sum up to a few thousand products); block energies of
the device sampler with identical fields against the float64 GPU run 1e-5 Eh."""

import numpy as np
import pytest
import torch

from cs_afqmc_amd import batch, hamiltonian, problems, propagation, rng, sampling, wavefunctions
from oracle import afqmc_oracle as o

pytestmark = pytest.mark.gpu

TOL = 2e-5
CASES = [
    ("rhf", 40, (8, 8), 100, 32),
    ("rhf", 100, (20, 20), 400, 16),
    ("uhf", 36, (7, 5), 90, 24),
    ("uhf", 136, (34, 30), 160, 8),   # streaming float64 Taylor kernel
    ("rhf", 57, (11, 11), 229, 12),   # odd sizes: unvectorised operand loads
]
IDS = [f"{c[0]}-N{c[1]}-n{c[2][0]}.{c[2][1]}-G{c[3]}-W{c[4]}" for c in CASES]


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


def build(walker_type, norb, nelec, nchol, W, precision, seed=7, dt=0.005):
    cs = 0.5 / np.sqrt(norb)
    if walker_type == "rhf":
        hd, wd, nelec = problems.synthetic_rhf(norb, nelec[0], nchol, seed=seed, chol_std=cs)
        trial = wavefunctions.rhf(norb, nelec)
        prop = propagation.propagator_restricted(dt=dt, n_walkers=W, precision=precision)
    else:
        hd, wd, nelec = problems.synthetic_uhf(norb, nelec[0], nelec[1], nchol, seed=seed,
                                               chol_std=cs)
        trial = wavefunctions.uhf(norb, nelec)
        prop = propagation.propagator_unrestricted(dt=dt, n_walkers=W, precision=precision)
    hd["h0"] = 0.37
    ham = hamiltonian.hamiltonian(norb)
    wave = dict(wd)
    hdat = ham.build_measurement_intermediates(dict(hd), trial, wave)
    hdat = ham.build_propagation_intermediates(hdat, prop, trial, wave)
    return hd, wd, nelec, trial, prop, ham, hdat, wave


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_fp32_step_force_bias_energy_match_oracle(gpu, case):
    walker_type, norb, nelec, nchol, W = case
    print(f"fp32 {case}", flush=True)
    hd, wd, nelec, trial, prop, ham, hdat, wave = build(*case, precision="fp32")
    dev = hdat["_afqmc_device"]
    rs = np.random.default_rng(3)
    walkers = problems.random_walkers(walker_type, norb, nelec, W, rs, near=wd["mo_coeff"],
                                      noise=0.12)
    oh = o.build_measurement_intermediates(walker_type, dict(hd), wd, norb)
    oh = o.build_propagation_intermediates(walker_type, oh, wd["rdm1"], norb, nelec, 0.005)
    # force bias and local energy of the same walkers
    fb = to_np(dev.force_bias(to_dev(walkers, gpu)))
    assert rel(fb, o.calc_force_bias(walker_type, walkers, oh, wd)) < TOL
    en = to_np(dev.energy(to_dev(walkers, gpu)))
    assert rel(en, o.calc_energy(walker_type, walkers, oh, wd)) < TOL
    # one phaseless step
    fields = rs.standard_normal((W, nchol))
    w0 = rs.uniform(0.5, 1.5, W)
    gpd = {"walkers": to_dev(walkers, gpu), "weights": to_dev(w0, gpu),
           "e_estimate": torch.tensor(-1.0, dtype=torch.float64, device=gpu),
           "pop_control_ene_shift": torch.tensor(-1.1, dtype=torch.float64, device=gpu)}
    gpd["overlaps"] = trial.calc_overlap(gpd["walkers"], wave)
    gpd = prop.propagate(trial, hdat, gpd, to_dev(fields, gpu), wave)
    opd = {"walkers": walkers, "weights": w0, "e_estimate": -1.0, "pop_control_ene_shift": -1.1,
           "overlaps": o.calc_overlap(walker_type, walkers, wd)}
    ref = o.propagate(walker_type, walker_type, oh, opd, fields, wd, 0.005, 6, norb)
    assert rel(to_np(gpd["walkers"]), ref["walkers"]) < TOL
    assert rel(to_np(gpd["weights"]), ref["weights"]) < TOL
    assert rel(to_np(gpd["overlaps"]), ref["overlaps"]) < TOL


@pytest.mark.parametrize("walker_type", ["rhf", "uhf"])
def test_fp32_sampler_blocks_match_fp64(gpu, walker_type):
    spec = ("rhf", 60, (12, 12), 240, 64) if walker_type == "rhf" else \
        ("uhf", 48, (10, 9), 190, 64)
    block = {}
    for precision in ("fp64", "fp32"):
        hd, wd, nelec, trial, prop, ham, hdat, wave = build(*spec, precision=precision, seed=11)
        pd = prop.init_prop_data(trial, wave, hdat)
        pd["key"] = rng.PRNGKey(5)
        smp = sampling.sampler(n_prop_steps=20, n_ene_blocks=2, n_sr_blocks=2)
        e, pd = smp.propagate_phaseless(ham, hdat, prop, pd, trial, wave)
        block[precision] = to_np(pd["block_energies"])
        print(precision, block[precision], flush=True)
    assert np.all(np.isfinite(block["fp64"]))
    assert np.max(np.abs(block["fp32"] - block["fp64"])) < 1e-5


def test_fp32_group_matches_single_context(gpu):
    """The grouped FP32 step (no split-K in the force bias) against each fragment's own
    context (split-K slices added in float64): block energies to 1e-6 relative."""
    specs = [("rhf", 20, (4, 4), 80, 32, 1), ("uhf", 30, (6, 5), 120, 16, 2),
             ("rhf", 45, (9, 9), 180, 24, 3)]

    def make(spec):
        wt, norb, nelec, nchol, W, seed = spec
        cs = 0.5 / np.sqrt(norb)
        if wt == "rhf":
            hd, wd, ne = problems.synthetic_rhf(norb, nelec[0], nchol, seed=seed, chol_std=cs)
        else:
            hd, wd, ne = problems.synthetic_uhf(norb, nelec[0], nelec[1], nchol, seed=seed,
                                                chol_std=cs)
        return batch.Fragment(hd, wd, wt, ne, W, dt=0.005, seed=seed, precision="fp32",
                              n_prop_steps=8)

    grouped = [make(s) for s in specs]
    bt = batch.FragmentBatch(grouped)
    for _ in range(2):
        bt.run_block()
    for spec, fg in zip(specs, grouped):
        fs = make(spec)
        for _ in range(2):
            fs.run_block_single(use_graph=False)
        a, b = fg.block_energies()[0], fs.block_energies()[0]
        print(spec, a, b, flush=True)
        assert np.all(np.isfinite(a))
        assert np.max(np.abs(a - b)) < 1e-6 * np.max(np.abs(b))
