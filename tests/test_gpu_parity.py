"""GPU parity: the HIP step against the CPU oracle on the same seeded inputs and the
same auxiliary fields.  Tolerance 1e-10 relative (max abs error / max abs value) for
the float64 path; SR indices bit-exact."""

import sys

import numpy as np
import pytest
import torch

from cs_afqmc_amd import hamiltonian, problems, propagation, wavefunctions
from oracle import afqmc_oracle as o

pytestmark = pytest.mark.gpu

TOL = 1e-10

# walker_type, norb, nelec, nchol, n_walkers
CASES = [
    ("rhf", 12, (4, 4), 20, 16),
    ("uhf", 12, (4, 3), 20, 16),
    ("rhf", 40, (8, 8), 100, 64),
    ("uhf", 36, (7, 5), 90, 48),
    ("rhf", 100, (20, 20), 400, 32),
    ("uhf", 70, (17, 13), 150, 24),
    # beyond the register-resident Taylor kernel (N > 128, n > 32): streaming form
    ("uhf", 136, (34, 30), 160, 8),
]
IDS = [f"{c[0]}-N{c[1]}-n{c[2][0]}.{c[2][1]}-G{c[3]}-W{c[4]}" for c in CASES]


def rel(a, b):
    if isinstance(b, (list, tuple)):
        return max(rel(x, y) for x, y in zip(a, b))
    a, b = np.asarray(a), np.asarray(b)
    scale = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / (scale if scale > 0 else 1.0)


def to_np(x):
    if isinstance(x, (list, tuple)):
        return [to_np(y) for y in x]
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    return np.asarray(x)


def to_dev(x, dev):
    if isinstance(x, (list, tuple)):
        return [to_dev(y, dev) for y in x]
    return torch.as_tensor(np.ascontiguousarray(x), device=dev)


class Case:
    def __init__(self, walker_type, norb, nelec, nchol, W, dev, seed=7, dt=0.005):
        print(f"case {walker_type} N={norb} nelec={nelec} G={nchol} W={W}", flush=True)
        sys.stdout.flush()
        self.wt, self.norb, self.W, self.dt = walker_type, norb, W, dt
        # Cholesky scale large enough that exp(VHS) is far from the identity
        cs = 0.5 / np.sqrt(norb)
        if walker_type == "rhf":
            hd, wd, nelec = problems.synthetic_rhf(norb, nelec[0], nchol, seed=seed, chol_std=cs)
            self.trial = wavefunctions.rhf(norb, nelec)
            self.prop = propagation.propagator_restricted(dt=dt, n_walkers=W)
        else:
            hd, wd, nelec = problems.synthetic_uhf(norb, nelec[0], nelec[1], nchol, seed=seed,
                                                   chol_std=cs)
            self.trial = wavefunctions.uhf(norb, nelec)
            self.prop = propagation.propagator_unrestricted(dt=dt, n_walkers=W)
        hd["h0"] = 0.37
        self.nelec = nelec
        self.nchol = nchol
        # oracle intermediates (independent host computation)
        oh = o.build_measurement_intermediates(walker_type, dict(hd), wd, norb)
        self.oham = o.build_propagation_intermediates(walker_type, oh, wd["rdm1"], norb,
                                                      nelec, dt)
        self.owave = {"mo_coeff": wd["mo_coeff"], "rdm1": wd["rdm1"]}
        # product objects
        ham = hamiltonian.hamiltonian(norb)
        self.wave = dict(wd)
        self.ham = ham.build_measurement_intermediates(dict(hd), self.trial, self.wave)
        self.ham = ham.build_propagation_intermediates(self.ham, self.prop, self.trial,
                                                       self.wave)
        rng = np.random.default_rng(seed + 1)
        base = wd["mo_coeff"]
        self.walkers_np = problems.random_walkers(walker_type, norb, nelec, W, rng, near=base,
                                                  noise=0.15)
        self.dev = dev

    def walkers(self):
        return to_dev(self.walkers_np, self.dev)


@pytest.fixture(scope="module", params=CASES, ids=IDS)
def case(request, gpu):
    return Case(*request.param, dev=gpu)


def test_intermediates_match_oracle(case):
    assert rel(case.ham["mf_shifts"], case.oham["mf_shifts"]) < 1e-13
    assert rel(case.ham["exp_h1"], case.oham["exp_h1"]) < 1e-13
    assert rel(to_np(case.ham["rot_chol"]), case.oham["rot_chol"]) < 1e-13
    assert abs(case.ham["h0_prop"] - case.oham["h0_prop"]) < 1e-12 * (1 + abs(case.oham["h0_prop"]))


def test_context_from_host_arrays(case):
    """afqmc_ctx_create copies the intermediates from host memory (numpy arrays) as it
    does from device memory (the Cholesky and rotated Cholesky tensors the set-up leaves
    on the device): same energies and force biases to the bit."""
    from cs_afqmc_amd.device import DeviceStep
    hd, dev = case.ham, case.ham["_afqmc_device"]
    host = DeviceStep(case.wt, case.norb, case.nelec, case.W, case.dt, 6, hd["h0"],
                      hd["h0_prop"], np.asarray(hd["chol"]), case.wave["mo_coeff"],
                      hd["rot_h1"], to_np(hd["rot_chol"]), hd["mf_shifts"], hd["exp_h1"])
    w = case.walkers()
    assert torch.equal(host.energy(w), dev.energy(w))
    assert torch.equal(host.force_bias(w), dev.force_bias(w))


def test_overlap(case):
    got = to_np(case.trial.calc_overlap(case.walkers(), case.wave))
    ref = o.calc_overlap(case.wt, case.walkers_np, case.owave)
    assert rel(got, ref) < TOL


def test_force_bias(case):
    got = to_np(case.trial.calc_force_bias(case.walkers(), case.ham, case.wave))
    ref = o.calc_force_bias(case.wt, case.walkers_np, case.oham, case.owave)
    assert got.shape == (case.W, case.nchol)
    assert rel(got, ref) < TOL


def test_energy(case):
    got = to_np(case.trial.calc_energy(case.walkers(), case.ham, case.wave))
    ref = o.calc_energy(case.wt, case.walkers_np, case.oham, case.owave)
    assert rel(got, ref) < TOL


def _prop_pair(case, rng):
    weights = 0.5 + rng.random(case.W)
    ovl = o.calc_overlap(case.wt, case.walkers_np, case.owave)
    e_est, pcs = -1.3, -1.25
    oprop = {"walkers": case.walkers_np, "weights": weights.copy(), "overlaps": ovl.copy(),
             "e_estimate": e_est, "pop_control_ene_shift": pcs}
    dev = case.dev
    gprop = {"walkers": case.walkers(), "weights": to_dev(weights, dev),
             "overlaps": to_dev(ovl, dev),
             "e_estimate": torch.tensor(e_est, dtype=torch.float64, device=dev),
             "pop_control_ene_shift": torch.tensor(pcs, dtype=torch.float64, device=dev)}
    return oprop, gprop


def test_propagate_one_step(case):
    rng = np.random.default_rng(3)
    oprop, gprop = _prop_pair(case, rng)
    fields = rng.standard_normal((case.W, case.nchol))
    fb_out = torch.empty((case.W, case.nchol), dtype=torch.complex128, device=case.dev)
    gprop = case.prop.propagate(case.trial, case.ham, gprop, to_dev(fields, case.dev),
                                case.wave, force_bias_out=fb_out)
    oprop, fb = o.propagate(case.wt, case.wt, case.oham, oprop, fields, case.owave, case.dt,
                            6, case.norb, return_force_bias=True)
    assert rel(to_np(fb_out), fb) < TOL
    gw, ow = to_np(gprop["walkers"]), oprop["walkers"]
    if case.wt == "rhf":
        assert rel(gw, ow) < TOL
    else:
        assert rel(gw[0], ow[0]) < TOL and rel(gw[1], ow[1]) < TOL
    assert rel(to_np(gprop["overlaps"]), oprop["overlaps"]) < TOL
    assert rel(to_np(gprop["weights"]), oprop["weights"]) < TOL
    assert abs(float(gprop["pop_control_ene_shift"]) - oprop["pop_control_ene_shift"]) < 1e-9


def test_propagate_five_steps(case):
    rng = np.random.default_rng(4)
    oprop, gprop = _prop_pair(case, rng)
    for _ in range(5):
        fields = rng.standard_normal((case.W, case.nchol))
        gprop = case.prop.propagate(case.trial, case.ham, gprop, to_dev(fields, case.dev),
                                    case.wave)
        oprop = o.propagate(case.wt, case.wt, case.oham, oprop, fields, case.owave, case.dt,
                            6, case.norb)
    gw, ow = to_np(gprop["walkers"]), oprop["walkers"]
    if case.wt == "rhf":
        assert rel(gw, ow) < TOL
    else:
        assert rel(gw[0], ow[0]) < TOL and rel(gw[1], ow[1]) < TOL
    assert rel(to_np(gprop["weights"]), oprop["weights"]) < TOL
    assert rel(to_np(gprop["overlaps"]), oprop["overlaps"]) < TOL


def test_orthonormalize(case):
    gprop = {"walkers": case.walkers(), "_afqmc_device": case.ham["_afqmc_device"]}
    gprop = case.prop.orthonormalize_walkers(gprop)
    ref = o.orthonormalize_walkers(case.wt, case.walkers_np)
    got = to_np(gprop["walkers"])
    if case.wt == "rhf":
        assert rel(got, ref) < 1e-12
    else:
        assert rel(got[0], ref[0]) < 1e-12 and rel(got[1], ref[1]) < 1e-12


@pytest.mark.parametrize("wt,norb,nelec", [("rhf", 200, (40, 40)), ("uhf", 256, (64, 57))],
                         ids=["rhf-N200-n40", "uhf-N256-n64.57"])
def test_orthonormalize_large_walkers(gpu, wt, norb, nelec):
    """QR of walkers copied to LDS (up to 156 KiB: N = 200, n = 40 is 125 KiB) and of
    walkers factorised in place in global memory (N = 256, n = 64: 256 KiB)."""
    case = Case(wt, norb, nelec, 8, 4, dev=gpu)
    test_orthonormalize(case)


@pytest.mark.parametrize("global_order", [False, True], ids=["local", "global"])
def test_sr_indices_bit_exact(case, global_order):
    rng = np.random.default_rng(9)
    weights = rng.random(case.W) * (rng.random(case.W) > 0.3)
    weights[0] = 0.7
    zeta = float(rng.random())
    dev_ctx = case.ham["_afqmc_device"]
    walkers = case.walkers()
    w_t = to_dev(weights, case.dev)
    idx = torch.empty(case.W, dtype=torch.int32, device=case.dev)
    dev_ctx.sr(walkers, w_t, zeta, idx_out=idx, global_order=global_order)
    ref_walkers, ref_w, ref_idx = o.stochastic_reconfiguration(case.wt, case.walkers_np,
                                                               weights, zeta,
                                                               global_order=global_order)
    assert np.array_equal(to_np(idx), ref_idx)
    got = to_np(walkers)
    if case.wt == "rhf":
        assert np.array_equal(got, ref_walkers)
    else:
        assert np.array_equal(got[0], ref_walkers[0]) and np.array_equal(got[1], ref_walkers[1])
    assert np.array_equal(to_np(w_t), ref_w)


def test_propagate_reuse_green(case):
    """Steps 2..5 reuse the Green's function the previous step left in the workspace."""
    rng = np.random.default_rng(5)
    oprop, gprop = _prop_pair(case, rng)
    dev = case.ham["_afqmc_device"]
    scalars = torch.tensor([oprop["e_estimate"], oprop["pop_control_ene_shift"]],
                           dtype=torch.float64, device=case.dev)
    for s in range(5):
        fields = rng.standard_normal((case.W, case.nchol))
        dev.propagate(gprop["walkers"], gprop["weights"], gprop["overlaps"],
                      to_dev(fields, case.dev), scalars, reuse_green=s > 0)
        oprop = o.propagate(case.wt, case.wt, case.oham, oprop, fields, case.owave, case.dt,
                            6, case.norb)
    assert rel(to_np(gprop["walkers"]), oprop["walkers"]) < TOL
    assert rel(to_np(gprop["weights"]), oprop["weights"]) < TOL
    assert rel(to_np(gprop["overlaps"]), oprop["overlaps"]) < TOL
    assert abs(float(scalars[1]) - oprop["pop_control_ene_shift"]) < 1e-9
