"""Correlated sampling: two systems with the same fields, against the oracle's
restatement of corr_sample.cs_block_scan + cs_steps_scan (float64, 1e-9 relative in
the local sums), and the device outlier rule against the oracle's (exact zero pattern)."""

import numpy as np
import pytest
import torch

from cs_afqmc_amd import corr_sample, hamiltonian, problems, propagation, wavefunctions
from oracle import afqmc_oracle as o

pytestmark = pytest.mark.gpu

NORB, NOCC, NCHOL, W, STEPS, DT = 24, 5, 60, 32, 6, 0.005


def systems(seed=3):
    cs = 0.5 / np.sqrt(NORB)
    hd1, wd, ne = problems.synthetic_rhf(NORB, NOCC, NCHOL, seed=seed, chol_std=cs)
    hd2 = dict(hd1)
    r = np.random.default_rng(seed + 1)
    dh = 0.02 * r.standard_normal((NORB, NORB))
    hd2["h1"] = hd1["h1"] + (dh + dh.T)[None]  # a perturbed one-body term
    hd2["chol"] = hd1["chol"] * 1.01
    out = []
    for hd in (hd1, hd2):
        trial = wavefunctions.rhf(NORB, ne)
        prop = propagation.propagator_restricted(dt=DT, n_walkers=W)
        ham = hamiltonian.hamiltonian(NORB)
        wave = dict(wd)
        h = ham.build_measurement_intermediates(dict(hd), trial, wave)
        h = ham.build_propagation_intermediates(h, prop, trial, wave)
        pd = prop.init_prop_data(trial, wave, h)
        oh = o.build_measurement_intermediates("rhf", dict(hd), wd, NORB)
        oh = o.build_propagation_intermediates("rhf", oh, wd["rdm1"], NORB, ne, DT)
        out.append(((h, prop, trial, wave, pd), (oh, wd)))
    return out


def to_np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("grouped", [True, False], ids=["grouped", "per-context"])
def test_cs_blocks_match_oracle(gpu, grouped):
    sy = systems()
    pair = corr_sample.CorrelatedPair(sy[0][0], sy[1][0], n_prop_steps=STEPS, grouped=grouped)
    osys = []
    for (h, prop, trial, wave, pd), (oh, wd) in sy:
        e = float(pd["e_estimate"])
        osys.append({"walker_type": "rhf", "ham_data": oh, "wave_data": wd, "norb": NORB,
                     "prop_data": {"walkers": to_np(pd["walkers"]), "weights": to_np(pd["weights"]),
                                   "overlaps": to_np(pd["overlaps"]), "e_estimate": e,
                                   "pop_control_ene_shift": e}})
    rng = np.random.default_rng(9)
    for b in range(2):
        fields = rng.standard_normal((STEPS, W, NCHOL))
        got = [float(x) for x in pair.block(torch.as_tensor(fields, device=gpu))]
        pds, ref = o.cs_block(osys, fields, DT, 6)
        for i in range(2):
            osys[i]["prop_data"] = pds[i]
        print(f"block {b}: device {got} oracle {ref}", flush=True)
        for g, r in zip(got, ref):
            assert abs(g - r) <= 1e-9 * max(1.0, abs(r))
        for i in range(2):
            wr = pds[i]["weights"]
            assert np.max(np.abs(to_np(pair.s[i].pd["weights"]) - wr)) <= 1e-9 * np.max(np.abs(wr))


def test_device_outlier_rule_matches_oracle(gpu):
    sy = systems(seed=5)
    dev = sy[0][0][0]["_afqmc_device"]
    r = np.random.default_rng(4)
    w1 = r.random(W) + 0.5
    w2 = r.random(W) + 0.5
    w1[3] = 40.0   # outlier in system 1
    w2[17] = 55.0  # outlier in system 2
    t1, t2 = torch.as_tensor(w1.copy(), device=gpu), torch.as_tensor(w2.copy(), device=gpu)
    dev.cs_outliers(t1, t2, 5.0)
    a, b = o.corr_outlier_rm(w1, w2, 5.0)
    assert np.array_equal(to_np(t1) == 0.0, a == 0.0) and np.array_equal(to_np(t2) == 0.0, b == 0.0)
    assert np.array_equal(to_np(t1), a) and np.array_equal(to_np(t2), b)
    assert a[3] == 0.0 and b[3] == 0.0 and a[17] == 0.0 and b[17] == 0.0
