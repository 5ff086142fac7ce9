"""Serving loop: requests that arrive while others run are admitted between ticks, every
request completes, and each result is bit-identical to the same request served alone
(tolerance 0: a fragment's kernels never mix data or summation order with its batch
neighbours)."""

import numpy as np
import pytest

from cs_afqmc_amd import problems, serving

pytestmark = pytest.mark.gpu

SPECS = [
    # norb, nocc, nchol, W, arrival (s)
    (20, 4, 80, 32, 0.00),
    (28, 5, 112, 16, 0.00),
    (36, 7, 144, 32, 0.02),
    (24, 4, 96, 64, 0.05),
    (44, 8, 176, 16, 0.05),
]


def requests(specs, steps=20):
    out = []
    for i, (norb, nocc, nchol, W, t) in enumerate(specs):
        hd, wd, ne = problems.synthetic_rhf(norb, nocc, nchol, seed=100 + i)
        out.append(serving.Request(arrival=t, rid=i, ham_data=hd, wave_data=wd, nelec=ne,
                                   n_walkers=W, n_steps=steps, seed=7 + i))
    return out


@pytest.mark.parametrize("precision,policy,max_active",
                         [("bf16x3", "fifo", 64), ("fp64", "fifo", 64), ("bf16x3", "sjf", 2)],
                         ids=["bf16x3-fifo", "fp64-fifo", "bf16x3-sjf2"])
def test_each_request_same_alone_and_in_mix(gpu, precision, policy, max_active):
    reqs = requests(SPECS)
    srv = serving.FragmentServer(precision=precision, n_prop_steps=10, policy=policy,
                                 max_active=max_active)
    mixed = {r.rid: r for r in srv.serve(reqs)}
    assert sorted(mixed) == list(range(len(SPECS)))
    stats = serving.latency_stats(list(mixed.values()))
    print(precision, policy, stats, flush=True)
    assert stats["p50_s"] > 0
    assert srv.peak_active <= max_active
    for r in reqs:
        alone = serving.FragmentServer(precision=precision, n_prop_steps=10).serve(
            [serving.Request(arrival=0.0, rid=r.rid, ham_data=r.ham_data,
                             wave_data=r.wave_data, nelec=r.nelec, n_walkers=r.n_walkers,
                             n_steps=r.n_steps, seed=r.seed)])[0]
        m = mixed[r.rid]
        assert np.array_equal(alone.block_energies, m.block_energies)
        assert np.array_equal(alone.block_weights, m.block_weights)
        assert alone.walker_steps == m.walker_steps == r.n_walkers * r.n_steps


def test_hbm_budget_limits_admission(gpu):
    """A budget that fits one fragment at a time serialises the requests; results are
    unchanged."""
    reqs = requests(SPECS[:3])
    one = min(serving.estimate_bytes(n, o, g, w, 10, "bf16x3") for n, o, g, w, _ in SPECS[:3])
    tight = serving.FragmentServer(precision="bf16x3", n_prop_steps=10, hbm_budget=one)
    res_t = {r.rid: r for r in tight.serve(reqs)}
    free = serving.FragmentServer(precision="bf16x3", n_prop_steps=10)
    res_f = {r.rid: r for r in free.serve(requests(SPECS[:3]))}
    for rid in res_f:
        assert np.array_equal(res_t[rid].block_energies, res_f[rid].block_energies)
    assert tight.peak_reserved == 1
    assert free.peak_reserved > 1
