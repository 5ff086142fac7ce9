"""Admission order of the serving loop (host logic, no GPU)."""

import heapq

import numpy as np
import pytest

from cs_afqmc_amd import serving


def request(rid, arrival, norb, nchol, W, steps=100):
    hd = {"h1": np.zeros((2, norb, norb)), "chol": np.zeros((nchol, norb * norb))}
    return serving.Request(arrival=arrival, rid=rid, ham_data=hd, wave_data={},
                           nelec=(norb // 5, norb // 5), n_walkers=W, n_steps=steps)


def order(policy, reqs):
    srv = serving.FragmentServer(policy=policy)
    q = []
    for r in reqs:
        heapq.heappush(q, (srv._priority(r), r))
    return [heapq.heappop(q)[1].rid for _ in range(len(reqs))]


def test_fifo_is_arrival_order():
    reqs = [request(0, 0.3, 200, 800, 1024), request(1, 0.1, 20, 80, 128),
            request(2, 0.2, 100, 400, 256)]
    assert order("fifo", reqs) == [1, 2, 0]


def test_sjf_is_smallest_work_first():
    reqs = [request(0, 0.1, 200, 800, 1024), request(1, 0.3, 20, 80, 128),
            request(2, 0.2, 100, 400, 256), request(3, 0.4, 100, 400, 128)]
    # work = W * steps * N^2 * (G + 8 n): 1 < 3 < 2 < 0
    assert order("sjf", reqs) == [1, 3, 2, 0]


def test_sjf_ties_keep_arrival_order():
    reqs = [request(5, 0.2, 50, 200, 256), request(4, 0.1, 50, 200, 256)]
    assert order("sjf", reqs) == [4, 5]


def test_unknown_policy_rejected():
    with pytest.raises(ValueError):
        serving.FragmentServer(policy="lifo")
