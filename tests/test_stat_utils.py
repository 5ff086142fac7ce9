"""Blocking analysis and outlier rejection against the reference's committed H10 run.

examples/samples_raw.dat holds the per-block (weight, energy, observable) samples of
the reference run whose result is examples/ene_err.txt (energy and error); both are
in tests/golden/h10_chain_r2.0_sto6g.npz (written by make_h10_fixture.py).  driver.py
(:375-421) rejects outliers on column 1 of the samples and runs the blocking analysis
on what is left.
"""

import os

import numpy as np
import pytest

from cs_afqmc_amd import stat_utils

FIX = os.path.join(os.path.dirname(__file__), "golden", "h10_chain_r2.0_sto6g.npz")


@pytest.fixture(scope="module")
def h10():
    with np.load(FIX) as z:
        return {k: z[k] for k in z.files}


def test_reject_outliers_keeps_reference_samples(h10):
    clean, keep = stat_utils.reject_outliers(h10["samples_raw"], 1)
    assert np.array_equal(clean, h10["samples"])
    assert keep.dtype == bool and keep.sum() == clean.shape[0]


def test_blocking_reproduces_reference_energy_and_error(h10):
    clean, _ = stat_utils.reject_outliers(h10["samples_raw"], 1)
    e, err = stat_utils.blocking_analysis(clean[:, 0], clean[:, 1], neql=0,
                                          block_dtype=np.float64)
    assert e == h10["ene_err"][0]
    assert err == h10["ene_err"][1]
    # the float32 blocks of the current stat_utils.py agree to float32 rounding
    e32, err32 = stat_utils.blocking_analysis(clean[:, 0], clean[:, 1], neql=0)
    assert e32 == h10["ene_err"][0]
    assert abs(err32 - h10["ene_err"][1]) < 1e-7 * h10["ene_err"][1]


def test_blocking_closed_form_alternating():
    """e = +1, -1, +1, ... with equal weights: blocks of 2 are all 0, so the search stops
    at b = 2 and the plateau error is the b = 1 error, sqrt(sum (e - m)^2) / (n - 1)
    = sqrt(n) / (n - 1) (v1 - v2/v1 = n - 1 for unit weights)."""
    n = 64
    e = np.where(np.arange(n) % 2 == 0, 1.0, -1.0)
    w = np.ones(n)
    for dt in (np.float64, np.float32):
        mean, err = stat_utils.blocking_analysis(w, e, block_dtype=dt)
        assert mean == 0.0
        assert np.isclose(err, np.sqrt(n) / (n - 1), rtol=1e-6, atol=0)


def test_blocking_too_few_samples_gives_no_error():
    mean, err = stat_utils.blocking_analysis(np.ones(2), np.array([1.0, 2.0]))
    assert mean == 1.5 and err is None
