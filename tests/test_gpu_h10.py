"""Whole-run statistics on a real molecule: the H10 chain (2.0 bohr, STO-6G, RHF trial)
of the reference's examples/ (FCIDUMP_chol, mo_coeff.npz; fixture written by
tests/golden/make_h10_fixture.py).

The reference's committed run of this system (examples/hchain.py options: 4 ranks x 50
walkers, n_eql 3, n_sr_blocks 50, n_ene_blocks 1, n_blocks 10, dt 0.01 by default,
seed 98) gave E = -5.376836518 +/- 0.001002 (examples/ene_err.txt).  The device driver
with the same options and the population on one GPU must agree within three combined
standard errors.  Jax's random streams are not reproduced (they are parity-unpinned),
so the agreement is statistical.

The mixed-precision mode is checked on the same molecule against the float64 mode with
identical device fields: short trajectories agree to 0.1 mEh per block.
"""

import os

import numpy as np
import pytest
import torch

from cs_afqmc_amd import driver, prep, rng, sampling

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "h10_chain_r2.0_sto6g.npz")

REF_OPTIONS = {"n_eql": 3, "n_ene_blocks": 1, "n_sr_blocks": 50, "n_blocks": 10,
               "n_walkers": 200, "seed": 98, "trial": "rhf", "walker_type": "rhf"}


def _fixture():
    with np.load(FIX) as z:
        return {k: np.array(z[k]) for k in z.files}


def test_h10_energy_matches_reference_run(gpu):
    fx = _fixture()
    e_ref, err_ref = fx["ene_err"]
    setup = prep.prep_afqmc(fx, fx["mo_coeff"], REF_OPTIONS)
    ham_data, ham, prop, trial, wave_data, smp, obs, options = setup
    print("h10 fp64 run", options, flush=True)
    e, err, samples, _ = driver.afqmc(ham_data, ham, prop, trial, wave_data, smp, obs,
                                      options, verbose=False, return_state=True)
    print(f"h10: E = {e:.6f} +/- {err:.6f}  (reference {e_ref:.6f} +/- {err_ref:.6f})",
          flush=True)
    assert np.all(np.isfinite(samples))
    err_used = err if err > 0 else err_ref
    assert abs(e - e_ref) < 3.0 * np.hypot(err_used, err_ref)


def test_h10_mixed_precision_blocks_match_fp64(gpu):
    """Same seed, same device fields: per-block energies of the split-bf16 mode and
    the float64 mode differ by less than 1e-4 Eh (0.1 mEh)."""
    fx = _fixture()
    block_e = {}
    for precision in ("fp64", "bf16x3"):
        opts = dict(REF_OPTIONS, precision=precision, n_walkers=512)
        ham_data, ham, prop, trial, wave_data, smp, obs, opts = prep.prep_afqmc(
            fx, fx["mo_coeff"], opts)
        wave_data["rdm1"] = trial.get_rdm1(wave_data)
        ham_data = ham.build_measurement_intermediates(ham_data, trial, wave_data)
        ham_data = ham.build_propagation_intermediates(ham_data, prop, trial, wave_data)
        pd = prop.init_prop_data(trial, wave_data, ham_data)
        pd["key"] = rng.PRNGKey(98)
        s = sampling.sampler(n_prop_steps=50, n_ene_blocks=5, n_sr_blocks=2)
        e, pd = s.propagate_phaseless(ham, ham_data, prop, pd, trial, wave_data)
        block_e[precision] = pd["block_energies"].cpu().numpy()
        print(precision, e, block_e[precision], flush=True)
        del pd, ham_data
        torch.cuda.synchronize()
    d = np.abs(block_e["bf16x3"] - block_e["fp64"])
    assert np.all(np.isfinite(block_e["fp64"]))
    assert np.max(d) < 1e-4
