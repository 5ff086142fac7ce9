"""AFQMC driver: equilibration and sampling blocks around the device sampler.

Mirrors ad_afqmc/driver.py ``afqmc`` (:18-133 setup and equilibration, :207-355 the
sampling loop with ad_mode None, :368-421 outlier rejection and blocking analysis) for
one process that holds the whole walker population on one GPU: the MPI reductions
and gathers of the reference reduce to the single population here, and the global
stochastic reconfiguration (propagation.py:243-249 -> sr.py:66-106) runs on the
device over all walkers.

Per block the host reads back two scalars (the block energy and the total weight);
walkers, weights and overlaps stay on the GPU.
"""

import time

import numpy as np
import torch

from . import rng, sampling, stat_utils


def _log(verbose, *a):
    if verbose:
        print(*a, flush=True)


def afqmc(ham_data, ham, propagator, trial, wave_data, sampler, observable=None,
          options=None, MPI=None, init_walkers=None, verbose=True, samples_file=None,
          return_state=False):
    """driver.py:18-421 with ad_mode None.  Returns (energy, stochastic error or None).

    options: "seed" (int), "n_eql" (equilibration blocks, default 1), "ad_mode" (only
    None is supported).  `observable` (op, constant) only enters the observable column
    of the samples, which is 0 with ad_mode None, as in the reference.  `MPI` is
    accepted for signature compatibility and ignored (one process, one GPU).
    `samples_file`, if given, receives the (weight, energy, observable) rows as
    driver.py:302-310 writes samples_raw.dat.  With return_state the call returns
    (energy, error, samples (n_blocks, 3), prop_data).
    """
    options = dict(options or {})
    if options.get("ad_mode") is not None:
        raise NotImplementedError("only ad_mode=None (energy only) is supported")
    t_init = time.time()
    seed = int(options.get("seed", 0))
    neql = int(options.get("n_eql", 1))
    observable_constant = 0.0 if observable is None else float(observable[1])

    trial_rdm1 = trial.get_rdm1(wave_data)
    if "rdm1" not in wave_data:
        wave_data["rdm1"] = trial_rdm1
    ham_data = ham.build_measurement_intermediates(ham_data, trial, wave_data)
    ham_data = ham.build_propagation_intermediates(ham_data, propagator, trial, wave_data)
    prop_data = propagator.init_prop_data(trial, wave_data, ham_data, init_walkers)
    if abs(complex(torch.sum(prop_data["overlaps"]).item())) < 1.0e-6:
        raise ValueError("Initial overlaps are zero. Pass walkers with non-zero overlap.")
    prop_data["key"] = rng.PRNGKey(seed)

    _log(verbose, "# Equilibration sweeps:")
    _log(verbose, "#   Iter        Block energy      Walltime")
    _log(verbose, f"# {0:5d}      {float(prop_data['e_estimate']):.9e}     "
                  f"{time.time() - t_init:.2e} ")
    sampler_eq = sampling.sampler(n_prop_steps=50, n_ene_blocks=5, n_sr_blocks=10)
    for n in range(1, neql + 1):
        block_energy, prop_data = sampler_eq.propagate_phaseless(
            ham, ham_data, propagator, prop_data, trial, wave_data)
        prop_data = propagator.orthonormalize_walkers(prop_data)
        prop_data = propagator.stochastic_reconfiguration_global(prop_data, None)
        prop_data["e_estimate"] = 0.9 * prop_data["e_estimate"] + 0.1 * block_energy
        _log(verbose, f"# {n:5d}      {block_energy:.9e}     {time.time() - t_init:.2e} ")

    _log(verbose, "#\n# Sampling sweeps:")
    _log(verbose, "#  Iter        Mean energy          Stochastic error       Walltime")
    nb = sampler.n_blocks
    weights = np.zeros(nb)
    energies = np.zeros(nb)
    observables = np.zeros(nb)
    for n in range(nb):
        block_energy, prop_data = sampler.propagate_phaseless(
            ham, ham_data, propagator, prop_data, trial, wave_data)
        weights[n] = float(torch.sum(prop_data["weights"]).item())
        energies[n] = block_energy
        observables[n] = 0.0 + observable_constant
        prop_data = propagator.orthonormalize_walkers(prop_data)
        prop_data = propagator.stochastic_reconfiguration_global(prop_data, None)
        prop_data["e_estimate"] = 0.9 * prop_data["e_estimate"] + 0.1 * block_energy
        if n % max(nb // 10, 1) == 0:
            e_n, err_n = stat_utils.blocking_analysis(weights[:n + 1], energies[:n + 1])
            err_s = f"{err_n:.9e}" if err_n is not None else "        -       "
            _log(verbose, f" {n:5d}      {e_n:.9e}        {err_s}        "
                          f"{time.time() - t_init:.2e} ")
            if samples_file:
                np.savetxt(samples_file, np.stack((weights[:n + 1], energies[:n + 1],
                                                   observables[:n + 1])).T)
    samples = np.stack((weights, energies, observables)).T
    if samples_file:
        np.savetxt(samples_file, samples)
    clean, _ = stat_utils.reject_outliers(samples, 1)
    _log(verbose, f"# Number of outliers in post: {nb - clean.shape[0]} ")
    e_afqmc, e_err = stat_utils.blocking_analysis(clean[:, 0], clean[:, 1], neql=0,
                                                  printQ=verbose)
    if e_err is not None:
        _log(verbose, f"AFQMC energy: {e_afqmc:.6f} +/- {e_err:.6f}\n")
    else:
        _log(verbose, f"AFQMC energy: {e_afqmc:.6f}\n")
        e_err = 0.0
    if return_state:
        return e_afqmc, e_err, samples, prop_data
    return e_afqmc, e_err
