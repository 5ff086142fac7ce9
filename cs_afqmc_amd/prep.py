"""Problem set-up with the reference's option API.

Mirrors ad_afqmc/mpi_jax.py ``_prep_afqmc`` (:39-140 integrals, trial orbitals and
options, :314-352 propagator / sampler selection) for the phaseless RHF / UHF path:
given the integrals of an FCIDUMP_chol file (pyscf_interface.py:578-598 writes them:
header [nelec, nmo, ms, nchol], energy_core, hcore, chol flat) and the trial orbitals
of mo_coeff.npz, it returns (ham_data, ham, prop, trial, wave_data, sampler,
observable, options) ready for driver.afqmc.

``read_fcidump_chol`` reads the HDF5 file with h5py when that is installed, or an
.npz holding the same datasets.
"""

import numpy as np

from . import hamiltonian, propagation, sampling, wavefunctions

OPTION_DEFAULTS = {
    # mpi_jax.py:62-90
    "dt": 0.01,
    "n_exp_terms": 6,
    "n_walkers": 50,
    "n_prop_steps": 50,
    "n_ene_blocks": 50,
    "n_sr_blocks": 1,
    "n_blocks": 50,
    "n_eql": 1,
    "ad_mode": None,
    "walker_type": "rhf",
    "trial": None,
    "ene0": 0.0,
    "n_batch": 1,
    # not in the reference: arithmetic of the Cholesky contractions ("fp64" or "bf16x3")
    "precision": "fp64",
}


def read_fcidump_chol(path):
    """Datasets of an FCIDUMP_chol file as numpy arrays (mpi_jax.py:39-51)."""
    if str(path).endswith(".npz"):
        with np.load(path) as z:
            return {k: np.array(z[k]) for k in z.files}
    try:
        import h5py
    except ImportError as e:
        raise ImportError("reading HDF5 FCIDUMP_chol needs h5py; convert it to .npz "
                          "with the same dataset names") from e
    with h5py.File(path, "r") as f:
        return {k: np.array(f[k]) for k in f.keys()}


def prep_afqmc(integrals, mo_coeff, options=None):
    """mpi_jax.py:39-140 and :314-352 for trial "rhf" / "uhf".

    integrals: dict with header [nelec, nmo, ms, nchol], energy_core, hcore (nmo*nmo),
    chol (nchol*nmo*nmo); mo_coeff: (2, nmo, nmo) trial orbitals (mo_coeff.npz)."""
    options = dict(OPTION_DEFAULTS, **(options or {}))
    if options["ad_mode"] is not None:
        raise NotImplementedError("only ad_mode=None is supported")
    nelec, nmo, ms, nchol = (int(x) for x in np.asarray(integrals["header"]).ravel())
    nelec_sp = ((nelec + abs(ms)) // 2, (nelec - abs(ms)) // 2)
    h0 = float(np.asarray(integrals["energy_core"]))
    h1 = np.asarray(integrals["hcore"], dtype=np.float64).reshape(nmo, nmo)
    chol = np.asarray(integrals["chol"], dtype=np.float64).reshape(nchol, nmo * nmo)
    ham_data = {"h0": h0, "h1": np.array([h1, h1]), "chol": chol, "ene0": options["ene0"]}
    mo_coeff = np.asarray(mo_coeff, dtype=np.float64)
    wave_data = {"rdm1": np.array([
        mo_coeff[0][:, :nelec_sp[0]] @ mo_coeff[0][:, :nelec_sp[0]].T,
        mo_coeff[1][:, :nelec_sp[1]] @ mo_coeff[1][:, :nelec_sp[1]].T])}
    trial_kind = options["trial"] or options["walker_type"]
    if trial_kind == "rhf":
        trial = wavefunctions.rhf(nmo, nelec_sp, n_batch=options["n_batch"])
        wave_data["mo_coeff"] = mo_coeff[0][:, :nelec_sp[0]]
    elif trial_kind == "uhf":
        trial = wavefunctions.uhf(nmo, nelec_sp, n_batch=options["n_batch"])
        wave_data["mo_coeff"] = [mo_coeff[0][:, :nelec_sp[0]], mo_coeff[1][:, :nelec_sp[1]]]
    else:
        raise NotImplementedError(f"trial {trial_kind!r}: only rhf / uhf are supported")
    ham = hamiltonian.hamiltonian(nmo)
    if options["walker_type"] == "rhf":
        prop = propagation.propagator_restricted(
            options["dt"], options["n_walkers"], options["n_exp_terms"], options["n_batch"],
            precision=options["precision"])
    elif options["walker_type"] == "uhf":
        prop = propagation.propagator_unrestricted(
            options["dt"], options["n_walkers"], options["n_exp_terms"], options["n_batch"],
            precision=options["precision"])
    else:
        raise ValueError(f"walker_type {options['walker_type']!r}")
    smp = sampling.sampler(options["n_prop_steps"], options["n_ene_blocks"],
                           options["n_sr_blocks"], options["n_blocks"])
    observable = None
    return ham_data, ham, prop, trial, wave_data, smp, observable, options
