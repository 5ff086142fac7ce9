"""On-disk inputs of the reference's run scripts: options.bin (mpi_jax.py:55-60, written
by run_afqmc.py:31-32 with pickle.dump) decoded without the pickle module; FCIDUMP_chol
(HDF5, pyscf_interface.py:578-598) read by the library's own HDF5 reader, bit for bit
against h5py's reading of the reference tests' file; and the file-based set-up
(mpi_jax.py:36-60)."""

import os
import pickle

import numpy as np
import pytest

from cs_afqmc_amd import prep

ROOT = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(ROOT, "golden", "h10_chain_r2.0_sto6g.npz")
H5 = os.path.join(ROOT, "golden", "FCIDUMP_chol_6orb")
H5_EXPECTED = os.path.join(ROOT, "golden", "FCIDUMP_chol_6orb_expected.npz")

OPTIONS = {
    "n_eql": 3, "n_ene_blocks": 1, "n_sr_blocks": 50, "n_blocks": 10, "n_walkers": 50,
    "seed": 98, "trial": "rhf", "walker_type": "rhf", "dt": 0.005, "ad_mode": None,
    "use_gpu": True, "free_projection": False, "max_error": 1.5e-3,
    "big": 2**70 + 3, "neg": -123456, "lst": [1, 2.5, "x", None, (True, False)],
    "unicode": "énergie", "empty": {}, "tup": (),
}


@pytest.mark.parametrize("protocol", [2, 3, 4, 5])
def test_options_round_trip(tmp_path, protocol):
    f = tmp_path / "options.bin"
    f.write_bytes(pickle.dumps(OPTIONS, protocol=protocol))
    assert prep.read_options(str(f)) == OPTIONS


@pytest.mark.parametrize("protocol", [3, 4, 5])
def test_options_bytes(tmp_path, protocol):
    # (protocol 2 writes bytes through a global call, which the decoder refuses)
    opts = {"blob": b"\x00\x01" * 200}
    f = tmp_path / "options.bin"
    f.write_bytes(pickle.dumps(opts, protocol=protocol))
    assert prep.read_options(str(f)) == opts


def test_options_memo_references(tmp_path):
    shared = [1, 2, 3]
    opts = {"a": shared, "b": shared, "c": "s" * 300}
    f = tmp_path / "options.bin"
    f.write_bytes(pickle.dumps(opts, protocol=4))
    got = prep.read_options(str(f))
    assert got == opts


class _Obj:
    pass


@pytest.mark.parametrize("value", [{1, 2}, _Obj(), np.int64(3), frozenset([1])])
def test_options_with_objects_are_refused(tmp_path, value):
    f = tmp_path / "options.bin"
    f.write_bytes(pickle.dumps({"x": value}, protocol=4))
    with pytest.raises(ValueError):
        prep.read_options(str(f))


def test_options_missing_file_gives_defaults(tmp_path):
    assert prep.read_options(str(tmp_path / "nothing.bin")) == {}


def test_options_truncated_file_is_an_error(tmp_path):
    f = tmp_path / "options.bin"
    f.write_bytes(pickle.dumps(OPTIONS, protocol=4)[:-5])
    with pytest.raises(ValueError):
        prep.read_options(str(f))


def test_prep_from_files(tmp_path):
    z = np.load(FIXTURE)
    chol = tmp_path / "FCIDUMP_chol.npz"
    np.savez(chol, **{k: z[k] for k in ("header", "energy_core", "hcore", "hcore_mod", "chol")})
    mo = tmp_path / "mo_coeff.npz"
    np.savez(mo, mo_coeff=z["mo_coeff"])
    opt = tmp_path / "options.bin"
    opt.write_bytes(pickle.dumps({"n_walkers": 24, "dt": 0.005, "walker_type": "rhf",
                                  "n_prop_steps": 5}, protocol=4))
    ham_data, ham, prop, trial, wave_data, smp, obs, opts = prep.prep_afqmc_files(
        option_file=str(opt), mo_file=str(mo), chol_file=str(chol))
    nelec, nmo, ms, nchol = (int(x) for x in z["header"])
    assert prop.n_walkers == 24 and prop.dt == 0.005 and smp.n_prop_steps == 5
    assert trial.norb == nmo and trial.nelec == (nelec // 2, nelec // 2)
    assert ham_data["chol"].shape == (nchol, nmo * nmo)
    assert opts["n_blocks"] == 50  # mpi_jax.py:62-90 defaults fill the rest
    np.testing.assert_allclose(wave_data["mo_coeff"], z["mo_coeff"][0][:, : nelec // 2])


def test_hdf5_fcidump_matches_h5py():
    ours = prep.read_fcidump_chol(H5)
    with np.load(H5_EXPECTED) as z:
        assert sorted(ours) == sorted(z.files)
        for k in z.files:
            assert ours[k].dtype == z[k].dtype and ours[k].shape == z[k].shape, k
            assert np.array_equal(ours[k], z[k]), k
    nelec, nmo, ms, nchol = (int(x) for x in ours["header"])
    assert ours["chol"].size == nchol * nmo * nmo and ours["hcore"].size == nmo * nmo


def test_hdf5_reader_rejects_other_files(tmp_path):
    f = tmp_path / "not_hdf5"
    f.write_bytes(b"\x00" * 128)
    with pytest.raises(ValueError):
        prep.hdf5_read.read_datasets(str(f))


def test_prep_from_hdf5_file(tmp_path):
    z = np.load(H5_EXPECTED)
    mo = tmp_path / "mo_coeff.npz"
    np.savez(mo, mo_coeff=np.array([z["mo_coeffs_up"], z["mo_coeffs_dn"]]))
    ham_data, ham, prop, trial, wave_data, smp, obs, opts = prep.prep_afqmc_files(
        options={"n_walkers": 8}, mo_file=str(mo), chol_file=H5)
    nelec, nmo, ms, nchol = (int(x) for x in z["header"])
    assert ham_data["chol"].shape == (nchol, nmo * nmo)
    assert ham_data["h0"] == float(z["energy_core"])
