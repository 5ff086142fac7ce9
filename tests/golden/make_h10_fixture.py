"""Write tests/golden/h10_chain_r2.0_sto6g.npz from the reference's H10 example files.

Run with an interpreter that has h5py (the image's /opt/conda/bin/python3.9):

    /opt/conda/bin/python3.9 tests/golden/make_h10_fixture.py

Inputs (data files of the reference, read as data, nothing executed):
  examples/FCIDUMP_chol   HDF5: header [nelec, nmo, ms, nchol], energy_core, hcore,
                          hcore_mod, chol (flat), mo_coeffs_up / mo_coeffs_dn
  examples/mo_coeff.npz   mo_coeff (2, nmo, nmo), the trial orbitals run_afqmc uses
  examples/samples_raw.dat, examples/samples.dat   per-block (weight, energy,
                          observable) of the reference run
  examples/ene_err.txt    that run's energy and blocking-analysis error
The layout of FCIDUMP_chol follows pyscf_interface.py:578-598 and its reader
mpi_jax.py:39-51.
"""

import os

import h5py
import numpy as np

REF = "/root/reference/examples"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "h10_chain_r2.0_sto6g.npz")


def main():
    with h5py.File(os.path.join(REF, "FCIDUMP_chol"), "r") as f:
        header = np.array(f["header"], dtype=np.int64)
        nelec, nmo, ms, nchol = (int(x) for x in header)
        data = {
            "header": header,
            "energy_core": np.array(f["energy_core"], dtype=np.float64),
            "hcore": np.array(f["hcore"], dtype=np.float64).reshape(nmo, nmo),
            "hcore_mod": np.array(f["hcore_mod"], dtype=np.float64).reshape(nmo, nmo),
            "chol": np.array(f["chol"], dtype=np.float64).reshape(nchol, nmo * nmo),
            "mo_coeffs_up": np.array(f["mo_coeffs_up"], dtype=np.float64),
            "mo_coeffs_dn": np.array(f["mo_coeffs_dn"], dtype=np.float64),
        }
    with np.load(os.path.join(REF, "mo_coeff.npz")) as z:
        data["mo_coeff"] = np.array(z["mo_coeff"], dtype=np.float64)
    data["samples_raw"] = np.loadtxt(os.path.join(REF, "samples_raw.dat"))
    data["samples"] = np.loadtxt(os.path.join(REF, "samples.dat"))
    data["ene_err"] = np.loadtxt(os.path.join(REF, "ene_err.txt"))
    np.savez(OUT, **data)
    print("wrote", OUT, {k: v.shape for k, v in data.items()})


if __name__ == "__main__":
    main()
