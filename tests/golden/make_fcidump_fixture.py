"""Expected datasets of tests/golden/FCIDUMP_chol_6orb (a copy of the FCIDUMP_chol file the
reference's own tests hold: 6 orbitals, 15 Cholesky vectors), read with h5py.

Run once with an interpreter that has h5py (this image: /opt/conda/bin/python3.9):
    /opt/conda/bin/python3.9 tests/golden/make_fcidump_fixture.py
It writes tests/golden/FCIDUMP_chol_6orb_expected.npz, against which the library's own
HDF5 reader is checked bit for bit."""

import os

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

with h5py.File(os.path.join(HERE, "FCIDUMP_chol_6orb"), "r") as f:
    arrays = {k: np.array(f[k]) for k in f.keys()}
np.savez(os.path.join(HERE, "FCIDUMP_chol_6orb_expected.npz"), **arrays)
print(sorted(arrays))
