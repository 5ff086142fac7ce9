"""The H-chain integrals, RHF and modified Cholesky (cs_afqmc_amd.hchain, standing in for
pyscf in pyscf_interface.prep_afqmc) against the reference's own H10 @ 2.0 bohr files
(examples/FCIDUMP_chol written by pyscf; fixture tests/golden/h10_chain_r2.0_sto6g.npz).

Quantities that do not depend on the orbital basis pin the integrals: the nuclear
repulsion, the spectrum of h1 in an orthonormal basis, the spectrum of the two-electron
supermatrix V = L^T L, the number of Cholesky vectors at the reference's 1e-5 cutoff,
and the closed-shell energy of the trial determinant.  The decomposition itself is
checked against what the mathematics fixes: it reconstructs the matrix to the cutoff,
its vectors are exact on the pivots, and it stops at the first residual diagonal at or
below the cutoff.
"""

import os

import numpy as np

from cs_afqmc_amd import hchain

FIX = os.path.join(os.path.dirname(__file__), "golden", "h10_chain_r2.0_sto6g.npz")


def _fixture():
    with np.load(FIX) as z:
        return {k: np.array(z[k]) for k in z.files}


def test_h10_r2_integrals_match_reference_fixture():
    fx = _fixture()
    hd, wd, nelec, info = hchain.prep(10, 2.0, 1e-5)
    assert nelec == (5, 5)
    assert abs(info["enuc"] - float(fx["energy_core"])) < 1e-12
    assert info["nchol"] == fx["chol"].shape[0] == 27
    e_fix = np.linalg.eigvalsh(fx["hcore"])
    e_own = np.linalg.eigvalsh(hd["h1"][0])
    assert np.max(np.abs(e_fix - e_own)) < 1e-10

    def vspec(chol):
        L = chol.reshape(chol.shape[0], -1)
        return np.sort(np.linalg.eigvalsh(L.T @ L))

    assert np.max(np.abs(vspec(fx["chol"]) - vspec(hd["chol"]))) < 1e-10
    # the trial of both is the first five orbitals of the RHF MO basis
    assert np.allclose(np.abs(fx["mo_coeff"][0][:, :5]), np.eye(10)[:, :5])
    assert np.allclose(np.abs(wd["mo_coeff"]), np.eye(10)[:, :5])
    e_fix_hf = hchain.hf_energy(fx["hcore"], fx["chol"], 5, float(fx["energy_core"]))
    e_own_hf = hchain.hf_energy(hd["h1"][0], hd["chol"], 5, info["enuc"])
    assert abs(e_fix_hf - e_own_hf) < 1e-9
    # the SCF energy from the exact integrals is the Cholesky one up to the cutoff
    assert abs(info["e_rhf"] - e_own_hf) < 1e-5


def test_rhf_is_stationary_and_aufbau():
    hd, wd, nelec, info = hchain.prep(6, 1.6, 1e-8)
    S, hcore, eri, enuc = hchain.ao_integrals(hchain.chain_coords(6, 1.6))
    # h1 in the MO basis: occupied-virtual Fock block vanishes at convergence
    n = hd["h1"].shape[-1]
    L = hd["chol"].reshape(-1, n, n)
    occ = 3
    J = 2.0 * np.einsum("gkk,gpq->pq", L[:, :occ, :occ], L)
    K = np.einsum("gpk,gkq->pq", L[:, :, :occ], L[:, :occ, :])
    F = hd["h1"][0] + J - K
    assert np.max(np.abs(F[:occ, occ:])) < 1e-6
    assert np.all(np.diff(np.diag(F)) > -1e-8)
    # overlap of the renormalised contracted functions has a unit diagonal
    assert np.allclose(np.diag(S), 1.0, atol=1e-14)
    # two hydrogens far apart: the core Hamiltonian element of one atom is the STO-6G 1s
    # energy (-0.47103 Eh, Hehre, Stewart and Pople 1969) plus the far nucleus' -1/R
    R = 80.0
    S2, h2, _, _ = hchain.ao_integrals(np.array([[0.0, 0.0, 0.0], [0.0, 0.0, R]]))
    assert abs(h2[0, 0] + 1.0 / R - (-0.47103)) < 2e-5
    assert abs(S2[0, 1]) < 1e-12


def test_modified_cholesky_reconstructs_and_pivots():
    rs = np.random.default_rng(5)
    A = rs.standard_normal((25, 12))
    M = A @ A.T + 1e-9 * np.eye(25)
    # a full-rank matrix stops at the vector cap (size - 1 vectors, as the reference's
    # nchol_max bound does) when the cutoff is below its smallest pivots
    assert hchain.modified_cholesky(np.eye(6) + 0.1, 0.0).shape[0] == 5
    for cut in (1e-2, 1e-6):
        L = hchain.modified_cholesky(M, cut)
        R = M - L.T @ L
        assert np.max(np.abs(np.diag(R))) <= cut + 1e-12
        # the residual diagonal just before the last vector was above the cutoff
        Lm = L[:-1]
        assert np.max(np.abs(np.diag(M - Lm.T @ Lm))) > cut
    # exact rank-deficient case: the decomposition is exact after `rank` vectors
    B = rs.standard_normal((20, 4))
    L = hchain.modified_cholesky(B @ B.T, 1e-10)
    assert L.shape[0] == 4
    assert np.max(np.abs(B @ B.T - L.T @ L)) < 1e-10
    # the first vector is the largest-diagonal column scaled by 1 / sqrt(diag)
    nu = int(np.argmax(np.diag(M)))
    L = hchain.modified_cholesky(M, 1e-6)
    assert np.allclose(L[0], M[nu] / np.sqrt(M[nu, nu]), rtol=0, atol=1e-14)
