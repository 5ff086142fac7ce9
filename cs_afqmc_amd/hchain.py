"""Linear hydrogen chains in STO-6G: integrals, RHF trial and Cholesky vectors.

The reference builds its molecular inputs with pyscf (``pyscf_interface.prep_afqmc``,
pyscf_interface.py:52-294, called by examples/hchain.py:11-59): AO integrals, an RHF
solution, the modified (pivoted) Cholesky decomposition of the AO two-electron integrals
(``chunked_cholesky`` :465-574 via ``generate_integrals`` :408-445, cutoff ``chol_cut``
= 1e-5 by default), and the rotation of h1 and the Cholesky vectors into the RHF
molecular-orbital basis (``ao2mo_chol`` :448-452), with the trial = the occupied MOs in
that basis.  pyscf is not available here, so this module computes the same quantities
for chains of hydrogen atoms, whose STO-6G basis is one contracted s function per atom:

* integrals over normalised s-type Gaussians in closed form (Szabo & Ostlund, Modern
  Quantum Chemistry, appendix A): overlap, kinetic energy, nuclear attraction and
  electron repulsion, the last two through the Boys function F0(t) = sqrt(pi/t)
  erf(sqrt t) / 2;
* the STO-6G contraction for hydrogen (Hehre, Stewart and Pople, J. Chem. Phys. 51, 2657
  (1969); exponents for zeta = 1.24), renormalised so that each contracted function has
  unit norm, as pyscf does;
* restricted Hartree-Fock by Roothaan iterations in the symmetric (Loewdin) orthogonal
  basis with DIIS;
* the Cholesky vectors by the pivoted decomposition of the AO matrix M[(ij),(kl)] =
  (ij|kl): pivot on the largest remaining diagonal, stop once it is at most chol_cut.

``prep`` returns the reference's ham_data / wave_data dicts (h0, h1 (2,N,N), chol
(G, N*N); mo_coeff, rdm1), the same content that mpi_jax.py:107-140 reads from
FCIDUMP_chol / mo_coeff.npz.
"""

import math

import numpy as np

# STO-6G hydrogen 1s (zeta = 1.24): exponents (bohr^-2) and contraction coefficients of
# normalised primitives
STO6G_H_EXP = np.array([35.52322122, 6.513143725, 1.822142904, 0.625955266, 0.243076747,
                        0.100112428])
STO6G_H_COEF = np.array([0.00916359628, 0.04936149294, 0.16853830490, 0.37056279970,
                         0.41649152980, 0.13033408410])


def boys0(t):
    """F0(t) = int_0^1 exp(-t u^2) du, elementwise."""
    t = np.asarray(t, dtype=np.float64)
    out = np.ones_like(t)
    big = t > 1e-12
    st = np.sqrt(t[big])
    out[big] = 0.5 * math.sqrt(math.pi) * np.vectorize(math.erf)(st) / st
    small = ~big
    out[small] = 1.0 - t[small] / 3.0
    return out


def chain_coords(n_atoms, spacing):
    """Atoms on the z axis at 0, spacing, 2 spacing, ... (bohr)."""
    return np.array([[0.0, 0.0, i * spacing] for i in range(n_atoms)])


def _primitives(coords):
    """Flattened primitives: centre, exponent, coefficient (including the primitive's
    normalisation (2a/pi)^(3/4)) and the index of the contracted function."""
    cen, alp, cof, fun = [], [], [], []
    for i, c in enumerate(coords):
        for a, d in zip(STO6G_H_EXP, STO6G_H_COEF):
            cen.append(c)
            alp.append(a)
            cof.append(d * (2.0 * a / math.pi) ** 0.75)
            fun.append(i)
    return np.array(cen), np.array(alp), np.array(cof), np.array(fun)


def _contract2(prim, nao, fun):
    """Sum primitive pair integrals (P, P) into contracted (nao, nao)."""
    out = np.zeros((nao, nao))
    np.add.at(out, (fun[:, None], fun[None, :]), prim)
    return out


def ao_integrals(coords, charges=None):
    """AO overlap S, core Hamiltonian T + V, ERI (nao, nao, nao, nao) in chemists'
    notation (ij|kl), and the nuclear repulsion energy, for one contracted STO-6G s
    function per hydrogen atom at `coords` (bohr)."""
    coords = np.asarray(coords, dtype=np.float64)
    nao = coords.shape[0]
    charges = np.ones(nao) if charges is None else np.asarray(charges, dtype=np.float64)
    cen, alp, cof, fun = _primitives(coords)
    # pair quantities
    a, b = alp[:, None], alp[None, :]
    p = a + b
    mu = a * b / p
    rab2 = np.sum((cen[:, None, :] - cen[None, :, :]) ** 2, axis=-1)
    kab = np.exp(-mu * rab2)
    cc = cof[:, None] * cof[None, :]
    s_prim = (math.pi / p) ** 1.5 * kab
    t_prim = mu * (3.0 - 2.0 * mu * rab2) * s_prim
    pc = (a[..., None] * cen[:, None, :] + b[..., None] * cen[None, :, :]) / p[..., None]
    v_prim = np.zeros_like(s_prim)
    for c, z in zip(coords, charges):
        rpc2 = np.sum((pc - c) ** 2, axis=-1)
        v_prim -= z * 2.0 * math.pi / p * kab * boys0(p * rpc2)
    S = _contract2(cc * s_prim, nao, fun)
    T = _contract2(cc * t_prim, nao, fun)
    V = _contract2(cc * v_prim, nao, fun)
    # renormalise the contracted functions (unit diagonal of S)
    nrm = 1.0 / np.sqrt(np.diag(S))
    cscale = nrm[fun]
    S = S * np.outer(nrm, nrm)
    T = T * np.outer(nrm, nrm)
    V = V * np.outer(nrm, nrm)
    # ERI over primitive quartets, contracted
    npr = alp.size
    pp = p.reshape(-1)
    kk = (cc * kab * np.outer(cscale, cscale)).reshape(-1)
    pcf = pc.reshape(-1, 3)
    P, Q = pp[:, None], pp[None, :]
    rpq2 = np.sum((pcf[:, None, :] - pcf[None, :, :]) ** 2, axis=-1)
    eri_prim = (2.0 * math.pi ** 2.5 / (P * Q * np.sqrt(P + Q)) * kk[:, None] * kk[None, :]
                * boys0(P * Q / (P + Q) * rpq2))
    fi = np.repeat(fun, npr)
    fj = np.tile(fun, npr)
    pair = fi * nao + fj
    eri2 = np.zeros((nao * nao, nao * nao))
    np.add.at(eri2, (pair[:, None], pair[None, :]), eri_prim)
    eri = eri2.reshape(nao, nao, nao, nao)
    # exact permutational symmetry (ij|kl) = (ji|kl) = (ij|lk) = (kl|ij)
    eri = (eri + eri.transpose(1, 0, 2, 3)) / 2.0
    eri = (eri + eri.transpose(0, 1, 3, 2)) / 2.0
    eri = (eri + eri.transpose(2, 3, 0, 1)) / 2.0
    enuc = 0.0
    for i in range(nao):
        for j in range(i):
            enuc += charges[i] * charges[j] / np.linalg.norm(coords[i] - coords[j])
    return S, T + V, eri, enuc


def rhf(S, hcore, eri, nocc, max_iter=200, tol=1e-11):
    """Closed-shell Roothaan SCF with DIIS.  Returns (energy without nuclear repulsion,
    MO coefficients (nao, nao) ordered by orbital energy, orbital energies)."""
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s ** -0.5) @ U.T
    f = X.T @ hcore @ X
    e, c = np.linalg.eigh(f)
    C = X @ c
    focks, errs = [], []
    e_old = 0.0
    for _ in range(max_iter):
        D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
        J = np.einsum("ijkl,kl->ij", eri, D)
        K = np.einsum("ikjl,kl->ij", eri, D)
        F = hcore + J - 0.5 * K
        energy = 0.5 * np.sum(D * (hcore + F))
        err = X.T @ (F @ D @ S - S @ D @ F) @ X
        focks.append(F)
        errs.append(err)
        if len(focks) > 8:
            focks.pop(0)
            errs.pop(0)
        nd = len(focks)
        B = -np.ones((nd + 1, nd + 1))
        B[nd, nd] = 0.0
        for i in range(nd):
            for j in range(nd):
                B[i, j] = np.sum(errs[i] * errs[j])
        rhs = np.zeros(nd + 1)
        rhs[nd] = -1.0
        try:
            w = np.linalg.solve(B, rhs)[:nd]
            F = sum(wi * fi for wi, fi in zip(w, focks))
        except np.linalg.LinAlgError:
            pass
        e, c = np.linalg.eigh(X.T @ F @ X)
        C = X @ c
        if abs(energy - e_old) < tol and np.max(np.abs(err)) < 1e-8:
            break
        e_old = energy
    D = 2.0 * C[:, :nocc] @ C[:, :nocc].T
    J = np.einsum("ijkl,kl->ij", eri, D)
    K = np.einsum("ikjl,kl->ij", eri, D)
    energy = 0.5 * np.sum(D * (2.0 * hcore + J - 0.5 * K))
    return energy, C, e


def modified_cholesky(mat, max_error):
    """Pivoted ("modified") Cholesky of a symmetric positive semidefinite matrix, in the
    order of pyscf_interface.chunked_cholesky (:465-574): pivot on the largest remaining
    diagonal |D_nu|, L_new = (M[:, nu] - sum_x L_x[nu] L_x) / sqrt(D_nu), until the
    largest remaining diagonal is at most max_error.  Returns (nchol, n)."""
    mat = np.asarray(mat, dtype=np.float64)
    n = mat.shape[0]
    diag = np.diagonal(mat).copy()
    vecs = np.zeros((n, n))
    nu = int(np.argmax(diag))
    delta_max = diag[nu]
    vecs[0] = mat[nu] / math.sqrt(delta_max)
    approx = np.zeros(n)
    nchol = 0
    while abs(delta_max) > max_error and nchol + 1 < n:
        approx += vecs[nchol] * vecs[nchol]
        delta = diag - approx
        nu = int(np.argmax(np.abs(delta)))
        delta_max = abs(delta[nu])
        r = vecs[: nchol + 1, nu] @ vecs[: nchol + 1, :]
        vecs[nchol + 1] = (mat[nu] - r) / math.sqrt(delta_max)
        nchol += 1
    return vecs[:nchol]


def prep(n_atoms=10, spacing=1.6, chol_cut=1e-5):
    """The reference's prep_afqmc(mf) for an RHF H chain (pyscf_interface.py:52-294 with
    basis_coeff = mf.mo_coeff): h1 and the Cholesky vectors in the RHF MO basis, the
    trial = the occupied MOs in that basis (QR of C^T S C, :269-276), and the energies.
    Returns (ham_data, wave_data, nelec, info)."""
    coords = chain_coords(n_atoms, spacing)
    S, hcore, eri, enuc = ao_integrals(coords)
    nao = S.shape[0]
    nocc = n_atoms // 2
    e_elec, C, eps = rhf(S, hcore, eri, nocc)
    chol_ao = modified_cholesky(eri.reshape(nao * nao, nao * nao), chol_cut)
    nchol = chol_ao.shape[0]
    chol = np.empty((nchol, nao * nao))
    for g in range(nchol):  # ao2mo_chol (:448-452)
        chol[g] = (C.T @ chol_ao[g].reshape(nao, nao) @ C).ravel()
    h1 = C.T @ hcore @ C
    q, _ = np.linalg.qr(C.T @ S @ C)
    mo = q[:, :nocc]
    ham_data = {"h0": enuc, "h1": np.array([h1, h1]), "chol": chol, "ene0": 0.0}
    wave_data = {"mo_coeff": mo, "rdm1": np.array([mo @ mo.T, mo @ mo.T])}
    info = {"e_rhf": e_elec + enuc, "enuc": enuc, "nchol": nchol, "orbital_energies": eps,
            "coords": coords}
    return ham_data, wave_data, (nocc, nocc), info


def hf_energy(h1, chol, nocc, enuc):
    """Closed-shell determinant energy of the lowest nocc orbitals of an orthonormal
    basis from h1 and (pq|rs) = sum_g L_g[pq] L_g[rs]."""
    n = h1.shape[-1]
    L = np.asarray(chol).reshape(-1, n, n)[:, :nocc, :nocc]
    j = np.einsum("gii->g", L)
    e2 = 2.0 * np.sum(j * j) - np.einsum("gij,gji->", L, L)
    return enuc + 2.0 * np.trace(h1[:nocc, :nocc]) + e2


def h10_config1(precision="fp64", n_walkers=200, dt=0.005):
    """The plain small configuration: linear H10, 1.6 bohr, STO-6G, RHF trial, modified
    Cholesky 1e-5, set up on the device with the reference's propagator / trial API."""
    from . import hamiltonian, propagation, wavefunctions

    hd, wd, nelec, info = prep(10, 1.6, 1e-5)
    norb = hd["h1"].shape[-1]
    trial = wavefunctions.rhf(norb, nelec)
    prop = propagation.propagator_restricted(dt=dt, n_walkers=n_walkers, precision=precision)
    ham = hamiltonian.hamiltonian(norb)
    wave = dict(wd)
    ham_data = ham.build_measurement_intermediates(dict(hd), trial, wave)
    ham_data = ham.build_propagation_intermediates(ham_data, prop, trial, wave)
    prop_data = prop.init_prop_data(trial, wave, ham_data)
    return {"ham": ham, "ham_data": ham_data, "prop": prop, "trial": trial,
            "wave_data": wave, "prop_data": prop_data, "nchol": info["nchol"],
            "e_hf": info["e_rhf"]}
