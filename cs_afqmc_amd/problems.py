"""Seeded synthetic AFQMC problems (host float64 arrays in the reference's layout).

``synthetic_rhf`` / ``synthetic_uhf`` build the workloads used for measurement:
  L_g = sym(A_g) exp(-g / 80), A_g entries N(0, sigma = 1/norb), sym(A) = (A + A^T) / sqrt(2)
  h1  = sym(B), B entries N(0, sigma = 0.1)
N(0, s) is read as standard deviation s.  Read as a variance instead (sigma =
1/sqrt(norb)), the VHS of one dt = 0.005 step has spectral norm ~0.9 at norb = 100 and
the phaseless weights of the whole population are clipped to zero within ~10 steps
(the CPU oracle shows the same), so no energy can be measured; ``chol_std`` /
``h1_std`` select other scales.
  trial = first nocc columns of a Haar-random orthogonal matrix (QR of a Gaussian
          matrix with the signs of diag(R) folded into Q); separate alpha / beta
          matrices for the unrestricted trial
  h0 = 0.
Returned as (ham_data, wave_data, nelec) with ham_data = {h0, h1 (2,N,N), chol (G,N*N),
ene0} and wave_data = {mo_coeff, rdm1}, the dicts mpi_jax.py:107-140 builds from
FCIDUMP_chol / mo_coeff.npz.
"""

import numpy as np

WORKLOADS = {
    # name: (walker_type, norb, nocc, nchol, n_walkers)
    "rhf_n100_o20_g400_w1024": ("rhf", 100, 20, 400, 1024),
    "uhf_n200_o40_g1000_w2048": ("uhf", 200, 40, 1000, 2048),
}


def _sym(a):
    return (a + a.T) / np.sqrt(2.0)


def haar_orthogonal(n, rng):
    z = rng.standard_normal((n, n))
    q, r = np.linalg.qr(z)
    return q * np.sign(np.diag(r))[None, :]


def _chol_h1(norb, nchol, rng, decay=80.0, chol_std=None, h1_std=0.1):
    chol_std = 1.0 / norb if chol_std is None else chol_std
    chol = np.empty((nchol, norb, norb))
    for g in range(nchol):
        a = rng.normal(0.0, chol_std, (norb, norb))
        chol[g] = _sym(a) * np.exp(-g / decay)
    h1 = _sym(rng.normal(0.0, h1_std, (norb, norb)))
    return chol.reshape(nchol, norb * norb), h1


def synthetic_rhf(norb, nocc, nchol, seed=0, decay=80.0, chol_std=None, h1_std=0.1):
    rng = np.random.default_rng(seed)
    chol, h1 = _chol_h1(norb, nchol, rng, decay, chol_std, h1_std)
    mo = haar_orthogonal(norb, rng)[:, :nocc]
    ham_data = {"h0": 0.0, "h1": np.array([h1, h1]), "chol": chol, "ene0": 0.0}
    wave_data = {"mo_coeff": mo, "rdm1": np.array([mo @ mo.T] * 2)}
    return ham_data, wave_data, (nocc, nocc)


def synthetic_uhf(norb, nocc_up, nocc_dn, nchol, seed=0, decay=80.0, chol_std=None,
                  h1_std=0.1):
    rng = np.random.default_rng(seed)
    chol, h1 = _chol_h1(norb, nchol, rng, decay, chol_std, h1_std)
    mo_a = haar_orthogonal(norb, rng)[:, :nocc_up]
    mo_b = haar_orthogonal(norb, rng)[:, :nocc_dn]
    ham_data = {"h0": 0.0, "h1": np.array([h1, h1]), "chol": chol, "ene0": 0.0}
    wave_data = {"mo_coeff": [mo_a, mo_b],
                 "rdm1": np.array([mo_a @ mo_a.T, mo_b @ mo_b.T])}
    return ham_data, wave_data, (nocc_up, nocc_dn)


def synthetic_uhf_spin(norb, nocc_up, nocc_dn, nchol, seed=0, decay=80.0, chol_std=None,
                       h1_std=0.1, mix=0.6):
    """Unrestricted problem with spin-resolved Cholesky vectors, in the layout of the
    reference's prop_unrestricted package: chol (2, G, N*N), h1 (2, N, N).
    L^a_g as in synthetic_uhf; L^b_g = sqrt(1 - mix^2) L^a_g + mix B_g with B_g an
    independent draw of the same scale and decay; h1[1] likewise from h1[0]."""
    rng = np.random.default_rng(seed)
    chol_a, h1_a = _chol_h1(norb, nchol, rng, decay, chol_std, h1_std)
    chol_n, h1_n = _chol_h1(norb, nchol, rng, decay, chol_std, h1_std)
    c = np.sqrt(1.0 - mix * mix)
    chol = np.array([chol_a, c * chol_a + mix * chol_n])
    h1 = np.array([h1_a, c * h1_a + mix * h1_n])
    mo_a = haar_orthogonal(norb, rng)[:, :nocc_up]
    mo_b = haar_orthogonal(norb, rng)[:, :nocc_dn]
    ham_data = {"h0": 0.0, "h1": h1, "chol": chol, "ene0": 0.0}
    wave_data = {"mo_coeff": [mo_a, mo_b],
                 "rdm1": np.array([mo_a @ mo_a.T, mo_b @ mo_b.T])}
    return ham_data, wave_data, (nocc_up, nocc_dn)


def random_walkers(walker_type, norb, nelec, n_walkers, rng, near=None, noise=0.1):
    """Complex walkers: ``near`` (trial coefficients) plus complex Gaussian noise, or
    purely random ones.  Used for parity tests away from the trial state."""
    def one(n, base):
        w = noise * (rng.standard_normal((n_walkers, norb, n))
                     + 1j * rng.standard_normal((n_walkers, norb, n)))
        if base is not None:
            w = w + base[None, :, :]
        return np.ascontiguousarray(w)
    if walker_type == "rhf":
        return one(nelec[0], near)
    base = near if near is not None else [None, None]
    return [one(nelec[0], base[0]), one(nelec[1], base[1])]


def fragment_specs(count, seed=0, norb_range=(20, 120), walkers=(256,)):
    """Synthetic LNO-style fragment set: norb uniform in norb_range, nocc = norb // 5,
    nchol = 4 norb, closed shell, walkers drawn from `walkers`.  Returns a list of
    (norb, nocc, nchol, n_walkers, seed) tuples."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        norb = int(rng.integers(norb_range[0], norb_range[1] + 1))
        nocc = max(1, norb // 5)
        out.append((norb, nocc, 4 * norb, int(rng.choice(walkers)), int(seed * 100003 + i)))
    return out
