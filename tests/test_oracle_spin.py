"""Pins the oracle's spin-resolved Cholesky functions (prop_unrestricted package of the
reference: chol (2, G, N*N), one set of vectors per spin).

The reference holds no golden values for this package, so the pins are what the
mathematics fixes: the reduction to the spin-summed unrestricted step when both spins
share their vectors; invariance under a rotation of the beta orbital basis alone and
under the exchange of the two spins (a beta determinant propagated or measured with
the alpha vectors breaks both); the Hartree-Fock energy from explicit spin-resolved
two-electron integrals; the force bias at the trial; and the Taylor propagator against
scipy's matrix exponential with the VHS summed by loops.
"""

import numpy as np
import pytest
import scipy.linalg

from cs_afqmc_amd import problems
from oracle import afqmc_oracle as o

NORB, NELEC, NCHOL, DT, W = 7, (3, 2), 9, 0.01, 4


def _problem(seed=5, mix=0.6, same=False):
    hd, wd, nelec = problems.synthetic_uhf_spin(NORB, NELEC[0], NELEC[1], NCHOL, seed=seed,
                                                chol_std=0.3, h1_std=0.2, mix=mix)
    if same:
        hd["chol"] = np.array([hd["chol"][0], hd["chol"][0]])
        hd["h1"] = np.array([hd["h1"][0], hd["h1"][0]])
    hd["h0"] = 0.41
    return hd, wd, nelec


def _walkers(wd, seed=9):
    rs = np.random.default_rng(seed)
    return problems.random_walkers("uhf", NORB, NELEC, W, rs, near=wd["mo_coeff"], noise=0.2)


def _build(hd, wd):
    oh = o.build_measurement_intermediates_spin(hd, wd, NORB)
    return o.build_propagation_intermediates_spin(oh, wd["rdm1"], NORB, DT)


def _step(oh, wd, walkers, fields):
    pd = {"walkers": walkers, "weights": np.linspace(0.7, 1.3, W), "e_estimate": -1.2,
          "pop_control_ene_shift": -1.25,
          "overlaps": o.calc_overlap("uhf", walkers, wd)}
    return o.propagate("uhf_spin", "uhf", oh, pd, fields, wd, DT, 8, NORB,
                       return_force_bias=True)


def test_shared_vectors_reduce_to_unrestricted_step():
    hd, wd, _ = _problem(same=True)
    oh = _build(hd, wd)
    flat = dict(hd, chol=hd["chol"][0])
    ref = o.build_measurement_intermediates("uhf", flat, wd, NORB)
    ref = o.build_propagation_intermediates("uhf", ref, wd["rdm1"], NORB, NELEC, DT)
    for k in ("mf_shifts", "h0_prop", "exp_h1"):
        np.testing.assert_allclose(oh[k], ref[k], rtol=1e-13, atol=1e-14)
    for s in range(2):
        np.testing.assert_allclose(oh["rot_chol"][s], ref["rot_chol"][s], atol=1e-14)
    walkers = _walkers(wd)
    fields = np.random.default_rng(1).standard_normal((W, NCHOL))
    a, fa = _step(oh, wd, walkers, fields)
    pd = {"walkers": walkers, "weights": np.linspace(0.7, 1.3, W), "e_estimate": -1.2,
          "pop_control_ene_shift": -1.25, "overlaps": o.calc_overlap("uhf", walkers, wd)}
    b, fb = o.propagate("uhf", "uhf", ref, pd, fields, wd, DT, 8, NORB,
                        return_force_bias=True)
    np.testing.assert_allclose(fa, fb, atol=1e-13)
    for s in range(2):
        np.testing.assert_allclose(a["walkers"][s], b["walkers"][s], atol=1e-13)
    np.testing.assert_allclose(a["weights"], b["weights"], rtol=1e-12)


def _rotate_beta(hd, wd, walkers, u):
    hd = dict(hd)
    chol = np.array(hd["chol"])
    cb = chol[1].reshape(-1, NORB, NORB)
    chol[1] = np.einsum("pi,gij,jq->gpq", u.T, cb, u).reshape(-1, NORB * NORB)
    hd["chol"] = chol
    h1 = np.array(hd["h1"])
    h1[1] = u.T @ h1[1] @ u
    hd["h1"] = h1
    mo = [wd["mo_coeff"][0], u.T @ wd["mo_coeff"][1]]
    wd = {"mo_coeff": mo, "rdm1": np.array([mo[0] @ mo[0].T, mo[1] @ mo[1].T])}
    return hd, wd, [walkers[0], np.einsum("pi,wiq->wpq", u.T, walkers[1])]


def test_beta_basis_rotation_invariance():
    hd, wd, _ = _problem()
    walkers = _walkers(wd)
    fields = np.random.default_rng(2).standard_normal((W, NCHOL))
    u = problems.haar_orthogonal(NORB, np.random.default_rng(4))
    hd2, wd2, walkers2 = _rotate_beta(hd, wd, walkers, u)
    oh, oh2 = _build(hd, wd), _build(hd2, wd2)
    np.testing.assert_allclose(oh2["mf_shifts"], oh["mf_shifts"], atol=1e-13)
    np.testing.assert_allclose(oh2["exp_h1"][1], u.T @ oh["exp_h1"][1] @ u, atol=1e-13)
    np.testing.assert_allclose(o.calc_energy("uhf", walkers2, oh2, wd2),
                               o.calc_energy("uhf", walkers, oh, wd), rtol=1e-12)
    a, fa = _step(oh, wd, walkers, fields)
    b, fb = _step(oh2, wd2, walkers2, fields)
    np.testing.assert_allclose(fb, fa, atol=1e-12)
    np.testing.assert_allclose(b["walkers"][0], a["walkers"][0], atol=1e-12)
    np.testing.assert_allclose(b["walkers"][1],
                               np.einsum("pi,wiq->wpq", u.T, a["walkers"][1]), atol=1e-12)
    np.testing.assert_allclose(b["weights"], a["weights"], rtol=1e-11)
    np.testing.assert_allclose(b["overlaps"], a["overlaps"], rtol=1e-11)


def test_spin_exchange_symmetry():
    hd, wd, _ = _problem(seed=6)
    walkers = _walkers(wd)
    # square both spins' occupations to one count so that the spins can be exchanged
    ne = NELEC[1]
    wd = {"mo_coeff": [wd["mo_coeff"][0][:, :ne], wd["mo_coeff"][1]]}
    wd["rdm1"] = np.array([m @ m.T for m in wd["mo_coeff"]])
    walkers = [walkers[0][:, :, :ne], walkers[1]]
    swapped_hd = dict(hd, chol=hd["chol"][::-1].copy(), h1=hd["h1"][::-1].copy())
    swapped_wd = {"mo_coeff": wd["mo_coeff"][::-1], "rdm1": wd["rdm1"][::-1].copy()}
    fields = np.random.default_rng(3).standard_normal((W, NCHOL))
    oh, oh2 = _build(hd, wd), _build(swapped_hd, swapped_wd)
    a, fa = _step(oh, wd, walkers, fields)
    b, fb = _step(oh2, swapped_wd, walkers[::-1], fields)
    np.testing.assert_allclose(fb, fa, atol=1e-12)
    np.testing.assert_allclose(b["walkers"][0], a["walkers"][1], atol=1e-12)
    np.testing.assert_allclose(b["walkers"][1], a["walkers"][0], atol=1e-12)
    np.testing.assert_allclose(b["weights"], a["weights"], rtol=1e-11)
    assert np.max(np.abs(oh["exp_h1"][0] - oh["exp_h1"][1])) > 1e-3  # spins differ


def test_energy_at_trial_from_spin_resolved_eris():
    hd, wd, _ = _problem(seed=8)
    oh = _build(hd, wd)
    mo = wd["mo_coeff"]
    rho = [mo[s] @ mo[s].T for s in range(2)]
    h1 = np.array([(h[:] + h.T) / 2.0 for h in hd["h1"]])
    # (pq|rs)_st = sum_g L^s_g[pq] L^t_g[rs], by loops
    L = np.asarray(hd["chol"]).reshape(2, NCHOL, NORB, NORB)
    eri = np.zeros((2, 2, NORB, NORB, NORB, NORB))
    for s in range(2):
        for t in range(2):
            for g in range(NCHOL):
                eri[s, t] += L[s, g][:, :, None, None] * L[t, g][None, None, :, :]
    e = hd["h0"] + sum(np.sum(h1[s] * rho[s]) for s in range(2))
    for s in range(2):
        for t in range(2):
            e += 0.5 * np.einsum("pqrs,pq,rs->", eri[s, t], rho[s], rho[t])
        e -= 0.5 * np.einsum("pqrs,ps,rq->", eri[s, s], rho[s], rho[s])
    walkers = [np.array([mo[0] + 0.0j]), np.array([mo[1] + 0.0j])]
    got = o.calc_energy("uhf", walkers, oh, wd)[0]
    assert abs(got - e) < 1e-12 * max(1.0, abs(e))


def test_force_bias_at_trial_is_minus_i_mf_shift():
    hd, wd, _ = _problem(seed=10)
    oh = _build(hd, wd)
    mo = wd["mo_coeff"]
    walkers = [np.array([mo[0] + 0.0j]), np.array([mo[1] + 0.0j])]
    fb = o.calc_force_bias("uhf", walkers, oh, wd)[0]
    np.testing.assert_allclose(fb, -1.0j * oh["mf_shifts"], atol=1e-13)


@pytest.mark.parametrize("spin", [0, 1])
def test_trotter_step_converges_to_expm(spin):
    hd, wd, _ = _problem(seed=12)
    oh = _build(hd, wd)
    L = np.asarray(hd["chol"]).reshape(2, NCHOL, NORB, NORB)
    h1 = np.array([(h + h.T) / 2.0 for h in hd["h1"]])
    rdm1 = wd["rdm1"]
    # intermediates by loops
    mf = np.zeros(NCHOL, dtype=complex)
    for g in range(NCHOL):
        for s in range(2):
            for p in range(NORB):
                for q in range(NORB):
                    mf[g] += 1.0j * L[s, g, p, q] * rdm1[s][p, q]
    v0 = np.zeros((NORB, NORB))
    v1 = np.zeros((NORB, NORB))
    for g in range(NCHOL):
        for i in range(NORB):
            for j in range(NORB):
                v0[i, j] += 0.5 * sum(L[spin, g, i, k] * L[spin, g, j, k] for k in range(NORB))
                v1[i, j] += np.real(1.0j * mf[g] * L[spin, g, i, j])
    half = scipy.linalg.expm(-DT * (h1[spin] - v0 - v1) / 2.0)
    np.testing.assert_allclose(oh["exp_h1"][spin], half, atol=1e-14)
    x = np.random.default_rng(5).standard_normal((W, NCHOL))
    walkers = _walkers(wd)
    hi = dict(oh)
    got = o.apply_trotprop("uhf_spin", hi, walkers, x, DT, 40, NORB)[spin]
    for w in range(W):
        vhs = sum(1.0j * np.sqrt(DT) * x[w, g] * L[spin, g] for g in range(NCHOL))
        want = half @ scipy.linalg.expm(vhs) @ half @ walkers[spin][w]
        np.testing.assert_allclose(got[w], want, atol=1e-12)
