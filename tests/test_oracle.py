"""Pins the CPU oracle (oracle/afqmc_oracle.py) before it is trusted as the checker.

Sources of truth, in order: the reference's own golden values
(tests/test_wavefunctions.py:127-183 of the reference, inputs regenerated with the same
legacy seed and call order), then closed forms and invariants that a dropped term, a
wrong sign or a transposed operand would break.
"""

import numpy as np
import pytest
import scipy.linalg

from oracle import afqmc_oracle as o

# ----------------------------------------------------------------- reference goldens


def _reference_fixture():
    """Inputs of tests/test_wavefunctions.py:12-85 (reference): np.random.seed(102) and
    the same sequence of np.random.rand calls (tuple elements left to right)."""
    np.random.seed(102)
    norb, nelec, nchol = 10, (5, 5), 5
    mo = np.eye(norb)[:, : nelec[0]]
    walker = np.random.rand(norb, nelec[0]) + 1.0j * np.random.rand(norb, nelec[0])
    h0 = np.random.rand(1)[0]
    rot_h1 = np.random.rand(nelec[0], norb)
    rot_chol = np.random.rand(nchol, nelec[0], norb)
    np.random.rand(norb, norb)  # h1 (unused by these trial tests)
    np.random.rand(nchol, norb, norb)  # chol
    nelec_sp = (3, 2)
    mo_u = [np.random.rand(norb, nelec_sp[0]), np.random.rand(norb, nelec_sp[1])]
    w_up = np.random.rand(norb, nelec_sp[0]) + 1.0j * np.random.rand(norb, nelec_sp[0])
    w_dn = np.random.rand(norb, nelec_sp[1]) + 1.0j * np.random.rand(norb, nelec_sp[1])
    return dict(mo=mo, walker=walker, h0=h0, rot_h1=rot_h1, rot_chol=rot_chol, mo_u=mo_u,
                w_up=w_up, w_dn=w_dn, nelec_sp=nelec_sp)


def test_reference_goldens_rhf():
    f = _reference_fixture()
    # reference tests/test_wavefunctions.py:127-146
    assert np.allclose(np.real(o.rhf_overlap_restricted(f["walker"], f["mo"])),
                       -0.10794844182417201, rtol=1e-12)
    g = o.rhf_green(f["walker"], f["mo"])
    assert g.shape == (5, 10)
    assert np.allclose(np.real(np.sum(g)), 12.181348093111438, rtol=1e-12)
    fb = o.rhf_force_bias_restricted(f["walker"], f["rot_chol"], f["mo"])
    assert fb.shape == (5,)
    assert np.allclose(np.real(np.sum(fb)), 66.13455680423321, rtol=1e-12)
    e = o.rhf_energy_restricted(f["walker"], f["h0"], f["rot_h1"], f["rot_chol"], f["mo"])
    assert np.allclose(np.real(e), 217.79874063608622, rtol=1e-12)


def test_reference_goldens_uhf():
    f = _reference_fixture()
    n0, n1 = f["nelec_sp"]
    rot_h1 = [f["rot_h1"][:n0], f["rot_h1"][:n1]]
    rot_chol = [f["rot_chol"][:, :n0], f["rot_chol"][:, :n1]]
    # reference tests/test_wavefunctions.py:156-183
    assert np.allclose(np.real(o.uhf_overlap(f["w_up"], f["w_dn"], f["mo_u"])),
                       -0.4029825074695857, rtol=1e-10)
    g = o.uhf_green(f["w_up"], f["w_dn"], f["mo_u"])
    assert np.allclose(np.real(np.sum(g[0]) + np.sum(g[1])), 3.6324117896217394, rtol=1e-12)
    fb = o.uhf_force_bias(f["w_up"], f["w_dn"], rot_chol, f["mo_u"])
    assert np.allclose(np.real(np.sum(fb)), 10.441272099672341, rtol=1e-12)
    e = o.uhf_energy(f["w_up"], f["w_dn"], f["h0"], rot_h1, rot_chol, f["mo_u"])
    assert np.allclose(np.real(e), -1.7203463308366032, rtol=1e-12)


# ----------------------------------------------------------------- closed forms


def _small_problem(walker_type, seed=3, norb=8, nelec=(3, 3), nchol=12):
    from cs_afqmc_amd import problems

    if walker_type == "rhf":
        ham, wave, nelec = problems.synthetic_rhf(norb, nelec[0], nchol, seed=seed, decay=6.0)
    else:
        ham, wave, nelec = problems.synthetic_uhf(norb, nelec[0], nelec[1], nchol, seed=seed,
                                                  decay=6.0)
    ham["h0"] = 0.7
    ham = o.build_measurement_intermediates(walker_type, ham, wave, norb)
    ham = o.build_propagation_intermediates(walker_type, ham, wave["rdm1"], norb, nelec,
                                            dt=0.01)
    return ham, wave, nelec, norb


@pytest.mark.parametrize("walker_type", ["rhf", "uhf"])
def test_invariance_under_walker_rotation(walker_type):
    ham, wave, nelec, norb = _small_problem(walker_type)
    rng = np.random.default_rng(1)
    if walker_type == "rhf":
        w = wave["mo_coeff"] + 0.3 * (rng.standard_normal((norb, nelec[0]))
                                      + 1j * rng.standard_normal((norb, nelec[0])))
        R = rng.standard_normal((nelec[0], nelec[0])) + 1j * rng.standard_normal((nelec[0],) * 2)
        ws, wr = w[None], (w @ R)[None]
        scale = np.linalg.det(R) ** 2
    else:
        w = [m + 0.3 * (rng.standard_normal(m.shape) + 1j * rng.standard_normal(m.shape))
             for m in wave["mo_coeff"]]
        R = [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for n in nelec]
        ws = [w[0][None], w[1][None]]
        wr = [(w[0] @ R[0])[None], (w[1] @ R[1])[None]]
        scale = np.linalg.det(R[0]) * np.linalg.det(R[1])
    ov, ovr = o.calc_overlap(walker_type, ws, wave), o.calc_overlap(walker_type, wr, wave)
    assert np.allclose(ovr, ov * scale, rtol=1e-11)
    fb, fbr = (o.calc_force_bias(walker_type, x, ham, wave) for x in (ws, wr))
    assert np.allclose(fbr, fb, rtol=1e-11, atol=1e-12)
    e, er = (o.calc_energy(walker_type, x, ham, wave) for x in (ws, wr))
    assert np.allclose(er, e, rtol=1e-11)


@pytest.mark.parametrize("walker_type", ["rhf", "uhf"])
def test_force_bias_at_trial_is_minus_i_mf_shift(walker_type):
    """At walker == trial the field shift vanishes (propagation.py:122 with
    mf_shifts from :298-300): fb = -i mf_shifts."""
    ham, wave, nelec, norb = _small_problem(walker_type)
    w = o.get_init_walkers(wave["rdm1"], nelec, 1, restricted=(walker_type == "rhf"))
    fb = o.calc_force_bias(walker_type, w, ham, wave)[0]
    assert np.allclose(fb, -1j * ham["mf_shifts"], atol=1e-12)


def _eri(chol, norb):
    L = chol.reshape(-1, norb, norb)
    return np.einsum("gpq,grs->pqrs", L, L)


@pytest.mark.parametrize("walker_type", ["rhf", "uhf"])
def test_energy_at_trial_is_hartree_fock(walker_type):
    """E(walker == trial) = textbook HF energy from h1 and (pq|rs) = sum_g L_pq L_rs."""
    ham, wave, nelec, norb = _small_problem(walker_type)
    eri = _eri(ham["chol"], norb)
    h = ham["h1"]
    if walker_type == "rhf":
        mo = wave["mo_coeff"]
        w = mo[None] + 0j
        D = [mo @ np.linalg.inv(mo.T @ mo) @ mo.T] * 2
    else:
        mo = wave["mo_coeff"]
        w = [mo[0][None] + 0j, mo[1][None] + 0j]
        D = [m @ np.linalg.inv(m.T @ m) @ m.T for m in mo]
    Dt = D[0] + D[1]
    e_hf = ham["h0"] + np.sum(h[0] * D[0]) + np.sum(h[1] * D[1])
    e_hf += 0.5 * np.einsum("pqrs,qp,sr->", eri, Dt, Dt)
    e_hf -= 0.5 * sum(np.einsum("pqrs,qr,sp->", eri, d, d) for d in D)
    e = o.calc_energy(walker_type, w, ham, wave)[0]
    assert abs(e.imag) < 1e-12
    assert np.isclose(e.real, e_hf, rtol=1e-12)


def test_taylor_converges_to_expm():
    rng = np.random.default_rng(5)
    n, k = 7, 3
    h = rng.standard_normal((n, n))
    h = (h + h.T) / 2
    E = scipy.linalg.expm(-0.01 * h / 2)
    V = 0.2 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    w = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    exact = E @ scipy.linalg.expm(V) @ E @ w
    assert np.allclose(o.apply_trotprop_det(E, V, w, 30), exact, rtol=1e-13, atol=1e-13)
    # truncated series error is bounded by the Taylor remainder
    nv = np.linalg.norm(V, 2)
    err = np.linalg.norm(o.apply_trotprop_det(E, V, w, 6) - exact)
    bound = np.linalg.norm(E, 2) ** 2 * np.linalg.norm(w) * nv**6 / 720 * np.exp(nv)
    assert err <= bound


def test_zero_cholesky_step_is_one_body_exponential():
    """chol == 0: VHS, mf shifts and force bias vanish, the walker becomes
    expm(-dt h1) walker and the weight factor is exp(dt (E_shift - h0)) O'/O."""
    norb, n, W, dt = 6, 2, 3, 0.05
    rng = np.random.default_rng(11)
    h1 = rng.standard_normal((norb, norb))
    h1 = (h1 + h1.T) / 2
    mo = np.linalg.qr(rng.standard_normal((norb, norb)))[0][:, :n]
    ham = {"h0": 0.3, "h1": np.array([h1, h1]), "chol": np.zeros((2, norb * norb)), "ene0": 0.0}
    wave = {"mo_coeff": mo, "rdm1": np.array([mo @ mo.T] * 2)}
    ham = o.build_measurement_intermediates("rhf", ham, wave, norb)
    ham = o.build_propagation_intermediates("rhf", ham, wave["rdm1"], norb, (n, n), dt)
    walkers = mo[None] + 0.1 * rng.standard_normal((W, norb, n)) + 0j
    prop = {"walkers": walkers, "weights": np.ones(W), "e_estimate": -1.0,
            "pop_control_ene_shift": -1.2}
    prop["overlaps"] = o.calc_overlap("rhf", walkers, wave)
    fields = rng.standard_normal((W, 2))
    new = o.propagate("rhf", "rhf", ham, prop, fields, wave, dt, 6, norb)
    U = scipy.linalg.expm(-dt * h1)
    assert np.allclose(new["walkers"], U @ walkers, rtol=1e-13, atol=1e-14)
    ratio = new["overlaps"] / prop["overlaps"]
    expect = np.exp(dt * (-1.2 - 0.3)) * np.abs(ratio) * np.cos(np.angle(ratio))
    expect = np.where((expect < 1e-3) | (expect > 100), 0.0, expect)
    assert np.allclose(new["weights"], expect, rtol=1e-13)
    # population control: E_est - 0.1 log(mean w) / dt
    assert np.isclose(new["pop_control_ene_shift"], -1.0 - 0.1 * np.log(np.mean(expect)) / dt)


def test_weight_clipping_branches():
    """NaN, < 1e-3 and > 100 importance factors and weights > 100 are zeroed
    (propagation.py:149-160)."""
    norb, n, dt = 4, 1, 0.01
    ham = {"h0": 0.0, "h1": np.zeros((2, norb, norb)), "chol": np.zeros((1, norb * norb)),
           "ene0": 0.0}
    mo = np.eye(norb)[:, :n]
    wave = {"mo_coeff": mo, "rdm1": np.array([mo @ mo.T] * 2)}
    ham = o.build_measurement_intermediates("rhf", ham, wave, norb)
    ham = o.build_propagation_intermediates("rhf", ham, wave["rdm1"], norb, (n, n), dt)
    walkers = np.array([mo] * 5) + 0j
    # overlap of each walker is 1; old overlaps chosen to hit each branch
    old = np.array([np.nan, 1e6, 1e-6, 1.0, 1.0], dtype=complex)
    weights = np.array([1.0, 1.0, 1.0, 150.0, 2.0])
    prop = {"walkers": walkers, "weights": weights, "overlaps": old, "e_estimate": 0.0,
            "pop_control_ene_shift": 0.0}
    new = o.propagate("rhf", "rhf", ham, prop, np.zeros((5, 1)), wave, dt, 6, norb)
    assert new["weights"][0] == 0.0  # NaN
    assert new["weights"][1] == 0.0  # < 1e-3
    assert new["weights"][2] == 0.0  # > 100
    assert new["weights"][3] == 0.0  # weight > 100
    assert np.isclose(new["weights"][4], 2.0, rtol=1e-14)


def test_qr_convention():
    rng = np.random.default_rng(2)
    w = rng.standard_normal((3, 9, 4)) + 1j * rng.standard_normal((3, 9, 4))
    q = o.orthonormalize_walkers("rhf", w)
    for qi, wi in zip(q, w):
        assert np.allclose(qi.conj().T @ qi, np.eye(4), atol=1e-13)
        r = qi.conj().T @ wi
        assert np.allclose(np.tril(r, -1), 0, atol=1e-13)
        assert np.allclose(np.diag(r).imag, 0, atol=1e-13)  # real diagonal (zlarfg)
        # zlarfg: beta = -sign(Re alpha) |x|; first column: alpha = w[0, 0]
        assert np.sign(np.diag(r)[0].real) == -np.sign(wi[0, 0].real)
        assert np.allclose(qi @ r, wi, atol=1e-13)


def test_sr_systematic_resampling_counts():
    rng = np.random.default_rng(4)
    for W in (1, 7, 64, 1000):
        w = rng.random(W) * (rng.random(W) > 0.2)
        if w.sum() == 0:
            w[0] = 1.0
        zeta = rng.random()
        for fn in (o.sr_indices, o.sr_indices_global):
            idx, total = fn(w, zeta)
            assert np.all(np.diff(idx) >= 0)
            counts = np.bincount(idx, minlength=W)
            assert counts.sum() == W
            expect = W * np.abs(w) / np.abs(w).sum()
            assert np.all(counts >= np.floor(expect) - 1e-9)
            assert np.all(counts <= np.ceil(expect) + 1e-9)
            assert np.all(counts[w == 0] == 0)
        # the two comb-point roundings agree away from ties
        assert np.array_equal(o.sr_indices(w, zeta)[0], o.sr_indices_global(w, zeta)[0])


def test_propagation_intermediates_by_loops():
    norb, nchol, dt = 4, 3, 0.02
    rng = np.random.default_rng(8)
    L = rng.standard_normal((nchol, norb, norb))
    L = L + L.transpose(0, 2, 1)
    h1 = rng.standard_normal((norb, norb))
    h1 = h1 + h1.T
    mo = np.linalg.qr(rng.standard_normal((norb, norb)))[0][:, :2]
    rdm1 = np.array([mo @ mo.T] * 2)
    ham = {"h0": 0.1, "h1": np.array([h1, h1]), "chol": L.reshape(nchol, -1), "ene0": 0.0}
    out = o.build_propagation_intermediates("rhf", ham, rdm1, norb, (2, 2), dt)
    mf = np.zeros(nchol, complex)
    v0 = np.zeros((norb, norb))
    for g in range(nchol):
        for i in range(norb):
            for j in range(norb):
                mf[g] += 1j * L[g, i, j] * (rdm1[0] + rdm1[1])[i, j]
                for k in range(norb):
                    v0[i, j] += 0.5 * L[g, i, k] * L[g, j, k]
    assert np.allclose(out["mf_shifts"], mf, rtol=1e-13)
    assert np.isclose(out["h0_prop"], -0.1 - np.sum(mf * mf) / 2)
    hmod = h1 - v0 - sum((1j * mf[g] * L[g]).real for g in range(nchol))
    # expm by its power series (independent of scipy's Pade)
    A = -dt * hmod / 2
    term, ser = np.eye(norb), np.eye(norb)
    for k in range(1, 30):
        term = term @ A / k
        ser = ser + term
    assert np.allclose(out["exp_h1"], ser, rtol=1e-13, atol=1e-15)


def test_corr_outlier_rule_closed_form():
    """corr_otler_rm: one weight far out in system 1 zeroes that walker in both systems;
    equal weights (std 0, z = 0/0) are kept, as jnp.where keeps them."""
    w1 = np.ones(30)
    w1[7] = 100.0
    w2 = np.linspace(0.5, 1.5, 30)
    m, sd = w1.mean(), w1.std()
    assert abs(100.0 - m) / sd > 5.0 and abs(1.0 - m) / sd < 5.0  # the case is what it claims
    a, b = o.corr_outlier_rm(w1, w2)
    assert a[7] == 0.0 and b[7] == 0.0
    keep = np.arange(30) != 7
    assert np.array_equal(a[keep], w1[keep]) and np.array_equal(b[keep], w2[keep])
    c, d = o.corr_outlier_rm(np.ones(8), np.ones(8))
    assert np.array_equal(c, np.ones(8)) and np.array_equal(d, np.ones(8))


# ------------------------------------------------------ restricted start-orbital fallback


def _rdm1(vecs, occs):
    """sum_i occ_i v_i v_i^T: natural orbitals v_i with distinct occupations, so the
    eigenvectors are fixed up to sign."""
    return sum(oc * np.outer(v, v) for v, oc in zip(vecs, occs))


def _unit(v):
    return v / np.linalg.norm(v)


def _fallback_case(n, eps):
    """Spin-up natural orbitals e_0 .. e_{n-1}, spin-down d_i = unit(e_{n+i} + eps e_i):
    det(up^T dn) = (eps / sqrt(1 + eps^2))^n, and the columns e_i + d_i are mutually
    orthogonal, so the fallback's QR leaves them as they are, normalised (up to sign)."""
    norb = 2 * n + 2
    e = np.eye(norb)
    up = [e[i] for i in range(n)]
    dn = [_unit(e[n + i] + eps * e[i]) for i in range(n)]
    occs = [0.95 - 0.1 * i for i in range(n)]
    rdm1 = np.array([_rdm1(up, occs), _rdm1(dn, occs)])
    expect = [_unit(u + d) for u, d in zip(up, dn)]
    return rdm1, (n, n), expect, up, dn


@pytest.mark.parametrize("n,eps", [(1, 5e-4), (2, 1e-2), (3, 0.05)])
def test_restricted_fallback_closed_form(n, eps):
    """wavefunctions.py:325-342: with |det(up^T dn)| <= 1e-3 the restricted start is
    QR(up_i + sign(up_i . dn_i) dn_i).  Closed form for orthogonal column sums: the
    columns unit(e_i + d_i) up to sign, and |det(Q^T up) det(Q^T dn)| =
    ((1 + c) / sqrt(2 + 2 c))^(2n) with c = eps / sqrt(1 + eps^2)."""
    rdm1, nelec, expect, up, dn = _fallback_case(n, eps)
    c = eps / np.sqrt(1 + eps**2)
    assert c**n <= 1e-3  # the fallback branch is taken
    w = o.get_init_walkers(rdm1, nelec, 3, restricted=True)
    assert w.shape == (3, rdm1.shape[1], n)
    q = w[0].real
    assert np.allclose(q.T @ q, np.eye(n), atol=1e-13)
    for j in range(n):
        assert abs(abs(q[:, j] @ expect[j]) - 1.0) < 1e-12
    det = np.linalg.det(q.T @ np.array(up).T) * np.linalg.det(q.T @ np.array(dn).T)
    assert abs(abs(det) - ((1 + c) / np.sqrt(2 + 2 * c)) ** (2 * n)) < 1e-12


def test_restricted_fallback_raises_without_overlap():
    """Orthogonal spin orbitals: sign(0) = 0 drops the spin-down columns, the QR gives the
    spin-up orbitals, det(Q^T dn) = 0, and the reference raises ValueError (:339-342)."""
    e = np.eye(4)
    rdm1 = np.array([np.outer(e[0], e[0]), np.outer(e[1], e[1])])
    with pytest.raises(ValueError):
        o.get_init_walkers(rdm1, (1, 1), 1, restricted=True)


def test_restricted_start_above_threshold_takes_spin_up():
    """|det(up^T dn)| > 1e-3 keeps the spin-up natural orbitals (:316-324)."""
    rdm1, nelec, _, up, _ = _fallback_case(1, 0.1)
    w = o.get_init_walkers(rdm1, nelec, 1, restricted=True)
    assert abs(abs(w[0][:, 0].real @ up[0]) - 1.0) < 1e-14


@pytest.mark.parametrize("n,eps", [(1, 5e-4), (2, 1e-2), (3, 0.05), (2, 0.5)])
def test_product_start_orbitals_match_oracle(n, eps):
    """cs_afqmc_amd.wavefunctions.init_orbitals (the host part of get_init_walkers) equals
    the oracle on the same rdm1, in both branches, and raises where the oracle does."""
    from cs_afqmc_amd.wavefunctions import init_orbitals

    rdm1, nelec, *_ = _fallback_case(n, eps)
    got = init_orbitals(rdm1, nelec, True)
    ref = o.get_init_walkers(rdm1, nelec, 1, restricted=True)[0]
    assert np.allclose(got, ref.real, atol=1e-14)
    up, dn = init_orbitals(rdm1, nelec, False)
    ref_up, ref_dn = o.get_init_walkers(rdm1, nelec, 1, restricted=False)
    assert np.allclose(up, ref_up[0].real, atol=1e-14)
    assert np.allclose(dn, ref_dn[0].real, atol=1e-14)
    e = np.eye(4)
    bad = np.array([np.outer(e[0], e[0]), np.outer(e[1], e[1])])
    with pytest.raises(ValueError):
        init_orbitals(bad, (1, 1), True)
