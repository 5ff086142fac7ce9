"""CPU restatement (numpy, float64/complex128) of ad_afqmc's phaseless AFQMC step.

TEST INFRASTRUCTURE ONLY.  This module is the oracle that the HIP path is checked
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or the timed CPU baseline) --
never as the code path that is shipped or measured on the GPU.

Every function follows one reference function and cites it (file:line, paths relative
to the reference repository root).  Contractions that XLA lowers to BLAS dot products in
the reference run through numpy's BLAS here too (einsum with optimize=True).  The reference is pure JAX; this restatement keeps
its operation order (e.g. ``exp(a) * o_new / o_old``) so that float64 results agree to
rounding.  Batching (``n_batch`` scans, ``vmap``) does not change results and is
dropped.

Walker conventions (as in the reference):
  * restricted walkers: complex array (W, norb, nocc)
  * unrestricted walkers: list [up (W, norb, nelec[0]), dn (W, norb, nelec[1])]

Pinning (what checks each function besides its own restatement; see tests/):
  * rhf/uhf overlap, green, force bias, energy: golden values of the reference's own
    tests (tests/test_wavefunctions.py:127-183), regenerated inputs; invariance under
    walker -> walker @ R; force bias at walker == trial equals -i * mf_shifts; energy
    at walker == trial equals the closed-shell Hartree-Fock energy from h1 and
    (pq|rs) = sum_g L_g[pq] L_g[rs].
  * apply_trotprop_det: n_exp_terms -> large converges to expm(-dt h/2) expm(V)
    expm(-dt h/2) (scipy); chol == 0 gives expm(-dt h1_mod) exactly.
  * propagate: field shift vanishes at walker == trial; closed-form single-walker
    weight update with chol == 0; clipping branches (NaN, < 1e-3, > 100, w > 100).
  * get_init_walkers: restricted fallback (|det(up^T dn)| <= 1e-3) against closed
    forms for orthogonal column sums (columns, det product), the ValueError case, and
    the spin-up branch above the threshold.
  * orthonormalize: numpy.linalg.qr is LAPACK geqrf + orgqr/ungqr, the routine
    jnp.linalg.qr calls on CPU; invariants Q^H Q = I, Q R = walker, real diag(R) with
    the zlarfg sign rule.
  * stochastic_reconfiguration: brute-force systematic resampling counts
    (copies_i in {floor, ceil} of W |w_i| / sum|w|, sum copies == W, sorted indices).
  * build_*_intermediates: scipy expm, explicit loops on tiny sizes.
  * Whole sampler: statistical agreement with the reference run committed in
    examples/ene_err.txt (H10 @ 2.0 bohr).
Parity unpinned (no reference vector exists, an import of jax would be needed):
jax.random field/zeta streams (fields and zeta are inputs here), XLA cumsum
association order (here: sequential left-to-right, as numpy.cumsum).
"""

import math

import numpy as np
import scipy.linalg


# ----------------------------------------------------------------------------------
# trial wave functions (ad_afqmc/wavefunctions.py)
# ----------------------------------------------------------------------------------


def rhf_overlap_restricted(walker, mo_coeff):
    """wavefunctions.py:478-480 (rhf._calc_overlap_restricted): det(C^H w)^2."""
    return np.linalg.det(mo_coeff.T.conj() @ walker) ** 2


def rhf_green(walker, mo_coeff):
    """wavefunctions.py:490-501 (rhf._calc_green): (w (C^H w)^-1)^T, shape (nocc, norb)."""
    return (walker.dot(np.linalg.inv(mo_coeff.T.conj() @ walker))).T


def rhf_force_bias_restricted(walker, rot_chol, mo_coeff):
    """wavefunctions.py:503-511: fb_g = 2 sum_ij rot_chol[g,i,j] G[i,j]."""
    green = rhf_green(walker, mo_coeff)
    return 2.0 * np.einsum("gij,ij->g", rot_chol, green)


def rhf_energy_restricted(walker, h0, rot_h1, rot_chol, mo_coeff):
    """wavefunctions.py:529-541 (rhf._calc_energy_restricted)."""
    ene0 = h0
    green = rhf_green(walker, mo_coeff)
    ene1 = 2.0 * np.sum(green * rot_h1)
    f = np.einsum("gij,jk->gik", rot_chol, green.T, optimize=True)
    c = np.trace(f, axis1=1, axis2=2)
    exc = np.sum(f * np.transpose(f, (0, 2, 1)))
    ene2 = 2.0 * np.sum(c * c) - exc
    return ene2 + ene1 + ene0


def rhf_rdm1(mo_coeff):
    """wavefunctions.py:563-565 (rhf._calc_rdm1)."""
    return np.array([mo_coeff @ mo_coeff.T] * 2)


def uhf_overlap(walker_up, walker_dn, mo_coeff):
    """wavefunctions.py:658-667 (uhf._calc_overlap)."""
    return np.linalg.det(mo_coeff[0].T.conj() @ walker_up) * np.linalg.det(
        mo_coeff[1].T.conj() @ walker_dn
    )


def uhf_green(walker_up, walker_dn, mo_coeff):
    """wavefunctions.py:669-692 (uhf._calc_green)."""
    g_up = (walker_up.dot(np.linalg.inv(mo_coeff[0].T.conj() @ walker_up))).T
    g_dn = (walker_dn.dot(np.linalg.inv(mo_coeff[1].T.conj() @ walker_dn))).T
    return [g_up, g_dn]


def uhf_force_bias(walker_up, walker_dn, rot_chol, mo_coeff):
    """wavefunctions.py:694-709 (uhf._calc_force_bias)."""
    g = uhf_green(walker_up, walker_dn, mo_coeff)
    fb_up = np.einsum("gij,ij->g", rot_chol[0], g[0])
    fb_dn = np.einsum("gij,ij->g", rot_chol[1], g[1])
    return fb_up + fb_dn


def uhf_energy(walker_up, walker_dn, h0, rot_h1, rot_chol, mo_coeff):
    """wavefunctions.py:711-743 (uhf._calc_energy)."""
    ene0 = h0
    g = uhf_green(walker_up, walker_dn, mo_coeff)
    ene1 = np.sum(g[0] * rot_h1[0]) + np.sum(g[1] * rot_h1[1])
    f_up = np.einsum("gij,jk->gik", rot_chol[0], g[0].T, optimize=True)
    f_dn = np.einsum("gij,jk->gik", rot_chol[1], g[1].T, optimize=True)
    c_up = np.trace(f_up, axis1=1, axis2=2)
    c_dn = np.trace(f_dn, axis1=1, axis2=2)
    exc_up = np.sum(f_up * np.transpose(f_up, (0, 2, 1)))
    exc_dn = np.sum(f_dn * np.transpose(f_dn, (0, 2, 1)))
    ene2 = (
        np.sum(c_up * c_up)
        + np.sum(c_dn * c_dn)
        + 2.0 * np.sum(c_up * c_dn)
        - exc_up
        - exc_dn
    ) / 2.0
    return ene2 + ene1 + ene0


def uhf_rdm1(mo_coeff):
    """wavefunctions.py:745-748 (uhf._calc_rdm1)."""
    return np.array(
        [mo_coeff[0] @ mo_coeff[0].T.conj(), mo_coeff[1] @ mo_coeff[1].T.conj()]
    )


# batched measurement entry points: wavefunctions.py:61-96 (calc_overlap),
# :127-165 (calc_force_bias), :203-241 (calc_energy)


def calc_overlap(trial, walkers, wave_data):
    if trial == "rhf":
        return np.array([rhf_overlap_restricted(w, wave_data["mo_coeff"]) for w in walkers])
    return np.array(
        [uhf_overlap(u, d, wave_data["mo_coeff"]) for u, d in zip(walkers[0], walkers[1])]
    )


def calc_green(trial, walkers, wave_data):
    if trial == "rhf":
        return np.array([rhf_green(w, wave_data["mo_coeff"]) for w in walkers])
    gs = [uhf_green(u, d, wave_data["mo_coeff"]) for u, d in zip(walkers[0], walkers[1])]
    return [np.array([g[0] for g in gs]), np.array([g[1] for g in gs])]


def calc_force_bias(trial, walkers, ham_data, wave_data):
    if trial == "rhf":
        return np.array(
            [
                rhf_force_bias_restricted(w, ham_data["rot_chol"], wave_data["mo_coeff"])
                for w in walkers
            ]
        )
    return np.array(
        [
            uhf_force_bias(u, d, ham_data["rot_chol"], wave_data["mo_coeff"])
            for u, d in zip(walkers[0], walkers[1])
        ]
    )


def calc_energy(trial, walkers, ham_data, wave_data):
    if trial == "rhf":
        return np.array(
            [
                rhf_energy_restricted(
                    w, ham_data["h0"], ham_data["rot_h1"], ham_data["rot_chol"],
                    wave_data["mo_coeff"],
                )
                for w in walkers
            ]
        )
    return np.array(
        [
            uhf_energy(
                u, d, ham_data["h0"], ham_data["rot_h1"], ham_data["rot_chol"],
                wave_data["mo_coeff"],
            )
            for u, d in zip(walkers[0], walkers[1])
        ]
    )


def build_measurement_intermediates(trial, ham_data, wave_data, norb):
    """rhf: wavefunctions.py:602-619; uhf: wavefunctions.py:817-841."""
    ham_data = dict(ham_data)
    h1 = np.array(ham_data["h1"], dtype=np.float64)
    h1 = np.array([(h1[0] + h1[0].T) / 2.0, (h1[1] + h1[1].T) / 2.0])
    ham_data["h1"] = h1
    chol3 = np.asarray(ham_data["chol"]).reshape(-1, norb, norb)
    if trial == "rhf":
        mo = wave_data["mo_coeff"]
        ham_data["rot_h1"] = mo.T.conj() @ ((h1[0] + h1[1]) / 2.0)
        ham_data["rot_chol"] = np.einsum("pi,gij->gpj", mo.T.conj(), chol3, optimize=True)
    else:
        mo = wave_data["mo_coeff"]
        ham_data["rot_h1"] = [mo[0].T.conj() @ h1[0], mo[1].T.conj() @ h1[1]]
        ham_data["rot_chol"] = [
            np.einsum("pi,gij->gpj", mo[0].T.conj(), chol3, optimize=True),
            np.einsum("pi,gij->gpj", mo[1].T.conj(), chol3, optimize=True),
        ]
    return ham_data


def get_init_walkers(rdm1, nelec, n_walkers, restricted):
    """wavefunctions.py:298-355 (wave_function.get_init_walkers): natural orbitals
    (descending occupation) of each spin rdm1, copied n_walkers times.  Restricted
    walkers (equal spin counts) take the spin-up orbitals when |det(up^T dn)| > 1e-3
    (:316-324); otherwise (:325-342) the columns up_i + sign(up_i . dn_i) dn_i,
    orthonormalised by QR, if |det(Q^T up) det(Q^T dn)| > 1e-3, else ValueError."""
    natorbs_up = np.linalg.eigh(rdm1[0])[1][:, ::-1][:, : nelec[0]]
    natorbs_dn = np.linalg.eigh(rdm1[1])[1][:, ::-1][:, : nelec[1]]
    if restricted:
        if nelec[0] != nelec[1]:
            raise NotImplementedError("restricted walkers need nelec[0] == nelec[1]")
        det_overlap = np.linalg.det(natorbs_up.T @ natorbs_dn)
        if np.abs(det_overlap) > 1e-3:
            return np.array([natorbs_up + 0.0j] * n_walkers)
        signs = np.sign(np.sum(natorbs_up * natorbs_dn, axis=0))
        q = np.linalg.qr(natorbs_up + natorbs_dn * signs[None, :])[0]
        det_overlap = np.linalg.det(q.T @ natorbs_up) * np.linalg.det(q.T @ natorbs_dn)
        if np.abs(det_overlap) > 1e-3:
            return np.array([q + 0.0j] * n_walkers)
        raise ValueError("no RHF orbitals with a trial overlap above 1e-3")
    return [
        np.array([natorbs_up + 0.0j] * n_walkers),
        np.array([natorbs_dn + 0.0j] * n_walkers),
    ]


# ----------------------------------------------------------------------------------
# propagator (ad_afqmc/propagation.py)
# ----------------------------------------------------------------------------------


def build_propagation_intermediates(walker_type, ham_data, rdm1, norb, nelec, dt):
    """restricted: propagation.py:292-325; unrestricted: propagation.py:464-509."""
    ham_data = dict(ham_data)
    chol = np.asarray(ham_data["chol"]).reshape(-1, norb * norb)
    chol3 = chol.reshape(-1, norb, norb)
    rdm1t = rdm1[0] + rdm1[1]
    mf_shifts = 1.0j * np.array([np.sum(x.reshape(norb, norb) * rdm1t) for x in chol])
    ham_data["mf_shifts"] = mf_shifts
    ham_data["h0_prop"] = -ham_data["h0"] - np.sum(mf_shifts**2) / 2.0
    v0 = 0.5 * np.einsum("gik,gjk->ij", chol3, chol3, optimize=True)
    v1 = np.real(1.0j * np.einsum("g,gik->ik", mf_shifts, chol3, optimize=True))
    h1 = np.asarray(ham_data["h1"])
    if walker_type == "rhf":
        h1_mod = (h1[0] + h1[1]) / 2.0 - v0
        h1_mod = h1_mod - v1
        ham_data["exp_h1"] = scipy.linalg.expm(-dt * h1_mod / 2.0)
    else:
        h1_mod = h1 - np.array([v0 + v1, v0 + v1])
        ham_data["exp_h1"] = np.array(
            [
                scipy.linalg.expm(-dt * h1_mod[0] / 2.0),
                scipy.linalg.expm(-dt * h1_mod[1] / 2.0),
            ]
        )
    return ham_data


def build_measurement_intermediates_spin(ham_data, wave_data, norb):
    """prop_unrestricted/wavefunctions_unrestricted.py:453-475 (uhf with spin-resolved
    Cholesky vectors, chol (2, G, N*N)): symmetrised h1 per spin,
    rot_h1[s] = C_s^H h1[s], rot_chol[s] = C_s^H L^s_g."""
    ham_data = dict(ham_data)
    h1 = np.array(ham_data["h1"], dtype=np.float64)
    h1 = np.array([(h1[0] + h1[0].T) / 2.0, (h1[1] + h1[1].T) / 2.0])
    ham_data["h1"] = h1
    chol = np.asarray(ham_data["chol"])
    mo = wave_data["mo_coeff"]
    ham_data["rot_h1"] = [mo[0].T.conj() @ h1[0], mo[1].T.conj() @ h1[1]]
    ham_data["rot_chol"] = [
        np.einsum("pi,gij->gpj", mo[s].T.conj(), chol[s].reshape(-1, norb, norb),
                  optimize=True)
        for s in range(2)
    ]
    return ham_data


def build_propagation_intermediates_spin(ham_data, rdm1, norb, dt):
    """prop_unrestricted/propagation.py:461-510: mf_shifts summed over both spins' vectors,
    v0 / v1 / exp_h1 per spin from that spin's vectors."""
    ham_data = dict(ham_data)
    chol = np.asarray(ham_data["chol"])
    mf_a = 1.0j * np.array([np.sum(x.reshape(norb, norb) * rdm1[0]) for x in chol[0]])
    mf_b = 1.0j * np.array([np.sum(x.reshape(norb, norb) * rdm1[1]) for x in chol[1]])
    ham_data["mf_shifts"] = mf_a + mf_b
    ham_data["h0_prop"] = -ham_data["h0"] - np.sum(ham_data["mf_shifts"] ** 2) / 2.0
    v0, v1 = [], []
    for s in range(2):
        c3 = chol[s].reshape(-1, norb, norb)
        v0.append(0.5 * np.einsum("gik,gjk->ij", c3, c3, optimize=True))
        v1.append(np.real(1.0j * np.einsum("g,gik->ik", ham_data["mf_shifts"], c3,
                                           optimize=True)))
    h1_mod = np.asarray(ham_data["h1"]) - np.array([v0[0] + v1[0], v0[1] + v1[1]])
    ham_data["exp_h1"] = np.array(
        [scipy.linalg.expm(-dt * h1_mod[0] / 2.0), scipy.linalg.expm(-dt * h1_mod[1] / 2.0)]
    )
    return ham_data


def apply_trotprop_det(exp_h1, vhs, walker, n_exp_terms):
    """propagation.py:67-92 (propagator._apply_trotprop_det):
    E (w + sum_{n=1}^{K} V^n w / n!) with E applied before and after."""
    walker = exp_h1.dot(walker)
    terms = []
    carry = walker
    for _ in range(1, n_exp_terms):
        carry = vhs.dot(carry)
        terms.append(carry)
    if terms:
        walker = walker + np.sum(
            np.stack([terms[n] / math.factorial(n + 1) for n in range(n_exp_terms - 1)]),
            axis=0,
        )
    walker = exp_h1.dot(walker)
    return walker


def build_vhs(chol, fields, dt, norb):
    """propagation.py:265-271: VHS = i sqrt(dt) fields . chol, reshaped (W, norb, norb)."""
    n_walkers = fields.shape[0]
    return (1.0j * np.sqrt(dt) * fields.dot(chol)).reshape(n_walkers, norb, norb)


def apply_trotprop(walker_type, ham_data, walkers, fields, dt, n_exp_terms, norb):
    """restricted: propagation.py:255-290; unrestricted: propagation.py:359-400;
    "uhf_spin" (spin-resolved vectors chol (2, G, N*N)):
    prop_unrestricted/propagation.py:361-411, one VHS per spin."""
    if walker_type == "uhf_spin":
        chol = np.asarray(ham_data["chol"])
        e = ham_data["exp_h1"]
        out = []
        for s in range(2):
            vhs = build_vhs(chol[s].reshape(-1, norb * norb), fields, dt, norb)
            out.append(np.array(
                [apply_trotprop_det(e[s], vhs[i], walkers[s][i], n_exp_terms)
                 for i in range(walkers[s].shape[0])]))
        return out
    chol = np.asarray(ham_data["chol"]).reshape(-1, norb * norb)
    vhs = build_vhs(chol, fields, dt, norb)
    if walker_type == "rhf":
        return np.array(
            [
                apply_trotprop_det(ham_data["exp_h1"], vhs[i], walkers[i], n_exp_terms)
                for i in range(walkers.shape[0])
            ]
        )
    e = ham_data["exp_h1"]
    return [
        np.array(
            [apply_trotprop_det(e[0], vhs[i], walkers[0][i], n_exp_terms)
             for i in range(walkers[0].shape[0])]
        ),
        np.array(
            [apply_trotprop_det(e[1], vhs[i], walkers[1][i], n_exp_terms)
             for i in range(walkers[1].shape[0])]
        ),
    ]


def propagate(walker_type, trial, ham_data, prop_data, fields, wave_data, dt,
              n_exp_terms, norb, n_walkers=None, return_force_bias=False):
    """propagation.py:100-165 (propagator.propagate), phaseless step; walker_type
    "uhf_spin" is the same step of prop_unrestricted/propagation.py:101-165 over
    spin-resolved Cholesky vectors (trial "uhf" with per-spin rot_chol)."""
    prop_data = dict(prop_data)
    if n_walkers is None:
        n_walkers = fields.shape[0]
    force_bias = calc_force_bias(trial, prop_data["walkers"], ham_data, wave_data)
    field_shifts = -np.sqrt(dt) * (1.0j * force_bias - ham_data["mf_shifts"])
    shifted_fields = fields - field_shifts
    shift_term = np.sum(shifted_fields * ham_data["mf_shifts"], axis=1)
    fb_term = np.sum(fields * field_shifts - field_shifts * field_shifts / 2.0, axis=1)

    prop_data["walkers"] = apply_trotprop(
        walker_type, ham_data, prop_data["walkers"], shifted_fields, dt, n_exp_terms, norb
    )

    overlaps_new = calc_overlap(trial, prop_data["walkers"], wave_data)
    with np.errstate(all="ignore"):
        imp_fun = (
            np.exp(
                -np.sqrt(dt) * shift_term
                + fb_term
                + dt * (prop_data["pop_control_ene_shift"] + ham_data["h0_prop"])
            )
            * overlaps_new
            / prop_data["overlaps"]
        )
        theta = np.angle(
            np.exp(-np.sqrt(dt) * shift_term) * overlaps_new / prop_data["overlaps"]
        )
        imp_fun_phaseless = np.abs(imp_fun) * np.cos(theta)
    imp_fun_phaseless = np.where(np.isnan(imp_fun_phaseless), 0.0, imp_fun_phaseless)
    imp_fun_phaseless = np.where(imp_fun_phaseless < 1.0e-3, 0.0, imp_fun_phaseless)
    imp_fun_phaseless = np.where(imp_fun_phaseless > 100.0, 0.0, imp_fun_phaseless)
    weights = imp_fun_phaseless * prop_data["weights"]
    weights = np.where(weights > 100, 0.0, weights)
    prop_data["weights"] = weights
    with np.errstate(all="ignore"):
        prop_data["pop_control_ene_shift"] = prop_data["e_estimate"] - 0.1 * (
            np.log(np.sum(weights) / n_walkers) / dt
        )
    prop_data["overlaps"] = overlaps_new
    if return_force_bias:
        return prop_data, force_bias
    return prop_data


def init_prop_data(walker_type, trial, wave_data, ham_data, n_walkers, nelec,
                   init_walkers=None):
    """restricted: propagation.py:210-232; unrestricted: propagation.py:333-357."""
    prop_data = {"weights": np.ones(n_walkers)}
    if init_walkers is not None:
        prop_data["walkers"] = init_walkers
    else:
        prop_data["walkers"] = get_init_walkers(
            wave_data["rdm1"], nelec, n_walkers, restricted=(walker_type == "rhf")
        )
    energy_samples = np.real(calc_energy(trial, prop_data["walkers"], ham_data, wave_data))
    e_estimate = np.sum(energy_samples) / n_walkers
    prop_data["e_estimate"] = e_estimate
    prop_data["pop_control_ene_shift"] = e_estimate
    prop_data["overlaps"] = calc_overlap(trial, prop_data["walkers"], wave_data)
    return prop_data


def orthonormalize_walkers(walker_type, walkers):
    """propagation.py:251-253 -> linalg_utils.py:51-55 (qr_vmap); unrestricted:
    propagation.py:422-424 -> linalg_utils.py:58-65.  Reduced QR, keep Q."""
    if walker_type == "rhf":
        return np.array([np.linalg.qr(w)[0] for w in walkers])
    return [
        np.array([np.linalg.qr(w)[0] for w in walkers[0]]),
        np.array([np.linalg.qr(w)[0] for w in walkers[1]]),
    ]


def sr_indices(weights, zeta):
    """sr.py:24-34 index part: sequential cumsum of |w|, z = total (i + zeta) / W,
    searchsorted side='left' (clamped to W-1 as jax gathers clamp)."""
    nwalkers = weights.shape[0]
    cumulative_weights = np.cumsum(np.abs(weights))
    total_weight = cumulative_weights[-1]
    z = total_weight * (np.arange(nwalkers) + zeta) / nwalkers
    idx = np.searchsorted(cumulative_weights, z, side="left")
    return np.minimum(idx, nwalkers - 1).astype(np.int32), total_weight


def sr_indices_global(weights, zeta, size=1):
    """sr.py:66-106 (stochastic_reconfiguration_mpi) index part on the gathered weights
    of `size` equal ranks: z = (i + zeta) / nwalkers / size, new_i =
    searchsorted(cumsum |w|, z * total), one walker at a time."""
    nwalkers = weights.shape[0] // size
    cumulative_weights = np.cumsum(np.abs(weights))
    total_weight = cumulative_weights[-1]
    idx = np.empty(nwalkers * size, dtype=np.int32)
    for i in range(nwalkers * size):
        z = (i + zeta) / nwalkers / size
        idx[i] = np.searchsorted(cumulative_weights, z * total_weight)
    return idx, total_weight


def stochastic_reconfiguration(walker_type, walkers, weights, zeta, global_order=False):
    """sr.py:24-34 (restricted) and sr.py:37-48 (unrestricted); global_order: the
    gather/scatter variant sr.py:66-106 / :109-156 on one rank."""
    nwalkers = weights.shape[0]
    if global_order:
        idx, total_weight = sr_indices_global(weights, zeta)
    else:
        idx, total_weight = sr_indices(weights, zeta)
    new_weights = np.ones(nwalkers) * (total_weight / nwalkers)
    if walker_type == "rhf":
        return walkers[idx], new_weights, idx
    return [walkers[0][idx], walkers[1][idx]], new_weights, idx


# ----------------------------------------------------------------------------------
# sampler (ad_afqmc/sampling.py) with injected fields
# ----------------------------------------------------------------------------------


def block_scan(walker_type, trial, ham_data, prop_data, wave_data, fields_block, dt,
               n_exp_terms, norb):
    """sampling.py:50-92 (_block_scan) with fields (n_prop_steps, W, G) passed in
    instead of drawn from jax.random."""
    prop_data = dict(prop_data)
    n_walkers = fields_block.shape[1]
    for step in range(fields_block.shape[0]):
        prop_data = propagate(walker_type, trial, ham_data, prop_data, fields_block[step],
                              wave_data, dt, n_exp_terms, norb)
    prop_data["n_killed_walkers"] = prop_data.get("n_killed_walkers", 0) + (
        prop_data["weights"].size - np.count_nonzero(prop_data["weights"])
    )
    prop_data["walkers"] = orthonormalize_walkers(walker_type, prop_data["walkers"])
    prop_data["overlaps"] = calc_overlap(trial, prop_data["walkers"], wave_data)
    energy_samples = np.real(calc_energy(trial, prop_data["walkers"], ham_data, wave_data))
    energy_samples = np.where(
        np.abs(energy_samples - prop_data["e_estimate"]) > np.sqrt(2.0 / dt),
        prop_data["e_estimate"],
        energy_samples,
    )
    block_weight = np.sum(prop_data["weights"])
    block_energy = np.sum(energy_samples * prop_data["weights"]) / block_weight
    prop_data["pop_control_ene_shift"] = (
        0.9 * prop_data["pop_control_ene_shift"] + 0.1 * block_energy
    )
    del n_walkers
    return prop_data, block_energy, block_weight


def corr_outlier_rm(w1, w2, z_thresh=5.0):
    """corr_sample.py:551-585 (corr_otler_rm): joint z-score rule on two weight arrays."""
    w1 = np.asarray(w1, dtype=np.float64)
    w2 = np.asarray(w2, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        z1 = np.abs(w1 - np.mean(w1)) / np.std(w1)
        z2 = np.abs(w2 - np.mean(w2)) / np.std(w2)
    z = np.where(z1 > z2, z1, z2)
    return np.where(z > z_thresh, 0.0, w1), np.where(z > z_thresh, 0.0, w2)


def cs_block(systems, fields_block, dt, n_exp_terms, z_thresh=5.0):
    """corr_sample.py:588-618 (cs_block_scan) with the fields passed in, followed by the
    local sums of cs_steps_scan (:667-690).  systems: two dicts with keys walker_type,
    ham_data, prop_data, wave_data, norb; both use fields_block (n_prop_steps, W, G).
    Returns (updated prop_data pair, (loc_en1, loc_w1, loc_en2, loc_w2))."""
    pds = []
    for sy in systems:
        pd, _, _ = block_scan(sy["walker_type"], sy["walker_type"], sy["ham_data"],
                              sy["prop_data"], sy["wave_data"], fields_block, dt, n_exp_terms,
                              sy["norb"])
        pds.append(pd)
    w1, w2 = corr_outlier_rm(pds[0]["weights"], pds[1]["weights"], z_thresh)
    pds[0]["weights"], pds[1]["weights"] = w1, w2
    sums = []
    for sy, pd in zip(systems, pds):
        e = np.real(calc_energy(sy["walker_type"], pd["walkers"], sy["ham_data"], sy["wave_data"]))
        e = np.where(np.abs(e - pd["e_estimate"]) > np.sqrt(2.0 / dt), pd["e_estimate"], e)
        sums += [np.sum(e * pd["weights"]), np.sum(pd["weights"])]
    return pds, tuple(sums)


def propagate_phaseless(walker_type, trial, ham_data, prop_data, wave_data, fields,
                        zetas, dt, n_exp_terms, norb, n_ene_blocks, n_sr_blocks):
    """sampling.py:319-341 (propagate_phaseless) -> :126-144 (_sr_block_scan) ->
    :50-92 (_block_scan).  fields: (n_sr_blocks, n_ene_blocks, n_prop_steps, W, G);
    zetas: (n_sr_blocks,) uniform draws used by each SR (propagation.py:234-241)."""
    prop_data = dict(prop_data)
    n_walkers = prop_data["weights"].shape[0]
    prop_data["overlaps"] = calc_overlap(trial, prop_data["walkers"], wave_data)
    prop_data["n_killed_walkers"] = 0
    prop_data["pop_control_ene_shift"] = prop_data["e_estimate"]
    energies, weights = [], []
    for s in range(n_sr_blocks):
        for b in range(n_ene_blocks):
            prop_data, be, bw = block_scan(walker_type, trial, ham_data, prop_data,
                                           wave_data, fields[s, b], dt, n_exp_terms, norb)
            energies.append(be)
            weights.append(bw)
        prop_data["walkers"], prop_data["weights"], _ = stochastic_reconfiguration(
            walker_type, prop_data["walkers"], prop_data["weights"], zetas[s]
        )
        prop_data["overlaps"] = calc_overlap(trial, prop_data["walkers"], wave_data)
    prop_data["n_killed_walkers"] /= n_sr_blocks * n_ene_blocks * n_walkers
    energies = np.array(energies)
    weights = np.array(weights)
    return np.sum(energies * weights) / np.sum(weights), prop_data, energies, weights
