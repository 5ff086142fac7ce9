"""Algorithmic work of the step's phases: flops and HBM bytes per walker.

These are the per-unit figures the measurement tools multiply by the walkers of a launch
to price a timed kernel against the chip's published peaks (real x complex MAC = 4 flop,
complex x complex MAC = 8 flop; bytes = each kernel's inputs read once and outputs
written once in the layouts the library keeps).
"""


def flops_per_walker(norb, nelec, nchol, n_exp_terms, walker_type, nv=None, spin_chol=False):
    """Algorithmic real flops per walker per phase (real x complex MAC = 4 flop,
    complex x complex MAC = 8 flop).  The measurement kernel ("green" at a block's start,
    "overlap" after each step) forms C^T w (4 N n^2), the inverse by elimination (8 n^3)
    and the Green's function w O^-1 (8 N n^2).  nv: entries of V the VHS product forms
    (N^2, or N(N+1)/2 when V is symmetric and kept as its upper triangle).  spin_chol:
    one set of Cholesky vectors, and so one V, per spin."""
    K = n_exp_terms - 1
    N, G = norb, nchol
    nv = N * N if nv is None else nv
    ns = [nelec[0]] if walker_type == "rhf" else [nelec[0], nelec[1]]
    return {
        "vhs_gemm": (2.0 if spin_chol else 1.0) * 4.0 * G * nv,
        "fb_gemm": sum(4.0 * G * N * n for n in ns),
        "green": sum(12.0 * N * n * n + 8.0 * n**3 for n in ns),
        "taylor": sum((8.0 + 8.0 * K) * N * N * n for n in ns),
        "overlap": sum(12.0 * N * n * n + 8.0 * n**3 for n in ns),
        "energy": sum(4.0 * G * N * n * n + 8.0 * G * n * n + 12.0 * N * n * n + 8.0 * n**3
                      for n in ns),
        "qr": sum(16.0 * N * n * n for n in ns),
        "fields": 40.0 * G,
        "weights": 1.0,
    }


def bytes_per_walker(norb, nelec, nchol, walker_type, precision, W, nv=None,
                     spin_chol=False):
    """Algorithmic HBM bytes per walker per phase: each kernel's inputs read once and
    outputs written once in the layouts the library keeps, with operands shared by all
    walkers (Cholesky vectors, trial, one-body factor) counted once per launch and
    divided by W.  nv: stored entries of one V plane.  spin_chol: one set of Cholesky
    vectors and one V per spin."""
    N, G = norb, nchol
    nv = N * N if nv is None else nv
    ns = [nelec[0]] if walker_type == "rhf" else [nelec[0], nelec[1]]
    nsum, nspin = sum(ns), len(ns)
    bf = precision == "bf16x3"
    Np32, Gp = -(-N // 32) * 32, -(-G // 32) * 32
    walker = 16.0 * N * nsum
    if bf:
        green, vhs, xs = 8.0 * nsum * Np32, 8.0 * nv, 8.0 * Gp
        chol, rot = 4.0 * nv * Gp / W, 4.0 * G * nsum * Np32 / W
        one_body = nspin * 4.0 * N * Np32 / W
    else:  # float64 operands (the FP32 mode rounds them to float32 on chip)
        green, vhs, xs = 16.0 * nsum * N, 16.0 * nv, 16.0 * G
        chol, rot = 8.0 * nv * G / W, 8.0 * G * nsum * N / W
        one_body = nspin * 8.0 * N * N / W
    trial = 8.0 * N * nsum / W
    if spin_chol:
        chol, vhs = 2.0 * chol, 2.0 * vhs
    tail = 32.0 + 32.0 + 16.0
    return {
        "vhs_gemm": (2.0 if spin_chol else 1.0) * xs + chol + vhs,
        "fb_gemm": green + rot + 16.0 * G,
        "fields": 16.0 * G + 8.0 * G + xs + 32.0,
        "taylor": 2.0 * walker + vhs + one_body,
        "green": walker + trial + green + 16.0 * nspin,
        "overlap": walker + trial + green + 16.0 * nspin + tail,
        "weights": 8.0,
        "qr": 2.0 * walker,
        "energy": walker + trial + green + rot,
    }


def phase_work(precision, members):
    """Algorithmic flops and bytes per phase for one step of a set of problems.
    members: (walker_type, norb, nelec, nchol, W, nv, n_exp_terms[, spin_chol]) tuples."""
    fl, by = {}, {}
    for m in members:
        wt, norb, nelec, nchol, W, nv, K = m[:7]
        spin = bool(m[7]) if len(m) > 7 else False
        f = flops_per_walker(norb, nelec, nchol, K, wt, nv, spin)
        b = bytes_per_walker(norb, nelec, nchol, wt, precision, W, nv, spin)
        for p in f:
            fl[p] = fl.get(p, 0.0) + f[p] * W
            by[p] = by.get(p, 0.0) + b[p] * W
    return fl, by
