"""Benchmark: walker-steps/s of the phaseless AFQMC step on one MI355X, mixed precision
against the plain float64 GPU form at the same fields.

Workload (default): synthetic closed-shell RHF, norb=100, nocc=20, nchol=400, 1024
complex128 walkers, dt=0.005; a "step" is one phaseless propagation of all walkers
(force bias, VHS, exp(VHS) Taylor + one-body half steps, overlaps, weights).  Steps
run in blocks of n_prop_steps=50 as sampler.propagate_phaseless does: auxiliary fields
are drawn on the device per block, and each block ends with QR re-orthonormalisation,
overlaps, local energies and block statistics; stochastic reconfiguration follows
every 4 blocks.  Block work is inside the timed region and amortised into the per-step
figure.

Default mode runs the same trajectory twice from the same seed (same device fields,
same SR draws): first in float64, then in the mixed mode (split-bf16 MFMA Cholesky
contractions and Taylor products, float64 overlaps / determinants / QR / weights /
energies).  `value` is the mixed run's walker-steps/s; the line also carries the
float64 rate, the speed-up and the largest block-energy difference |dE| between the
two runs.  `--precision fp64|bf16x3` runs one mode only.

Prints ONE JSON line.  `value` = walker-steps/s over the timed region (inputs already
resident in HBM).  `phases` / `rooflines` come from a profiled pass of up to 50 steps
just before the timed region, with a HIP event pair around every phase (each pair costs
device time); the timed region records events only around the dominant phase, and
`roofline` describes that kernel from those events.  `cpu_baseline` times
the numpy restatement of the reference algorithm (the oracle) on a bounded sample on
the host cores, one single-threaded process per core (the reference's own model: one
MPI rank per core with one intra-op thread).
"""

import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# published MI355X dense peaks
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}
HBM_PEAK_GBS = 8000.0

N_PROP_STEPS = 50
N_ENE_BLOCKS = 4
PROFILE_STEPS = 50  # steps of the all-phase profiled pass before the timed region

def phase_kernel(precision, phase, norb, nelec):
    """Device kernel behind a timed phase (symbol prefix as rocprofv3 lists it), by the
    library's dispatch rules."""
    reg = norb <= 128 and max(nelec) <= 32
    if phase == "taylor":
        if precision == "fp64":
            return "taylor_reg_kernel" if reg else "taylor_kernel"
        return "taylor_bf_kernel" if reg else "taylor_bfs_kernel"
    if phase in ("vhs_gemm", "fb_gemm"):
        return "gemm_nt_f64_kernel" if precision == "fp64" else "gemm_nt_bf16x3_kernel"
    if phase in ("overlap", "green"):
        return "measure_kernel"
    if phase == "energy":
        return "exchange_f64_kernel" if precision == "fp64" else "exchange_bf16x3_kernel"
    return {"fields": "fields_kernel", "weights": "pcs_kernel", "qr": "qr_kernel",
            "sr": "sr_index_kernel", "rng": "normal_kernel"}.get(phase, phase)


def flops_per_walker(norb, nelec, nchol, n_exp_terms, walker_type, nv=None):
    """Algorithmic real flops per walker per phase (real x complex MAC = 4 flop,
    complex x complex MAC = 8 flop).  The measurement kernel ("green" at a block's start,
    "overlap" after each step) forms C^T w (4 N n^2), the inverse by elimination (8 n^3)
    and the Green's function w O^-1 (8 N n^2).  nv: entries of V the VHS product forms
    (N^2, or N(N+1)/2 when V is symmetric and kept as its upper triangle)."""
    K = n_exp_terms - 1
    N, G = norb, nchol
    nv = N * N if nv is None else nv
    ns = [nelec[0]] if walker_type == "rhf" else [nelec[0], nelec[1]]
    f = {
        "vhs_gemm": 4.0 * G * nv,
        "fb_gemm": sum(4.0 * G * N * n for n in ns),
        "green": sum(12.0 * N * n * n + 8.0 * n**3 for n in ns),
        "taylor": sum((8.0 + 8.0 * K) * N * N * n for n in ns),
        "overlap": sum(12.0 * N * n * n + 8.0 * n**3 for n in ns),
        "energy": sum(4.0 * G * N * n * n + 8.0 * G * n * n + 12.0 * N * n * n + 8.0 * n**3
                      for n in ns),
        "qr": sum(16.0 * N * n * n for n in ns),
        "fields": 40.0 * G,  # shifts, shifted fields and the two field sums, complex
        "weights": 1.0,      # the population-control sum
    }
    return f


def bytes_per_walker(norb, nelec, nchol, walker_type, precision, W, nv=None):
    """Algorithmic HBM bytes per walker per phase: each kernel's inputs read once and
    outputs written once in the layouts the library keeps (complex128 walkers; float64
    or split-bf16 Green's function / shifted fields / Cholesky planes; float64 or float32
    VHS), with operands shared by all walkers (Cholesky vectors, trial, one-body factor)
    counted once per launch and divided by W.  nv: stored entries of one V plane."""
    N, G = norb, nchol
    nv = N * N if nv is None else nv
    ns = [nelec[0]] if walker_type == "rhf" else [nelec[0], nelec[1]]
    nsum, nspin = sum(ns), len(ns)
    mixed = precision != "fp64"
    Np32, Gp = -(-N // 32) * 32, -(-G // 32) * 32
    walker = 16.0 * N * nsum
    green = 8.0 * nsum * Np32 if mixed else 16.0 * nsum * N     # hi/lo bf16 or float64
    vhs = 8.0 * nv if mixed else 16.0 * nv                       # float32 or float64 re/im
    xs = 8.0 * Gp if mixed else 16.0 * G
    chol = (4.0 * nv * Gp if mixed else 8.0 * nv * G) / W
    rot = (4.0 * G * nsum * Np32 if mixed else 8.0 * G * nsum * N) / W
    one_body = nspin * (4.0 * N * Np32 if mixed else 8.0 * N * N) / W
    trial = 8.0 * N * nsum / W
    tail = 32.0 + 32.0 + 16.0  # terms read, overlap and weight read + written
    return {
        "vhs_gemm": xs + chol + vhs,
        "fb_gemm": green + rot + 16.0 * G,
        "fields": 16.0 * G + 8.0 * G + xs + 32.0,
        "taylor": 2.0 * walker + vhs + one_body,
        "green": walker + trial + green + 16.0 * nspin,
        "overlap": walker + trial + green + 16.0 * nspin + tail,
        "weights": 8.0,
        "qr": 2.0 * walker,
        "energy": walker + trial + green + rot,
    }


def roofline_of(phase, precision, norb, nelec, nchol, n_exp_terms, walker_type, W, avg_s,
                nv=None):
    """Roofline record of one phase's kernel: the bound is whichever of MFMA time and
    HBM time at published peaks is longer for the algorithmic flops and bytes of one
    launch; achieved = those flops (or bytes) / the measured average launch time."""
    fl = flops_per_walker(norb, nelec, nchol, n_exp_terms, walker_type, nv).get(phase)
    by = bytes_per_walker(norb, nelec, nchol, walker_type, precision, W, nv).get(phase)
    if fl is None or by is None:
        return None
    flop, byt = fl * W, by * W
    peak_tf = phase_peak(precision, phase, norb, nelec)
    mfma_bound = flop / (peak_tf * 1e12) >= byt / (HBM_PEAK_GBS * 1e9)
    if mfma_bound:
        achieved, peak, unit = flop / avg_s / 1e12, peak_tf, "TFLOP/s"
    else:
        achieved, peak, unit = byt / avg_s / 1e9, HBM_PEAK_GBS, "GB/s"
    return {
        "phase": phase,
        "kernel": phase_kernel(precision, phase, norb, nelec),
        "bound": "mfma" if mfma_bound else "hbm",
        "achieved": round(achieved, 3),
        "peak": peak,
        "unit": unit,
        "frac": round(achieved / peak, 4),
        "flop_per_launch": flop,
        "bytes_per_launch": byt,
        "avg_launch_ms": round(avg_s * 1e3, 5),
    }


def phase_peak(precision, phase, norb, nelec):
    """Published peak of the arithmetic a phase runs on."""
    if precision == "fp64":
        return PEAK_TFLOPS["fp64"]
    if phase in ("vhs_gemm", "fb_gemm", "energy"):  # split-bf16 contractions
        return PEAK_TFLOPS["bf16"]
    if phase == "taylor":  # split-bf16 register or streaming kernel
        return PEAK_TFLOPS["bf16"]
    return PEAK_TFLOPS["fp64"]


def traffic_from_profiles(kernel, workload, precision):
    """HBM bytes per launch of `kernel` from the PMC table in profiles/ (FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by the copy-kernel factor), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 7 and cells[0] == kernel and cells[1] == workload \
                    and cells[2] == precision:
                try:
                    return float(cells[6])
                except ValueError:
                    return None
    return None


# ------------------------------------------------------------------ CPU baseline


def _cpu_worker(args):
    """One single-threaded process: oracle steps on a slice of the workload."""
    (walker_type, norb, nocc, nchol, n_walk, budget_s, seed, dt) = args
    import numpy as np

    from cs_afqmc_amd import problems
    from oracle import afqmc_oracle as o

    if walker_type == "rhf":
        hd, wd, nelec = problems.synthetic_rhf(norb, nocc, nchol, seed=seed)
    else:
        hd, wd, nelec = problems.synthetic_uhf(norb, nocc, nocc, nchol, seed=seed)
    oh = o.build_measurement_intermediates(walker_type, hd, wd, norb)
    oh = o.build_propagation_intermediates(walker_type, oh, wd["rdm1"], norb, nelec, dt)
    rng = np.random.default_rng(seed + 17)
    walkers = o.get_init_walkers(wd["rdm1"], nelec, n_walk, restricted=walker_type == "rhf")
    prop = {"walkers": walkers, "weights": np.ones(n_walk), "e_estimate": 0.0,
            "pop_control_ene_shift": 0.0,
            "overlaps": o.calc_overlap(walker_type, walkers, wd)}
    steps = 0
    t_prop = 0.0
    while True:
        t0 = time.perf_counter()
        fields = rng.standard_normal((n_walk, nchol))
        prop = o.propagate(walker_type, walker_type, oh, prop, fields, wd, dt, 6, norb,
                           n_walkers=n_walk)
        t_prop += time.perf_counter() - t0
        steps += 1
        if steps % N_PROP_STEPS == 0:  # keep the walkers orthonormal as the blocks do
            prop["walkers"] = o.orthonormalize_walkers(walker_type, prop["walkers"])
            prop["overlaps"] = o.calc_overlap(walker_type, prop["walkers"], wd)
        if t_prop > budget_s:
            break
    t_step = t_prop / steps
    t1 = time.perf_counter()
    prop["walkers"] = o.orthonormalize_walkers(walker_type, prop["walkers"])
    o.calc_overlap(walker_type, prop["walkers"], wd)
    o.calc_energy(walker_type, prop["walkers"], oh, wd)
    t_block = time.perf_counter() - t1
    return n_walk, steps, t_step, t_block


def cpu_baseline(workload, budget_s):
    walker_type, norb, nocc, nchol, W = workload
    cores = len(os.sched_getaffinity(0))
    procs = max(1, min(cores, 16))
    n_walk = max(1, min(16, W // procs))
    env_before = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS",
                                                 "MKL_NUM_THREADS")}
    for k in env_before:
        os.environ[k] = "1"
    try:
        ctx = mp.get_context("spawn")
        with ctx.Pool(procs) as pool:
            res = pool.map(_cpu_worker, [(walker_type, norb, nocc, nchol, n_walk, budget_s,
                                          11 + i, 0.005) for i in range(procs)])
    finally:
        for k, v in env_before.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    # per-process walker-steps/s with block work (QR, overlap, energy) amortised over
    # N_PROP_STEPS steps; processes run concurrently, so rates add
    rate = 0.0
    for n_w, steps, t_step, t_block in res:
        rate += n_w / (t_step + t_block / N_PROP_STEPS)
    steps_done = sum(r[1] for r in res)
    return {
        "value": round(rate, 1),
        "unit": "walker-steps/s",
        "cores": procs,
        "kind": "port",
        "sample": (f"{procs} single-threaded processes x {n_walk} walkers of the same workload, "
                   f"{steps_done} oracle steps total (~{budget_s:.0f} s each) + one "
                   f"QR/overlap/energy per process amortised over {N_PROP_STEPS} steps"),
    }


# ------------------------------------------------------------------ GPU run


def copy_probe(nbytes=1 << 30):
    """A plain device copy of known size, for scaling PMC byte counters."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def run(args, precision):
    import numpy as np
    import torch

    from cs_afqmc_amd import hamiltonian, problems, propagation, rng, sampling, wavefunctions

    workload = problems.WORKLOADS[args.workload]
    walker_type, norb, nocc, nchol, W = workload
    dt = 0.005
    torch.cuda.set_device(0)
    if walker_type == "rhf":
        hd, wd, nelec = problems.synthetic_rhf(norb, nocc, nchol, seed=0)
        trial = wavefunctions.rhf(norb, nelec)
        prop = propagation.propagator_restricted(dt=dt, n_walkers=W, precision=precision)
    else:
        hd, wd, nelec = problems.synthetic_uhf(norb, nocc, nocc, nchol, seed=0)
        trial = wavefunctions.uhf(norb, nelec)
        prop = propagation.propagator_unrestricted(dt=dt, n_walkers=W, precision=precision)
    ham = hamiltonian.hamiltonian(norb)
    ham_data = ham.build_measurement_intermediates(hd, trial, wd)
    ham_data = ham.build_propagation_intermediates(ham_data, prop, trial, wd)
    prop_data = prop.init_prop_data(trial, wd, ham_data)
    prop_data["key"] = rng.PRNGKey(0)
    dev = ham_data["_afqmc_device"]

    # device state as in sampler.propagate_phaseless
    ddev = dev.device
    e0 = prop_data["e_estimate"].reshape(())
    scalars = torch.stack([e0, e0]).contiguous()
    prop_data["_scalars"] = scalars
    prop_data["_energies"] = torch.empty(W, dtype=torch.complex128, device=ddev)
    fields_buf = torch.empty((N_PROP_STEPS, W, nchol), dtype=torch.float64, device=ddev)
    total_steps = args.warmup + args.steps
    stats = torch.zeros((total_steps // N_PROP_STEPS + 1, 3), dtype=torch.float64,
                        device=ddev)

    state = {"block": 0, "in_block": 0, "blocks_since_sr": 0, "fields": None}

    def one_step():
        """Advance one propagation step; block and SR work happen at the same points
        as in sampler.propagate_phaseless."""
        if state["in_block"] == 0:
            prop_data["key"], subkey = rng.split(prop_data["key"])
            state["fields"] = dev.fields_normal(subkey[0], subkey[1], fields_buf.shape,
                                                out=fields_buf)
        dev.propagate(prop_data["walkers"], prop_data["weights"], prop_data["overlaps"],
                      state["fields"][state["in_block"]], scalars,
                      reuse_green=state["in_block"] > 0)
        state["in_block"] += 1
        if state["in_block"] == N_PROP_STEPS:
            state["in_block"] = 0
            dev.orthonormalize(prop_data["walkers"])
            dev.overlap(prop_data["walkers"], out=prop_data["overlaps"])
            energies = dev.energy(prop_data["walkers"], out=prop_data["_energies"])
            dev.block_stats(energies, prop_data["weights"], scalars, stats[state["block"]])
            state["block"] += 1
            state["blocks_since_sr"] += 1
            if state["blocks_since_sr"] == N_ENE_BLOCKS:
                state["blocks_since_sr"] = 0
                prop.stochastic_reconfiguration_local(prop_data)
                dev.overlap(prop_data["walkers"], out=prop_data["overlaps"])

    for _ in range(args.warmup):
        one_step()
    torch.cuda.synchronize()
    print(f"# {precision}: warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
    if args.copy_probe:
        copy_probe()

    # phase breakdown: a short pass with an event pair around every phase (the pairs
    # cost device time, so the timed region below records only the dominant phase)
    nv = dev.vhs_entries()  # V kept whole or as its upper triangle
    fl = flops_per_walker(norb, nelec, nchol, prop.n_exp_terms, walker_type, nv)
    dev.prof_collect()  # drop anything recorded before
    dev.prof_enable(True)
    for _ in range(min(args.steps, PROFILE_STEPS)):
        one_step()
    torch.cuda.synchronize()
    dev.prof_enable(False)
    phases = dev.prof_collect()
    # the population-control sum ("weights") is one workgroup reducing W doubles; its
    # event pair directly follows the measurement kernel's and has been seen to take
    # that kernel's start as its own, so it is not a dominant-kernel candidate
    dom = max((p for p, (ms, n) in phases.items() if n and p in fl and p != "weights"),
              key=lambda p: phases[p][0])

    dev.prof_enable(True, phases=[dom])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        one_step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    dev.prof_enable(False)
    timed = dev.prof_collect()
    dev.prof_enable(False, phases=None)
    print(f"# {precision}: timed {args.steps} steps in {elapsed:.3f} s", file=sys.stderr,
          flush=True)

    nb = state["block"]
    st = stats[:nb].cpu().numpy() if nb else np.zeros((0, 3))
    e_mixed = float(np.sum(st[:, 0] * st[:, 1]) / np.sum(st[:, 1])) if nb else float("nan")

    value = W * args.steps / elapsed
    per_phase = {p: {"ms_total": round(ms, 4), "launches": n,
                     "ms_avg": round(ms / n, 5) if n else None}
                 for p, (ms, n) in phases.items() if n}
    # every phase's roofline from the profiled pass; the dominant kernel's from its
    # events over the timed region
    rooflines = {}
    for p, (ms, n) in phases.items():
        if n:
            r = roofline_of(p, precision, norb, nelec, nchol, prop.n_exp_terms, walker_type, W,
                            ms / n / 1e3, nv)
            if r is not None:
                rooflines[p] = r
    ms_dom, n_dom = timed[dom]
    roofline = roofline_of(dom, precision, norb, nelec, nchol, prop.n_exp_terms, walker_type,
                           W, ms_dom / n_dom / 1e3, nv)
    roofline["launches_timed"] = n_dom
    roofline["traffic"] = traffic_from_profiles(roofline["kernel"], args.workload, precision)
    step_flops = sum(v for k, v in fl.items() if k not in ("energy", "qr")) + (
        fl["energy"] + fl["qr"]) / N_PROP_STEPS
    res = {
        "value": value,
        "elapsed": elapsed,
        "block_energies": st[:, 0].tolist(),
        "block_weights": st[:, 1].tolist(),
        "mixed_energy": e_mixed,
        "phases": per_phase,
        "roofline": roofline,
        "rooflines": {p: {k: r[k] for k in ("kernel", "bound", "achieved", "unit", "frac")}
                      for p, r in rooflines.items()},
        "achieved_tflops_step": round(step_flops * value / 1e12, 3),
        "config": {
            "workload": args.workload,
            "walker_type": walker_type, "norb": norb, "nocc": nocc, "nchol": nchol,
            "n_walkers": W, "dt": dt, "n_exp_terms": prop.n_exp_terms,
            "n_prop_steps": N_PROP_STEPS, "n_ene_blocks": N_ENE_BLOCKS,
        },
    }
    del dev, ham_data, prop_data, fields_buf
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return res, workload


# ------------------------------------------------------------------ fragment batch


FRAG_WORKLOAD = "fragments64_n20-120_w256"


def _make_fragments(specs, precision, n_prop_steps):
    from cs_afqmc_amd import batch, problems
    frags = []
    for norb, nocc, nchol, W, seed in specs:
        hd, wd, ne = problems.synthetic_rhf(norb, nocc, nchol, seed=seed)
        frags.append(batch.Fragment(hd, wd, "rhf", ne, W, dt=0.005, seed=seed,
                                    precision=precision, n_prop_steps=n_prop_steps,
                                    tag=(norb, nocc, nchol, W)))
    return frags


def _single_block(f):
    """The plain (ungrouped) form: one fragment's block through afqmc_propagate."""
    pd = f.prop_data
    f.begin_block()
    for s in range(f.n_prop_steps):
        f.dev.propagate(pd["walkers"], pd["weights"], pd["overlaps"], f.fields[s], f.scalars,
                        reuse_green=s > 0)
    f.end_block()


def run_fragments(args):
    """Batched-fragment workload: 64 closed-shell fragments, norb uniform in [20, 120], nocc =
    norb / 5, nchol = 4 norb, 256 walkers each.  value = walker-steps/s of the grouped
    mixed-precision step; fp64_value = the plain float64 form (each fragment through
    afqmc_propagate in turn); also the grouped float64 rate."""
    import numpy as np
    import torch

    from cs_afqmc_amd import batch, problems

    torch.cuda.set_device(0)
    n_prop = N_PROP_STEPS
    blocks = max(1, args.steps // n_prop)
    wblocks = max(1, args.warmup // n_prop)
    specs = problems.fragment_specs(64, seed=0)
    total_w = sum(s[3] for s in specs)
    res = {}
    energies = {}
    for mode in ("bf16x3", "fp64", "fp64_plain"):
        prec = "fp64" if mode.startswith("fp64") else mode
        frags = _make_fragments(specs, prec, n_prop)
        bt = None if mode == "fp64_plain" else batch.FragmentBatch(frags, tune=True,
                                                                   use_graph=True)
        for _ in range(wblocks):
            if bt is None:
                for f in frags:
                    _single_block(f)
            else:
                bt.run_block()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(blocks):
            if bt is None:
                for f in frags:
                    _single_block(f)
            else:
                bt.run_block()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[mode] = total_w * n_prop * blocks / el
        energies[mode] = [f.block_energies()[0] for f in frags]
        print(f"# {mode}: {blocks} blocks in {el:.3f} s", file=sys.stderr, flush=True)
        if bt is not None and bt.tuned:
            res[mode + "_tuned"] = bt.tuned
        del frags, bt
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
    # |dE| of the first block (before the first resampling: the precision's own effect)
    # and of every block (later blocks can also differ by a discrete event, a walker's
    # weight clipped or resampled in one precision and not the other)
    d1 = [float(abs(a[0] - b[0])) for a, b in zip(energies["bf16x3"], energies["fp64_plain"])
          if a.size]
    d = [float(np.max(np.abs(a - b))) for a, b in zip(energies["bf16x3"], energies["fp64_plain"])
         if a.size]
    dg = [float(np.max(np.abs(a - b))) for a, b in zip(energies["fp64"], energies["fp64_plain"])
          if a.size]
    finite = [bool(np.all(np.isfinite(a))) for a in energies["bf16x3"]]
    # per fragment: block-by-block |dE| and the energy scale, largest first
    per = sorted(((float(np.max(np.abs(a - b))), i) for i, (a, b) in
                  enumerate(zip(energies["bf16x3"], energies["fp64_plain"])) if a.size),
                 reverse=True)
    worst = [{"fragment": i, "norb": specs[i][0], "n_walkers": specs[i][3],
              "delta_e_blocks_eh": np.abs(energies["bf16x3"][i] - energies["fp64_plain"][i]).tolist(),
              "energy_scale_eh": float(np.max(np.abs(energies["fp64_plain"][i])))}
             for _, i in per[:4]]
    return {
        "metric": "walker-steps/s mixed vs FP64 GPU at |dE|<0.1 mEh (batched fragments)",
        "value": round(res["bf16x3"], 1),
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": blocks * n_prop,
        "warmup": wblocks * n_prop,
        "ms_per_step": round(total_w / res["bf16x3"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE["bf16x3"],
        "data": "synthetic fragments (seeded Cholesky vectors, h1, Haar trial; device Philox "
                "fields)",
        "config": {"workload": FRAG_WORKLOAD, "n_fragments": len(specs),
                   "norb": [s[0] for s in specs], "n_walkers_total": total_w,
                   "n_prop_steps": n_prop, "dt": 0.005, "precision": "bf16x3"},
        "fp64_value": round(res["fp64_plain"], 1),
        "fp64_grouped_value": round(res["fp64"], 1),
        "speedup_vs_fp64": round(res["bf16x3"] / res["fp64_plain"], 3),
        "delta_e_max_eh": max(d1) if d1 else None,
        "delta_e_max_all_blocks_eh": max(d) if d else None,
        "delta_e_fp64_grouped_vs_plain_eh": max(dg) if dg else None,
        "delta_e_worst_fragments": worst,
        "fragments_finite": sum(finite),
        "gemm_tuning_ms": {k: v for k, v in res.items() if k.endswith("_tuned")},
        "roofline": None,
        "cpu_baseline": None,
    }


# ------------------------------------------------------------------ serving


SERVE_WORKLOAD = "serve_poisson200_n20-250"
# shortest estimated solve first among queued requests, at most 8 fragments stepping
# together (measured on this workload: p50 0.69 s, p99 1.84 s against 1.68 s / 2.03 s
# for arrival order with every request admitted at once)
SERVE_POLICY, SERVE_MAX_ACTIVE = "sjf", 8


def run_serving(args):
    """Serving workload: Poisson arrivals (mean 200 requests/s) of fragment solves with
    norb uniform in [20, 250], nocc = norb / 5, nchol = 4 norb, walkers in {128, 256,
    512, 1024}, 100 steps each; a bounded sample of --requests requests.  Integrals are
    the requests' payload (made before the clock starts); intermediates, context
    creation and the solve are the server's work.  value = walker-steps/s of the whole
    run (completed walker-steps / makespan); p50 / p99 latency from arrival to GPU
    completion."""
    import numpy as np
    import torch

    from cs_afqmc_amd import problems, serving

    torch.cuda.set_device(0)
    rng = np.random.default_rng(2024)
    t = 0.0
    reqs = []
    for i in range(args.requests):
        t += rng.exponential(1.0 / 200.0)
        norb = int(rng.integers(20, 251))
        nocc = max(1, norb // 5)
        W = int(rng.choice([128, 256, 512, 1024]))
        hd, wd, ne = problems.synthetic_rhf(norb, nocc, 4 * norb, seed=5000 + i)
        reqs.append(serving.Request(arrival=t, rid=i, ham_data=hd, wave_data=wd, nelec=ne,
                                    n_walkers=W, n_steps=100, seed=17 + i))
    print(f"# serving: {len(reqs)} requests generated", file=sys.stderr, flush=True)
    srv = serving.FragmentServer(precision="bf16x3", n_prop_steps=N_PROP_STEPS,
                                 hbm_budget=200 << 30, max_active=SERVE_MAX_ACTIVE,
                                 policy=SERVE_POLICY)
    t0 = time.perf_counter()
    res = srv.serve(reqs)
    wall = time.perf_counter() - t0
    ws = sum(r.walker_steps for r in res)
    makespan = max(r.done for r in res)
    st = serving.latency_stats(res)
    return {
        "metric": "p50/p99 fragment latency under load, with walker-steps/s (serving)",
        "value": round(ws / makespan, 1),
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": 100,
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE["bf16x3"],
        "data": "synthetic fragment requests (seeded integrals, Haar trial), Poisson arrivals",
        "config": {"workload": SERVE_WORKLOAD, "n_requests": len(reqs),
                   "arrival_rate_per_s": 200.0, "n_prop_steps": N_PROP_STEPS,
                   "steps_per_request": 100, "hbm_budget_gib": 200, "precision": "bf16x3",
                   "admission": SERVE_POLICY, "max_active_fragments": SERVE_MAX_ACTIVE},
        "latency_p50_s": st["p50_s"],
        "latency_p99_s": st["p99_s"],
        "latency_mean_s": st["mean_s"],
        "latency_max_s": st.get("max_s"),
        "makespan_s": makespan,
        "wall_s": wall,
        "build_s_mean": float(np.mean([r.build_s for r in res])),
        "build_s_max": float(np.max([r.build_s for r in res])),
        "admit_to_done_s_mean": float(np.mean([r.done - r.admitted for r in res])),
        "arrival_to_admit_s_mean": float(np.mean([r.admitted - r.arrival for r in res])),
        "peak_active_fragments": srv.peak_active,
        "peak_reserved_fragments": srv.peak_reserved,
        "completed": len(res),
        "roofline": None,
        "cpu_baseline": None,
    }


DTYPE = {"fp64": "f64",
         "bf16x3": "mixed: bf16x3 MFMA contractions and Taylor products (f32 accumulation), "
                   "f64 overlaps/QR/weights/energy accumulation"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="rhf_n100_o20_g400_w1024")
    ap.add_argument("--precision", default="compare", choices=["compare", "fp64", "bf16x3"],
                    help="compare: float64 then mixed at the same fields (default)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--requests", type=int, default=32,
                    help="number of requests of the serving workload")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device copy before the timed region (PMC scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the walker population runs on one GPU; --gpus {args.gpus} "
              "is not supported (sharding across GPUs is not implemented).", file=sys.stderr)
        sys.exit(2)
    if args.workload == FRAG_WORKLOAD:
        print(json.dumps(run_fragments(args)), flush=True)
        return
    if args.workload == SERVE_WORKLOAD:
        print(json.dumps(run_serving(args)), flush=True)
        return
    modes = ["fp64", "bf16x3"] if args.precision == "compare" else [args.precision]
    cpu = None
    if not args.no_cpu_baseline:
        # before anything touches the GPU: the worker processes are started by spawn
        from cs_afqmc_amd import problems
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(problems.WORKLOADS[args.workload], args.cpu_seconds)
    runs = {}
    for m in modes:
        runs[m], _ = run(args, m)
    main_mode = modes[-1]
    r = runs[main_mode]
    out = {
        "metric": "walker-steps/s mixed vs FP64 GPU at |dE|<0.1 mEh (phaseless AFQMC step)",
        "value": round(r["value"], 1),
        "unit": "walker-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["elapsed"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[main_mode],
        "data": "synthetic (seeded Cholesky vectors, h1, Haar trial; device Philox fields)",
        "config": dict(r["config"], precision=main_mode),
    }
    if "fp64" in runs and "bf16x3" in runs:
        b64 = runs["fp64"]["block_energies"]
        bmx = runs["bf16x3"]["block_energies"]
        d = [abs(a - b) for a, b in zip(bmx, b64)]
        out["fp64_value"] = round(runs["fp64"]["value"], 1)
        out["speedup_vs_fp64"] = round(runs["bf16x3"]["value"] / runs["fp64"]["value"], 3)
        out["delta_e_max_eh"] = max(d) if d else None
        out["delta_e_mean_eh"] = abs(runs["bf16x3"]["mixed_energy"] - runs["fp64"]["mixed_energy"])
        out["fp64_phases"] = runs["fp64"]["phases"]
        out["fp64_roofline"] = runs["fp64"]["roofline"]
    out["mixed_energy"] = r["mixed_energy"]
    out["block_energies"] = r["block_energies"]
    out["achieved_tflops_step"] = r["achieved_tflops_step"]
    out["phases"] = r["phases"]
    out["roofline"] = r["roofline"]
    out["rooflines"] = r["rooflines"]
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
