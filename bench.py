"""Benchmark: walker-steps/s of the mixed-precision batched AFQMC step against the plain
float64 GPU form on one MI355X, with the energy difference between the two and the
p50 / p99 latency of fragment solves served under load.

Default run (``python bench.py``, also under ``--steps K --warmup W``), one JSON line:

* headline, ``fragments64_n20-120_w256``: 64 synthetic closed-shell fragments, norb
  uniform in [20, 120], nocc = norb / 5, nchol = 4 norb, 256 walkers each, dt = 0.005.
  ``value`` = walker-steps/s of the batched mixed-precision form (every fragment's
  block -- device field draws, 50 grouped steps, QR, overlaps, local energies, block
  statistics, stochastic reconfiguration, e_estimate update -- replayed from one
  captured HIP graph); ``fp64_value`` = the plain float64 form (each fragment's block
  through its own context, one fragment after another, each block a captured graph).
* ``single``: the single-problem synthetic RHF workload (norb 100, nocc 20, nchol 400,
  1024 walkers), float64 and mixed, graph-captured blocks.
* ``serving``: a Poisson stream (mean 200 requests/s) of fragment solves, norb uniform in
  [20, 250], walkers in {128, 256, 512, 1024}, 100 steps each: p50 / p99 latency.
* ``h10``: the linear H10 chain (1.6 bohr, STO-6G, RHF trial, modified Cholesky 1e-5),
  200 walkers, dt 0.005, QR every 5 steps: mixed against float64 block energies.

Timing.  Steps run in whole blocks of 50 (``--steps`` / ``--warmup`` are rounded up to
whole blocks): the warm-up blocks include the eager first block and the graph capture;
then each of 5 repeats replays ceil(steps / 50) blocks between two HIP events on the
stream the blocks run on, and the median repeat gives the rate.  A block's QR, energy,
statistics and resampling are inside the timed region, amortised into the per-step
figure.

Energy difference.  The mixed and float64 runs start from the same state with the same
seeds, so they draw the same fields and resampling uniforms.  |dE| is taken per block
only where both runs' block energies are finite and fewer than 1% of either run's
walkers had their local energy clamped (|E_L - e_estimate| > sqrt(2 / dt)); clamp and
zero-weight fractions are reported beside it, and non-finite blocks are counted and
flagged.

``roofline``: the dominant kernel phase of the headline form, timed with HIP event
pairs around that phase only, over one eager block after the timed repeats (the pairs
cost device time, so the repeats run without them).  ``cpu_baseline``: the numpy
restatement of the reference algorithm (the oracle) on a bounded sample of the same
fragments, one single-threaded process per host core (the reference's one rank per
core with one intra-op thread).
"""

import argparse
import json
import math
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
REPEATS = 5
DT = 0.005
CPU_PROCS_MAX = 16  # the host-core share of one GPU on the measurement boxes

FRAG_WORKLOAD = "fragments64_n20-120_w256"
SINGLE_WORKLOADS = {
    "rhf_n100_o20_g400_w1024": ("rhf", 100, 20, 400, 1024),
    "uhf_n200_o40_g1000_w2048": ("uhf", 200, 40, 1000, 2048),
}
SERVE_WORKLOAD = "serve_poisson200_n20-250"
H10_WORKLOAD = "h10_chain_r1.6_sto6g_w200"
H10_CALLS = 40  # sampler calls of 50 steps (10 blocks) per precision
# admission: shortest estimated solve first among queued requests, at most 8 fragments
# stepping together
SERVE_POLICY, SERVE_MAX_ACTIVE = "sjf", 8

from cs_afqmc_amd.work import bytes_per_walker, flops_per_walker, phase_work  # noqa: E402

DTYPE = {"fp64": "f64",
         "fp32": "mixed: f32 MFMA Cholesky contractions (f32 accumulation), f64 Taylor, "
                 "overlaps, QR, weights, energy accumulation",
         "bf16x3": "mixed: split-bf16 (bf16x3) MFMA contractions and Taylor products "
                   "(f32 accumulation), f64 overlaps/QR/weights/energy accumulation"}


def blocks_for(steps):
    return max(1, -(-int(steps) // N_PROP_STEPS))


# ------------------------------------------------------------------ algorithmic work


def phase_kernel(precision, phase, norb, nelec):
    """Device kernel behind a timed phase (symbol prefix as rocprofv3 lists it)."""
    reg = norb <= 128 and max(nelec) <= 32
    if phase == "taylor":
        if precision in ("fp64", "fp32"):
            return "taylor_reg_kernel" if reg else "taylor_kernel"
        return "taylor_bf_kernel" if reg else "taylor_bfs_kernel"
    if phase in ("vhs_gemm", "fb_gemm"):
        return {"fp64": "gemm_nt_f64_kernel", "fp32": "gemm_nt_f32_kernel"}.get(
            precision, "gemm_nt_bf16x3_kernel")
    if phase in ("overlap", "green"):
        return "measure_kernel"
    if phase == "energy":
        return {"fp64": "exchange_f64_kernel", "fp32": "exchange_f64_kernel<F32>"}.get(
            precision, "exchange_bf16x3_kernel")
    return {"fields": "fields_kernel", "weights": "pcs_kernel", "qr": "qr_kernel",
            "sr": "sr_index_kernel", "rng": "normal_key_kernel"}.get(phase, phase)


def phase_peak(precision, phase):
    """Published peak of the arithmetic a phase runs on."""
    if precision == "fp64":
        return PEAK_TFLOPS["fp64"]
    if phase in ("vhs_gemm", "fb_gemm", "energy") or (phase == "taylor" and
                                                      precision == "bf16x3"):
        return PEAK_TFLOPS["bf16" if precision == "bf16x3" else "fp32"]
    return PEAK_TFLOPS["fp64"]


def roofline_record(phase, precision, kernel, flop, byt, avg_s):
    """The bound is whichever of MFMA time and HBM time at published peaks is longer for
    the algorithmic flops and bytes of one launch; achieved = those / launch time."""
    peak_tf = phase_peak(precision, phase)
    mfma_bound = flop / (peak_tf * 1e12) >= byt / (HBM_PEAK_GBS * 1e9)
    if mfma_bound:
        achieved, peak, unit = flop / avg_s / 1e12, peak_tf, "TFLOP/s"
    else:
        achieved, peak, unit = byt / avg_s / 1e9, HBM_PEAK_GBS, "GB/s"
    return {"phase": phase, "kernel": kernel, "bound": "mfma" if mfma_bound else "hbm",
            "achieved": round(achieved, 3), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "flop_per_launch": flop,
            "bytes_per_launch": byt, "avg_launch_ms": round(avg_s * 1e3, 5)}


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


# ------------------------------------------------------------------ energy comparison


def compare_blocks(stats_a, stats_b, n_walkers, clamp_max=0.01, idx_a=None, idx_b=None,
                   sr_every=1):
    """Block-by-block energy comparison of two runs with the same draws.  stats_*:
    (blocks, 4) arrays of block energy, block weight, zero weights, clamped energies.
    Returns the usable blocks' differences (both finite, both under clamp_max of the
    walkers clamped) with the bookkeeping that says how many blocks were usable.
    idx_a / idx_b: the resampling indices of each SR of the two runs (an SR after every
    sr_every blocks); blocks before the first SR whose indices differ hold the same
    walker populations in both runs, so their difference is the arithmetic's alone."""
    import numpy as np

    a, b = np.asarray(stats_a), np.asarray(stats_b)
    nb = min(len(a), len(b))
    a, b = a[:nb], b[:nb]
    finite = np.isfinite(a[:, 0]) & np.isfinite(b[:, 0]) & (a[:, 1] > 0) & (b[:, 1] > 0)
    unclamped = (a[:, 3] < clamp_max * n_walkers) & (b[:, 3] < clamp_max * n_walkers)
    ok = finite & unclamped
    same = np.zeros(nb, dtype=bool)
    if idx_a is not None and idx_b is not None:
        first = 0
        while (first < min(len(idx_a), len(idx_b))
               and np.array_equal(np.asarray(idx_a[first]), np.asarray(idx_b[first]))):
            first += 1
        same = np.array([bi // sr_every <= first for bi in range(nb)], dtype=bool)
    return {
        "same_history": int(np.sum(same)),
        "diff_same": (a[ok & same, 0] - b[ok & same, 0]).tolist(),
        "blocks": int(nb),
        "nonfinite": int(np.sum(~finite)),
        "clamped_blocks": int(np.sum(finite & ~unclamped)),
        "usable": int(np.sum(ok)),
        "diff": (a[ok, 0] - b[ok, 0]).tolist(),
        "first_usable_is_block0": bool(nb > 0 and ok[0]),
        "clamp_frac_a": float(np.mean(a[:, 3]) / n_walkers) if nb else None,
        "clamp_frac_b": float(np.mean(b[:, 3]) / n_walkers) if nb else None,
        "zero_weight_frac_a": float(np.mean(a[:, 2]) / n_walkers) if nb else None,
        "zero_weight_frac_b": float(np.mean(b[:, 2]) / n_walkers) if nb else None,
    }


def summarize_delta(comps):
    """Aggregate compare_blocks results: max |dE| over usable blocks, max |dE| of the
    first block (before any resampling could separate the runs), the mean difference
    over usable blocks with its standard error."""
    import numpy as np

    diffs = [d for c in comps for d in c["diff"]]
    first = [c["diff"][0] for c in comps if c["first_usable_is_block0"]]
    same = [d for c in comps for d in c.get("diff_same", [])]
    nonfinite = sum(c["nonfinite"] for c in comps)
    out = {
        "delta_e_same_history_max_eh": float(np.max(np.abs(same))) if same else None,
        "blocks_same_history": sum(c.get("same_history", 0) for c in comps),
        "blocks_same_history_usable": len(same),
        "delta_e_max_eh": float(np.max(np.abs(diffs))) if diffs else None,
        "delta_e_first_block_max_eh": float(np.max(np.abs(first))) if first else None,
        "delta_e_mean_eh": float(np.mean(diffs)) if diffs else None,
        "delta_e_mean_stderr_eh": (float(np.std(diffs, ddof=1) / math.sqrt(len(diffs)))
                                   if len(diffs) > 1 else None),
        "blocks_usable": len(diffs),
        "blocks_total": sum(c["blocks"] for c in comps),
        "blocks_clamped": sum(c["clamped_blocks"] for c in comps),
        "blocks_nonfinite": nonfinite,
        "clamp_frac_mean": float(np.mean([c["clamp_frac_a"] for c in comps
                                          if c["clamp_frac_a"] is not None] or [0.0])),
        "zero_weight_frac_mean": float(np.mean([c["zero_weight_frac_a"] for c in comps
                                                if c["zero_weight_frac_a"] is not None]
                                               or [0.0])),
    }
    out["energy_check"] = ("ok" if nonfinite == 0 and diffs else
                           f"FAILED: {nonfinite} non-finite blocks" if nonfinite else
                           "FAILED: no usable block")
    if nonfinite:
        print(f"# WARNING: {nonfinite} non-finite block energies", file=sys.stderr, flush=True)
    return out


# ------------------------------------------------------------------ CPU baseline


def _cpu_fragment_worker(args):
    """One single-threaded process: oracle steps of each given fragment on a few walkers
    (time per walker-step, with one QR / overlap / energy amortised over 50 steps)."""
    specs, n_walk, n_steps = args
    import numpy as np

    from cs_afqmc_amd import problems
    from oracle import afqmc_oracle as o

    out = []
    for wt, norb, nocc, nchol, W, seed in specs:
        if wt == "h10":
            from cs_afqmc_amd import hchain
            hd, wd, nelec, info = hchain.prep(10, 1.6, 1e-5)
            wt, norb, nchol = "rhf", 10, info["nchol"]
        elif wt == "rhf":
            hd, wd, nelec = problems.synthetic_rhf(norb, nocc, nchol, seed=seed)
        else:
            hd, wd, nelec = problems.synthetic_uhf(norb, nocc, nocc, nchol, seed=seed)
        oh = o.build_measurement_intermediates(wt, hd, wd, norb)
        oh = o.build_propagation_intermediates(wt, oh, wd["rdm1"], norb, nelec, DT)
        rng = np.random.default_rng(seed + 17)
        pd = o.init_prop_data(wt, wt, wd, oh, n_walk, nelec)
        t0 = time.perf_counter()
        for _ in range(n_steps):
            pd = o.propagate(wt, wt, oh, pd, rng.standard_normal((n_walk, nchol)), wd, DT, 6,
                             norb, n_walkers=n_walk)
        t_step = (time.perf_counter() - t0) / (n_steps * n_walk)
        t1 = time.perf_counter()
        wk = o.orthonormalize_walkers(wt, pd["walkers"])
        o.calc_overlap(wt, wk, wd)
        o.calc_energy(wt, wk, oh, wd)
        t_block = (time.perf_counter() - t1) / n_walk
        out.append((W, t_step + t_block / N_PROP_STEPS))
    return out


def cpu_baseline(specs, n_walk, n_steps, label):
    """Whole-job CPU rate of the oracle for a list of problems (walker_type, norb, nocc,
    nchol, W, seed): each problem's time per walker-step is measured on n_walk walkers x
    n_steps steps; the P processes share the job's walker-steps, so the rate is
    P * sum(W) / sum(W * t_walker_step)."""
    cores = len(os.sched_getaffinity(0))
    procs = max(1, min(cores, CPU_PROCS_MAX, len(specs)))
    env_before = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS",
                                                 "MKL_NUM_THREADS")}
    for k in env_before:
        os.environ[k] = "1"
    t0 = time.perf_counter()
    try:
        ctx = mp.get_context("spawn")
        chunks = [specs[i::procs] for i in range(procs)]
        with ctx.Pool(procs) as pool:
            res = pool.map(_cpu_fragment_worker, [(c, n_walk, n_steps) for c in chunks])
    finally:
        for k, v in env_before.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    rows = [r for part in res for r in part]
    work = sum(W for W, _ in rows)
    secs = sum(W * t for W, t in rows)
    return {
        "value": round(procs * work / secs, 1),
        "unit": "walker-steps/s",
        "cores": procs,
        "host_cpus": os.cpu_count(),
        "kind": "port",
        "sample": (f"{label}: oracle on every listed problem, {n_walk} walkers x {n_steps} "
                   f"steps each + one QR/overlap/energy amortised over {N_PROP_STEPS} steps, "
                   f"{procs} single-threaded processes ({time.perf_counter() - t0:.0f} s wall)"),
    }


# ------------------------------------------------------------------ GPU helpers


def event_repeats(stream, fn, repeats):
    """Milliseconds of `fn` (which enqueues work on `stream`) per repeat, from a HIP
    event pair on that stream."""
    import torch

    ms = []
    for _ in range(repeats):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    return ms


def copy_probe(nbytes=1 << 30):
    """A plain device copy of known size, for scaling PMC byte counters."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def median(xs):
    s = sorted(xs)
    n = len(s)
    return s[n // 2] if n % 2 else 0.5 * (s[n // 2 - 1] + s[n // 2])


def dominant_roofline(precision, phases, fl, by, kernel_of, workload):
    """Roofline of the phase with the most device time among the step phases."""
    cands = [p for p, (ms, n) in phases.items() if n and p in fl and p != "weights"]
    if not cands:
        return None, None
    dom = max(cands, key=lambda p: phases[p][0])
    ms, n = phases[dom]
    r = roofline_record(dom, precision, kernel_of(dom), fl[dom], by[dom], ms / n / 1e3)
    r["launches_timed"] = n
    r["traffic"] = traffic_from_profiles(r["kernel"], workload, precision)
    return dom, r


def all_rooflines(precision, phases, fl, by, kernel_of):
    out = {}
    for p, (ms, n) in phases.items():
        if n and p in fl:
            r = roofline_record(p, precision, kernel_of(p), fl[p], by[p], ms / n / 1e3)
            out[p] = {k: r[k] for k in ("kernel", "bound", "achieved", "unit", "frac",
                                        "avg_launch_ms")}
    return out


# ------------------------------------------------------------------ 64 fragments, N 20-120, W 256


def _make_fragments(specs, precision):
    from cs_afqmc_amd import batch, problems
    frags = []
    for wt, norb, nocc, nchol, W, seed in specs:
        hd, wd, ne = problems.synthetic_rhf(norb, nocc, nchol, seed=seed)
        frags.append(batch.Fragment(hd, wd, wt, ne, W, dt=DT, seed=seed, precision=precision,
                                    n_prop_steps=N_PROP_STEPS, tag=(norb, nocc, nchol, W)))
    return frags


def fragment_specs():
    from cs_afqmc_amd import problems
    return [("rhf", n, o, g, w, s) for n, o, g, w, s in problems.fragment_specs(64, seed=0)]


def run_fragments_mode(specs, mode, warm_blocks, timed_blocks, side_streams=4):
    """One form of the fragment workload.  mode: "<precision>_grouped" (one captured
    graph per block of every fragment) or "<precision>_plain" (each fragment's own
    captured block, one fragment after another)."""
    import torch

    from cs_afqmc_amd import batch

    prec, form = mode.split("_")
    t0 = time.perf_counter()
    frags = _make_fragments(specs, prec)
    setup_s = time.perf_counter() - t0
    walkers = sum(f.n_walkers for f in frags)
    stream = torch.cuda.Stream()
    bt = None
    if form == "grouped":
        bt = batch.FragmentBatch(frags, tune=True, use_graph=True, stream=stream,
                                 keep_sr_indices=True, side_streams=side_streams)

        def block():
            bt.run_block()
    else:
        def block():
            with torch.cuda.stream(stream):
                for f in frags:
                    f.run_block_single(use_graph=True)
    for _ in range(warm_blocks):
        block()
    torch.cuda.synchronize()
    print(f"# fragments {mode}: set-up {setup_s:.1f} s, {warm_blocks} warm-up blocks done",
          file=sys.stderr, flush=True)

    def timed():
        for _ in range(timed_blocks):
            block()

    ms = event_repeats(stream, timed, REPEATS)
    torch.cuda.synchronize()
    med = median(ms)
    rate = walkers * N_PROP_STEPS * timed_blocks / (med / 1e3)
    print(f"# fragments {mode}: {REPEATS} x {timed_blocks} blocks, median {med:.2f} ms "
          f"-> {rate / 1e6:.3f} M walker-steps/s", file=sys.stderr, flush=True)
    res = {"value": rate, "ms_per_step": med / (timed_blocks * N_PROP_STEPS),
           "repeat_ms": [round(x, 3) for x in ms], "set_up_s": round(setup_s, 2),
           "stats": [f.block_stats() for f in frags],
           "sr_idx": [[i.cpu().numpy() for i in f.sr_indices] for f in frags]}
    if bt is not None and bt.tuned:
        res["gemm_tile_ms"] = bt.tuned
    if bt is not None:
        # per-phase device time of one eager block: all phases, then the dominant one only
        members = [(f.walker_type, f.norb, f.nelec, f.dev.nchol, f.n_walkers,
                    f.dev.vhs_entries(), f.prop.n_exp_terms) for f in frags]
        fl, by = phase_work(prec, members)
        frag_n = [f.norb for f in frags]
        kern = (lambda p: phase_kernel(prec, p, max(frag_n), (max(f.nelec[0] for f in frags),)
                                       * 2))
        bt.use_graph = False
        bt._drop_replay()
        with torch.cuda.stream(stream):
            bt.group.prof_collect()
            bt.group.prof_enable(True)
        bt.run_block()
        with torch.cuda.stream(stream):
            phases = bt.group.prof_collect()
            dom, _ = dominant_roofline(prec, phases, fl, by, kern, FRAG_WORKLOAD)
            bt.group.prof_enable(True, phases=[dom])
        bt.run_block()
        with torch.cuda.stream(stream):
            timed_ph = bt.group.prof_collect()
            bt.group.prof_enable(False, phases=None)
        _, roof = dominant_roofline(prec, {dom: timed_ph[dom]}, fl, by, kern, FRAG_WORKLOAD)
        res["phases"] = {p: {"ms_total": round(v[0], 4), "launches": v[1],
                             "ms_avg": round(v[0] / v[1], 5)}
                         for p, v in phases.items() if v[1]}
        res["roofline"] = roof
        res["rooflines"] = all_rooflines(prec, phases, fl, by, kern)
        res["phase_sum_ms_per_step"] = round(
            sum(v[0] for v in phases.values()) / N_PROP_STEPS, 5)
    del frags, bt
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return res


def run_fragments(args, modes):
    specs = fragment_specs()
    warm = blocks_for(args.warmup) + 1  # the eager block and the capture
    timed = blocks_for(args.steps)
    runs = {m: run_fragments_mode(specs, m, warm, timed, args.side_streams) for m in modes}
    return specs, runs


# ------------------------------------------------------------------ one problem (RHF N100 / UHF N200)


def run_single(workload, precision, warm_blocks, timed_blocks, n_ene_blocks=4):
    """One problem, the sampler's graph-captured block (sampling._BlockState): fields,
    50 steps, QR, overlaps, energies, block statistics; SR and the e_estimate update
    every n_ene_blocks blocks (one sampler call of driver.afqmc)."""
    import numpy as np
    import torch

    from cs_afqmc_amd import hamiltonian, problems, propagation, rng, sampling, wavefunctions

    walker_type, norb, nocc, nchol, W = SINGLE_WORKLOADS[workload]
    if walker_type == "rhf":
        hd, wd, nelec = problems.synthetic_rhf(norb, nocc, nchol, seed=0)
        trial = wavefunctions.rhf(norb, nelec)
        prop = propagation.propagator_restricted(dt=DT, n_walkers=W, precision=precision)
    else:
        hd, wd, nelec = problems.synthetic_uhf(norb, nocc, nocc, nchol, seed=0)
        trial = wavefunctions.uhf(norb, nelec)
        prop = propagation.propagator_unrestricted(dt=DT, n_walkers=W, precision=precision)
    ham = hamiltonian.hamiltonian(norb)
    ham_data = ham.build_measurement_intermediates(hd, trial, wd)
    ham_data = ham.build_propagation_intermediates(ham_data, prop, trial, wd)
    pd = prop.init_prop_data(trial, wd, ham_data)
    pd["key"] = rng.PRNGKey(0)
    dev = ham_data["_afqmc_device"]
    st = sampling._BlockState(dev, pd, N_PROP_STEPS, W, use_graph=True)
    e0 = pd["e_estimate"].reshape(())
    st.scalars[0].copy_(e0)
    st.scalars[1].copy_(e0)
    rng.device_key(pd["key"], dev.device, out=st.key)
    stream = torch.cuda.Stream()
    hist = []
    sr_hist = []
    since_sr = [0]

    def block():
        with torch.cuda.stream(stream):
            st.block()
            hist.append(st.slot.clone())
            since_sr[0] += 1
            if since_sr[0] == n_ene_blocks:
                since_sr[0] = 0
                st.sr()
                sr_hist.append(st.idx.clone())
                dev.estimate_mix(st.scalars, torch.stack(hist[-n_ene_blocks:]))

    # the warm-up covers two resamplings too (the eager one and the capture)
    total_warm = max(warm_blocks, 2 * n_ene_blocks)
    for _ in range(total_warm):
        block()
    torch.cuda.synchronize()

    def timed():
        for _ in range(timed_blocks):
            block()

    ms = event_repeats(stream, timed, REPEATS)
    med = median(ms)
    rate = W * N_PROP_STEPS * timed_blocks / (med / 1e3)
    print(f"# single {workload} {precision}: median {med:.2f} ms / {timed_blocks} blocks -> "
          f"{rate / 1e6:.3f} M walker-steps/s", file=sys.stderr, flush=True)
    stats = torch.stack(hist).cpu().numpy()
    # phases: one eager block with all event pairs, one with the dominant phase only
    nv = dev.vhs_entries()
    fl, by = phase_work(precision, [(walker_type, norb, nelec, nchol, W, nv,
                                     prop.n_exp_terms)])
    kern = lambda p: phase_kernel(precision, p, norb, nelec)  # noqa: E731
    with torch.cuda.stream(stream):
        dev.prof_collect()
        dev.prof_enable(True)
        st._block()
        phases = dev.prof_collect()
        dom, _ = dominant_roofline(precision, phases, fl, by, kern, workload)
        dev.prof_enable(True, phases=[dom])
        st._block()
        timed_ph = dev.prof_collect()
        dev.prof_enable(False, phases=None)
    _, roof = dominant_roofline(precision, {dom: timed_ph[dom]}, fl, by, kern, workload)
    step_phases = ("green", "fb_gemm", "fields", "vhs_gemm", "taylor", "overlap", "weights")
    block_phases = ("qr", "energy", "rng")
    phase_sum = (sum(phases[p][0] for p in step_phases) +
                 sum(phases[p][0] for p in block_phases)) / N_PROP_STEPS
    res = {"value": rate, "ms_per_step": med / (timed_blocks * N_PROP_STEPS),
           "repeat_ms": [round(x, 3) for x in ms], "stats": stats,
           "sr_idx": [i.cpu().numpy() for i in sr_hist], "sr_every": n_ene_blocks,
           "phases": {p: {"ms_total": round(v[0], 4), "launches": v[1],
                          "ms_avg": round(v[0] / v[1], 5)} for p, v in phases.items() if v[1]},
           "phase_sum_ms_per_step": round(phase_sum, 5),
           "roofline": roof, "rooflines": all_rooflines(precision, phases, fl, by, kern),
           "n_walkers": W}
    del st, dev, ham_data, pd
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    _ = np
    return res


# ------------------------------------------------------------------ H10 chain, 1.6 bohr, STO-6G


def run_h10(precision, n_calls):
    """Linear H10, 1.6 bohr, STO-6G, RHF trial, modified Cholesky 1e-5 (own integrals),
    200 walkers, dt 0.005, QR every 5 steps (n_prop_steps 5, n_ene_blocks 10: 50 steps
    per sampler call), seed 0; n_calls sampler calls with the driver's e_estimate update
    after each.  Returns the block statistics, the SR indices of each call, and the
    walker-steps/s of the calls (events around the whole loop, host synchronisation of
    each call's energy included)."""
    import numpy as np
    import torch

    from cs_afqmc_amd import hchain, rng, sampling

    setup = hchain.h10_config1(precision=precision)
    smp = sampling.sampler(n_prop_steps=5, n_ene_blocks=10, n_sr_blocks=1)
    pd = setup["prop_data"]
    pd["key"] = rng.PRNGKey(0)
    rows, idx = [], []
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    for c in range(n_calls):
        if c == 2:  # calls 0 and 1 run eagerly and capture the graphs
            e0.record(stream)
        e, pd = smp.propagate_phaseless(setup["ham"], setup["ham_data"], setup["prop"], pd,
                                        setup["trial"], setup["wave_data"])
        rows.append(pd["block_stats"].cpu().numpy())
        idx.append(pd["_block_state"].idx.cpu().numpy())
        pd["e_estimate"] = 0.9 * pd["e_estimate"] + 0.1 * e
    e1.record(stream)
    e1.synchronize()
    ms = e0.elapsed_time(e1)
    # roofline of the dominant step phase: one eager 5-step block with every phase's event
    # pair, then one with the dominant phase's only (after the timed calls)
    dev, st, prop, trial = pd["_afqmc_device"], pd["_block_state"], setup["prop"], setup["trial"]
    norb, nelec = trial.norb, tuple(trial.nelec)
    fl, by = phase_work(precision, [("rhf", norb, nelec, setup["nchol"], prop.n_walkers,
                                     dev.vhs_entries(), prop.n_exp_terms)])
    kern = lambda p: phase_kernel(precision, p, norb, nelec)  # noqa: E731
    dev.prof_collect()
    dev.prof_enable(True)
    st._block()
    phases = dev.prof_collect()
    dom, _ = dominant_roofline(precision, phases, fl, by, kern, H10_WORKLOAD)
    dev.prof_enable(True, phases=[dom])
    st._block()
    timed_ph = dev.prof_collect()
    dev.prof_enable(False, phases=None)
    _, roof = dominant_roofline(precision, {dom: timed_ph[dom]}, fl, by, kern, H10_WORKLOAD)
    return {"stats": np.concatenate(rows), "sr_idx": idx, "nchol": setup["nchol"],
            "e_hf": setup["e_hf"], "roofline": roof,
            "value": 200 * 50 * (n_calls - 2) / (ms / 1e3)}


# ------------------------------------------------------------------ serving: Poisson 200 req/s, N 20-250


def serving_draws(n_requests):
    """(arrival, norb, nocc, W, problem seed) of each request: Poisson arrivals at 200/s,
    norb ~ U[20, 250], nocc = norb / 5, W from {128, 256, 512, 1024}."""
    import numpy as np

    rng = np.random.default_rng(2024)
    t = 0.0
    out = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / 200.0)
        norb = int(rng.integers(20, 251))
        nocc = max(1, norb // 5)
        W = int(rng.choice([128, 256, 512, 1024]))
        out.append((t, norb, nocc, W, 5000 + i))
    return out


def serving_requests(n_requests):
    from cs_afqmc_amd import problems, serving

    reqs = []
    for i, (t, norb, nocc, W, seed) in enumerate(serving_draws(n_requests)):
        hd, wd, ne = problems.synthetic_rhf(norb, nocc, 4 * norb, seed=seed)
        reqs.append(serving.Request(arrival=t, rid=i, ham_data=hd, wave_data=wd, nelec=ne,
                                    n_walkers=W, n_steps=100, seed=17 + i))
    return reqs


def run_serving(reqs, profile=True):
    import numpy as np
    import torch

    from cs_afqmc_amd import serving

    torch.cuda.set_device(0)
    srv = serving.FragmentServer(precision="bf16x3", n_prop_steps=N_PROP_STEPS,
                                 hbm_budget=200 << 30, max_active=SERVE_MAX_ACTIVE,
                                 policy=SERVE_POLICY, profile=profile)
    t0 = time.perf_counter()
    res = srv.serve(reqs)
    wall = time.perf_counter() - t0
    ws = sum(r.walker_steps for r in res)
    makespan = max(r.done for r in res)
    st = serving.latency_stats(res)
    last_arrival = max(r.arrival for r in reqs)
    out = {
        "value": round(ws / makespan, 1),
        "unit": "walker-steps/s",
        "latency_p50_s": st["p50_s"], "latency_p99_s": st["p99_s"],
        "latency_mean_s": st["mean_s"], "latency_max_s": st.get("max_s"),
        "n_requests": len(reqs), "completed": len(res),
        "arrival_window_s": round(last_arrival, 3),
        "offered_walker_steps_per_s": round(
            sum(r.n_walkers * r.n_steps for r in reqs) / max(last_arrival, 1e-9), 1),
        "makespan_s": round(makespan, 3), "wall_s": round(wall, 3),
        "build_s_mean": float(np.mean([r.build_s for r in res])),
        "arrival_to_admit_s_mean": float(np.mean([r.admitted - r.arrival for r in res])),
        "admit_to_done_s_mean": float(np.mean([r.done - r.admitted for r in res])),
        "peak_active_fragments": srv.peak_active,
        "peak_device_bytes_counted": srv.peak_used,
        "admission": SERVE_POLICY, "max_active_fragments": SERVE_MAX_ACTIVE,
    }
    if srv.phase_ms:
        fl, by, ms = srv.phase_flops, srv.phase_bytes, srv.phase_ms
        # requests up to norb 128 run the register kernels, larger ones the streaming ones
        kern = lambda p: (phase_kernel("bf16x3", p, 100, (20, 20)) + " / "  # noqa: E731
                          + phase_kernel("bf16x3", p, 250, (50, 50))
                          if p in ("taylor",) else phase_kernel("bf16x3", p, 250, (50, 50)))
        cands = [p for p in ms if ms[p][1] and p in fl and p != "weights"]
        if cands:
            dom = max(cands, key=lambda p: ms[p][0])
            # work and time summed over every tick: achieved = total work / total time
            tot_ms, n = ms[dom]
            r = roofline_record(dom, "bf16x3", kern(dom), fl[dom] / n, by[dom] / n,
                                tot_ms / n / 1e3)
            r["launches_timed"] = n
            r["traffic"] = None
            out["roofline"] = r
    return out


# ------------------------------------------------------------------ main


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50,
                    help="timed steps per repeat, rounded up to whole 50-step blocks")
    ap.add_argument("--warmup", type=int, default=50,
                    help="warm-up steps, rounded up to whole blocks (plus the capture)")
    ap.add_argument("--workload", default="all",
                    choices=["all", FRAG_WORKLOAD, SERVE_WORKLOAD, H10_WORKLOAD,
                             *SINGLE_WORKLOADS])
    ap.add_argument("--precision", default="compare",
                    choices=["compare", "fp64", "fp32", "bf16x3"])
    ap.add_argument("--requests", type=int, default=64,
                    help="requests of the serving workload")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--side-streams", type=int, default=4,
                    help="side streams for the fragments' draws and block ends")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device copy first (scales PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the walker population runs on one GPU; --gpus {args.gpus} is not "
              "supported (sharding across GPUs is not implemented).", file=sys.stderr)
        sys.exit(2)

    warm, timed = blocks_for(args.warmup) + 1, blocks_for(args.steps)
    mixed = "bf16x3" if args.precision in ("compare", "bf16x3") else args.precision
    want = args.workload

    # CPU baselines first: the worker processes are spawned before anything touches the GPU
    cpu = {}
    if not args.no_cpu_baseline:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        if want in ("all", FRAG_WORKLOAD):
            cpu["fragments"] = cpu_baseline(fragment_specs(), 16, 8, FRAG_WORKLOAD)
        if want in ("all", H10_WORKLOAD):
            # 200 walkers over the processes, as the reference's ranks share them
            cpu["h10"] = cpu_baseline([("h10", 10, 5, 27, 200, i) for i in range(16)], 13, 50,
                                      H10_WORKLOAD)
        if want in SINGLE_WORKLOADS:
            wt, n, o, g, W = SINGLE_WORKLOADS[want]
            cpu["single"] = cpu_baseline([(wt, n, o, g, W, 11 + i) for i in range(16)], 4, 4,
                                         want)
        if want in ("all", SERVE_WORKLOAD) and not args.no_serving:
            cpu["serving"] = cpu_baseline(
                [("rhf", norb, nocc, 4 * norb, W, seed)
                 for _, norb, nocc, W, seed in serving_draws(args.requests)], 2, 2,
                SERVE_WORKLOAD)
    import numpy as np
    import torch
    torch.cuda.set_device(0)
    if args.copy_probe:
        copy_probe()

    out = {}
    if want in ("all", FRAG_WORKLOAD):
        modes = ([f"{mixed}_grouped", "fp32_grouped", "fp64_grouped", "fp64_plain"]
                 if args.precision == "compare" else [f"{mixed}_grouped"])
        specs, runs = run_fragments(args, modes)
        m = runs[f"{mixed}_grouped"]
        total_w = sum(s[4] for s in specs)
        out.update({
            "metric": "walker-steps/s mixed vs FP64 GPU at |dE|<0.1 mEh; p50/p99 fragment "
                      "latency",
            "value": round(m["value"], 1),
            "unit": "walker-steps/s",
            "n_gpus": 1,
            "steps": timed * N_PROP_STEPS,
            "warmup": warm * N_PROP_STEPS,
            "ms_per_step": round(m["ms_per_step"], 5),
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": DTYPE[mixed],
            "data": "synthetic fragments (seeded Cholesky vectors N(0, 1/norb^2) x exp(-g/80), "
                    "h1, Haar trial; device Philox fields)",
            "config": {"workload": FRAG_WORKLOAD, "n_fragments": len(specs),
                       "norb_range": [20, 120], "nocc": "norb/5", "nchol": "4 norb",
                       "n_walkers_total": total_w, "walkers_per_fragment": 256,
                       "n_prop_steps": N_PROP_STEPS, "dt": DT, "precision": mixed,
                       "form": "grouped step, whole block replayed from one HIP graph",
                       "side_streams": args.side_streams,
                       "repeats": REPEATS, "blocks_per_repeat": timed},
            "repeat_ms": m["repeat_ms"],
        })
        if "fp64_plain" in runs:
            out["fp64_value"] = round(runs["fp64_plain"]["value"], 1)
            out["fp64_form"] = "plain: each fragment's block through its own context in turn"
            out["speedup_vs_fp64"] = round(m["value"] / runs["fp64_plain"]["value"], 3)
            out["fp64_grouped_value"] = round(runs["fp64_grouped"]["value"], 1)
            out["speedup_vs_fp64_grouped"] = round(m["value"] / runs["fp64_grouped"]["value"], 3)
            comps = [compare_blocks(a, b, s[4], idx_a=ia, idx_b=ib) for a, b, s, ia, ib in
                     zip(m["stats"], runs["fp64_plain"]["stats"], specs, m["sr_idx"],
                         runs["fp64_plain"]["sr_idx"])]
            out.update(summarize_delta(comps))
            gp = [compare_blocks(a, b, s[4]) for a, b, s in
                  zip(runs["fp64_grouped"]["stats"], runs["fp64_plain"]["stats"], specs)]
            out["delta_e_fp64_grouped_vs_plain_eh"] = summarize_delta(gp)["delta_e_max_eh"]
            out["fp64_phases"] = runs["fp64_grouped"].get("phases")
            out["fp64_roofline"] = runs["fp64_grouped"].get("roofline")
            f32 = runs["fp32_grouped"]
            c32 = [compare_blocks(a, b, s[4], idx_a=ia, idx_b=ib) for a, b, s, ia, ib in
                   zip(f32["stats"], runs["fp64_plain"]["stats"], specs, f32["sr_idx"],
                       runs["fp64_plain"]["sr_idx"])]
            d32 = summarize_delta(c32)
            out["fp32"] = {"form": "grouped step, FP32-MFMA Cholesky contractions",
                           "dtype": DTYPE["fp32"], "value": round(f32["value"], 1),
                           "speedup_vs_fp64": round(f32["value"] / runs["fp64_plain"]["value"], 3),
                           "roofline": f32.get("roofline"), "phases": f32.get("phases"),
                           **{k: d32[k] for k in ("delta_e_same_history_max_eh",
                                                  "blocks_same_history_usable",
                                                  "delta_e_first_block_max_eh",
                                                  "delta_e_max_eh", "energy_check")}}
        out["phases"] = m.get("phases")
        out["phase_sum_ms_per_step"] = m.get("phase_sum_ms_per_step")
        out["gemm_tile_ms"] = m.get("gemm_tile_ms")
        out["roofline"] = m.get("roofline")
        out["rooflines"] = m.get("rooflines")
        out["cpu_baseline"] = cpu.get("fragments")

    if want == "all" or want in SINGLE_WORKLOADS:
        wl = want if want in SINGLE_WORKLOADS else "rhf_n100_o20_g400_w1024"
        precs = ["fp64", "fp32", mixed] if args.precision == "compare" else [mixed]
        sr = {p: run_single(wl, p, warm, timed) for p in precs}
        W = SINGLE_WORKLOADS[wl][4]
        single = {"workload": wl,
                  "value": round(sr[mixed]["value"], 1),
                  "ms_per_step": round(sr[mixed]["ms_per_step"], 5),
                  "phase_sum_ms_per_step": sr[mixed]["phase_sum_ms_per_step"],
                  "repeat_ms": sr[mixed]["repeat_ms"],
                  "phases": sr[mixed]["phases"], "roofline": sr[mixed]["roofline"],
                  "rooflines": sr[mixed]["rooflines"]}
        if "fp64" in sr and mixed != "fp64":
            single["fp64_value"] = round(sr["fp64"]["value"], 1)
            single["fp64_ms_per_step"] = round(sr["fp64"]["ms_per_step"], 5)
            single["fp64_phase_sum_ms_per_step"] = sr["fp64"]["phase_sum_ms_per_step"]
            single["speedup_vs_fp64"] = round(sr[mixed]["value"] / sr["fp64"]["value"], 3)
            single["fp64_roofline"] = sr["fp64"]["roofline"]
            single.update(summarize_delta([compare_blocks(
                sr[mixed]["stats"], sr["fp64"]["stats"], W, idx_a=sr[mixed]["sr_idx"],
                idx_b=sr["fp64"]["sr_idx"], sr_every=sr[mixed]["sr_every"])]))
        if "fp64" in sr and "fp32" in sr and mixed != "fp32":
            d32 = summarize_delta([compare_blocks(
                sr["fp32"]["stats"], sr["fp64"]["stats"], W, idx_a=sr["fp32"]["sr_idx"],
                idx_b=sr["fp64"]["sr_idx"], sr_every=sr["fp32"]["sr_every"])])
            single["fp32"] = {"dtype": DTYPE["fp32"], "value": round(sr["fp32"]["value"], 1),
                              "ms_per_step": round(sr["fp32"]["ms_per_step"], 5),
                              "speedup_vs_fp64": round(sr["fp32"]["value"] / sr["fp64"]["value"],
                                                       3),
                              "roofline": sr["fp32"]["roofline"],
                              **{k: d32[k] for k in ("delta_e_same_history_max_eh",
                                                     "delta_e_first_block_max_eh",
                                                     "blocks_usable", "energy_check")}}
        if want == "all":
            out["single"] = single
        else:
            out.update({"metric": "walker-steps/s mixed vs FP64 GPU (one problem)",
                        "value": single.pop("value"), "unit": "walker-steps/s", "n_gpus": 1,
                        "steps": timed * N_PROP_STEPS, "warmup": warm * N_PROP_STEPS,
                        "ms_per_step": single.pop("ms_per_step"), "higher_is_better": True,
                        "scaling": "weak", "vs_baseline": None, "dtype": DTYPE[mixed],
                        "data": "synthetic", "config": {"workload": wl, "precision": mixed}})
            out.update(single)
            out["cpu_baseline"] = cpu.get("single")

    if want in ("all", H10_WORKLOAD):
        from cs_afqmc_amd import stat_utils
        h = {p: run_h10(p, H10_CALLS) for p in ("fp64", mixed)}
        hd = summarize_delta([compare_blocks(h[mixed]["stats"], h["fp64"]["stats"], 200,
                                             idx_a=h[mixed]["sr_idx"],
                                             idx_b=h["fp64"]["sr_idx"], sr_every=10)])
        h10 = {"workload": H10_WORKLOAD, "nchol": h["fp64"]["nchol"],
               "e_rhf": h["fp64"]["e_hf"], "calls": H10_CALLS, "blocks_per_call": 10,
               "value": round(h[mixed]["value"], 1), "fp64_value": round(h["fp64"]["value"], 1),
               "roofline": h[mixed]["roofline"], "fp64_roofline": h["fp64"]["roofline"]}
        for p, tag in (("fp64", "fp64"), (mixed, "mixed")):
            st = h[p]["stats"][10 * 2:]  # the first two calls equilibrate
            e, err = stat_utils.blocking_analysis(st[:, 1], st[:, 0], block_dtype=np.float64)
            h10[f"energy_{tag}"] = float(e)
            h10[f"energy_{tag}_err"] = float(err) if err is not None else None
        if h10["energy_fp64_err"] and h10["energy_mixed_err"]:
            h10["delta_e_run_eh"] = h10["energy_mixed"] - h10["energy_fp64"]
            h10["delta_e_run_sigma"] = abs(h10["delta_e_run_eh"]) / math.hypot(
                h10["energy_fp64_err"], h10["energy_mixed_err"])
        h10.update(hd)
        h10["cpu_baseline"] = cpu.get("h10")
        if want == "all":
            out["h10"] = h10
        else:
            out.update({"metric": "walker-steps/s mixed vs FP64 GPU (H10 chain)",
                        "value": h10["value"], "unit": "walker-steps/s", "n_gpus": 1,
                        "steps": 50 * (H10_CALLS - 2), "warmup": 100, "ms_per_step": None,
                        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                        "dtype": DTYPE[mixed], "data": "H10 chain, own STO-6G integrals",
                        "config": {"workload": H10_WORKLOAD}})
            out.update(h10)
            out["cpu_baseline"] = cpu.get("h10")

    if want in ("all", SERVE_WORKLOAD) and not args.no_serving:
        t0 = time.perf_counter()
        reqs = serving_requests(args.requests)
        print(f"# serving: {len(reqs)} requests generated in {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        sv = run_serving(reqs)
        print(f"# serving: p50 {sv['latency_p50_s']:.3f} s, p99 {sv['latency_p99_s']:.3f} s",
              file=sys.stderr, flush=True)
        # The Poisson-200/s stream offers an order of magnitude more work than one GPU
        # serves, so its latencies are queueing.  The same requests again, arrivals
        # stretched so the offered walker-steps/s is 70% of the rate just served: the
        # latencies of a sustainable stream.
        import dataclasses
        stretch = sv["offered_walker_steps_per_s"] / (0.7 * sv["value"])
        if stretch > 1.0:
            slow = [dataclasses.replace(r, arrival=r.arrival * stretch) for r in reqs]
            sus = run_serving(slow, profile=False)
            sv["sustainable"] = {k: sus[k] for k in (
                "value", "latency_p50_s", "latency_p99_s", "latency_mean_s", "latency_max_s",
                "arrival_window_s", "offered_walker_steps_per_s", "makespan_s",
                "arrival_to_admit_s_mean", "admit_to_done_s_mean")}
            sv["sustainable"]["arrival_stretch"] = round(stretch, 3)
            print(f"# serving at 70% load: p50 {sus['latency_p50_s']:.3f} s, "
                  f"p99 {sus['latency_p99_s']:.3f} s", file=sys.stderr, flush=True)
        sv["cpu_baseline"] = cpu.get("serving")
        if want == "all":
            out["serving"] = dict(sv, workload=SERVE_WORKLOAD)
            out["latency_p50_s"] = sv["latency_p50_s"]
            out["latency_p99_s"] = sv["latency_p99_s"]
        else:
            out.update({"metric": "p50/p99 fragment latency under load (serving)",
                        "value": sv.pop("value"), "unit": "walker-steps/s", "n_gpus": 1,
                        "steps": 100, "warmup": 0, "ms_per_step": None,
                        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                        "dtype": DTYPE["bf16x3"], "data": "synthetic fragment requests",
                        "config": {"workload": SERVE_WORKLOAD}})
            out.update(sv)
            out["cpu_baseline"] = sv.get("cpu_baseline")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
