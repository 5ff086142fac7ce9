"""Serving loop for fragment solves with continuous batching.

Requests (one AFQMC fragment solve each: integrals, trial, walker count, number of
steps, seed) arrive over time.  The server keeps every admitted fragment's walker
state resident on the GPU and advances all of them together, one block of
n_prop_steps grouped steps per tick (``batch.FragmentBatch``).  Between ticks it
  * retires fragments whose last block has been issued (their result is read once the
    GPU has finished it, without stalling the other fragments),
  * admits queued requests while the admitted set stays within the HBM budget and the
    fragment limit, oldest first ("fifo") or smallest estimated solve first ("sjf",
    which lowers the median latency when requests queue; under sustained overload it
    can hold large requests back): prep threads build each admitted request's
    intermediates and device context on their own streams while the GPU keeps ticking,
  * joins the requests whose set-up has finished.
So requests that arrive while others are on the GPU join at the next tick and their
work overlaps on the device.  A fragment's result does not depend on what else is in
the batch (see batch.py), so a request gives the same energy alone and in the mix.

Latency of a request = GPU completion of its last block - arrival (both on the
server's clock, which starts at the first call of ``serve``).
"""

import heapq
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np
import torch

from . import batch, work


@dataclass(order=True)
class Request:
    arrival: float
    rid: int
    ham_data: dict = field(compare=False)
    wave_data: dict = field(compare=False)
    nelec: tuple = field(compare=False)
    n_walkers: int = field(compare=False)
    n_steps: int = field(compare=False, default=100)
    seed: int = field(compare=False, default=0)
    walker_type: str = field(compare=False, default="rhf")


@dataclass
class Result:
    rid: int
    energy: float
    block_energies: np.ndarray
    block_weights: np.ndarray
    arrival: float
    admitted: float
    done: float
    walker_steps: int
    build_s: float = 0.0  # host + device set-up time of the request (prep thread)

    @property
    def latency(self):
        return self.done - self.arrival


def _spin_chol(ham_data):
    return len(getattr(ham_data["chol"], "shape", np.shape(ham_data["chol"]))) == 3


def _nchol(ham_data):
    shape = getattr(ham_data["chol"], "shape", np.shape(ham_data["chol"]))
    return int(shape[1] if len(shape) == 3 else shape[0])


def estimate_bytes(norb, nelec, nchol, n_walkers, n_prop_steps, precision,
                   walker_type="rhf", spin_chol=False):
    """Upper estimate of one fragment's device bytes before it is built: context
    intermediates and workspace (as afqmc_ctx_create sizes them), walker state, fields
    block, and the rotated Cholesky vectors ham_data keeps on the device.  nelec:
    (n_up, n_dn) or one count for closed shells.  Once built, the server replaces this
    with the fragment's measured bytes (Fragment.device_bytes).  spin_chol: one set of
    Cholesky vectors per spin (chol (2, G, N*N)), which doubles the vectors and the VHS."""
    if np.ndim(nelec) == 0:
        nelec = (int(nelec), int(nelec))
    N, G, W = norb, nchol, n_walkers
    ns = [nelec[0]] if walker_type == "rhf" else [nelec[0], nelec[1]]
    mixed = precision != "fp64"
    gp, np32 = -(-G // 32) * 32, -(-N // 32) * 32
    cs = 2 if spin_chol else 1
    b = cs * 8 * N * N * G + 16 * G         # chol_t, mf_shifts
    b += cs * 8 * N * N * G                 # the caller's chol upload, freed at creation
    fb_slices = 0
    for n in ns:
        b += 2 * 8 * G * n * N              # rot_chol in the context and in ham_data
        b += 8 * N * n + 8 * n * N + 8 * N * N  # trial, rot_h1, exp_h1
        b += 16 * W * n * N                 # float64 Green's function
        b += 2 * 16 * W * N * n             # walkers and the SR gather buffer
        kd = n * np32 if mixed else n * N
        fb_slices += min(8, max(1, kd // 256))
        b += 16 * W * -(-G // max(1, (128 if mixed else 64) // n))  # exchange partials
        if mixed:
            b += 2 * 2 * G * n * np32 + 2 * 2 * W * 2 * n * np32  # rot_chol, Green planes
            b += 2 * 2 * N * np32           # exp_h1 - I planes
    b += 8 * W * 2 * G * fb_slices + 8 * W * 2 * G  # force-bias slices, shifted fields
    if mixed:
        b += cs * 2 * 2 * N * N * gp + 2 * 2 * W * 2 * gp  # Cholesky, shifted-field planes
        b += cs * 4 * W * 2 * N * N + 32    # float32 VHS
    else:
        b += cs * 16 * W * N * N            # float64 VHS
    b += 8 * n_prop_steps * W * G          # fields block
    b += 64 * W + (1 << 20)                 # scalars, energies, SR, exchange partials, slack
    return int(b)


class FragmentServer:
    def __init__(self, precision="bf16x3", n_prop_steps=50, hbm_budget=64 << 30,
                 max_active=64, dt=0.005, tune=False, prep_workers=4, max_inflight=2,
                 policy="fifo", profile=False, profile_ticks=48):
        """profile: record HIP event pairs around every phase of the grouped step for the
        first `profile_ticks` ticks (the event pool holds 16384 pairs, about 300 per tick)
        and the algorithmic work of those ticks: phase_ms / phase_flops / phase_bytes."""
        if policy not in ("fifo", "sjf"):
            raise ValueError(f"policy {policy!r}: 'fifo' or 'sjf'")
        self.policy = policy
        self.prep_workers = max(1, int(prep_workers))
        self.max_inflight = max(1, int(max_inflight))
        self.precision = precision
        self.n_prop_steps = int(n_prop_steps)
        self.hbm_budget = int(hbm_budget)
        self.max_active = int(max_active)
        self.dt = dt
        self.tune = tune
        self.profile = bool(profile)
        self.profile_ticks = int(profile_ticks)
        self.phase_ms, self.phase_flops, self.phase_bytes = {}, {}, {}

    def _priority(self, req):
        """Admission order among queued requests: arrival ("fifo"), or the estimated
        work of the whole solve first ("sjf": walkers x steps x N^2 (G + 8 n), the
        VHS and Taylor products that dominate a step)."""
        if self.policy == "fifo":
            return (req.arrival, req.rid)
        norb = int(np.asarray(req.ham_data["h1"]).shape[-1])
        nchol = _nchol(req.ham_data)
        work = req.n_walkers * req.n_steps * norb * norb * (nchol + 8 * max(req.nelec))
        return (work, req.arrival, req.rid)

    def _build(self, req, side):
        """Host intermediates + device context of one request, on the calling prep
        thread's stream; returns (fragment, event after its set-up work)."""
        t = time.perf_counter()
        with torch.cuda.stream(side):
            frag = batch.Fragment(req.ham_data, req.wave_data, req.walker_type, req.nelec,
                                  req.n_walkers, dt=self.dt, seed=req.seed,
                                  precision=self.precision, n_prop_steps=self.n_prop_steps,
                                  tag=req.rid)
            ev = torch.cuda.Event()
            ev.record(side)
        frag.build_s = time.perf_counter() - t
        return frag, ev

    def serve(self, requests, poll_s=0.0005):
        """Run until every request is done; returns [Result] in completion order."""
        pending = sorted(requests)
        queue = []
        building = []  # (future, request, bytes)
        active = {}    # id(frag) -> (frag, request, admitted time, bytes)
        retiring = []  # (frag, request, admitted, event, host stats, bytes)
        results = []
        # finished fragments and their bytes: freeing a context waits for the whole device
        # (hipFree), so they are freed when the GPU is idle, or when admission needs their
        # memory; until then their bytes stay counted in `used`
        graveyard = []
        used = 0
        self.peak_active = 0    # fragments stepping together
        self.peak_reserved = 0  # fragments holding budget (stepping or being built)
        self.peak_used = 0      # largest device bytes counted against the budget
        self.graveyard_frees = 0  # times admission freed retired fragments early
        dev = torch.cuda.current_device()
        tick = torch.cuda.Stream()  # the grouped steps
        local = threading.local()   # one set-up stream per prep thread

        def prep(req):
            torch.cuda.set_device(dev)
            if not hasattr(local, "side"):
                local.side = torch.cuda.Stream()
            return self._build(req, local.side)

        pool = ThreadPoolExecutor(max_workers=self.prep_workers)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev0.record(tick)
        nxt = 0
        prof_ticks = 0
        inflight = []  # events after the queued ticks
        with torch.cuda.stream(tick):
            bt = batch.FragmentBatch(max_fragments=max(self.max_active, 1))
            while nxt < len(pending) or queue or building or active or retiring:
                now = time.perf_counter() - t0
                while nxt < len(pending) and pending[nxt].arrival <= now:
                    heapq.heappush(queue, (self._priority(pending[nxt]), pending[nxt]))
                    nxt += 1
                # retire: results of fragments whose last block the GPU has finished
                still = []
                for frag, req, adm, ev, host, nb in retiring:
                    if ev.query():
                        be, bw = host[:, 0].numpy().copy(), host[:, 1].numpy().copy()
                        done = ev0.elapsed_time(ev) / 1e3
                        energy = float(np.sum(be * bw) / np.sum(bw))
                        results.append(Result(req.rid, energy, be, bw, req.arrival, adm, done,
                                              frag.walker_steps, frag.build_s))
                        graveyard.append((frag, nb))
                    else:
                        still.append((frag, req, adm, ev, host, nb))
                retiring = still
                # admit: reserve the budget and build on the prep thread, oldest first
                while queue and len(active) + len(building) < self.max_active:
                    req = queue[0][1]
                    nchol = _nchol(req.ham_data)
                    norb = int(np.asarray(req.ham_data["h1"]).shape[-1])
                    need = estimate_bytes(norb, req.nelec, nchol, req.n_walkers,
                                          self.n_prop_steps, self.precision, req.walker_type,
                                          spin_chol=_spin_chol(req.ham_data))
                    if used + need > self.hbm_budget and graveyard:
                        tick.synchronize()
                        used -= sum(nb for _, nb in graveyard)
                        graveyard.clear()
                        self.graveyard_frees += 1
                    if (active or building or retiring) and used + need > self.hbm_budget:
                        break
                    heapq.heappop(queue)
                    used += need
                    building.append((pool.submit(prep, req), req, need))
                    self.peak_reserved = max(self.peak_reserved, len(active) + len(building))
                # join built fragments at this tick boundary
                joined = []
                for fut, req, need in building:
                    if fut.done():
                        frag, ev = fut.result()
                        tick.wait_event(ev)
                        frag.n_blocks_total = max(1, -(-req.n_steps // self.n_prop_steps))
                        # the measured footprint replaces the estimate
                        real = frag.device_bytes() + _tensor_bytes(frag.ham_data)
                        used += real - need
                        self.peak_used = max(self.peak_used, used)
                        active[id(frag)] = (frag, req, time.perf_counter() - t0, real)
                        self.peak_active = max(self.peak_active, len(active))
                        bt.add(frag)
                        joined.append(fut)
                building = [b for b in building if b[0] not in joined]
                if joined and self.tune and len(bt.fragments) > 1:
                    bt.tune()
                if not active:
                    if graveyard and not retiring and not building:
                        torch.cuda.synchronize()
                        used -= sum(nb for _, nb in graveyard)
                        graveyard.clear()
                    if building or retiring:
                        time.sleep(poll_s)
                    elif nxt < len(pending):
                        wait = pending[nxt].arrival - (time.perf_counter() - t0)
                        if wait > 0:
                            time.sleep(wait)
                    continue
                # one tick: a block of every active fragment.  At most `max_inflight`
                # ticks are queued on the GPU: fragments that join meanwhile enter the
                # next tick instead of waiting behind a deep queue, and the loop sleeps
                # (releasing the interpreter to the prep threads) rather than spinning.
                inflight = [e for e in inflight if not e.query()]
                if len(inflight) >= self.max_inflight:
                    time.sleep(poll_s)
                    continue
                if self.profile and prof_ticks < self.profile_ticks:
                    members = [(f.walker_type, f.norb, f.nelec, f.dev.nchol, f.n_walkers,
                                f.dev.vhs_entries(), f.prop.n_exp_terms, f.dev.spin_chol)
                               for f in bt.fragments]
                    fl, by = work.phase_work(self.precision, members)
                    for ph in fl:  # the first step computes the Green's function, later ones reuse it
                        k = 1 if ph == "green" else self.n_prop_steps
                        self.phase_flops[ph] = self.phase_flops.get(ph, 0.0) + k * fl[ph]
                        self.phase_bytes[ph] = self.phase_bytes.get(ph, 0.0) + k * by[ph]
                    bt.group.prof_enable(True)
                    bt.run_block()
                    bt.group.prof_enable(False)
                    prof_ticks += 1
                else:
                    bt.run_block()
                tev = torch.cuda.Event()
                tev.record(tick)
                inflight.append(tev)
                done = [f for f, *_ in active.values() if f.blocks_done >= f.n_blocks_total]
                if done:
                    for f in done:
                        _, req, adm, nb = active.pop(id(f))
                        st = torch.stack(f.stats)
                        host = torch.empty(st.shape, dtype=st.dtype, pin_memory=True)
                        host.copy_(st, non_blocking=True)
                        ev = torch.cuda.Event(enable_timing=True)
                        ev.record(tick)
                        retiring.append((f, req, adm, ev, host, nb))
                    bt.remove(done)
        pool.shutdown(wait=True)
        torch.cuda.synchronize()
        if prof_ticks:
            self.phase_ms = {k: v for k, v in bt.group.prof_collect().items() if v[1]}
        graveyard.clear()
        return results


def _tensor_bytes(d):
    """Bytes of the device tensors a ham_data dict keeps (rotated Cholesky vectors)."""
    n = 0
    for v in d.values():
        for t in (v if isinstance(v, (list, tuple)) else [v]):
            if isinstance(t, torch.Tensor) and t.is_cuda:
                n += t.numel() * t.element_size()
    return n


def latency_stats(results):
    lat = np.array([r.latency for r in results])
    if lat.size == 0:
        return {"p50_s": None, "p99_s": None, "mean_s": None}
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
