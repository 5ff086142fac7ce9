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

from . import batch


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


def estimate_bytes(norb, nocc, nchol, n_walkers, n_prop_steps, precision):
    """Device bytes of one fragment (context, workspace, walker state, fields block, and
    the rotated Cholesky vectors ham_data keeps on the device)."""
    N, n, G, W = norb, nocc, nchol, n_walkers
    b = 8 * N * N * G + 2 * 8 * G * n * N + 16 * G  # chol_t, rot_chol (ctx, ham_data), mf
    if precision != "fp64":
        gp, np32 = -(-G // 32) * 32, -(-N // 32) * 32
        b += 4 * N * N * gp + 4 * G * n * np32 + 4 * W * 2 * n * np32 + 4 * W * 2 * gp
        b += 4 * W * 2 * N * N  # float32 VHS
    else:
        b += 16 * W * N * N
    b += 16 * W * n * N * 3  # walkers, Green, temporaries
    b += 8 * W * 2 * G * 9   # force-bias slices, shifted fields
    b += 8 * n_prop_steps * W * G  # fields block
    return int(b)


class FragmentServer:
    def __init__(self, precision="bf16x3", n_prop_steps=50, hbm_budget=64 << 30,
                 max_active=64, dt=0.005, tune=False, prep_workers=4, max_inflight=2,
                 policy="fifo"):
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

    def _priority(self, req):
        """Admission order among queued requests: arrival ("fifo"), or the estimated
        work of the whole solve first ("sjf": walkers x steps x N^2 (G + 8 n), the
        VHS and Taylor products that dominate a step)."""
        if self.policy == "fifo":
            return (req.arrival, req.rid)
        norb = int(np.asarray(req.ham_data["h1"]).shape[-1])
        nchol = int(np.asarray(req.ham_data["chol"]).shape[0])
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
        retiring = []  # (frag, request, admitted, event, host stats)
        results = []
        graveyard = []  # finished fragments, freed when the GPU is idle
        used = 0
        self.peak_active = 0    # fragments stepping together
        self.peak_reserved = 0  # fragments holding budget (stepping or being built)
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
                for frag, req, adm, ev, host in retiring:
                    if ev.query():
                        be, bw = host[:, 0].numpy().copy(), host[:, 1].numpy().copy()
                        done = ev0.elapsed_time(ev) / 1e3
                        energy = float(np.sum(be * bw) / np.sum(bw))
                        results.append(Result(req.rid, energy, be, bw, req.arrival, adm, done,
                                              frag.walker_steps, frag.build_s))
                        graveyard.append(frag)
                    else:
                        still.append((frag, req, adm, ev, host))
                retiring = still
                # admit: reserve the budget and build on the prep thread, oldest first
                while queue and len(active) + len(building) < self.max_active:
                    req = queue[0][1]
                    nchol = int(np.asarray(req.ham_data["chol"]).shape[0])
                    norb = int(np.asarray(req.ham_data["h1"]).shape[-1])
                    need = estimate_bytes(norb, req.nelec[0], nchol, req.n_walkers,
                                          self.n_prop_steps, self.precision)
                    if (active or building) and used + need > self.hbm_budget:
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
                        active[id(frag)] = (frag, req, time.perf_counter() - t0, need)
                        self.peak_active = max(self.peak_active, len(active))
                        bt.add(frag)
                        joined.append(fut)
                building = [b for b in building if b[0] not in joined]
                if joined and self.tune and len(bt.fragments) > 1:
                    bt.group.tune()
                if not active:
                    if graveyard and not retiring and not building:
                        torch.cuda.synchronize()
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
                bt.run_block()
                tev = torch.cuda.Event()
                tev.record(tick)
                inflight.append(tev)
                done = [f for f, *_ in active.values() if f.blocks_done >= f.n_blocks_total]
                if done:
                    for f in done:
                        _, req, adm, need = active.pop(id(f))
                        used -= need
                        st = torch.stack(f.stats)
                        host = torch.empty(st.shape, dtype=st.dtype, pin_memory=True)
                        host.copy_(st, non_blocking=True)
                        ev = torch.cuda.Event(enable_timing=True)
                        ev.record(tick)
                        retiring.append((f, req, adm, ev, host))
                    bt.remove(done)
        pool.shutdown(wait=True)
        torch.cuda.synchronize()
        graveyard.clear()
        return results


def latency_stats(results):
    lat = np.array([r.latency for r in results])
    if lat.size == 0:
        return {"p50_s": None, "p99_s": None, "mean_s": None}
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max())}
