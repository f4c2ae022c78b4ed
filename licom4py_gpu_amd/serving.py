"""Serving barotropic solves: continuous batching under an HBM budget.

A request asks for `steps` RK2 substeps of one field at resolution C{nx}, in
FP64 or FP32, from a given initial (h, u, v).  The server

  * admits waiting requests in arrival order while the device bytes they will
    hold (their workspace and two resident state buffers, plus the consts of a
    resolution not yet resident) fit the HBM budget;
  * runs ticks: every admitted request advances by min(chunk, remaining)
    substeps, all requests of one precision in one batched run
    (``sw_batch_run``: one launch per kernel class and stage), the state staying
    resident on the device between ticks;
  * retires finished requests and admits newcomers between ticks, so the batch
    composition changes continuously.

Each request's result is bit-identical to the same solve run alone (substeps
are deterministic functions of the state, whatever batch they run in).
Latency = completion - arrival, queueing included.  This is the caller one
step up from ``schedule_core`` (mymodule/schedule.py:130-313: a chunked loop
over the compiled core) turned into a multi-tenant loop.

    python -m licom4py_gpu_amd.serving --requests 200 --rate 200 --budget-gb 200

replays a seeded Poisson trace (C in {48, ..., 768} uniform, steps
log-uniform in [50, 5000], FP32 / FP64 half each) and prints one JSON line with
p50 / p99 latency and the throughput of the busy period.
"""

from __future__ import annotations

import argparse
import collections
import json
import math
import sys
import time
from dataclasses import dataclass, field
from typing import Callable, Deque, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import barotr, ensemble, initial

SCRATCH_FIELDS = 12  # stage scratch and substep ping-pong of a context workspace
STATE_FIELDS = 6     # two resident (h, u, v) buffers per request
CONST_PLANES = 26    # context planes of a resolution (23) + index tables + halo plan, rounded up


@dataclass
class Request:
    rid: int
    nx: int
    steps: int
    dtype: torch.dtype
    dt: float
    init: Tuple[torch.Tensor, torch.Tensor, torch.Tensor]  # (h, u, v), (6, N, N), any device
    arrival: float = 0.0
    keep_result: bool = True
    t_admit: Optional[float] = None
    t_done: Optional[float] = None
    result: Optional[Tuple[torch.Tensor, ...]] = None
    checksum: Optional[float] = None


def field_bytes(nx: int, dtype: torch.dtype) -> int:
    n = nx + 6
    return 6 * n * n * (8 if dtype == torch.float64 else 4)


def request_bytes(nx: int, dtype: torch.dtype) -> int:
    """Device bytes a request holds while admitted."""
    return (SCRATCH_FIELDS + STATE_FIELDS) * field_bytes(nx, dtype)


class HBMLedger:
    """Device bytes held under the budget, per kind (nx, dtype) of request:
    admitted requests hold a workspace and two state buffers; a retired
    request's workspace stays idle for the next request of its kind until
    room is needed; a kind's consts stay resident once loaded."""

    def __init__(self, budget_bytes: float):
        self.budget = float(budget_bytes)
        self.used = 0.0
        self.peak = 0.0
        self.resident = set()
        self.idle: Dict[tuple, int] = collections.defaultdict(int)

    def need(self, key) -> float:
        """Bytes admitting a request of `key` adds now."""
        if self.idle[key]:
            return STATE_FIELDS * field_bytes(*key)
        extra = 0 if key in self.resident else CONST_PLANES * field_bytes(*key)
        return request_bytes(*key) + extra

    def fits(self, key) -> bool:
        return self.used + self.need(key) <= self.budget

    def admit(self, key) -> bool:
        """Account an admission; True if it reuses an idle workspace."""
        self.used += self.need(key)
        self.peak = max(self.peak, self.used)
        self.resident.add(key)
        if self.idle[key]:
            self.idle[key] -= 1
            return True
        return False

    def retire(self, key) -> None:
        self.used -= STATE_FIELDS * field_bytes(*key)
        self.idle[key] += 1

    def drop_idle(self):
        """Release the largest idle workspace; returns its kind, or None."""
        keys = [k for k, n in self.idle.items() if n]
        if not keys:
            return None
        k = max(keys, key=lambda k: field_bytes(*k))
        self.idle[k] -= 1
        self.used -= SCRATCH_FIELDS * field_bytes(*k)
        return k


@dataclass
class _Active:
    req: Request
    ctx: "barotr.SWContext"
    bufs: List[List[torch.Tensor]]
    cur: int = 0
    remaining: int = 0


# Band heights of the served contexts.  A batch of many members fills the
# device, so taller bands than a lone context's default (fewer rows marched
# twice where bands overlap) are faster.
BATCH_ROWS = (32, 32, 32)


class Server:
    def __init__(self, budget_bytes: float = 200e9, chunk: int = 100, device=None, rows=BATCH_ROWS):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.budget = float(budget_bytes)
        self.chunk = int(chunk)
        self.pool = ensemble.GridPool(self.device, rows=rows)
        self.batches = {torch.float64: ensemble.Batch(), torch.float32: ensemble.Batch()}
        self.queue: Deque[Request] = collections.deque()
        self.active: List[_Active] = []
        self.done: List[Request] = []
        self.ledger = HBMLedger(budget_bytes)
        # retired requests' contexts, reused by later requests of the same kind
        self._idle: Dict[Tuple[int, torch.dtype], List["barotr.SWContext"]] = collections.defaultdict(list)
        self.ticks = 0
        self.batch_sizes: List[int] = []
        self.cell_substeps = 0

    # -- admission ------------------------------------------------------------
    def submit(self, req: Request) -> None:
        if req.steps < 0:
            raise ValueError("steps must be >= 0")
        if req.dtype not in (torch.float64, torch.float32):
            raise ValueError("dtype must be float64 or float32")
        need = request_bytes(req.nx, req.dtype) + CONST_PLANES * field_bytes(req.nx, req.dtype)
        if need > self.budget:
            raise ValueError(f"request {req.rid} (C{req.nx}) does not fit the HBM budget")
        self.queue.append(req)

    @property
    def used(self) -> float:
        return self.ledger.used

    @property
    def peak_used(self) -> float:
        return self.ledger.peak

    def _admit(self, now: float) -> None:
        while self.queue:
            req = self.queue[0]
            key = (req.nx, req.dtype)
            if not self.ledger.fits(key):
                dropped = self.ledger.drop_idle()
                if dropped is None:
                    break
                self._idle[dropped].pop()  # its workspace is freed with the context
                continue
            self.queue.popleft()
            reuse = self.ledger.admit(key)
            ctx = self._idle[key].pop() if reuse else self.pool.context(req.nx, req.dtype)
            bufs = [ctx.empty(3), ctx.empty(3)]
            for k in range(3):
                bufs[0][k].copy_(req.init[k].to(device=self.device, dtype=req.dtype))
            req.t_admit = now
            self.active.append(_Active(req, ctx, bufs, 0, req.steps))

    # -- one tick ---------------------------------------------------------------
    def _retire(self, a: _Active, now: float) -> None:
        h, u, v = a.bufs[a.cur]
        a.req.t_done = now
        a.req.checksum = float(h[:, 3:-3, 3:-3].double().sum())
        a.req.result = (h, u, v) if a.req.keep_result else None
        self.ledger.retire((a.req.nx, a.req.dtype))  # the workspace stays, idle
        self._idle[(a.req.nx, a.req.dtype)].append(a.ctx)
        self.done.append(a.req)

    def tick(self, clock: Callable[[], float]) -> None:
        """Admit what fits, advance every admitted request by up to `chunk`
        substeps, retire the finished ones."""
        self._admit(clock())
        if not self.active:
            return
        for dt_, batch in self.batches.items():
            group = [a for a in self.active if a.req.dtype == dt_ and a.remaining > 0]
            if not group:
                continue
            members = []
            for a in group:
                h, u, v = a.bufs[a.cur]
                n = min(self.chunk, a.remaining)
                members.append(ensemble.Member(a.ctx, (h, h, u, v, u, v), n, a.req.dt,
                                               out=a.bufs[1 - a.cur]))
                self.cell_substeps += 6 * a.req.nx * a.req.nx * n
            batch.run(members, diagnostics=False)
            for a, m in zip(group, members):
                a.cur = 1 - a.cur
                a.remaining -= m.nbb
            self.batch_sizes.append(len(group))
        torch.cuda.synchronize(self.device)
        now = clock()
        self.ticks += 1
        still = []
        for a in self.active:
            if a.remaining <= 0:
                self._retire(a, now)
            else:
                still.append(a)
        self.active = still

    def run_trace(self, requests: Sequence[Request], clock: Callable[[], float] = time.perf_counter):
        """Replay `requests` (arrival times in seconds from the start) in real
        time; returns when all have completed."""
        pending = sorted(requests, key=lambda r: r.arrival)
        t0 = clock()
        rel = lambda: clock() - t0  # noqa: E731
        i = 0
        while i < len(pending) or self.queue or self.active:
            now = rel()
            while i < len(pending) and pending[i].arrival <= now:
                self.submit(pending[i])
                i += 1
            if not self.active and not self.queue:
                if i < len(pending):
                    time.sleep(max(0.0, min(pending[i].arrival - rel(), 0.01)))
                continue
            self.tick(rel)
        return self.done


def latency_stats(done: Sequence[Request]) -> dict:
    lat = np.array([r.t_done - r.arrival for r in done])
    wait = np.array([r.t_admit - r.arrival for r in done])
    return {"n": int(lat.size), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()),
            "admission_wait_p50_s": float(np.percentile(wait, 50)),
            "admission_wait_p99_s": float(np.percentile(wait, 99))}


def fcfs_latencies(arrivals: Sequence[float], service: Sequence[float]):
    """Latencies of a single first-come-first-served server (arrivals sorted)
    and the time it finishes."""
    end = 0.0
    lat = []
    for t, s in zip(arrivals, service):
        end = max(end, t) + s
        lat.append(end - t)
    return np.array(lat), end


def one_at_a_time(requests: Sequence[Request], pool: "ensemble.GridPool", force_fp64: bool = False):
    """The plain path as the comparison: every request run alone (sw_rk2 with
    recorded substep graphs) in arrival order.  Service times are measured;
    latencies follow from a first-come-first-served single queue."""
    requests = sorted(requests, key=lambda r: r.arrival)
    service = []
    ctxs: Dict[Tuple[int, torch.dtype], "barotr.SWContext"] = {}
    for r in requests:
        dtype = torch.float64 if force_fp64 else r.dtype
        key = (r.nx, dtype)
        if key not in ctxs:
            ctxs[key] = pool.context(r.nx, dtype)
        h, u, v = (x.to(device=pool.device, dtype=dtype) for x in r.init)
        torch.cuda.synchronize(pool.device)
        t0 = time.perf_counter()
        ctxs[key].rk((h, h, u, v, u, v), r.steps, r.dt, diagnostics=False, graph=True)
        torch.cuda.synchronize(pool.device)
        service.append(time.perf_counter() - t0)
    lat, end = fcfs_latencies([r.arrival for r in requests], service)
    work = sum(6 * r.nx * r.nx * r.steps for r in requests)
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "makespan_s": float(end), "busy_s": float(sum(service)),
            "throughput_cell_substeps_per_s": work / sum(service)}


def seeded_trace(n: int, rate: float, resolutions: Sequence[int], steps_range=(50, 5000),
                 fp32_share: float = 0.5, seed: int = 7, device=None, keep_result: bool = False):
    """Poisson arrivals of the seeded balanced recipe: each request scales one
    per-resolution base field (h, u, v) by 1 + 0.01 z (geostrophic balance is
    linear, so the scaled field stays balanced) with a Courant dt."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    nxs = rng.choice(resolutions, size=n)
    lo, hi = math.log(steps_range[0]), math.log(steps_range[1])
    steps = np.exp(rng.uniform(lo, hi, size=n)).astype(int)
    fp32 = rng.uniform(size=n) < fp32_share
    scale = 1.0 + 0.01 * rng.normal(size=n)
    base = {}
    pool = ensemble.GridPool(device)
    out = []
    for k in range(n):
        nx = int(nxs[k])
        if nx not in base:
            g, _ = pool.consts(nx)
            h, u, v, sp = initial.seeded_balanced(g)
            dt = initial.courant_dt(g, h, sp)
            base[nx] = ([torch.as_tensor(x).to(device) for x in (h, u, v)], dt)
        (h, u, v), dt = base[nx]
        f = float(scale[k])
        init = (h * f, u * f, v * f)
        out.append(Request(rid=k, nx=nx, steps=int(steps[k]),
                           dtype=torch.float32 if fp32[k] else torch.float64,
                           dt=dt / max(f, 1.0), init=init, arrival=float(t[k]), keep_result=keep_result))
    return out


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--requests", type=int, default=200)
    p.add_argument("--rate", type=float, default=200.0)
    p.add_argument("--res", type=int, nargs="+", default=[48, 96, 192, 384, 768])
    p.add_argument("--steps-min", type=int, default=50)
    p.add_argument("--steps-max", type=int, default=5000)
    p.add_argument("--budget-gb", type=float, default=200.0)
    p.add_argument("--chunk", type=int, default=100)
    p.add_argument("--seed", type=int, default=7)
    p.add_argument("--no-baseline", action="store_true",
                   help="skip the one-request-at-a-time comparison runs")
    a = p.parse_args(argv)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    line = serve_report(a.requests, a.rate, a.res, (a.steps_min, a.steps_max), a.budget_gb, a.chunk,
                        a.seed, dev, baselines=() if a.no_baseline else ("mixed", "fp64"))
    print(json.dumps(line), flush=True)
    return 0 if line["finite"] else 1


def serve_report(requests: int, rate: float, res: Sequence[int], steps_range, budget_gb: float, chunk: int,
                 seed: int, dev, baselines=("mixed", "fp64")) -> dict:
    """Serve a seeded Poisson trace on a warm server and report latency and
    throughput; `baselines` adds the plain path one request at a time with
    each request's own precision ("mixed") and all in FP64 ("fp64")."""
    reqs = seeded_trace(requests, rate, res, steps_range, seed=seed, device=dev)
    # make each resolution's consts resident and load the kernels before the
    # timed trace (a warm service), then reset the statistics
    srv = Server(budget_gb * 1e9, chunk, dev)
    warm = []
    for dt_ in (torch.float64, torch.float32):
        for nx in sorted(set(r.nx for r in reqs)):
            r0 = next(r for r in reqs if r.nx == nx)
            warm.append(Request(-1, nx, 2, dt_, r0.dt, r0.init, keep_result=False))
    srv.run_trace(warm)
    srv.done, srv.ticks, srv.batch_sizes, srv.cell_substeps = [], 0, [], 0
    t0 = time.perf_counter()
    done = srv.run_trace(reqs)
    makespan = time.perf_counter() - t0
    work = sum(6 * r.nx * r.nx * r.steps for r in reqs)
    line = {
        "workload": f"{requests} requests, Poisson {rate}/s, C in {sorted(set(res))}, steps log-uniform "
                    f"[{steps_range[0]}, {steps_range[1]}], FP32/FP64 half each, RK2",
        "offered_cell_substeps_per_s": work / (reqs[-1].arrival if reqs else 1.0),
        "latency": latency_stats(done),
        "makespan_s": makespan,
        "throughput_cell_substeps_per_s": work / makespan,
        "ticks": srv.ticks, "chunk": chunk,
        "mean_batch": float(np.mean(srv.batch_sizes)) if srv.batch_sizes else 0.0,
        "max_batch": int(max(srv.batch_sizes)) if srv.batch_sizes else 0,
        "peak_hbm_gb": srv.peak_used / 1e9, "budget_gb": budget_gb,
        "finite": bool(all(math.isfinite(r.checksum) for r in done)),
    }
    # the plain path runs each request in a context of the library's default
    # band heights (chosen for one context alone on the device)
    plain = srv.pool.sibling()
    if "mixed" in baselines:
        line["one_at_a_time"] = one_at_a_time(reqs, plain)
    if "fp64" in baselines:
        line["one_at_a_time_fp64"] = one_at_a_time(reqs, plain, force_fp64=True)
    return line


if __name__ == "__main__":
    sys.exit(main())
