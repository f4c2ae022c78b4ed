"""Serving barotropic solves: continuous batching under an HBM budget.

A request asks for `steps` RK2 substeps of one field at resolution C{nx}, in
FP64 or FP32, from a given initial (h, u, v).  The server

  * admits waiting requests in arrival order while the device bytes they will
    hold fit the HBM budget: two resident state buffers, a context workspace
    (reused from a retired request of the same kind when one is idle) and, for
    a kind (resolution, precision) not yet resident, its const planes;
  * runs ticks: the admitted requests with the least remaining work (cells x
    substeps) join the tick until it spans `width` cells, enough to fill the
    device, and each advances by min(chunk, remaining) substeps (policy
    "srpt"; policy "all" advances every admitted request), all requests of one
    precision in one batched run
    (``sw_batch_run``: one launch per kernel class and stage, recorded
    hipGraphs of substep pairs), the state staying resident on the device
    between ticks.  FP64 and FP32 groups go to two streams (default priority),
    so their launches overlap on the device;
  * never waits for the device between ticks: each group's last launch records
    an event, the host admits, builds the next tick's tables and launches it
    while the device still runs the previous one, and retires a request when
    the event of its last tick has completed.  Its completion time is that
    event's device timestamp, so latency does not include the host's polling;
  * frees room when the request at the head of the queue does not fit: idle
    workspaces first, then whole kinds with nothing admitted.  A request that
    cannot fit even then (kept results hold the rest of the budget) fails with
    an error instead of waiting forever.

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
from dataclasses import dataclass
from typing import Callable, Deque, Dict, List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import barotr, duogrid, ensemble, initial

SCRATCH_FIELDS = 12  # stage scratch and substep ping-pong of a context workspace
STATE_FIELDS = 6     # two resident (h, u, v) buffers per admitted request
RESULT_FIELDS = 3    # a kept result: one (h, u, v) buffer
CONST_PLANES = 29    # field-sized const planes of a kind: 23 planes + 3 interleaved pair planes of 2
STRIP_PLANES = 11    # planes of the W / E face-edge strip copy (9 columns each)
PLAN_ROT = 12        # rotation coefficients per halo-plan entry

Init = Union[Tuple[torch.Tensor, torch.Tensor, torch.Tensor], Callable[[], Tuple[torch.Tensor, ...]]]


@dataclass
class Request:
    rid: int
    nx: int
    steps: int
    dtype: torch.dtype
    dt: float
    init: Init  # (h, u, v) of shape (6, N, N) on any device, or a callable forming them at admission
    arrival: float = 0.0
    keep_result: bool = True
    t_admit: Optional[float] = None
    t_done: Optional[float] = None
    result: Optional[Tuple[torch.Tensor, ...]] = None
    checksum: Optional[float] = None
    error: Optional[str] = None

    def initial(self):
        return self.init() if callable(self.init) else self.init


def field_bytes(nx: int, dtype: torch.dtype) -> int:
    n = nx + 6
    return 6 * n * n * (8 if dtype == torch.float64 else 4)


def request_bytes(nx: int, dtype: torch.dtype) -> int:
    """Device bytes a request holds while admitted (its own workspace)."""
    return (SCRATCH_FIELDS + STATE_FIELDS) * field_bytes(nx, dtype)


def kind_bytes(nx: int, dtype: torch.dtype) -> int:
    """Device bytes of a kind's consts, held while any of its contexts lives:
    the field-sized planes, the W / E strip copy, the two int32 index planes
    and the halo plan (uint32 destination and sources, two weights and the
    rotation coefficients per entry, each array padded to 256 bytes)."""
    n = nx + 6
    esize = 8 if dtype == torch.float64 else 4
    strip = 12 * STRIP_PLANES * n * 9 * esize
    locs = 2 * 6 * n * n * 4
    entries = 6 * (12 * nx + 36)
    pad = lambda b: (b + 255) // 256 * 256  # noqa: E731
    plan = 3 * pad(4 * entries) + 2 * pad(esize * entries) + PLAN_ROT * entries * esize + 256
    return CONST_PLANES * field_bytes(nx, dtype) + strip + locs + plan


class HBMLedger:
    """Device bytes held under the budget."""

    def __init__(self, budget_bytes: float):
        self.budget = float(budget_bytes)
        self.used = 0.0
        self.peak = 0.0

    def fits(self, nbytes: float) -> bool:
        return self.used + nbytes <= self.budget

    def take(self, nbytes: float) -> None:
        self.used += nbytes
        self.peak = max(self.peak, self.used)

    def give(self, nbytes: float) -> None:
        self.used -= nbytes


@dataclass
class _Kind:
    """Contexts of one kind (nx, dtype), each with a workspace.  They share the
    kind's const planes, which stay allocated while any of them lives."""
    live: List          # every live context; live[0] is cloned for new ones
    idle: List          # live contexts with no request


@dataclass
class _Active:
    req: Request
    ctx: object
    bufs: List
    cur: int = 0
    remaining: int = 0


@dataclass
class _Tick:
    events: List        # one per precision group launched
    finishing: List     # _Active entries whose last substeps this tick runs


# Band heights of the served contexts.  A batch of many members fills the
# device, so taller bands than a lone context's default (fewer rows marched
# twice where bands overlap) are faster: on the 200-request C48-C768 trace,
# 48 rows gave p50 0.95 s / p99 8.6 s at 1.99e10 cell-substeps/s against
# 0.97 s / 8.8 s / 1.95e10 with 32 rows and 0.98 s / 8.9 s / 1.93e10 with 40.
BATCH_ROWS = (48, 48, 48)


class GPUBackend:
    """Device side of the server: contexts, state buffers, batched runs on one
    stream per precision, events."""

    def __init__(self, device=None, rows=BATCH_ROWS, graph: bool = True):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.rows = tuple(rows) if rows is not None else None
        self.graph = graph
        self.streams = {torch.float64: torch.cuda.Stream(self.device), torch.float32: torch.cuda.Stream(self.device)}
        self.batches = {torch.float64: ensemble.Batch(), torch.float32: ensemble.Batch()}
        self._host_consts: Dict[int, tuple] = {}
        self._ref = None

    def host_consts(self, nx: int):
        if nx not in self._host_consts:
            self._host_consts[nx] = duogrid.barotr_consts(duogrid.grid(nx), 5000.0)
        return self._host_consts[nx]

    def new_context(self, nx: int, dtype: torch.dtype, base=None):
        if base is not None:
            return base.like()
        # from host consts: the device copy sw_create reads is temporary
        ctx = barotr.SWContext(self.host_consts(nx), dtype)
        if self.rows is not None:
            ctx.set_tuning(*self.rows)
            ctx.set_band_split(0)  # batched launches use uniform bands
        return ctx

    def free_context(self, ctx) -> None:
        del ctx  # sw_destroy when the last reference goes (waits for its stream)

    def admit(self, ctx, dtype: torch.dtype, init):
        """Two state buffers; the first receives `init` on the precision's stream."""
        s = self.streams[dtype]
        s.wait_stream(torch.cuda.current_stream(self.device))  # the caller's init tensors are ready
        with torch.cuda.stream(s):
            bufs = [ctx.empty(3), ctx.empty(3)]
            x = init() if callable(init) else init
            for k in range(3):
                bufs[0][k].copy_(x[k].to(device=self.device, dtype=dtype), non_blocking=True)
        return bufs

    def run(self, dtype: torch.dtype, items):
        """items: (ctx, (h, u, v) in, (h, u, v) out, substeps, dt); returns an event."""
        s = self.streams[dtype]
        with torch.cuda.stream(s):
            members = [ensemble.Member(ctx, (x[0], x[0], x[1], x[2], x[1], x[2]), n, dt, out=list(y))
                       for ctx, x, y, n, dt in items]
            self.batches[dtype].run(members, diagnostics=False, graph=self.graph)
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(s)
        return ev

    def start(self) -> None:
        """Reference point of event times: the device is idle when it is taken."""
        torch.cuda.synchronize(self.device)
        self._ref = torch.cuda.Event(enable_timing=True)
        self._ref.record(torch.cuda.current_stream(self.device))
        self._ref.synchronize()

    def done(self, ev) -> bool:
        return ev.query()

    def wait(self, ev) -> None:
        ev.synchronize()

    def seconds(self, ev) -> float:
        """Device time of `ev` since start()."""
        return self._ref.elapsed_time(ev) / 1e3

    def finish(self) -> None:
        torch.cuda.synchronize(self.device)


class Server:
    def __init__(self, budget_bytes: float = 200e9, chunk: int = 100, device=None, rows=BATCH_ROWS,
                 backend=None, depth: int = 2, graph: bool = True, policy: str = "srpt",
                 width: float = 2e7):
        if policy not in ("srpt", "all"):
            raise ValueError("policy must be 'srpt' or 'all'")
        self.backend = backend or GPUBackend(device, rows, graph)
        self.budget = float(budget_bytes)
        self.chunk = int(chunk)
        self.policy = policy
        self.width = float(width)  # interior cells a tick spans under "srpt" (C768: 3.5e6)
        self.depth = max(1, int(depth))  # ticks the host may have queued on the device
        self.queue: Deque[Request] = collections.deque()
        self.active: List[_Active] = []
        self.inflight: Deque[_Tick] = collections.deque()
        self.done: List[Request] = []
        self.failed: List[Request] = []
        self.ledger = HBMLedger(budget_bytes)
        self.kinds: Dict[Tuple[int, torch.dtype], _Kind] = {}
        self.ticks = 0
        self.batch_sizes: List[int] = []
        self.cell_substeps = 0
        self.evictions = 0
        self._t0_host = 0.0

    # -- admission ------------------------------------------------------------
    def submit(self, req: Request) -> None:
        if req.steps < 0:
            raise ValueError("steps must be >= 0")
        if req.dtype not in (torch.float64, torch.float32):
            raise ValueError("dtype must be float64 or float32")
        if request_bytes(req.nx, req.dtype) + kind_bytes(req.nx, req.dtype) > self.budget:
            raise ValueError(f"request {req.rid} (C{req.nx}) does not fit the HBM budget")
        self.queue.append(req)

    @property
    def used(self) -> float:
        return self.ledger.used

    @property
    def peak_used(self) -> float:
        return self.ledger.peak

    def _need(self, key) -> float:
        fb = field_bytes(*key)
        k = self.kinds.get(key)
        if k is None:
            return kind_bytes(*key) + (SCRATCH_FIELDS + STATE_FIELDS) * fb
        if k.idle:
            return STATE_FIELDS * fb
        return (SCRATCH_FIELDS + STATE_FIELDS) * fb

    def _drop_context(self, key) -> None:
        """Free one idle context of `key`; the kind's planes go with its last."""
        k = self.kinds[key]
        ctx = k.idle.pop()
        k.live = [c for c in k.live if c is not ctx]
        self.ledger.give(SCRATCH_FIELDS * field_bytes(*key))
        if not k.live:
            self.ledger.give(kind_bytes(*key))
            del self.kinds[key]
        self.backend.free_context(ctx)
        self.evictions += 1

    def _make_room(self, key, need: float) -> bool:
        """Free idle contexts until `need` bytes fit: the largest workspaces
        first, whole kinds (planes too) once their last context goes; the
        request's own kind keeps the idle context it will reuse."""
        while not self.ledger.fits(need):
            cands = [q for q, k in self.kinds.items() if len(k.idle) > (1 if q == key else 0)]
            if not cands:
                return False
            self._drop_context(max(cands, key=lambda q: field_bytes(*q)))
        return True

    def _admit(self, now: float) -> None:
        while self.queue:
            req = self.queue[0]
            key = (req.nx, req.dtype)
            need = self._need(key)
            if not self.ledger.fits(need) and not self._make_room(key, need):
                if self.active or self.inflight:
                    break  # room comes back as admitted requests finish
                # nothing runs and nothing idle is left to free: it never fits
                self.queue.popleft()
                req.error = (f"request {req.rid} (C{req.nx}, {req.dtype}) needs {need / 1e9:.3f} GB with "
                             f"{self.ledger.used / 1e9:.3f} GB of {self.budget / 1e9:.3f} GB held by kept results "
                             "and resident kinds")
                self.failed.append(req)
                continue
            self.queue.popleft()
            self.ledger.take(need)
            k = self.kinds.get(key)
            if k is None:
                ctx = self.backend.new_context(req.nx, req.dtype)
                self.kinds[key] = _Kind(live=[ctx], idle=[])
            elif k.idle:
                ctx = k.idle.pop()
            else:
                ctx = self.backend.new_context(req.nx, req.dtype, base=k.live[0])
                k.live.append(ctx)
            bufs = self.backend.admit(ctx, req.dtype, req.init)
            req.t_admit = now
            a = _Active(req, ctx, bufs, 0, req.steps)
            if req.steps == 0:
                self.inflight.append(_Tick([], [a]))  # nothing to run; retired with the next poll
            else:
                self.active.append(a)

    # -- retirement -------------------------------------------------------------
    def _retire(self, a: _Active, t_done: float) -> None:
        key = (a.req.nx, a.req.dtype)
        fb = field_bytes(*key)
        h, u, v = a.bufs[a.cur]
        a.req.t_done = t_done
        a.req.checksum = float(h[:, 3:-3, 3:-3].double().sum())
        self.ledger.give(STATE_FIELDS * fb)
        if a.req.keep_result:
            a.req.result = (h, u, v)
            self.ledger.take(RESULT_FIELDS * fb)
        self.kinds[key].idle.append(a.ctx)
        self.done.append(a.req)

    def release(self, req: Request) -> None:
        """Drop a kept result and its bytes from the budget."""
        if req.result is not None:
            req.result = None
            self.ledger.give(RESULT_FIELDS * field_bytes(req.nx, req.dtype))

    def _complete(self, t: _Tick) -> None:
        when = max((self.backend.seconds(e) for e in t.events), default=None)
        for a in t.finishing:
            self._retire(a, when if when is not None else self._host_now())
        self.inflight.remove(t)

    def _poll(self) -> None:
        """Retire the requests of every tick the device has finished."""
        for t in list(self.inflight):
            if all(self.backend.done(e) for e in t.events):
                self._complete(t)
            else:
                break  # ticks finish in order

    def _wait_oldest(self) -> None:
        t = self.inflight[0]
        for e in t.events:
            self.backend.wait(e)
        self._complete(t)

    # -- one tick ---------------------------------------------------------------
    def plan(self) -> List[Tuple[_Active, int]]:
        """The admitted requests this tick advances and by how many substeps
        (min(chunk, remaining) each).  "all": every one (processor sharing).
        "srpt": in order of least remaining work (cells x remaining substeps;
        arrival breaks ties) until the tick spans `width` interior cells, so
        its launches still fill the device.  A request with less left than the
        others therefore finishes first, which shortens the median latency
        without narrowing the launches; under sustained overload the largest
        requests wait for smaller ones."""
        if self.policy == "all":
            return [(a, min(self.chunk, a.remaining)) for a in self.active]
        cells = lambda a: 6 * a.req.nx * a.req.nx  # noqa: E731
        order = sorted(self.active, key=lambda a: (a.remaining * cells(a), a.req.arrival, a.req.rid))
        out, span = [], 0.0
        for a in order:
            if out and span >= self.width:
                break
            out.append((a, min(self.chunk, a.remaining)))
            span += cells(a)
        return out

    def tick(self, clock: Callable[[], float]) -> bool:
        """Retire finished work, admit what fits, launch the planned substeps
        of the admitted requests (one batched run per precision) and return
        without waiting for them; True if anything was launched."""
        self._poll()
        self._admit(clock())
        if not self.active:
            return False
        events, finishing = [], []
        plan = self.plan()
        for dt_ in (torch.float64, torch.float32):
            group = [(a, n) for a, n in plan if a.req.dtype == dt_]
            if not group:
                continue
            items = []
            for a, n in group:
                items.append((a.ctx, a.bufs[a.cur], a.bufs[1 - a.cur], n, a.req.dt))
                self.cell_substeps += 6 * a.req.nx * a.req.nx * n
            events.append(self.backend.run(dt_, items))
            for a, n in group:
                a.cur = 1 - a.cur
                a.remaining -= n
            self.batch_sizes.append(len(group))
        finishing = [a for a in self.active if a.remaining <= 0]
        self.active = [a for a in self.active if a.remaining > 0]
        self.inflight.append(_Tick(events, finishing))
        self.ticks += 1
        while len(self.inflight) > self.depth:
            self._wait_oldest()
        return True

    def _host_now(self) -> float:
        return time.perf_counter() - self._t0_host

    def run_trace(self, requests: Sequence[Request], clock: Callable[[], float] = time.perf_counter):
        """Replay `requests` (arrival times in seconds from the start) in real
        time; returns the completed requests when all have completed or
        failed (``failed`` lists those that could never fit)."""
        pending = sorted(requests, key=lambda r: r.arrival)
        self.backend.start()
        t0 = clock()
        self._t0_host = time.perf_counter()
        rel = lambda: clock() - t0  # noqa: E731
        i = 0
        while i < len(pending) or self.queue or self.active or self.inflight:
            now = rel()
            while i < len(pending) and pending[i].arrival <= now:
                self.submit(pending[i])
                i += 1
            if self.queue or self.active:
                if self.tick(rel):
                    continue
            if self.inflight:
                self._wait_oldest()  # nothing could be launched: room or results come from finished work
                continue
            if self.queue:
                continue  # the next admission pass frees room or fails the head request
            if i < len(pending):
                time.sleep(max(0.0, min(pending[i].arrival - rel(), 0.01)))
        self.backend.finish()
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
        h, u, v = (x.to(device=pool.device, dtype=dtype) for x in r.initial())
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
    linear, so the scaled field stays balanced) with a Courant dt.  The scaled
    field is formed on the device at admission, so a queued request holds no
    device memory; the base fields (three FP64 fields per resolution) belong
    to the trace."""
    device = device or torch.device("cuda", torch.cuda.current_device())
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=n))
    nxs = rng.choice(resolutions, size=n)
    lo, hi = math.log(steps_range[0]), math.log(steps_range[1])
    steps = np.exp(rng.uniform(lo, hi, size=n)).astype(int)
    fp32 = rng.uniform(size=n) < fp32_share
    scale = 1.0 + 0.01 * rng.normal(size=n)
    base = {}
    out = []
    for k in range(n):
        nx = int(nxs[k])
        if nx not in base:
            g = duogrid.grid(nx)
            h, u, v, sp = initial.seeded_balanced(g)
            dt = initial.courant_dt(g, h, sp)
            base[nx] = ([torch.as_tensor(x).to(device) for x in (h, u, v)], dt)
        fields, dt = base[nx]
        f = float(scale[k])
        init = (lambda fields=fields, f=f: tuple(x * f for x in fields))
        out.append(Request(rid=k, nx=nx, steps=int(steps[k]),
                           dtype=torch.float32 if fp32[k] else torch.float64,
                           dt=dt / max(f, 1.0), init=init, arrival=float(t[k]), keep_result=keep_result))
    torch.cuda.synchronize(device)
    return out


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--requests", type=int, default=200)
    p.add_argument("--rate", type=float, default=200.0)
    p.add_argument("--res", type=int, nargs="+", default=[48, 96, 192, 384, 768])
    p.add_argument("--steps-min", type=int, default=50)
    p.add_argument("--steps-max", type=int, default=5000)
    p.add_argument("--budget-gb", type=float, default=200.0)
    p.add_argument("--chunk", type=int, default=400)
    p.add_argument("--policy", choices=("srpt", "all"), default="srpt",
                   help="srpt: least remaining work first until a tick spans --width cells; all: every "
                        "admitted request advances each tick")
    p.add_argument("--width", type=float, default=2e7)
    p.add_argument("--seed", type=int, default=7)
    p.add_argument("--no-baseline", action="store_true",
                   help="skip the one-request-at-a-time comparison runs")
    a = p.parse_args(argv)
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    line = serve_report(a.requests, a.rate, a.res, (a.steps_min, a.steps_max), a.budget_gb, a.chunk,
                        a.seed, dev, baselines=() if a.no_baseline else ("mixed", "fp64"), policy=a.policy,
                        width=a.width)
    print(json.dumps(line), flush=True)
    return 0 if line["finite"] else 1


def serve_report(requests: int, rate: float, res: Sequence[int], steps_range, budget_gb: float, chunk: int,
                 seed: int, dev, baselines=("mixed", "fp64"), policy: str = "srpt", width: float = 2e7) -> dict:
    """Serve a seeded Poisson trace on a warm server and report latency and
    throughput; `baselines` adds the plain path one request at a time with
    each request's own precision ("mixed") and all in FP64 ("fp64")."""
    reqs = seeded_trace(requests, rate, res, steps_range, seed=seed, device=dev)
    # make each kind resident and load the kernels before the timed trace (a
    # warm service), then reset the statistics
    srv = Server(budget_gb * 1e9, chunk, dev, policy=policy, width=width)
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
        "ticks": srv.ticks, "chunk": chunk, "policy": policy, "width": width,
        "mean_batch": float(np.mean(srv.batch_sizes)) if srv.batch_sizes else 0.0,
        "max_batch": int(max(srv.batch_sizes)) if srv.batch_sizes else 0,
        "peak_hbm_gb": srv.peak_used / 1e9, "budget_gb": budget_gb,
        "failed": len(srv.failed),
        "finite": bool(all(math.isfinite(r.checksum) for r in done)) and not srv.failed,
    }
    # the plain path runs each request in a context of the library's default
    # band heights (chosen for one context alone on the device)
    plain = ensemble.GridPool(dev)
    if "mixed" in baselines:
        line["one_at_a_time"] = one_at_a_time(reqs, plain)
    if "fp64" in baselines:
        line["one_at_a_time_fp64"] = one_at_a_time(reqs, plain, force_fp64=True)
    return line


if __name__ == "__main__":
    sys.exit(main())
