"""Host-side logic of the serving loop (CPU): HBM accounting, admission,
eviction, failure of requests that can never fit, retirement from events, and
the FCFS comparison model.  A stub backend stands in for the device: a
"substep" adds 1 to every value of the state, and events complete when the
test says so."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from licom4py_gpu_amd import serving
from licom4py_gpu_amd.serving import (CONST_PLANES, RESULT_FIELDS, SCRATCH_FIELDS, STATE_FIELDS, HBMLedger,
                                      field_bytes, kind_bytes, request_bytes)


class _Event:
    def __init__(self, backend):
        self.backend = backend
        self.t = None  # device time once complete

    def ready(self):
        return self.t is not None


class StubBackend:
    """Contexts are tokens, buffers are small CPU tensors, a batched run adds
    the substep count to the state; events complete immediately unless
    `hold` is set (then on wait / release_all)."""

    def __init__(self, hold=False):
        self.hold = hold
        self.created = []   # (nx, dtype, cloned)
        self.freed = 0
        self.pending = []
        self.clock = 0.0
        self.runs = []

    def new_context(self, nx, dtype, base=None):
        ctx = {"nx": nx, "dtype": dtype, "id": len(self.created)}
        self.created.append((nx, dtype, base is not None))
        return ctx

    def free_context(self, ctx):
        self.freed += 1

    def admit(self, ctx, dtype, init):
        x = init() if callable(init) else init
        n = ctx["nx"] + 6
        bufs = [[torch.zeros((6, n, n), dtype=dtype) for _ in range(3)] for _ in range(2)]
        for k in range(3):
            bufs[0][k].copy_(x[k].to(dtype))
        return bufs

    def run(self, dtype, items):
        self.runs.append(len(items))
        for ctx, x, y, n, dt in items:
            for k in range(3):
                y[k].copy_(x[k] + n)
        ev = _Event(self)
        self.clock += 1.0
        if self.hold:
            self.pending.append(ev)
        else:
            ev.t = self.clock
        return ev

    def start(self):
        pass

    def done(self, ev):
        return ev.ready()

    def wait(self, ev):
        if ev.t is None:
            ev.t = self.clock
            self.pending.remove(ev)

    def seconds(self, ev):
        return ev.t

    def finish(self):
        for ev in list(self.pending):
            self.wait(ev)


def _req(rid, nx, steps, dtype=torch.float64, keep=True, arrival=0.0, val=0.0):
    n = nx + 6
    init = tuple(torch.full((6, n, n), val + k, dtype=torch.float64) for k in range(3))
    return serving.Request(rid, nx, steps, dtype, 1.0, init, arrival=arrival, keep_result=keep)


def test_field_bytes():
    assert field_bytes(48, torch.float64) == 6 * 54 * 54 * 8
    assert field_bytes(48, torch.float32) == 6 * 54 * 54 * 4
    assert request_bytes(48, torch.float64) == (SCRATCH_FIELDS + STATE_FIELDS) * field_bytes(48, torch.float64)
    # the consts' device allocations: planes, W / E strip copy, index planes, halo plan
    n, entries = 54, 6 * (12 * 48 + 36)
    strip, locs = 12 * 11 * n * 9 * 4, 2 * 6 * n * n * 4
    plan = 3 * ((4 * entries + 255) // 256 * 256) + 2 * ((4 * entries + 255) // 256 * 256) + 12 * entries * 4 + 256
    assert kind_bytes(48, torch.float32) == CONST_PLANES * field_bytes(48, torch.float32) + strip + locs + plan


def test_ledger():
    L = HBMLedger(100.0)
    assert L.fits(100.0) and not L.fits(100.5)
    L.take(60.0)
    L.take(30.0)
    L.give(50.0)
    assert L.used == 40.0 and L.peak == 90.0


def test_results_are_the_substeps_and_bytes_return():
    b = StubBackend()
    srv = serving.Server(1e12, chunk=4, backend=b, policy="all")
    reqs = [_req(0, 8, 10), _req(1, 8, 3), _req(2, 12, 0), _req(3, 12, 7, torch.float32)]
    done = srv.run_trace(reqs)
    assert sorted(r.rid for r in done) == [0, 1, 2, 3] and not srv.failed
    for r in done:
        for k in range(3):
            assert torch.all(r.result[k][:, 3:-3, 3:-3] == k + r.steps)
    # chunks of 4: request 0 runs 4 + 4 + 2 substeps, request 1 one tick of 3
    assert srv.ticks == 3
    # what stays held: the kinds' planes and workspaces, and the kept results
    fb8, fb12, fb12s = field_bytes(8, torch.float64), field_bytes(12, torch.float64), field_bytes(12, torch.float32)
    held = (kind_bytes(8, torch.float64) + kind_bytes(12, torch.float64) + kind_bytes(12, torch.float32)
            + 2 * SCRATCH_FIELDS * fb8 + SCRATCH_FIELDS * (fb12 + fb12s)
            + RESULT_FIELDS * (2 * fb8 + fb12 + fb12s))
    assert srv.used == pytest.approx(held)
    for r in done:
        srv.release(r)
    assert srv.used == pytest.approx(held - RESULT_FIELDS * (2 * fb8 + fb12 + fb12s))


def test_idle_workspace_is_reused():
    b = StubBackend()
    srv = serving.Server(1e12, chunk=8, backend=b, policy="all")
    srv.run_trace([_req(0, 8, 2), _req(1, 8, 2, arrival=0.0)])
    srv.run_trace([_req(2, 8, 2)])
    # two contexts of C8 (the second cloned), the third request reuses one
    assert b.created == [(8, torch.float64, False), (8, torch.float64, True)]


def test_head_request_evicts_unused_kinds_instead_of_stalling():
    """The stall the ledger used to allow: consts of other kinds resident,
    nothing admitted and nothing idle of the head request's kind.  The server
    must free the unused kinds and admit it."""
    b = StubBackend()
    big = (48, torch.float64)
    fb = field_bytes(*big)
    budget = kind_bytes(*big) + request_bytes(*big)  # exactly one C48 FP64 kind and request
    srv = serving.Server(budget, chunk=4, backend=b, policy="all")
    srv.run_trace([_req(0, 16, 2, keep=False), _req(1, 8, 2, torch.float32, keep=False)])
    assert srv.used > 0 and len(srv.kinds) == 2     # two small kinds resident and idle
    assert not srv.ledger.fits(kind_bytes(*big) + request_bytes(*big))
    done = srv.run_trace([_req(2, 48, 3, keep=False)])
    assert [r.rid for r in done[-1:]] == [2] and not srv.failed
    assert srv.evictions == 2 and list(srv.kinds) == [big]
    assert srv.used == pytest.approx(kind_bytes(*big) + SCRATCH_FIELDS * fb)
    assert srv.peak_used <= budget


def test_request_that_can_never_fit_fails_instead_of_looping():
    """Kept results hold most of the budget: with nothing running and nothing
    idle to free, the head request fails with an error and the trace ends."""
    b = StubBackend()
    k = (16, torch.float64)
    fb = field_bytes(*k)
    budget = kind_bytes(*k) + request_bytes(*k) + 2 * RESULT_FIELDS * fb
    srv = serving.Server(budget, chunk=4, backend=b, policy="all")
    first = srv.run_trace([_req(0, 16, 1), _req(1, 16, 1, arrival=0.0)])
    assert len(first) == 2  # two kept results
    # the kind and its workspaces can be freed, but the results stay: a C16
    # request needs kind + workspace + state, more than what is left
    done = srv.run_trace([_req(2, 16, 1), _req(3, 16, 1)])
    assert [r.rid for r in done[2:]] == [2]
    assert [r.rid for r in srv.failed] == [3] and "GB" in srv.failed[0].error
    for r in list(srv.done):
        srv.release(r)
    srv.failed.clear()
    srv.run_trace([_req(4, 16, 1)])
    assert srv.done[-1].rid == 4 and not srv.failed


def test_ticks_run_ahead_of_retirement():
    """The host launches the next tick before the device finishes the last
    one (up to `depth` ticks queued); a request is retired, and its bytes
    returned, only when the event of its last tick has completed, with that
    event's time as its completion time."""
    b = StubBackend(hold=True)
    srv = serving.Server(1e12, chunk=2, backend=b, depth=2, policy="all")
    srv.submit(_req(0, 8, 2))
    srv.submit(_req(1, 8, 6))
    clock = iter(np.arange(0.0, 100.0, 0.5))
    assert srv.tick(lambda: next(clock))          # tick 1: both requests; request 0 finishes here
    assert len(srv.inflight) == 1 and srv.done == []
    assert srv.tick(lambda: next(clock))          # tick 2 queued behind tick 1, nothing waited for
    assert len(srv.inflight) == 2 and srv.done == [] and b.pending
    srv.tick(lambda: next(clock))                 # tick 3: over the depth, tick 1 is waited for
    assert [r.rid for r in srv.done] == [0]
    assert srv.done[0].t_done == 3.0              # the event's time (stub clock at the wait)


def test_fcfs_latencies():
    lat, end = serving.fcfs_latencies([0.0, 0.5, 3.0], [1.0, 1.0, 1.0])
    assert np.allclose(lat, [1.0, 1.5, 1.0])
    assert end == 4.0


def test_latency_stats():
    reqs = [serving.Request(k, 8, 1, torch.float64, 1.0, (None, None, None), arrival=float(k),
                            t_admit=float(k), t_done=float(k) + 1 + k) for k in range(10)]
    st = serving.latency_stats(reqs)
    assert st["n"] == 10 and st["max_s"] == 10.0 and st["p50_s"] == pytest.approx(5.5)


def test_global_diag_and_format():
    from licom4py_gpu_amd.schedule import format_diag, global_diag
    h = np.zeros((6, 10, 10))
    h[:, 3:-3, 3:-3] = np.arange(6 * 16).reshape(6, 4, 4)
    h[0, 0, 0] = 1e9  # halo values are ignored
    st = global_diag(h, -h, h * 2)
    assert st["h0"] == (95.0, 0.0, 47.5)
    assert st["ub"] == (0.0, -95.0, -47.5)
    text = format_diag("Final time", st)
    assert text.splitlines()[0] == "Final time"
    assert "Global h0:" in text and "47.5000000000000000" in text


def test_surplus_idle_workspaces_of_the_same_kind_are_freed():
    b = StubBackend()
    k = (16, torch.float64)
    fb = field_bytes(*k)
    srv = serving.Server(1e12, chunk=4, backend=b, policy="all")
    srv.run_trace([_req(i, 16, 1, keep=False) for i in range(3)])  # three contexts, all idle now
    assert len(srv.kinds[k].idle) == 3
    srv.ledger.budget = srv.used + STATE_FIELDS * fb - 1.0  # one byte short of admitting into an idle one
    srv.run_trace([_req(9, 16, 1, keep=False)])
    assert srv.done[-1].rid == 9 and not srv.failed and srv.evictions == 1


def _active(srv, rid, nx, remaining, arrival=0.0, dtype=torch.float64):
    a = serving._Active(req=serving.Request(rid, nx, remaining, dtype, 1.0, (None, None, None), arrival=arrival),
                        ctx=None, bufs=[None, None], remaining=remaining)
    srv.active.append(a)
    return a


def test_srpt_plan_least_remaining_work_first_until_the_width():
    """Requests join the tick in order of cells x remaining substeps until it
    spans `width` cells; each advances min(chunk, remaining) substeps."""
    srv = serving.Server(1e12, chunk=10, backend=StubBackend(), width=6 * 16 * 16 + 6 * 8 * 8)
    _active(srv, 0, 48, 1000)              # 6*48^2 cells x 1000: the most work
    _active(srv, 1, 8, 5, arrival=1.0)     # 6*8^2 cells x 5: the least
    _active(srv, 2, 16, 500, arrival=2.0)  # 6*16^2 cells x 500
    _active(srv, 3, 8, 700, arrival=3.0)   # 6*8^2 cells x 700
    plan = srv.plan()
    assert [(a.req.rid, n) for a, n in plan] == [(1, 5), (3, 10), (2, 10)]  # width reached: request 0 waits
    srv.width = 1.0
    assert [(a.req.rid, n) for a, n in srv.plan()] == [(1, 5)]  # at least one request per tick


def test_all_policy_takes_every_request_and_policy_is_checked():
    srv = serving.Server(1e12, chunk=5, backend=StubBackend(), policy="all")
    for k in range(4):
        _active(srv, k, 16, 100 if k else 3, arrival=float(k))
    assert [(a.req.rid, n) for a, n in srv.plan()] == [(0, 3), (1, 5), (2, 5), (3, 5)]
    with pytest.raises(ValueError, match="policy"):
        serving.Server(1e12, backend=StubBackend(), policy="fifo")


def test_srpt_trace_results_and_small_requests_first():
    """Under "srpt" every result is still its own substeps, and of requests
    that arrive together the ones with less work finish first."""
    b = StubBackend()
    srv = serving.Server(1e12, chunk=2, backend=b, width=6 * 12 * 12)
    reqs = [_req(0, 12, 9), _req(1, 8, 3), _req(2, 12, 2), _req(3, 8, 8, torch.float32)]
    done = srv.run_trace(reqs)
    assert sorted(r.rid for r in done) == [0, 1, 2, 3] and not srv.failed
    for r in done:
        for k in range(3):
            assert torch.all(r.result[k][:, 3:-3, 3:-3] == k + r.steps)
    order = [r.rid for r in sorted(done, key=lambda r: (r.t_done, r.rid))]
    assert order.index(1) < order.index(0) and order.index(2) < order.index(0)
