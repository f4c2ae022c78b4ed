"""Serving loop: every result equals the same solve alone (GPU).

A small trace of mixed resolutions, precisions and step counts runs through
the continuous-batching server with a budget that admits only a few requests
at a time (so requests queue, join running batches mid-way and leave early);
each request's final (h, u, v) must be bit-identical to sw_rk2 on its own.
"""

from __future__ import annotations

import numpy as np
import pytest

from tests._util import grid_consts, seeded_state

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _trace(torch):
    from licom4py_gpu_amd import serving
    rng = np.random.default_rng(11)
    reqs = []
    for k in range(14):
        nx = int(rng.choice([8, 12, 16]))
        state, _, dt = seeded_state(nx)
        f = 1.0 + 0.05 * rng.normal()
        init = tuple(torch.as_tensor(state[i] * f).cuda() for i in (0, 2, 3))
        steps = int(rng.integers(0, 26))
        dtype = torch.float32 if rng.uniform() < 0.5 else torch.float64
        reqs.append(serving.Request(rid=k, nx=nx, steps=steps, dtype=dtype, dt=dt * 0.9, init=init,
                                    arrival=0.0))
    return reqs


@pytest.mark.parametrize("policy", ["srpt", "all"])
def test_served_results_equal_alone(policy):
    torch = _torch()
    from licom4py_gpu_amd import serving
    from licom4py_gpu_amd.barotr import SWContext
    reqs = _trace(torch)
    # room for the six resolutions' consts and about four requests at a time
    budget = 4 * serving.request_bytes(16, torch.float64) + sum(
        serving.kind_bytes(nx, dt) for nx in (8, 12, 16) for dt in (torch.float32, torch.float64))
    srv = serving.Server(budget_bytes=budget, chunk=4, policy=policy, width=6 * 16 * 16 * 2)
    done = srv.run_trace(reqs)
    assert len(done) == len(reqs)
    assert srv.peak_used <= budget
    assert max(srv.batch_sizes) >= 2                      # requests did share launches
    assert any(r.t_admit > r.arrival + 1e-4 for r in done)  # and some had to queue
    ctxs = {}
    for r in done:
        key = (r.nx, r.dtype)
        if key not in ctxs:
            ctxs[key] = SWContext(grid_consts(r.nx)[1], r.dtype)
        h, u, v = (x.to(r.dtype) for x in r.init)
        alone = ctxs[key].rk((h, h, u, v, u, v), r.steps, r.dt, diagnostics=False, graph=False)
        for name, got, want in zip("huv", r.result, alone[:3]):
            assert torch.equal(got, want), f"request {r.rid} (C{r.nx}, {r.dtype}, {r.steps} steps): {name}"
    stats = serving.latency_stats(done)
    assert stats["p99_s"] >= stats["p50_s"] > 0


def test_request_larger_than_budget_is_refused():
    torch = _torch()
    from licom4py_gpu_amd import serving
    srv = serving.Server(budget_bytes=1e6, chunk=4)
    r = serving.Request(0, 48, 10, torch.float64, 100.0, (None, None, None))
    with pytest.raises(ValueError, match="budget"):
        srv.submit(r)
