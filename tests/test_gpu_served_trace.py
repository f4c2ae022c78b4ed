"""A served trace across the full resolution range (GPU).

Twelve requests spanning C48, C96, C192, C384 and C768, 3 to 30 substeps
each, FP32 and FP64 mixed, all arriving at once under an HBM budget of 2.2 GB:
the C768 FP64 kind alone takes 1.27 GB with one request, so requests queue and
are admitted as others finish, kinds are evicted to make room, and requests of
both precisions run in overlapping batches on two streams.  Each result must
be bit-identical (torch.equal) to the same solve run alone.
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RES = (48, 96, 192, 384, 768)


def test_served_trace_c48_to_c768_equals_alone():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from licom4py_gpu_amd import barotr, duogrid, initial, serving
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(5)
    base = {}
    for nx in RES:
        g = duogrid.grid(nx)
        h, u, v, sp = initial.seeded_balanced(g)
        base[nx] = ([torch.as_tensor(x).to(dev) for x in (h, u, v)], initial.courant_dt(g, h, sp),
                    duogrid.barotr_consts(g, 5000.0))
    spec = [(nx, torch.float64 if k % 2 == 0 else torch.float32) for k, nx in enumerate(RES)]
    spec += [(768, torch.float64), (768, torch.float32), (384, torch.float64)]
    spec += [(int(rng.choice(RES)), torch.float64 if rng.uniform() < 0.5 else torch.float32) for _ in range(4)]
    reqs = []
    for k, (nx, dtype) in enumerate(spec):
        fields, dt, _ = base[nx]
        f = 1.0 + 0.01 * rng.normal()
        init = tuple(x * f for x in fields)
        reqs.append(serving.Request(rid=k, nx=nx, steps=int(rng.integers(3, 31)), dtype=dtype, dt=dt / max(f, 1.0),
                                    init=init, arrival=0.0))
    budget = 2.2e9
    srv = serving.Server(budget_bytes=budget, chunk=8)
    done = srv.run_trace(reqs)
    assert not srv.failed and len(done) == len(reqs)
    assert srv.peak_used <= budget
    assert max(srv.batch_sizes) >= 2                  # requests shared launches
    first_done = min(r.t_done for r in done)
    assert sum(r.t_admit >= first_done for r in done) >= 2  # and some waited for room
    assert srv.evictions >= 1
    print(f"\n[served] {len(done)} requests, {srv.ticks} ticks, batches {srv.batch_sizes}, "
          f"evictions {srv.evictions}, peak {srv.peak_used / 1e9:.2f} GB", flush=True)
    ctxs = {}
    for r in done:
        key = (r.nx, r.dtype)
        if key not in ctxs:
            ctxs[key] = barotr.SWContext(base[r.nx][2], r.dtype)
        h, u, v = (x.to(r.dtype) for x in r.init)
        alone = ctxs[key].rk((h, h, u, v, u, v), r.steps, r.dt, diagnostics=False, graph=False)
        for name, got, want in zip("huv", r.result, alone[:3]):
            assert torch.equal(got, want), f"request {r.rid} (C{r.nx}, {r.dtype}, {r.steps} steps): {name}"
        srv.release(r)
