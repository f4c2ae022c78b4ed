"""The C768 workload itself (GPU).

Three substeps of the benchmark field are compared with the NumPy oracle bit
for bit (a few seconds of oracle time).  Longer runs are checked through
properties that hold at any size: graph replay and eager launches give
identical bits, the band height does not change a bit, the binary64 mass
integral is conserved to round-off, and the state stays finite.
"""

from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu

NX = 768
SUBSTEPS = 8


@pytest.fixture(scope="module")
def c768_host():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from licom4py_gpu_amd import duogrid, initial
    g = duogrid.grid(NX)
    h, u, v, sp = initial.seeded_balanced(g)
    dt = initial.courant_dt(g, h, sp)
    return duogrid.barotr_consts(g, 5000.0), initial.state_tuple(h, u, v), dt


@pytest.fixture(scope="module")
def c768(c768_host):
    import torch
    from licom4py_gpu_amd import barotr
    host_consts, state, dt = c768_host
    consts = barotr.as_device_consts(host_consts, torch.device("cuda", 0))
    core = [torch.as_tensor(x).cuda() for x in state[:6]]
    return barotr.SWContext(consts, torch.float64), core, dt


def test_c768_three_substeps_match_oracle(c768_host):
    """The benchmark configuration at its own size, through the drop-in
    entry point (a context with the band height sw_create picks, recorded
    graphs), against the oracle's bits."""
    import numpy as np
    import torch
    from licom4py_gpu_amd import barotr
    from oracle import barotr_np as B
    host_consts, state, dt = c768_host
    dev = torch.device("cuda", 0)
    dconsts = barotr.as_device_consts(host_consts, dev)
    dstate = tuple(torch.as_tensor(x).to(dev) for x in state)
    out = barotr.barotr_rk2_core(dstate, dconsts, 3, dt)
    got = [x.cpu().numpy() for x in out]
    want = B.barotr_rk2_core(state, host_consts, 3, dt)
    assert len(got) == len(want) == 23
    for k in range(23):
        assert np.array_equal(got[k], want[k]), k
    # the FP32 form against the same FP64 oracle: relative L2 of h, u, v
    # within 1e-5 (the bar for FP32 heights)
    ctx32 = barotr.SWContext(dconsts, torch.float32)
    core32 = [torch.as_tensor(x).to(dev, torch.float32) for x in state[:6]]
    out32 = ctx32.rk(core32, 3, dt, diagnostics=False, graph=True)
    for k, idx in enumerate((0, 2, 3)):
        a = out32[k].double().cpu().numpy()[:, 3:-3, 3:-3]
        b = want[idx][:, 3:-3, 3:-3]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-5, idx


def test_c768_graph_equals_eager_and_band_height_is_invisible(c768):
    import torch
    ctx, core, dt = c768
    ctx.set_tuning(30, 30, 32)
    ctx.set_band_split(0)
    eager = [x.clone() for x in ctx.rk(core, SUBSTEPS, dt, graph=False)]
    graph = ctx.rk(core, SUBSTEPS, dt, graph=True)
    for k in range(13):
        assert torch.equal(eager[k], graph[k]), k
    for rows, split in (((16, 44, 0), 0), ((30, 30, 20), 0), ((30, 30, 36), 0), ((30, 30, 48), 0),
                        ((30, 30, 48), 22), ((30, 30, 50), 22), ((30, 30, 40), 30)):
        ctx.set_tuning(*rows)
        ctx.set_band_split(split)
        other = ctx.rk(core, SUBSTEPS, dt, graph=True)
        for k in range(3):
            assert torch.equal(eager[k], other[k]), (rows, split, k)
    assert bool(torch.isfinite(eager[0][:, 3:-3, 3:-3]).all())


def test_c768_mass_conserved(c768):
    ctx, core, dt = c768
    m0, e0 = ctx.integrals(core[0], core[2], core[3])
    out = ctx.rk(core, SUBSTEPS, dt, diagnostics=False, graph=True)
    m1, e1 = ctx.integrals(out[0], out[1], out[2])
    assert abs(m1 - m0) / m0 < 1e-14
    assert abs(e1 - e0) / e0 < 1e-3
