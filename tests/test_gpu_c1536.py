"""The C1536 mixed-precision workload at its own size (GPU).

6 x 1536^2 = 14,155,776 interior cells.  One FP64 substep of the seeded
balanced field through the drop-in entry point is compared with the NumPy
oracle bit for bit (a few seconds of oracle time per stage on the host).
Then the FP32 form runs 300 substeps beside the FP64 path: its height must
stay within 1e-5 relative L2 of the FP64 height (velocity within 1e-4), and
the binary64 compensated mass integral must drift by less than 1e-14
(FP64) and 1e-10 (FP32) of the total.  Mass is conserved by the scheme up to
round-off (the flux through every cube edge is averaged, so both faces see the
same flux), which is what the tolerances bound.
"""

from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu

NX = 1536
SUBSTEPS = 300


@pytest.fixture(scope="module")
def c1536_host():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from licom4py_gpu_amd import duogrid, initial
    print(f"\n[c1536] generating the C{NX} grid and seeded field", flush=True)
    g = duogrid.grid(NX)
    h, u, v, sp = initial.seeded_balanced(g)
    dt = initial.courant_dt(g, h, sp)
    return duogrid.barotr_consts(g, 5000.0), initial.state_tuple(h, u, v), dt


def test_c1536_one_substep_matches_oracle(c1536_host):
    import numpy as np
    import torch
    from licom4py_gpu_amd import barotr
    from oracle import barotr_np as B
    host_consts, state, dt = c1536_host
    dev = torch.device("cuda", 0)
    dconsts = barotr.as_device_consts(host_consts, dev)
    dstate = tuple(torch.as_tensor(x).to(dev) for x in state)
    out = barotr.barotr_rk2_core(dstate, dconsts, 1, dt)
    got = [x.cpu().numpy() for x in out]
    del out, dstate
    print("[c1536] oracle substep", flush=True)
    want = B.barotr_rk2_core(state, host_consts, 1, dt)
    for k in range(23):
        assert np.array_equal(got[k], want[k]), barotr.STATE_KEYS[k]


def test_c1536_fp32_against_fp64_and_mass_drift(c1536_host):
    import torch
    from licom4py_gpu_amd import barotr
    host_consts, state, dt = c1536_host
    dev = torch.device("cuda", 0)
    dconsts = barotr.as_device_consts(host_consts, dev)
    finals, drift = {}, {}
    for tdt in (torch.float64, torch.float32):
        ctx = barotr.SWContext(dconsts, tdt)
        core = [torch.as_tensor(x).to(dev, tdt) for x in state[:6]]
        m0, _ = ctx.integrals(core[0], core[2], core[3])
        out = ctx.rk(core, SUBSTEPS, dt, diagnostics=False, graph=True)
        m1, _ = ctx.integrals(out[0], out[1], out[2])
        drift[tdt] = abs(m1 - m0) / m0
        finals[tdt] = [x[:, 3:-3, 3:-3].double() for x in out[:3]]
        assert bool(torch.isfinite(finals[tdt][0]).all())
        del ctx, core, out
    rel = [float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
           for a, b in zip(finals[torch.float32], finals[torch.float64])]
    print(f"[c1536] {SUBSTEPS} substeps: FP32 vs FP64 rel L2 h {rel[0]:.3e} u {rel[1]:.3e} v {rel[2]:.3e}; "
          f"mass drift FP64 {drift[torch.float64]:.3e} FP32 {drift[torch.float32]:.3e}", flush=True)
    assert rel[0] <= 1e-5
    assert rel[1] <= 1e-4 and rel[2] <= 1e-4
    assert drift[torch.float64] < 1e-14
    assert drift[torch.float32] < 1e-10


def test_c1536_default_bands_fit_one_round(c1536_host):
    """Without the tuner a C1536 context marches bands of the lowest height
    whose waves fit one round of resident waves (140 rows FP64, 91 FP32 on
    256 CUs), uniform: the fastest heights measured there."""
    import torch
    from licom4py_gpu_amd import barotr
    from licom4py_gpu_amd.autotune import one_round_rows
    host_consts, _, _ = c1536_host
    dev = torch.device("cuda", 0)
    dconsts = barotr.as_device_consts(host_consts, dev)
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    for dtype in (torch.float64, torch.float32):
        ctx = barotr.SWContext(dconsts, dtype)
        assert ctx.tuning[2] == one_round_rows(NX, dtype == torch.float64, cus) > 64, (dtype, ctx.tuning)
        assert ctx.split == 0
        del ctx
