"""The graph-mode kernel timer computes what the plain path computes (GPU).

sw_profile_graph records the substep-pair graph with the fused kernel's
launch stamps on and replays it; its outputs must be the bits of the same
number of substeps through sw_rk2 (the stamps change nothing else), and it
must time every fused launch of every replay."""

from __future__ import annotations

import pytest

from tests._util import seeded_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_profile_graph_outputs_and_counts(dtype):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from licom4py_gpu_amd import barotr
    tdt = torch.float64 if dtype == "f64" else torch.float32
    state, consts, dt = seeded_state(20)
    ctx = barotr.SWContext(consts, tdt)
    core = [torch.as_tensor(x).cuda().to(tdt) for x in state[:6]]
    pairs, replays = 3, 2
    prof = ctx.profile_graph(core, pairs, replays, dt)
    got = ctx.last_profile_out
    want = ctx.rk(core, 1 + 2 * pairs * 4 * replays, dt, diagnostics=False, graph=False)  # 3 lead replays each
    for a, b in zip(got, want):
        assert torch.equal(a, b)
    assert prof["fused"][1] == 2 * 2 * pairs * replays and prof["graph"][1] == replays
    fused_ms, graph_ms = prof["fused"][0], prof["graph"][0]
    assert 0 < fused_ms < graph_ms  # the fused launches are part of each replay
