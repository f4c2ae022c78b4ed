"""The 64-member ensemble at its own size (GPU).

64 members with resolutions drawn uniformly (seeded) from {48, 96, 192, 384},
each the seeded balanced field with 1 % Gaussian noise on its harmonic
coefficients and its own Courant dt, in FP64 and in FP32.  The batch's band
height and stage-kernel form are picked by the project's autotuner
(`autotune.tune_batch`, timing the whole batch), then every member advances 4
substeps in one batched run with recorded substep-pair graphs.  Every output
of every member must be bit-identical to the same member run alone, and in
FP64 the C48 members must be bit-identical to the NumPy oracle (all 23 fields).
"""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MEMBERS = 64
RES = (48, 96, 192, 384)
SEED = 20261203
SUBSTEPS = 4


@pytest.fixture(scope="module")
def host_members():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from licom4py_gpu_amd import duogrid, initial
    rng = np.random.default_rng(SEED)
    res = [int(x) for x in rng.choice(RES, size=MEMBERS)]
    grids = {nx: duogrid.grid(nx) for nx in sorted(set(res))}
    consts = {nx: duogrid.barotr_consts(g, 5000.0) for nx, g in grids.items()}
    out = []
    print(f"\n[ensemble] generating {MEMBERS} members {dict((r, res.count(r)) for r in RES)}", flush=True)
    for k, nx in enumerate(res):
        h, u, v, sp = initial.seeded_balanced(grids[nx], seed=SEED, perturb=0.01, member_seed=k)
        out.append((nx, initial.state_tuple(h, u, v), initial.courant_dt(grids[nx], h, sp)))
    return out, consts


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_ensemble64_batched_equals_alone(host_members, dtype):
    import torch
    from licom4py_gpu_amd import autotune, barotr, ensemble
    from oracle import barotr_np as B
    members_host, consts = host_members
    tdt = torch.float64 if dtype == "f64" else torch.float32
    dev = torch.device("cuda", 0)
    base = {}
    members = []
    for k, (nx, state, dt) in enumerate(members_host):
        if nx not in base:
            base[nx] = ctx = barotr.SWContext(consts[nx], tdt)
        else:
            ctx = base[nx].like()
        core = [torch.as_tensor(x).to(dev, tdt) for x in state[:6]]
        members.append(ensemble.Member(ctx, core, SUBSTEPS, dt, tag={"k": k}))
    batch = ensemble.Batch()
    tuning = autotune.tune_batch(batch, members, candidates=(16, 24, 32, 48))
    print(f"[ensemble] {dtype}: tuned {tuning['rows']} rows, fused {tuning['fused']}", flush=True)
    batch.run(members, diagnostics=True, graph=True)
    torch.cuda.synchronize()
    for m in members:
        alone = m.ctx.rk(m.core6, SUBSTEPS, m.dtb, diagnostics=True, graph=True)
        for i in range(13):
            assert torch.equal(m.out[i], alone[i]), (m.tag["k"], m.ctx.nx, i)
    if tdt != torch.float64:
        return
    checked = 0
    for m, (nx, state, dt) in zip(members, members_host):
        if nx != 48:
            continue
        want = B.barotr_rk2_core(state, consts[nx], SUBSTEPS, dt)
        got = [m.out[0], m.out[0], m.out[1], m.out[2], m.out[1], m.out[2]] + list(m.out[3:13])
        for i in range(16):
            assert np.array_equal(got[i].cpu().numpy(), want[i]), (m.tag["k"], barotr.STATE_KEYS[i])
        checked += 1
    assert checked > 0
