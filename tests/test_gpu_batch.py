"""Batched runs (sw_batch_run): each member equals the same run alone (GPU).

Members of different resolutions, substep counts and lengths are advanced by
the same launches; every output of every member must be bit-identical to
sw_rk2 / sw_rk3 on that member's context alone, and (FP64) to the oracle.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import barotr_np as B
from tests._util import generic_consts, grid_consts, random_core

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _members(torch, dtype, spec, seed=0):
    """spec: list of (nx, nbb, dtb); two entries of one nx share a grid."""
    from licom4py_gpu_amd import ensemble
    from licom4py_gpu_amd.barotr import SWContext
    base = {}
    out = []
    for k, (nx, nbb, dtb) in enumerate(spec):
        _, consts = grid_consts(nx)
        if nx in base:
            ctx = base[nx].like()
        else:
            ctx = base[nx] = SWContext(consts, dtype)
        core_np = random_core(nx, seed=seed + k, distinct_p=False)
        core = [torch.as_tensor(x).cuda().to(dtype) for x in core_np]
        out.append(ensemble.Member(ctx, core, nbb, dtb, tag={"consts": consts, "core": core_np}))
    return out


def _check_alone(torch, members, rk=2, diagnostics=True):
    for k, m in enumerate(members):
        alone = m.ctx.rk(m.core6, m.nbb, m.dtb, rk=rk, diagnostics=diagnostics, graph=False)
        n = 13 if (diagnostics and m.nbb > 0) else 3
        for i in range(n):
            a, b = m.out[i].cpu().numpy(), alone[i].cpu().numpy()
            assert np.array_equal(a, b), f"member {k} (C{m.ctx.nx}, nbb {m.nbb}) output {i} differs"


SPEC = [(8, 5, 40.0), (13, 3, 25.0), (16, 2, 30.0), (16, 4, 31.0), (24, 1, 12.5), (10, 4, 33.0)]
# with a member that takes no substep (diagnostics off: nbb = 0 computes none)
SPEC0 = SPEC[:2] + [(16, 0, 30.0)] + SPEC[3:]


@pytest.mark.parametrize("rk", [2, 3])
def test_batch_fp64_matches_alone_and_oracle(rk):
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    members = _members(torch, torch.float64, SPEC)
    ensemble.Batch().run(members, rk=rk)
    _check_alone(torch, members, rk=rk)
    # the C8 member against the oracle
    m = members[0]
    core = m.tag["core"]
    st = list(core) + [np.zeros_like(core[0])] * 17
    fn = B.barotr_rk2_core if rk == 2 else B.barotr_rk3_core
    want = fn(tuple(st), m.tag["consts"], m.nbb, m.dtb)
    assert np.array_equal(m.out[0].cpu().numpy(), want[0])
    assert np.array_equal(m.out[1].cpu().numpy(), want[2])
    assert np.array_equal(m.out[2].cpu().numpy(), want[3])


def test_batch_fp32_matches_alone():
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    members = _members(torch, torch.float32, SPEC, seed=7)
    ensemble.Batch().run(members)
    _check_alone(torch, members)


@pytest.mark.parametrize("nbb", [3, 8])
def test_batch_graph_replay_matches_alone(nbb):
    """Equal substep counts: substep pairs replay a recorded graph; a second
    run with other members of the same layout reuses it."""
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    b = ensemble.Batch()
    spec = [(8, nbb, 40.0), (13, nbb, 25.0), (16, nbb, 31.0), (16, nbb, 30.0)]
    for seed in (1, 2):
        members = _members(torch, torch.float64, spec, seed=seed)
        b.run(members, graph=True)
        _check_alone(torch, members)


def test_batch_object_reused_with_other_members():
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    b = ensemble.Batch()
    first = _members(torch, torch.float64, SPEC0, seed=3)
    b.run(first, diagnostics=False)
    second = _members(torch, torch.float64, [(12, 6, 20.0), (9, 2, 44.0)], seed=5)
    b.run(second, diagnostics=False)
    _check_alone(torch, first, diagnostics=False)
    _check_alone(torch, second, diagnostics=False)


def test_batch_rejects_shared_context_and_mixed_dtypes():
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    members = _members(torch, torch.float64, [(8, 2, 10.0), (8, 2, 10.0)])
    members[1].ctx = members[0].ctx
    with pytest.raises(RuntimeError, match="own context"):
        ensemble.Batch().run(members)
    m64 = _members(torch, torch.float64, [(8, 2, 10.0)])
    m32 = _members(torch, torch.float32, [(9, 2, 10.0)])
    with pytest.raises(RuntimeError, match="dtype"):
        ensemble.Batch().run(m64 + m32)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_separate_kernels_match_alone(dtype):
    """Members whose contexts select the separate height / velocity kernels
    (rows_f = 0) are batched with them, and still equal the fused path alone."""
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    tdt = torch.float64 if dtype == "f64" else torch.float32
    members = _members(torch, tdt, SPEC, seed=11)
    for m in members:
        m.ctx.set_tuning(30, 30, 0)
    ensemble.Batch().run(members)
    for m in members:
        m.ctx.set_tuning(30, 30, 32)  # alone: the fused kernel
    _check_alone(torch, members)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_mixed_substeps_graph_replay_matches_alone(dtype):
    """Different substep counts under SW_GRAPH: the substeps are replayed
    from graphs recorded per number of members still stepping (with eager
    substeps to reach odd starts), and every member still equals its run
    alone; a second run of the same layout reuses the graphs."""
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    tdt = torch.float64 if dtype == "f64" else torch.float32
    b = ensemble.Batch()
    spec = [(8, 11, 40.0), (13, 8, 25.0), (16, 8, 31.0), (16, 5, 30.0), (10, 2, 33.0), (12, 1, 20.0)]
    for seed in (21, 22):
        members = _members(torch, tdt, spec, seed=seed)
        b.run(members, graph=True)
        _check_alone(torch, members)


def test_batch_many_graph_layouts():
    """More recorded layouts than the batch object keeps: evicted graphs are
    destroyed only after the work queued before their eviction, and every run
    stays equal to the members alone."""
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    b = ensemble.Batch()
    for count in range(1, 21):
        spec = [(8 + (k % 3), 3 + (k % 4), 30.0 + k) for k in range(count)]
        members = _members(torch, torch.float64, spec, seed=100 + count)
        b.run(members, diagnostics=False, graph=True)
        if count % 5 == 0:
            _check_alone(torch, members, diagnostics=False)
    torch.cuda.synchronize()


def test_nbb_zero_needs_skip_diag():
    """nbb = 0 computes no diagnostics: with diagnostics requested the C-ABI
    refuses (the reference's empty loop returns the input's diagnostics, which
    the six input fields do not carry); without them h, u, v are copied."""
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    members = _members(torch, torch.float64, [(8, 0, 10.0), (9, 3, 10.0)])
    with pytest.raises(RuntimeError, match="SW_SKIP_DIAG"):
        ensemble.Batch().run(members)
    with pytest.raises(RuntimeError, match="SW_SKIP_DIAG"):
        members[0].ctx.rk(members[0].core6, 0, 10.0)
    out = members[0].ctx.rk(members[0].core6, 0, 10.0, diagnostics=False)
    for got, want in zip(out, (members[0].core6[0], members[0].core6[2], members[0].core6[3])):
        assert torch.equal(got, want)


@pytest.mark.parametrize("graph", [False, True])
def test_batch_frame_forms_match_alone_and_oracle(graph):
    """Every member on the cube-edge frame launch (k_frame_edge_batch), and a
    batch where one member's context uses the two-launch form (the whole
    batch then runs k_frame_batch + k_halo_uv_edge_batch): each member equals
    its run alone, and the C8 member equals the oracle."""
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    spec = [(8, 5, 40.0), (16, 4, 30.0), (48, 3, 20.0), (16, 2, 31.0)]
    for mixed in (False, True):
        members = _members(torch, torch.float64, spec, seed=11)
        assert all(m.ctx.frame_form() for m in members)
        if mixed:
            members[2].ctx.set_frame_form(False)
        ensemble.Batch().run(members, graph=graph)
        _check_alone(torch, members)
        m = members[0]
        core = m.tag["core"]
        st = list(core) + [np.zeros_like(core[0])] * 17
        want = B.barotr_rk2_core(tuple(st), m.tag["consts"], m.nbb, m.dtb)
        for i, idx in enumerate((0, 2, 3)):
            assert np.array_equal(m.out[i].cpu().numpy(), want[idx]), (mixed, idx)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_batch_generic_layout_matches_alone(dtype):
    """Members whose consts lack the grid's uniform / face-shared / symmetric
    structure run the kernels' general paths (k_frame_edge_batch<T, false>
    among them): each equals its run alone, and FP64 members the oracle."""
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    from licom4py_gpu_amd.barotr import SWContext
    tdt = torch.float64 if dtype == "f64" else torch.float32
    members = []
    for k, (nx, nbb, dtb) in enumerate([(8, 3, 40.0), (13, 2, 25.0), (16, 3, 30.0)]):
        _, base = grid_consts(nx)
        consts = generic_consts(base, ("dz", "shared", "gsym"))
        ctx = SWContext(consts, tdt)
        f = ctx.layout_flags()
        assert not (f["dz_uniform"] or f["shared_metrics"] or f["gct_symmetric"])
        assert ctx.frame_form()
        core_np = random_core(nx, seed=20 + k, distinct_p=False)
        core = [torch.as_tensor(x).cuda().to(tdt) for x in core_np]
        members.append(ensemble.Member(ctx, core, nbb, dtb, tag={"consts": consts, "core": core_np}))
    ensemble.Batch().run(members, graph=True)
    _check_alone(torch, members)
    if dtype == "f64":
        for m in members:
            core = m.tag["core"]
            st = list(core) + [np.zeros_like(core[0])] * 17
            want = B.barotr_rk2_core(tuple(st), m.tag["consts"], m.nbb, m.dtb)
            for i, idx in enumerate((0, 2, 3)):
                assert np.array_equal(m.out[i].cpu().numpy(), want[idx]), (m.ctx.nx, idx)


def test_batch_ragged_frame_segments_match_alone_and_oracle():
    """Resolutions whose cube edges leave a one-position last segment of the
    batched frame launch (C29, C57: 28 + 1, 28 + 28 + 1) and the smallest one
    (C7): each member equals its run alone and the oracle."""
    torch = _torch()
    from licom4py_gpu_amd import ensemble
    members = _members(torch, torch.float64, [(29, 2, 25.0), (57, 2, 20.0)], seed=31)
    assert all(m.ctx.frame_form() for m in members)  # the batch runs k_frame_edge_batch
    members += _members(torch, torch.float64, [(7, 3, 30.0)], seed=41)
    ensemble.Batch().run(members, graph=True)
    _check_alone(torch, members)
    for m in members:
        core = m.tag["core"]
        st = list(core) + [np.zeros_like(core[0])] * 17
        want = B.barotr_rk2_core(tuple(st), m.tag["consts"], m.nbb, m.dtb)
        for i, idx in enumerate((0, 2, 3)):
            assert np.array_equal(m.out[i].cpu().numpy(), want[idx]), (m.ctx.nx, idx)
