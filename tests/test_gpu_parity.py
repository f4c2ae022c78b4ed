"""HIP path vs the NumPy FP64 oracle (parity proper; needs a GPU).

FP64 results must be bit-identical to the oracle (the kernels evaluate every
expression in the reference's order with contraction disabled); the bar the
step must meet in any case is relative L2 <= 1e-12.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import barotr_np as B
from tests._util import generic_consts, grid_consts, max_ulp_diff, random_core, rel_l2, seeded_state

pytestmark = pytest.mark.gpu

NAMES13 = ("h0", "ub", "vb", "celerity_x", "celerity_y", "ub_ct", "vb_ct",
           "ub_cx", "ub_cy", "vb_cx", "vb_cy", "advx", "advy")


def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _dev(x, torch, dtype=None):
    t = torch.as_tensor(np.ascontiguousarray(x)).cuda()
    return t if dtype is None else t.to(dtype)


def _host(t):
    return t.detach().cpu().numpy()


def _assert_exact(name, got, want):
    if not np.array_equal(got, want):
        raise AssertionError(
            f"{name}: not bit-identical (max ulp {max_ulp_diff(got, want)}, "
            f"rel L2 {rel_l2(got, want):.3e})")


@pytest.mark.parametrize("nx", [8, 48])
def test_halo_scalar_matches_oracle(nx):
    torch = _torch()
    from licom4py_gpu_amd.barotr import SWContext
    g, consts = grid_consts(nx)
    ctx = SWContext(consts, torch.float64)
    rng = np.random.default_rng(nx)
    f = rng.normal(size=(6, nx + 6, nx + 6))
    want = B.ext_scalar(f, consts[5], consts[6], consts[7])
    got = _host(ctx.halo_scalar_(_dev(f, torch)))
    _assert_exact("ext_scalar", got, want)
    # stacked (6, 2, N, N) as ext_vector packs it
    f2 = rng.normal(size=(6, 2, nx + 6, nx + 6))
    want2 = B.ext_scalar(f2, consts[5], consts[6], consts[7])
    got2 = _host(ctx.halo_scalar_(_dev(f2, torch)))
    _assert_exact("ext_scalar stacked", got2, want2)


@pytest.mark.parametrize("nx", [8, 48])
def test_halo_vector_matches_oracle(nx):
    torch = _torch()
    from licom4py_gpu_amd.barotr import SWContext
    g, consts = grid_consts(nx)
    ctx = SWContext(consts, torch.float64)
    rng = np.random.default_rng(nx + 1)
    u = rng.normal(size=(6, nx + 6, nx + 6))
    v = rng.normal(size=(6, nx + 6, nx + 6))
    wu, wv = B.ext_vector(u, v, consts[5], consts[6], consts[7], consts[8], consts[9],
                          consts[10], consts[11])
    gu, gv = ctx.halo_vector_(_dev(u, torch), _dev(v, torch))
    _assert_exact("ext_vector u", _host(gu), wu)
    _assert_exact("ext_vector v", _host(gv), wv)


@pytest.mark.parametrize("nx", [8, 48])
def test_edge_avg_matches_oracle(nx):
    torch = _torch()
    from licom4py_gpu_amd.barotr import SWContext
    g, consts = grid_consts(nx)
    ctx = SWContext(consts, torch.float64)
    rng = np.random.default_rng(nx + 2)
    u = rng.normal(size=(6, nx + 6, nx + 6))
    v = rng.normal(size=(6, nx + 6, nx + 6))
    wu, wv = B.boundary_communication(u, v)
    gu, gv = ctx.edge_avg_(_dev(u, torch), _dev(v, torch))
    _assert_exact("edge avg u", _host(gu), wu)
    _assert_exact("edge avg v", _host(gv), wv)


def test_stage_matches_golden_fixture():
    """One C8 stage against the frozen oracle output in tests/golden/stage_c8.npz."""
    import os
    torch = _torch()
    from licom4py_gpu_amd.barotr import SWContext
    path = os.path.join(os.path.dirname(__file__), "golden", "stage_c8.npz")
    with np.load(path, allow_pickle=False) as z:
        core = [z[f"in_{k}"] for k in ("h", "hp", "u", "v", "up", "vp")]
        want = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
        dt, beta, last = float(z["dt"]), float(z["beta"]), bool(z["last"])
    _, consts = grid_consts(8)
    got = SWContext(consts, torch.float64).stage(*[_dev(x, torch) for x in core], dt, beta, last)
    for name, gv in zip(NAMES13, got):
        _assert_exact(f"golden {name}", _host(gv), want[name])


@pytest.mark.parametrize("nx", [7, 8, 13, 48, 52, 53, 57, 58, 70, 104, 105, 140])
@pytest.mark.parametrize("last", [False, True])
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_stage_matches_oracle(nx, last, beta):
    torch = _torch()
    from licom4py_gpu_amd.barotr import SWContext
    g, consts = grid_consts(nx)
    ctx = SWContext(consts, torch.float64)
    core = random_core(nx, seed=nx * 7 + int(last))
    want = B.step_rk_logic(*core, consts, 37.5, beta, last)
    got = ctx.stage(*[_dev(x, torch) for x in core], 37.5, beta, last)
    for name, gv, wv in zip(NAMES13, got, want):
        _assert_exact(f"stage {name}", _host(gv), wv)


def test_rk2_c48_seeded_200_substeps():
    """The C48 parity workload: seeded balanced field, Courant 0.5, 200 substeps
    as 4 calls of nbb = 50, all 23 state arrays compared after every call."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(48)
    dconsts = barotr.as_device_consts(consts)
    gstate = tuple(_dev(x, torch) for x in state)
    ostate = state
    for call in range(4):
        ostate = B.barotr_rk2_core(ostate, consts, 50, dt)
        gstate = barotr.barotr_rk2_core(gstate, dconsts, 50, dt)
        for k, (gv, wv) in enumerate(zip(gstate, ostate)):
            got = _host(gv)
            assert rel_l2(got, wv) <= 1e-12, f"call {call} field {B.STATE_KEYS[k]}"
            _assert_exact(f"call {call} {B.STATE_KEYS[k]}", got, wv)


def test_rk3_matches_oracle():
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(24)
    dconsts = barotr.as_device_consts(consts)
    want = B.barotr_rk3_core(state, consts, 3, dt * 0.5)
    got = barotr.barotr_rk3_core(tuple(_dev(x, torch) for x in state), dconsts, 3, dt * 0.5)
    for k, (gv, wv) in enumerate(zip(got, want)):
        _assert_exact(f"rk3 {B.STATE_KEYS[k]}", _host(gv), wv)


def test_create_rejects_invalid_consts():
    """sw_create validates what its kernels index with: the interior / halo
    masks and the local halo indices of the strips (a RuntimeError naming the
    table); shapes and the dtype are checked on the Python side."""
    torch = _torch()
    from licom4py_gpu_amd.barotr import SWContext
    _, consts = grid_consts(8)

    def broken(k, edit):
        c = [np.array(x, copy=True) for x in consts]
        edit(c[k])
        return tuple(c)

    def set_(idx, val):
        def f(a):
            a[idx] = val
        return f

    with pytest.raises(RuntimeError, match="inner/outer"):
        SWContext(broken(10, set_((0, 4, 4), 0.0)), torch.float64)   # an interior cell marked halo
    with pytest.raises(RuntimeError, match="loc_i_local"):
        SWContext(broken(6, set_((0, 5, 0), 0)), torch.float64)      # W strip index out of range
    with pytest.raises(RuntimeError, match="loc_j_local"):
        SWContext(broken(7, set_((2, 0, 5), 99)), torch.float64)     # S strip index out of range
    c = list(consts)
    c[3] = np.array(c[3][:, :-1, :-1], copy=True)
    with pytest.raises(ValueError, match="shape"):
        SWContext(tuple(c), torch.float64)
    with pytest.raises(ValueError, match="dtype"):
        SWContext(consts, torch.float16)
    SWContext(consts, torch.float64)  # the unbroken tables still build a context


def test_distinct_previous_fields():
    """h0p != h0 at entry (allowed by the signature): stage 1 must use h0p."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(16)
    rng = np.random.default_rng(5)
    state = list(state)
    state[1] = state[0] + rng.normal(size=state[0].shape)
    state[4] = state[2] + 0.01 * rng.normal(size=state[0].shape)
    state = tuple(state)
    want = B.barotr_rk2_core(state, consts, 2, dt)
    got = barotr.barotr_rk2_core(tuple(_dev(x, torch) for x in state),
                                 barotr.as_device_consts(consts), 2, dt)
    for k, (gv, wv) in enumerate(zip(got, want)):
        _assert_exact(f"{B.STATE_KEYS[k]}", _host(gv), wv)


def test_schedule_core_chunks():
    torch = _torch()
    from licom4py_gpu_amd import barotr, make_schedule_core
    state, consts, dt = seeded_state(12)
    extra = np.arange(5.0)
    want = B.schedule_core(state + (extra,), consts, 3, 2, dt)
    core = make_schedule_core(2)
    got = core(tuple(_dev(x, torch) for x in state) + (extra,), barotr.as_device_consts(consts), 3, 2, dt)
    assert got[-1] is extra or np.array_equal(got[-1], extra)
    for k in range(23):
        _assert_exact(f"schedule {B.STATE_KEYS[k]}", _host(got[k]), want[k])


def test_graph_replays_match_eager():
    """Recorded substep-pair graphs replayed several times, with eager work and
    host synchronisation between replays, give the eager results bit for bit."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(24)
    ctx = barotr.SWContext(consts, torch.float64)
    core = [_dev(x, torch) for x in state[:6]]
    for nbb in (3, 7, 36, 37):
        eager = ctx.rk(core, nbb, dt, graph=False)
        for rep in range(3):
            got = ctx.rk(core, nbb, dt, graph=True)
            torch.cuda.synchronize()
            _ = (core[0] * 2.0).sum().item()  # unrelated eager work between replays
            for k in range(13):
                _assert_exact(f"graph nbb={nbb} rep={rep} out{k}", _host(got[k]), _host(eager[k]))


def test_profile_pass_matches_and_counts():
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(16)
    ctx = barotr.SWContext(consts, torch.float64)
    core = [_dev(x, torch) for x in state[:6]]
    prof = ctx.profile(core, 4, dt)  # fused stage kernel (the default)
    # the frame launch also fills the new height's halo
    assert prof["fused"][1] == 8 and prof["frame"][1] == 8 and prof["halo_h"][1] == 0
    assert prof["march_a"][1] == 0 and prof["march_b"][1] == 0 and prof["diag"][1] == 3  # 3 passes
    assert all(ms > 0 for ms, n in prof.values() if n)
    want = ctx.rk(core, 4, dt, graph=False)
    # repeated (idempotent) launches time the kernels and leave the results alone
    prof = ctx.profile(core, 4, dt, reps=3)
    assert prof["fused"][1] == 24 and prof["diag"][1] == 9
    for k in range(13):
        assert torch.equal(ctx.last_profile_out[k], want[k]), k
    ctx.set_tuning(32, 32, 0)  # separate height and velocity kernels
    prof = ctx.profile(core, 4, dt)
    assert prof["march_a"][1] == 8 and prof["march_b"][1] == 8 and prof["fused"][1] == 0
    for k in range(13):
        assert torch.equal(ctx.last_profile_out[k], want[k]), k
    flags = ctx.layout_flags()
    assert flags == {"dz_uniform": True, "shared_metrics": True, "gct_symmetric": True}


@pytest.mark.parametrize("rows", [(4, 4, 0), (7, 13, 0), (64, 100, 0), (4, 4, 4), (7, 7, 5), (8, 8, 13),
                                  (32, 32, 64), (30, 30, 4096)])
def test_band_height_does_not_change_results(rows):
    """Band heights of the separate kernels (rows_f = 0) and of the fused
    kernel (rows_f > 0, bands shorter and longer than its 6-row stencil reach)
    give the oracle's bits; nx = 70 puts two strips on a face row."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    nx = 70
    _, consts = grid_consts(nx)
    ctx = barotr.SWContext(consts, torch.float64)
    core = random_core(nx, seed=3)
    want = B.step_rk_logic(*core, consts, 20.0, 0.0, True)
    ctx.set_tuning(*rows)
    got = ctx.stage(*[_dev(x, torch) for x in core], 20.0, 0.0, True)
    for name, gv, wv in zip(NAMES13, got, want):
        _assert_exact(f"rows={rows} {name}", _host(gv), wv)


@pytest.mark.parametrize("nx,rows_f,split", [(70, 8, 4), (70, 12, 5), (70, 40, 30), (70, 64, 8), (70, 4, 4),
                                             (140, 66, 66), (140, 100, 40), (140, 20, 13), (48, 30, 12)])
def test_band_split_does_not_change_results(nx, rows_f, split):
    """Split bands of the fused kernel (the first blocks dispatched march
    rows_f rows, the rest `split` rows; bands longer than 64 rows reload the
    cube-edge rows they keep in LDS, marching down and up) give the oracle's
    bits for a stage and for 3 RK2 substeps in graph replays."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(nx)
    ctx = barotr.SWContext(consts, torch.float64)
    ctx.set_tuning(32, 32, rows_f)
    ctx.set_band_split(split)
    assert ctx.band_split() == split
    core = random_core(nx, seed=5)
    want = B.step_rk_logic(*core, consts, 20.0, 0.0, True)
    got = ctx.stage(*[_dev(x, torch) for x in core], 20.0, 0.0, True)
    for name, gv, wv in zip(NAMES13, got, want):
        _assert_exact(f"split {rows_f}/{split} {name}", _host(gv), wv)
    got = ctx.rk([_dev(x, torch) for x in state[:6]], 3, dt, graph=True)
    want = B.barotr_rk2_core(state, consts, 3, dt)
    for k, idx in enumerate((0, 2, 3)):
        _assert_exact(f"split {rows_f}/{split} rk2 field {idx}", _host(got[k]), want[idx])


def _poison_cache(torch, nbytes):
    """Leave NaN-filled blocks in torch's caching allocator, so outputs a run
    allocates next start as NaN rather than as an earlier run's values."""
    x = torch.full((nbytes // 8,), float("nan"), dtype=torch.float64, device="cuda")
    del x


@pytest.mark.parametrize("nx", [8, 15, 16, 22, 48, 140])
@pytest.mark.parametrize("cube_edge", [True, False])
def test_frame_forms_match_oracle(nx, cube_edge):
    """Both forms of the launch after the fused kernel (one cube-edge launch
    for frame, velocity halo and next edge pass; or the frame launch and a
    halo / edge launch) give the oracle's bits for a stage and for 3 RK2
    substeps, eager and in graph replays.  nx = 15, 16 and 22 put the edge of
    a block's 15 positions at and next to the face end; every halo cell
    (corners included) is compared."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(nx)
    ctx = barotr.SWContext(consts, torch.float64)
    assert ctx.frame_form()  # the default where the grid allows it
    ctx.set_frame_form(cube_edge)
    assert ctx.frame_form() == cube_edge
    core = random_core(nx, seed=9)
    want = B.step_rk_logic(*core, consts, 20.0, 0.0, True)
    _poison_cache(torch, 64 * 6 * (nx + 6) ** 2 * 8)
    got = ctx.stage(*[_dev(x, torch) for x in core], 20.0, 0.0, True)
    for name, gv, wv in zip(NAMES13, got, want):
        _assert_exact(f"C{nx} cube_edge={cube_edge} {name}", _host(gv), wv)
    want = B.barotr_rk2_core(state, consts, 3, dt)
    for graph in (False, True):
        _poison_cache(torch, 64 * 6 * (nx + 6) ** 2 * 8)
        got = ctx.rk([_dev(x, torch) for x in state[:6]], 3, dt, graph=graph)
        for k, idx in enumerate((0, 2, 3)):
            _assert_exact(f"C{nx} cube_edge={cube_edge} graph={graph} rk2 field {idx}", _host(got[k]), want[idx])


def test_frame_forms_agree_in_fp32():
    """The FP32 forms evaluate the same fused multiply-adds: both frame forms
    give the same bits over 4 RK2 substeps at C48."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(48)
    ctx = barotr.SWContext(consts, torch.float32)
    runs = []
    for cube_edge in (True, False):
        ctx.set_frame_form(cube_edge)
        runs.append(ctx.rk([_dev(x, torch, torch.float32) for x in state[:6]], 4, dt, graph=True))
    for a, b in zip(*runs):
        assert torch.equal(a, b)


def test_default_tuning_read_back():
    """sw_create picks the fused kernel's band height from the resolution and
    the device; the context reports what the library holds."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    for nx, dtype in ((48, torch.float64), (140, torch.float64), (140, torch.float32)):
        _, consts = grid_consts(nx)
        ctx = barotr.SWContext(consts, dtype)
        ra, rb, rf = ctx.tuning
        assert (ra, rb) == (32, 32) and (rf in (8, 12, 16, 20, 24, 28, 32, 36, 40, 48, 64) or ctx.split > 0)
        ctx.set_tuning(16, 20, 0)
        assert ctx.tuning == (16, 20, 0)
        assert ctx.like().tuning == (16, 20, 0)


@pytest.mark.parametrize("graph,rows_f", [(True, 32), (False, 32), (True, 0), (False, 0)])
def test_rk2_c140(graph, rows_f):
    """A size with three strips per face row and several bands, eager and
    graph, fused stage kernel and separate kernels."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(140)
    ctx = barotr.SWContext(consts, torch.float64)
    ctx.set_tuning(32, 32, rows_f)
    core = [_dev(x, torch) for x in state[:6]]
    got = ctx.rk(core, 5, dt, graph=graph)
    want = B.barotr_rk2_core(state, consts, 5, dt)
    for k, idx in enumerate((0, 2, 3)):
        _assert_exact(f"C140 graph={graph} field {idx}", _host(got[k]), want[idx])
    for k in range(10):
        _assert_exact(f"C140 diagnostic {k}", _host(got[3 + k]), want[6 + k])


@pytest.mark.parametrize("which,flags", [
    (("dz", "shared", "gsym"), (False, False, False)),
    (("shared",), (True, False, True)),
])
@pytest.mark.parametrize("last", [False, True])
def test_generic_layout_matches_oracle(which, flags, last):
    """Kernel paths for consts without the grid's uniform / shared structure."""
    torch = _torch()
    from licom4py_gpu_amd.barotr import SWContext
    nx = 13
    _, base = grid_consts(nx)
    consts = generic_consts(base, which)
    ctx = SWContext(consts, torch.float64)
    f = ctx.layout_flags()
    assert (f["dz_uniform"], f["shared_metrics"], f["gct_symmetric"]) == flags
    core = random_core(nx, seed=11 + int(last))
    want = B.step_rk_logic(*core, consts, 30.0, 0.5, last)
    got = ctx.stage(*[_dev(x, torch) for x in core], 30.0, 0.5, last)
    for name, gv, wv in zip(NAMES13, got, want):
        _assert_exact(f"{which} {name}", _host(gv), wv)


# ---------------------------------------------------------------------------
# FP32 form and the binary64 compensated integrals
# ---------------------------------------------------------------------------
def _f32_round(arrs):
    return tuple(np.asarray(a, np.float32).astype(np.float64) for a in arrs)


def test_fp32_stage_close_to_oracle():
    """One FP32 stage vs the FP64 oracle on the same (float32-rounded) inputs:
    relative L2 <= 1e-5 on every output (FP32 rounding is ~6e-8 per operation)."""
    torch = _torch()
    from licom4py_gpu_amd.barotr import SWContext
    nx = 48
    _, consts = grid_consts(nx)
    ctx = SWContext(consts, torch.float32)
    core = _f32_round(random_core(nx, seed=12))
    want = B.step_rk_logic(*core, consts, 30.0, 0.0, True)
    got = ctx.stage(*[_dev(x, torch, torch.float32) for x in core], 30.0, 0.0, True)
    for name, gv, wv in zip(NAMES13, got, want):
        err = rel_l2(_host(gv).astype(np.float64), wv)
        assert err <= 1e-5, f"fp32 stage {name}: rel L2 {err:.3e}"


def test_fp32_rk2_c48_vs_fp64_oracle():
    """FP32 state and stencils, 200 Courant-0.5 substeps at C48: height and
    velocity relative L2 vs the FP64 oracle within 1e-5."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(48)
    state = _f32_round(state)
    want = B.barotr_rk2_core(state, consts, 200, dt)
    got = barotr.barotr_rk2_core(tuple(_dev(x, torch, torch.float32) for x in state),
                                 barotr.as_device_consts(consts), 200, dt)
    for k in (0, 2, 3):
        err = rel_l2(_host(got[k]).astype(np.float64)[:, 3:-3, 3:-3], want[k][:, 3:-3, 3:-3])
        assert err <= 1e-5, f"fp32 {B.STATE_KEYS[k]}: rel L2 {err:.3e}"


def _integrals_reference(state, consts, dtype):
    import math
    h, u, v = (np.asarray(x, dtype).astype(np.float64)[:, 3:-3, 3:-3] for x in (state[0], state[2], state[3]))
    sl = (slice(None), slice(3, -3), slice(3, -3))
    rda = np.asarray(consts[14], dtype).astype(np.float64)[sl]
    dz = np.asarray(consts[0], dtype).astype(np.float64)[sl]
    G = np.asarray(consts[12], dtype).astype(np.float64)[sl]
    da = 1.0 / rda
    depth = dz + h
    q2 = u * (G[..., 0, 0] * u + G[..., 0, 1] * v) + v * (G[..., 1, 0] * u + G[..., 1, 1] * v)
    mass = math.fsum((depth * da).ravel())
    energy = math.fsum((da * (0.5 * depth * q2 + 0.5 * 9.8 * h * h)).ravel())
    return mass, energy


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_integrals_match_exact_sum(dtype):
    torch = _torch()
    from licom4py_gpu_amd.barotr import SWContext
    tdt = torch.float64 if dtype == "f64" else torch.float32
    ndt = np.float64 if dtype == "f64" else np.float32
    state, consts, dt = seeded_state(96)
    ctx = SWContext(consts, tdt)
    m, e = ctx.integrals(*(_dev(np.asarray(state[k], ndt), torch) for k in (0, 2, 3)))
    rm, re_ = _integrals_reference(state, consts, ndt)
    assert abs(m - rm) <= 1e-14 * abs(rm)
    assert abs(e - re_) <= 1e-13 * abs(re_)


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-13), ("f32", 1e-6)])
def test_mass_drift_small(dtype, tol):
    torch = _torch()
    from licom4py_gpu_amd import barotr
    tdt = torch.float64 if dtype == "f64" else torch.float32
    state, consts, dt = seeded_state(48)
    dconsts = barotr.as_device_consts(consts)
    ctx = barotr.context_for(dconsts, tdt)
    s = tuple(_dev(x, torch, tdt) for x in state)
    m0, _ = ctx.integrals(s[0], s[2], s[3])
    s = barotr.barotr_rk2_core(s, dconsts, 100, dt)
    m1, _ = ctx.integrals(s[0], s[2], s[3])
    assert abs(m1 - m0) / m0 <= tol


@pytest.mark.parametrize("rk", [2, 3])
def test_donate_updates_state_in_place(rk):
    """donate=True writes the results into the state's own tensors (eager for
    nbb < 3, substep-pair graphs above) and equals the default call bit for bit;
    aliased outputs without the in-place flag are refused."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    state, consts, dt = seeded_state(16)
    dconsts = barotr.as_device_consts(consts)
    fn = barotr.barotr_rk2_core if rk == 2 else barotr.barotr_rk3_core
    for nbb in (1, 2, 3, 6):
        want = fn(tuple(_dev(x, torch) for x in state), dconsts, nbb, dt)
        mine = tuple(_dev(x, torch) for x in state)
        got = fn(mine, dconsts, nbb, dt, donate=True)
        assert all(g is m for g, m in zip(got, mine))
        for k in range(16):
            _assert_exact(f"donate rk{rk} nbb={nbb} {B.STATE_KEYS[k]}", _host(got[k]), _host(want[k]))
    ctx = barotr.context_for(dconsts, torch.float64)
    core = [_dev(x, torch) for x in state[:6]]
    ref = ctx.rk(core, 5, dt, rk=rk, graph=False)
    outs = [core[0], core[2], core[3]] + ctx.empty(10)
    ctx.rk(core, 5, dt, rk=rk, out=outs, graph=False, inplace=True)
    for k in range(13):
        _assert_exact(f"inplace eager out{k}", _host(outs[k]), _host(ref[k]))
    with pytest.raises(RuntimeError, match="alias"):
        ctx.rk(core, 2, dt, rk=rk, out=[core[0], core[2], core[3]] + ctx.empty(10))


@pytest.mark.parametrize("donate", [False, True])
def test_driver_snapshots_match_oracle(donate):
    """The chunked driver's host snapshots (taken on a side stream while the
    next chunk runs) equal the oracle's fields after each chunk."""
    torch = _torch()
    from licom4py_gpu_amd import barotr
    from licom4py_gpu_amd.schedule import Driver, global_diag
    state, consts, dt = seeded_state(12)
    snaps = []
    drv = Driver(tuple(_dev(x, torch) for x in state), barotr.as_device_consts(consts), 3, dt,
                 donate=donate, on_snapshot=lambda tag, h, u, v: snaps.append((tag, h.copy(), u.copy(), v.copy())))
    final = drv.run(5, diag_interval=2)
    tags = [s[0] for s in snaps]
    assert tags == ["Initial time", "Time: step 2", "Time: step 4", "Time: step 5"]
    s = state
    for i, steps in enumerate((0, 2, 2, 1)):
        if steps:
            s = B.schedule_core(s, consts, steps, 3, dt)
        for name, j, got in (("h0", 0, snaps[i][1]), ("ub", 2, snaps[i][2]), ("vb", 3, snaps[i][3])):
            _assert_exact(f"snapshot {tags[i]} {name}", got, s[j])
    for k in range(23):
        _assert_exact(f"driver final {B.STATE_KEYS[k]}", _host(final[k]), s[k])
    st = global_diag(final[0], final[2], final[3])
    assert st["h0"][0] == float(np.max(s[0][..., 3:-3, 3:-3]))
