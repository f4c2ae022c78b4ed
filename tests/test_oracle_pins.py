"""Pins of the NumPy oracle against facts independent of its own formulas.

Each test is chosen so that a plausible transcription slip in the oracle (a
wrong neighbour, a reversed edge, a shifted stencil index, a dropped term)
fails it: known-answer tables, closed-form index maps with unique ids, exact
polynomial reproduction of the stencils, the symmetry the edge exchange must
produce, invariants of the continuous equations (rest, mass), second-order
convergence of the halo remap on a smooth field, and bit-for-bit agreement
with the independent per-point loop formulation.
"""

from __future__ import annotations

import numpy as np
import pytest

from licom4py_gpu_amd import duogrid, initial
from oracle import barotr_np as B
from oracle import loop_ref as L
from tests._util import grid_consts, random_core

E, W, N, S = B.E, B.W, B.N, B.S

# Known answer: neighbour face, its edge, transform (0 = ID, 1 = FLIP).
ROUTING_KAT = [
    [[2, S, 1], [5, E, 0], [1, S, 0], [4, E, 1]],
    [[2, W, 0], [5, N, 1], [3, W, 1], [0, N, 0]],
    [[4, S, 1], [1, E, 0], [3, S, 0], [0, E, 1]],
    [[4, W, 0], [1, N, 1], [5, W, 1], [2, N, 0]],
    [[0, S, 1], [3, E, 0], [5, S, 0], [2, E, 1]],
    [[0, W, 0], [3, N, 1], [1, W, 1], [4, N, 0]],
]


def _fortran_neighbours(n):
    """get_neighbor_tile_num of the grid generator (1-based faces), returned as
    (west, east, south, north)."""
    if n % 2 == 0:
        nn, ne, ns, nw = (n + 0) % 6 + 1, (n + 1) % 6 + 1, (n + 3) % 6 + 1, (n + 4) % 6 + 1
    else:
        ne, nn, nw, ns = (n + 0) % 6 + 1, (n + 1) % 6 + 1, (n + 3) % 6 + 1, (n + 4) % 6 + 1
    return nw, ne, ns, nn


def test_routing_known_answer():
    assert B.ROUTING.tolist() == ROUTING_KAT


def test_routing_matches_grid_generator_neighbours():
    # Python E/W are along the last axis (Fortran j: north/south), Python N/S
    # along axis -2 (Fortran i: east/west).
    for t in range(6):
        nw, ne, ns, nn = _fortran_neighbours(t + 1)
        assert B.ROUTING[t, E, 0] == nn - 1
        assert B.ROUTING[t, W, 0] == ns - 1
        assert B.ROUTING[t, N, 0] == ne - 1
        assert B.ROUTING[t, S, 0] == nw - 1


def test_routing_is_an_involution():
    for t in range(6):
        for d in range(4):
            tb, db, tid = B.ROUTING[t, d]
            assert B.ROUTING[tb, db].tolist() == [t, d, tid]
    assert len(B.cube_edges()) == 12


def _closed_form_halo_source(t, d, layer, r, n):
    nx = n - 6
    tb, db, tid = (int(x) for x in B.ROUTING[t, d])
    rr = nx - 1 - r if tid else r
    if db == E:
        return tb, 3 + rr, n - 3 - layer
    if db == W:
        return tb, 3 + rr, 2 + layer
    if db == N:
        return tb, n - 3 - layer, 3 + rr
    return tb, 2 + layer, 3 + rr


def _halo_cell(d, layer, r, n):
    if d == E:
        return 3 + r, n - 4 + layer
    if d == W:
        return 3 + r, 3 - layer
    if d == N:
        return n - 4 + layer, 3 + r
    return 3 - layer, 3 + r


@pytest.mark.parametrize("nx", [5, 8, 11])
def test_update_domain_unique_ids(nx):
    n = nx + 6
    ids = np.arange(6 * n * n, dtype=np.float64).reshape(6, n, n)
    marked = ids.copy()
    marked[:, :3, :] = -1.0
    marked[:, -3:, :] = -1.0
    marked[:, :, :3] = -1.0
    marked[:, :, -3:] = -1.0
    out = B.update_domain(marked)
    for t in range(6):
        for d in (E, W, N, S):
            for layer in (1, 2, 3):
                for r in range(nx):
                    hi, hj = _halo_cell(d, layer, r, n)
                    st, si, sj = _closed_form_halo_source(t, d, layer, r, n)
                    assert out[t, hi, hj] == ids[st, si, sj]
    # corners untouched, interior untouched
    assert np.all(out[:, :3, :3] == -1.0)
    assert np.array_equal(out[:, 3:-3, 3:-3], ids[:, 3:-3, 3:-3])


def _poly_cell_averages(coef, k):
    """Cell averages over [k-1/2, k+1/2] of the polynomial with coefficients."""
    P = np.polynomial.Polynomial(coef).integ()
    return P(k + 0.5) - P(k - 0.5)


def test_poly_faces_exact_on_quartic():
    rng = np.random.default_rng(3)
    coef = rng.normal(size=5)
    k = np.arange(12, dtype=np.float64)
    avg = _poly_cell_averages(coef, k)
    p = np.polynomial.Polynomial(coef)
    field = np.tile(avg[:, None], (1, 3))
    e, w = B.vector_ew(field)
    for i in range(2, 10):
        assert abs(e[i, 0] - p(i + 0.5)) < 1e-10 * max(1.0, abs(p(i + 0.5)))
        assert abs(w[i, 0] - p(i - 0.5)) < 1e-10 * max(1.0, abs(p(i - 0.5)))
    n_, s_ = B.vector_ns(field.T)
    assert np.allclose(n_[0, 2:10], e[2:10, 0], rtol=0, atol=1e-12)
    assert np.allclose(s_[0, 2:10], w[2:10, 0], rtol=0, atol=1e-12)


def test_centred_face_exact_on_quintic():
    rng = np.random.default_rng(4)
    coef = rng.normal(size=6)
    n = 14
    k = np.arange(n, dtype=np.float64)
    avg = _poly_cell_averages(coef, k)
    p = np.polynomial.Polynomial(coef)
    field = np.tile(avg[:, None], (1, n))
    sx = B.scalar_x(field)
    for i in range(3, n - 2):
        assert abs(sx[i, 5] - p(i - 0.5)) < 1e-10 * max(1.0, abs(p(i - 0.5)))
    sy = B.scalar_y(field.T)
    assert np.allclose(sy[5, 3:n - 2], sx[3:n - 2, 5], rtol=0, atol=1e-12)
    # outside the valid window both are zero-padded
    assert np.all(sx[:3] == 0) and np.all(sx[:, -2:] == 0)


def _edge_line(d, r, n):
    return {E: (3 + r, n - 3), W: (3 + r, 3), N: (n - 3, 3 + r), S: (3, 3 + r)}[d]


def test_edge_average_makes_both_sides_equal():
    nx = 9
    n = nx + 6
    rng = np.random.default_rng(9)
    u = rng.normal(size=(6, n, n))
    v = rng.normal(size=(6, n, n))
    u2, v2 = B.boundary_communication(u, v)
    for t in range(6):
        for d in (E, W, N, S):
            tb, db, tid = (int(x) for x in B.ROUTING[t, d])
            for r in range(nx):
                rr = nx - 1 - r if tid else r
                a = (v2 if d in (E, W) else u2)[(t,) + _edge_line(d, r, n)]
                b = (v2 if db in (E, W) else u2)[(tb,) + _edge_line(db, rr, n)]
                a0 = (v if d in (E, W) else u)[(t,) + _edge_line(d, r, n)]
                b0 = (v if db in (E, W) else u)[(tb,) + _edge_line(db, rr, n)]
                assert a == b == (a0 + b0) / 2.0


def test_resting_state_stays_at_rest():
    nx = 12
    _, consts = grid_consts(nx)
    z = np.zeros((6, nx + 6, nx + 6))
    st = initial.state_tuple(z, z, z)
    out = B.barotr_rk2_core(st, consts, 3, 600.0)
    for k in (0, 2, 3):
        assert np.all(out[k] == 0.0)


@pytest.mark.parametrize("nx", [16, 24])
def test_mass_is_conserved(nx):
    g, consts = grid_consts(nx)
    h, u, v, sp = initial.seeded_balanced(g, seed=11)
    dt = initial.courant_dt(g, h, sp)
    st = initial.state_tuple(h, u, v)
    area = 1.0 / g["rda"][:, 3:-3, 3:-3]
    m0 = np.sum(h[:, 3:-3, 3:-3] * area)
    scale = np.sum(np.abs(h[:, 3:-3, 3:-3]) * area)
    out = B.barotr_rk2_core(st, consts, 5, dt)
    m1 = np.sum(out[0][:, 3:-3, 3:-3] * area)
    assert abs(m1 - m0) / scale < 1e-12


def _smooth(xyz):
    x, y, z = xyz[..., 0], xyz[..., 1], xyz[..., 2]
    return np.sin(1.3 * x + 0.4) * np.cos(0.7 * y) + 0.5 * z * z + 0.3 * x * y


def _halo_error(nx):
    g = duogrid.grid(nx)
    lon, lat = g["a_pt"][..., 0], g["a_pt"][..., 1]
    xyz = np.stack([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)], -1)
    f = _smooth(xyz)
    src = np.zeros_like(f)
    src[:, 3:-3, 3:-3] = f[:, 3:-3, 3:-3]
    out = B.ext_scalar(src, g["k2e_coef"], g["k2e_loc_i"], g["k2e_loc_j"])
    mask = np.ones(f.shape, bool)
    mask[:, 3:-3, 3:-3] = False
    for c in ((slice(None, 3), slice(None, 3)), (slice(None, 3), slice(-3, None)),
              (slice(-3, None), slice(None, 3)), (slice(-3, None), slice(-3, None))):
        mask[(slice(None),) + c] = False  # corners are copies, not samples
    return np.max(np.abs(out - f)[mask])


def test_halo_remap_second_order_on_smooth_field():
    e1, e2 = _halo_error(24), _halo_error(48)
    assert e2 < 2e-3
    assert e1 / e2 > 3.0  # ~4 for second order


@pytest.mark.parametrize("nx", [8, 10])
@pytest.mark.parametrize("last", [False, True])
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_loop_formulation_matches_vectorised(nx, last, beta):
    _, consts = grid_consts(nx)
    core = random_core(nx, seed=nx + int(last))
    a = B.step_rk_logic(*core, consts, 30.0, beta, last)
    b = L.stage(*core, consts, 30.0, beta, last)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_rk2_is_two_stages():
    nx = 8
    _, consts = grid_consts(nx)
    core = random_core(nx, seed=1, distinct_p=False)
    st = initial.state_tuple(core[0], core[2], core[3])
    out = B.barotr_rk2_core(st, consts, 1, 40.0)
    r1 = B.step_rk_logic(core[0], core[0], core[2], core[3], core[2], core[3], consts, 20.0, 0.0, False)
    r2 = B.step_rk_logic(r1[0], core[0], r1[1], r1[2], core[2], core[3], consts, 40.0, 0.0, True)
    assert np.array_equal(out[0], r2[0]) and np.array_equal(out[1], r2[0])
    assert np.array_equal(out[2], r2[1]) and np.array_equal(out[5], r2[2])
    for k in range(3, 13):
        assert np.array_equal(out[k + 3], r2[k])


def test_oracle_reproduces_golden_stage():
    """The oracle still produces the committed C8 stage (tests/golden/stage_c8.npz)."""
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "stage_c8.npz")
    with np.load(path, allow_pickle=False) as z:
        core = [z[f"in_{k}"] for k in ("h", "hp", "u", "v", "up", "vp")]
        want = [z[k] for k in z.files if k.startswith("out_")]
        names = [k for k in z.files if k.startswith("out_")]
        dt, beta, last = float(z["dt"]), float(z["beta"]), bool(z["last"])
    _, consts = grid_consts(8)
    got = B.step_rk_logic(*core, consts, dt, beta, last)
    order = ("h0", "ub", "vb", "celerity_x", "celerity_y", "ub_ct", "vb_ct",
             "ub_cx", "ub_cy", "vb_cx", "vb_cy", "advx", "advy")
    for k, g in zip(order, got):
        assert np.array_equal(g, want[names.index(f"out_{k}")]), k
