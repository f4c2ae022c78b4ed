"""Vectorised NumPy FP64 oracle of the reference barotropic step.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or the CPU baseline.

Restates, for one device holding all six faces (px = py = pdev = 1):
  operators/poly.py:25-155, operators/agrid.py:32-69, operators/remap.py:35-109,
  kernel/communication.py:25-87, 193-235, 270-338, 343-474, 485-570,
  kernel/cube.py:26-153, momentum/barotr_pure.py:21-168,
  momentum/barotr.py:41-87, mymodule/schedule.py:62-120, 148-190.
Arrays carry the face axis first: (6, N, N) with N = nx + 6, or (6, 2, N, N)
for the stacked vector halo exchange.  NumPy never fuses multiply-add, so every
expression below is evaluated with the reference's association order.

How each part is pinned (tests/test_oracle_pins.py unless noted).  The
reference ships no golden vectors and its JAX path cannot run here, so the
pins are known answers and what the mathematics fixes:
  routing_table / cube_edges   known-answer table, cross-checked against the
                               Fortran neighbour rule; involution
  update_domain (halo copy)    unique-id fill checked against the closed form
  cube_rmp / ext_scalar        second-order reproduction of a smooth field;
                               per-point loop restatement (loop_ref.py)
  ext_vector                   loop restatement; rotation pairs inverse
                               (tests/test_grid.py)
  Poly (vector_*, scalar_*)    exact on polynomial cell averages (degree 4 / 5)
  boundary_communication       both sides of every cube edge equal afterwards
  step_rk_logic (all operators)  per-point loop restatement bit-identical at
                               C8 / C10; resting state stays at rest; mass
                               conserved to round-off with flux averaging
  barotr_rk2_core / rk3_core   two (three) explicit stages
Parity unpinned (no import or fixture can fix them here): the rounding
sequence of the reference's XLA CPU trajectory (FMA contraction or reordered
sums), for which the 1e-12 relative-L2 bar is the tolerance.
"""

from __future__ import annotations

import numpy as np

H = 3  # halo width, kernel/gpu_mesh.py:27

# ---------------------------------------------------------------------------
# Poly (operators/poly.py:29-37)
# ---------------------------------------------------------------------------
EP23 = -1.0 / 20.0
EP13 = 9.0 / 20.0
EP03 = 47.0 / 60.0
EM13 = -13.0 / 60.0
EM23 = 1.0 / 30.0
M13 = 37.0 / 60.0
M23 = -2.0 / 15.0
M33 = 1.0 / 60.0


def _pad(core, axis_pads):
    """Zero-pad the last two axes (the jnp.pad calls of the reference)."""
    pads = [(0, 0)] * (core.ndim - 2) + list(axis_pads)
    return np.pad(core, pads)


def vector_ew(e):
    """poly.py:40-65 — 5-point faces along axis -2, valid on [2, N-2)."""
    ce = (EM23 * e[..., 0:-4, :] + EM13 * e[..., 1:-3, :] + EP03 * e[..., 2:-2, :]
          + EP13 * e[..., 3:-1, :] + EP23 * e[..., 4:, :])
    cw = (EP23 * e[..., 0:-4, :] + EP13 * e[..., 1:-3, :] + EP03 * e[..., 2:-2, :]
          + EM13 * e[..., 3:-1, :] + EM23 * e[..., 4:, :])
    return _pad(ce, [(2, 2), (0, 0)]), _pad(cw, [(2, 2), (0, 0)])


def vector_ns(e):
    """poly.py:68-93 — 5-point faces along axis -1, valid on [2, N-2)."""
    cn = (EM23 * e[..., :, 0:-4] + EM13 * e[..., :, 1:-3] + EP03 * e[..., :, 2:-2]
          + EP13 * e[..., :, 3:-1] + EP23 * e[..., :, 4:])
    cs = (EP23 * e[..., :, 0:-4] + EP13 * e[..., :, 1:-3] + EP03 * e[..., :, 2:-2]
          + EM13 * e[..., :, 3:-1] + EM23 * e[..., :, 4:])
    return _pad(cn, [(0, 0), (2, 2)]), _pad(cs, [(0, 0), (2, 2)])


def vector(e):
    """poly.py:96-107."""
    ee, ew = vector_ew(e)
    en, es = vector_ns(e)
    return ee, ew, en, es


def scalar_x(s):
    """poly.py:110-126 — 6-point centred value, valid on [3, N-2)^2."""
    core = (M13 * (s[..., 2:-3, 3:-2] + s[..., 3:-2, 3:-2])
            + M23 * (s[..., 1:-4, 3:-2] + s[..., 4:-1, 3:-2])
            + M33 * (s[..., :-5, 3:-2] + s[..., 5:, 3:-2]))
    return _pad(core, [(3, 2), (3, 2)])


def scalar_y(s):
    """poly.py:129-145."""
    core = (M13 * (s[..., 3:-2, 2:-3] + s[..., 3:-2, 3:-2])
            + M23 * (s[..., 3:-2, 1:-4] + s[..., 3:-2, 4:-1])
            + M33 * (s[..., 3:-2, :-5] + s[..., 3:-2, 5:]))
    return _pad(core, [(3, 2), (3, 2)])


# ---------------------------------------------------------------------------
# AGrid (operators/agrid.py:32-69)
# ---------------------------------------------------------------------------
def vorticity(vv, uu, rda, d_dx, c_dy):
    """agrid.py:32-41."""
    core = rda[..., :-1, :-1] * (
        uu[..., :-1, :-1] * d_dx[..., :-1, :-1]
        + vv[..., 1:, :-1] * c_dy[..., 1:, :-1]
        - uu[..., :-1, 1:] * d_dx[..., :-1, 1:]
        - vv[..., :-1, :-1] * c_dy[..., :-1, :-1])
    return _pad(core, [(0, 1), (0, 1)])


def div(uu, vv, rda, c_dy, d_dx):
    """agrid.py:44-53."""
    core = rda[..., :-1, :-1] * (
        uu[..., 1:, :-1] * c_dy[..., 1:, :-1]
        - uu[..., :-1, :-1] * c_dy[..., :-1, :-1]
        + vv[..., :-1, 1:] * d_dx[..., :-1, 1:]
        - vv[..., :-1, :-1] * d_dx[..., :-1, :-1])
    return _pad(core, [(0, 1), (0, 1)])


def grad(eta, rdx, rdy):
    """agrid.py:56-69."""
    uh = scalar_x(eta)
    vh = scalar_y(eta)
    cx = rdx[..., 3:-3, 3:-3] * (uh[..., 4:-2, 3:-3] - uh[..., 3:-3, 3:-3])
    cy = rdy[..., 3:-3, 3:-3] * (vh[..., 3:-3, 4:-2] - vh[..., 3:-3, 3:-3])
    return _pad(cx, [(3, 3), (3, 3)]), _pad(cy, [(3, 3), (3, 3)])


# ---------------------------------------------------------------------------
# Cube topology and routing (communication.py:25-87, 270-291)
# ---------------------------------------------------------------------------
E, W, N, S = 0, 1, 2, 3
ID, FLIP = 0, 1
NTILE = 6


def _joint_transform(d_a, d_b):
    """communication.py:40-55: same-axis joints keep order, cross-axis flip."""
    same = {(E, W), (W, E), (N, S), (S, N)}
    flip = {(E, S), (W, N), (S, E), (N, W)}
    if (d_a, d_b) in same:
        return ID
    if (d_a, d_b) in flip:
        return FLIP
    raise KeyError((d_a, d_b))


def cube_edges():
    """communication.py:57-87 — the 12 face joints (tA, dA, tB, dB, tid)."""
    edges = []
    for t in range(NTILE):
        if t % 2 == 0:
            rules = {S: ((t + 4) % 6, E), E: ((t + 2) % 6, S),
                     N: ((t + 1) % 6, S), W: ((t + 5) % 6, E)}
        else:
            rules = {E: ((t + 1) % 6, W), N: ((t + 2) % 6, W),
                     W: ((t + 4) % 6, N), S: ((t + 5) % 6, N)}
        for d_a, (nb, d_b) in rules.items():
            if t < nb:
                edges.append((t, d_a, nb, d_b, _joint_transform(d_a, d_b)))
    return edges


def routing_table():
    """communication.py:277-284 — int32 (6, 4, 3): (neighbour, its edge, tid)."""
    r = np.zeros((NTILE, 4, 3), np.int32)
    for ta, da, tb, db, tid in cube_edges():
        r[ta, da] = (tb, db, tid)
        r[tb, db] = (ta, da, tid)
    return r


ROUTING = routing_table()


# ---------------------------------------------------------------------------
# Halo exchange, px = py = 1 (communication.py:193-223, 326-338, 343-430)
# ---------------------------------------------------------------------------
def _pack_edge(u, d):
    """communication.py:193-207 (full=False), 343-353: 3 interior layers."""
    if d == E:
        return np.swapaxes(u[..., H:-H, -2 * H:-H], -1, -2)
    if d == W:
        return np.swapaxes(u[..., H:-H, H:2 * H], -1, -2)
    if d == N:
        return u[..., -2 * H:-H, H:-H]
    return u[..., H:2 * H, H:-H]


def _transform(data, recv_dir, tid, n):
    """communication.py:138-172 — rotate an incoming pack into the receiver's
    orientation (E/W receive: swap axes, FLIP reverses along the edge)."""
    raw = data[..., :, :n]
    if recv_dir in (E, W):
        out = np.swapaxes(raw, -1, -2)
        return np.flip(out, axis=-2) if tid == FLIP else out
    return np.flip(raw, axis=-1) if tid == FLIP else raw


_TARGET = {
    E: (Ellipsis, slice(H, -H), slice(-H, None)),
    W: (Ellipsis, slice(H, -H), slice(None, H)),
    N: (Ellipsis, slice(-H, None), slice(H, -H)),
    S: (Ellipsis, slice(None, H), slice(H, -H)),
}


def update_domain(u):
    """communication.py:387-430 — single global pass (the only one run when
    px = py = 1).  u: (6, ..., N, N); packs are taken before any write."""
    nx = u.shape[-1] - 2 * H
    packs = {d: _pack_edge(u, d) for d in (E, W, N, S)}
    out = u.copy()
    for t in range(NTILE):
        for d in (E, W, N, S):
            tb, db, tid = (int(x) for x in ROUTING[t, d])
            out[(t,) + _TARGET[d]] = _transform(packs[db][tb], d, tid, nx)
    return out


def boundary_communication(u, v):
    """communication.py:485-564 — average the C-grid normal component on every
    face edge with the neighbour's (values packed before any update)."""
    nx = u.shape[-1] - 2 * H
    pk = {
        E: v[..., H:-H, -3].copy(),
        W: v[..., H:-H, 3].copy(),
        N: u[..., -3, H:-H].copy(),
        S: u[..., 3, H:-H].copy(),
    }

    def line_from(t, d):
        tb, db, tid = (int(x) for x in ROUTING[t, d])
        padded = np.zeros(pk[db][tb].shape[:-1] + (H, nx))
        padded[..., 0, :] = pk[db][tb]
        arr = _transform(padded, d, tid, nx)
        return arr[..., :, 0] if d in (E, W) else arr[..., 0, :]

    u = u.copy()
    v = v.copy()
    for t in range(NTILE):
        v[t, ..., H:-H, -3] = (v[t, ..., H:-H, -3] + line_from(t, E)) / 2.0
        v[t, ..., H:-H, 3] = (v[t, ..., H:-H, 3] + line_from(t, W)) / 2.0
        u[t, ..., -3, H:-H] = (u[t, ..., -3, H:-H] + line_from(t, N)) / 2.0
        u[t, ..., 3, H:-H] = (u[t, ..., 3, H:-H] + line_from(t, S)) / 2.0
    return u, v


# ---------------------------------------------------------------------------
# Duo-grid remap (cube.py:40-153)
# ---------------------------------------------------------------------------
def cube_rmp(var, coef, loc_i, loc_j):
    """cube.py:40-127 for px = py = 1.  var: (6, N, N) or (6, k, N, N)."""
    if var.ndim > 3:
        return np.stack([cube_rmp(var[:, c], coef, loc_i, loc_j)
                         for c in range(var.shape[1])], axis=1)
    var = var.copy()
    g = H
    # west
    vw = var[..., :, :g]
    lw = loc_i[..., g:-g, :g]
    new = (coef[..., g:-g, :g, 0] * np.take_along_axis(vw, lw, axis=-2)
           + coef[..., g:-g, :g, 1] * np.take_along_axis(vw, lw + 1, axis=-2))
    var[..., g:-g, :g] = new
    # east
    ve = var[..., :, -g:]
    le = loc_i[..., g:-g, -g:]
    new = (coef[..., g:-g, -g:, 0] * np.take_along_axis(ve, le, axis=-2)
           + coef[..., g:-g, -g:, 1] * np.take_along_axis(ve, le + 1, axis=-2))
    var[..., g:-g, -g:] = new
    # south
    vs = var[..., :g, :]
    ls = loc_j[..., :g, g:-g]
    new = (coef[..., :g, g:-g, 0] * np.take_along_axis(vs, ls, axis=-1)
           + coef[..., :g, g:-g, 1] * np.take_along_axis(vs, ls + 1, axis=-1))
    var[..., :g, g:-g] = new
    # north
    vn = var[..., -g:, :]
    ln = loc_j[..., -g:, g:-g]
    new = (coef[..., -g:, g:-g, 0] * np.take_along_axis(vn, ln, axis=-1)
           + coef[..., -g:, g:-g, 1] * np.take_along_axis(vn, ln + 1, axis=-1))
    var[..., -g:, g:-g] = new
    # corners from the edge-updated rows
    sw = var[..., g, :g][..., None, :].copy()
    se = var[..., g, -g:][..., None, :].copy()
    nw = var[..., -g - 1, :g][..., None, :].copy()
    ne = var[..., -g - 1, -g:][..., None, :].copy()
    var[..., :g, :g] = sw
    var[..., :g, -g:] = se
    var[..., -g:, :g] = nw
    var[..., -g:, -g:] = ne
    return var


def ext_scalar(var, coef, loc_i, loc_j):
    """cube.py:130-134."""
    return cube_rmp(update_domain(var), coef, loc_i, loc_j)


def ext_vector(u, v, coef, loc_i, loc_j, c2l, l2c, inner, outer):
    """cube.py:137-153."""
    ull = (c2l[..., 0, 0] * u + c2l[..., 0, 1] * v) * inner
    vll = (c2l[..., 1, 0] * u + c2l[..., 1, 1] * v) * inner
    p = ext_scalar(np.stack([ull, vll], axis=1), coef, loc_i, loc_j)
    un = (l2c[..., 0, 0] * p[:, 0] + l2c[..., 0, 1] * p[:, 1]) * outer
    vn = (l2c[..., 1, 0] * p[:, 0] + l2c[..., 1, 1] * p[:, 1]) * outer
    return un + u * inner, vn + v * inner


# ---------------------------------------------------------------------------
# Remap (remap.py:35-109)
# ---------------------------------------------------------------------------
def to_a_grid(u, v, gct, sina):
    """remap.py:35-39."""
    uct = (gct[..., 0, 0] * u + gct[..., 0, 1] * v) * sina
    vct = (gct[..., 1, 0] * u + gct[..., 1, 1] * v) * sina
    return uct, vct


def to_c_grid(u, v):
    """remap.py:42-55."""
    ue, uw = vector_ew(u)
    un, us = vector_ns(v)
    uc = _pad(0.5 * (ue[..., 2:-3, :] + uw[..., 3:-2, :]), [(3, 2), (0, 0)])
    vc = _pad(0.5 * (un[..., :, 2:-3] + us[..., :, 3:-2]), [(0, 0), (3, 2)])
    return uc, vc


def to_d_grid_upwind(u, v, uc, vc):
    """remap.py:74-95 (strict > 0 selects the east/north face)."""
    un, us = vector_ns(u)
    ve, vw = vector_ew(v)
    vd = _pad(np.where(uc[..., 3:-2, :] > 0.0, ve[..., 2:-3, :], vw[..., 3:-2, :]),
              [(3, 2), (0, 0)])
    ud = _pad(np.where(vc[..., :, 3:-2] > 0.0, un[..., :, 2:-3], us[..., :, 3:-2]),
              [(0, 0), (3, 2)])
    return ud, vd


def vector_trans_2d(u, v, gct, sina):
    """remap.py:98-109."""
    uct, vct = to_a_grid(u, v, gct, sina)
    ucx, vcy = to_c_grid(uct, vct)
    ucx, vcy = boundary_communication(ucx, vcy)
    ucy, vcx = to_d_grid_upwind(u, v, ucx, vcy)
    return uct, vct, ucx, ucy, vcx, vcy


# ---------------------------------------------------------------------------
# Stage physics (barotr_pure.py:21-168)
# ---------------------------------------------------------------------------
def get_celerity(h, dzph_x, dzph_y):
    """barotr_pure.py:21-26."""
    hx, hy = scalar_x(h), scalar_y(h)
    cx = np.sqrt(9.8 * (hx + dzph_x))
    cy = np.sqrt(9.8 * (hy + dzph_y))
    return np.maximum(cx, 10.0), np.maximum(cy, 10.0)


def get_vel_vis_2d(cx, cy, h, u, v):
    """barotr_pure.py:29-34."""
    he, hw, hn, hs = vector(h)
    u = u.copy()
    v = v.copy()
    u[..., 3:-2, :] += 4.9 / cx[..., 3:-2, :] * (he[..., 2:-3, :] - hw[..., 3:-2, :])
    v[..., :, 3:-2] += 4.9 / cy[..., :, 3:-2] * (hn[..., :, 2:-3] - hs[..., :, 3:-2])
    return u, v


def flux_calculation(h, ucx, vcy, dzph_x, dzph_y):
    """barotr_pure.py:37-49."""
    he, hw, hn, hs = vector(h)
    core_x = np.where(ucx[..., 1:, :] > 0.0, he[..., :-1, :], hw[..., 1:, :])
    core_y = np.where(vcy[..., :, 1:] > 0.0, hn[..., :, :-1], hs[..., :, 1:])
    fu = dzph_x * ucx
    fv = dzph_y * vcy
    fu[..., 1:, :] += core_x * ucx[..., 1:, :]
    fv[..., :, 1:] += core_y * vcy[..., :, 1:]
    return fu, fv


def fb_scheme(h_new, h_old, beta):
    """barotr_pure.py:52-54."""
    return (1.0 - beta) * h_new + beta * h_old


def calculate_pgf(h, rdx, rdy):
    """barotr_pure.py:57-60."""
    gx, gy = grad(h, rdx, rdy)
    return -9.80 * gx, -9.80 * gy


def calculate_pgf_vis_2d(cx, cy, rdx, rdy, u, v):
    """barotr_pure.py:63-75."""
    ue, uw = vector_ew(u)
    vn, vs = vector_ns(v)
    core_x = 0.5 * cx[..., 3:-2, :] * (ue[..., 2:-3, :] - uw[..., 3:-2, :])
    core_y = 0.5 * cy[..., :, 3:-2] * (vn[..., :, 2:-3] - vs[..., :, 3:-2])
    dx = rdx[..., 3:-3, 3:-3] * (core_x[..., 1:, 3:-3] - core_x[..., :-1, 3:-3])
    dy = rdy[..., 3:-3, 3:-3] * (core_y[..., 3:-3, 1:] - core_y[..., 3:-3, :-1])
    return _pad(-dx, [(3, 3), (3, 3)]), _pad(-dy, [(3, 3), (3, 3)])


def calculate_advection(vcx, ucy, ucx, vcy, vct, uct, rdx, rdy, rda, d_dx, c_dy):
    """barotr_pure.py:78-97."""
    vort = vorticity(vcx, ucy, rda, d_dx, c_dy)
    ku = 0.5 * (ucx ** 2 + vcx ** 2)
    kv = 0.5 * (vcy ** 2 + ucy ** 2)
    ax = vort[..., :-1, :] * vct[..., :-1, :] - (ku[..., 1:, :] - ku[..., :-1, :]) * rdx[..., :-1, :]
    ay = -vort[..., :, :-1] * uct[..., :, :-1] - (kv[..., :, 1:] - kv[..., :, :-1]) * rdy[..., :, :-1]
    return _pad(ax, [(0, 1), (0, 0)]), _pad(ay, [(0, 0), (0, 1)])


def step_rk_logic(h, hp, u, v, up, vp, consts, dt, beta, last):
    """barotr_pure.py:100-168 — one RK stage.  consts: the 17 barotropic consts
    (coupling fields are unused by the step and may be appended or omitted)."""
    (dzph_x, dzph_y, rdx, rdy, a_f, coef, loc_i, loc_j, c2l, l2c, inner, outer,
     gct, sina, rda, c_dy, d_dx) = consts[:17]
    cx, cy = get_celerity(h, dzph_x, dzph_y)
    uct, vct, ucx, ucy, vcx, vcy = vector_trans_2d(u, v, gct, sina)
    ucx, vcy = get_vel_vis_2d(cx, cy, h, ucx, vcy)
    fu, fv = flux_calculation(h, ucx, vcy, dzph_x, dzph_y)
    if last:
        fu, fv = boundary_communication(fu, fv)
    d = div(fu, fv, rda, c_dy, d_dx)
    hn = hp - d * dt
    hn = ext_scalar(hn, coef, loc_i, loc_j)
    advx, advy = calculate_advection(vcx, ucy, ucx, vcy, vct, uct, rdx, rdy, rda, d_dx, c_dy)
    vis_x, vis_y = calculate_pgf_vis_2d(cx, cy, rdx, rdy, uct, vct)
    ht = fb_scheme(hn, h, beta)
    pgf_u, pgf_v = calculate_pgf(ht, rdx, rdy)
    un = up + (pgf_u + advx + a_f * vct + vis_x) * dt
    vn = vp + (pgf_v + advy - a_f * uct + vis_y) * dt
    un, vn = ext_vector(un, vn, coef, loc_i, loc_j, c2l, l2c, inner, outer)
    return (hn, un, vn, cx, cy, uct, vct, ucx, ucy, vcx, vcy, advx, advy)


# ---------------------------------------------------------------------------
# RK cores and schedule (barotr.py:41-87, schedule.py:62-120, 148-190)
# ---------------------------------------------------------------------------
STATE_KEYS = ("h0", "h0p", "ub", "vb", "ubp", "vbp", "celerity_x", "celerity_y",
              "ub_ct", "vb_ct", "ub_cx", "ub_cy", "vb_cx", "vb_cy", "advx", "advy",
              "pax", "pxb", "whx", "pay", "pyb", "why", "wgp")
N_BAROTR_STATE = 23
N_BAROTR_CONSTS = 17


def barotr_rk2_core(state, consts, nbb, dtb):
    """barotr.py:41-61 — nbb RK2-midpoint substeps; coupling fields pass through."""
    coupling = tuple(state[16:23])
    s = tuple(state)
    for _ in range(nbb):
        h0, h0p, ub, vb, ubp, vbp = s[:6]
        r1 = step_rk_logic(h0, h0p, ub, vb, ubp, vbp, consts, dtb / 2.0, 0.0, False)
        r2 = step_rk_logic(r1[0], h0p, r1[1], r1[2], ubp, vbp, consts, dtb, 0.0, True)
        s = (r2[0], r2[0], r2[1], r2[2], r2[1], r2[2]) + tuple(r2[3:]) + coupling
    return s


def barotr_rk3_core(state, consts, nbb, dtb):
    """barotr.py:64-87 — three stages at dt/3, dt/2, dt with beta = 1."""
    coupling = tuple(state[16:23])
    s = tuple(state)
    for _ in range(nbb):
        h0, h0p, ub, vb, ubp, vbp = s[:6]
        r1 = step_rk_logic(h0, h0p, ub, vb, ubp, vbp, consts, dtb / 3.0, 1.0, False)
        r2 = step_rk_logic(r1[0], h0p, r1[1], r1[2], ubp, vbp, consts, dtb / 2.0, 1.0, False)
        r3 = step_rk_logic(r2[0], h0p, r2[1], r2[2], ubp, vbp, consts, dtb, 1.0, True)
        s = (r3[0], r3[0], r3[1], r3[2], r3[1], r3[2]) + tuple(r3[3:]) + coupling
    return s


def schedule_core(state, consts, chunk_size, nbb, dtb, rk=2):
    """schedule.py:148-190 — chunk_size baroclinic steps, each running the
    barotropic core on state[:23] with consts[:17]; extra fields pass through."""
    core = barotr_rk2_core if rk == 2 else barotr_rk3_core
    s = tuple(state)
    for _ in range(chunk_size):
        s = tuple(core(s[:N_BAROTR_STATE], consts[:N_BAROTR_CONSTS], nbb, dtb)) + s[N_BAROTR_STATE:]
    return s
