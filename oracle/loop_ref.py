"""Per-point loop restatement of one barotropic RK stage — TEST INFRASTRUCTURE.

An independent second formulation used to cross-check ``barotr_np`` on tiny
grids (C4..C8).  Where ``barotr_np`` follows the reference's array program
(pads, packs, routing transforms, take_along_axis), this module writes every
quantity as a scalar function of (face, i, j) with explicit validity ranges,
and the cross-face exchanges in closed index form:

  halo copy (communication.py:404-430, px = py = 1): halo layer l of edge d of
  face t, along-edge index r, copies layer l of the neighbour (t', d') from
  the routing table, at along-edge index r' = nx-1-r when the joint flips;
  edge averaging (communication.py:485-564): the C-grid normal component on
  the edge line averages with the neighbour's edge line at r';
  k2e remap (cube.py:40-127): two-point interpolation along the halo strip,
  then the 3x3 corners copy the adjacent strip row.

Python floats are IEEE binary64 and no multiply-add is fused, so the values
must agree bit for bit with ``barotr_np``.
"""

from __future__ import annotations

import numpy as np

from .barotr_np import (EM13, EM23, EP03, EP13, EP23, M13, M23, M33, ROUTING,
                        E, W, N, S, FLIP)


def _zeros(n):
    return [[[0.0] * n for _ in range(n)] for _ in range(6)]


def _arr(a):
    return [[[float(a[t, i, j]) for j in range(a.shape[2])] for i in range(a.shape[1])]
            for t in range(6)]


def _np(a):
    return np.array(a, dtype=np.float64)


class _Face:
    """Lazy per-point evaluators over plain nested lists."""

    def __init__(self, n):
        self.n = n


def _e_face(f, t, i, j, n, axis):
    """East (axis 0) or north (axis 1) 5-point face value; 0 outside [2, n-2)."""
    k = i if axis == 0 else j
    if not (2 <= k < n - 2):
        return 0.0
    g = (lambda d: f[t][i + d][j]) if axis == 0 else (lambda d: f[t][i][j + d])
    return EM23 * g(-2) + EM13 * g(-1) + EP03 * g(0) + EP13 * g(1) + EP23 * g(2)


def _w_face(f, t, i, j, n, axis):
    k = i if axis == 0 else j
    if not (2 <= k < n - 2):
        return 0.0
    g = (lambda d: f[t][i + d][j]) if axis == 0 else (lambda d: f[t][i][j + d])
    return EP23 * g(-2) + EP13 * g(-1) + EP03 * g(0) + EM13 * g(1) + EM23 * g(2)


def _centred(s, t, i, j, n, axis):
    if not (3 <= i < n - 2 and 3 <= j < n - 2):
        return 0.0
    g = (lambda d: s[t][i + d][j]) if axis == 0 else (lambda d: s[t][i][j + d])
    return M13 * (g(-1) + g(0)) + M23 * (g(-2) + g(1)) + M33 * (g(-3) + g(2))


def _edge_line_coords(d, r, n):
    """Cell of the C-grid edge line of edge d at along-edge index r."""
    if d == E:
        return 3 + r, n - 3
    if d == W:
        return 3 + r, 3
    if d == N:
        return n - 3, 3 + r
    return 3, 3 + r


def _edge_average(u, v, n):
    """Closed-form edge averaging; u: normal component on N/S edges (axis -2),
    v on E/W edges.  Returns new lists."""
    nx = n - 6
    pre_u = [[row[:] for row in face] for face in u]
    pre_v = [[row[:] for row in face] for face in v]
    out_u = [[row[:] for row in face] for face in u]
    out_v = [[row[:] for row in face] for face in v]

    def val(arr_u, arr_v, t, d, r):
        i, j = _edge_line_coords(d, r, n)
        return arr_v[t][i][j] if d in (E, W) else arr_u[t][i][j]

    for t in range(6):
        for d in (E, W, N, S):
            tb, db, tid = (int(x) for x in ROUTING[t, d])
            for r in range(nx):
                rr = nx - 1 - r if tid == FLIP else r
                mine = val(out_u, out_v, t, d, r)
                nb = val(pre_u, pre_v, tb, db, rr)
                i, j = _edge_line_coords(d, r, n)
                if d in (E, W):
                    out_v[t][i][j] = (mine + nb) / 2.0
                else:
                    out_u[t][i][j] = (mine + nb) / 2.0
    return out_u, out_v


def _halo_cell(d, layer, r, n):
    """Halo cell at distance `layer` (1..3) outside edge d, along-edge index r."""
    if d == E:
        return 3 + r, n - 4 + layer
    if d == W:
        return 3 + r, 3 - layer
    if d == N:
        return n - 4 + layer, 3 + r
    return 3 - layer, 3 + r


def _interior_cell(d, layer, r, n):
    """Interior cell at distance `layer` (1..3) inside edge d."""
    if d == E:
        return 3 + r, n - 3 - layer
    if d == W:
        return 3 + r, 2 + layer
    if d == N:
        return n - 3 - layer, 3 + r
    return 2 + layer, 3 + r


def ext_scalar(f, coef, loc_i, loc_j, n):
    """Halo fill of one scalar list-field (closed forms; see module doc)."""
    nx = n - 6
    x = [[row[:] for row in face] for face in f]
    for t in range(6):
        for d in (E, W, N, S):
            tb, db, tid = (int(q) for q in ROUTING[t, d])
            for layer in (1, 2, 3):
                for r in range(nx):
                    rr = nx - 1 - r if tid == FLIP else r
                    hi, hj = _halo_cell(d, layer, r, n)
                    si, sj = _interior_cell(db, layer, rr, n)
                    x[t][hi][hj] = f[tb][si][sj]
    y = [[row[:] for row in face] for face in x]
    for t in range(6):
        for i in range(3, n - 3):
            for j in list(range(0, 3)) + list(range(n - 3, n)):
                L = int(loc_i[t, i, j])
                y[t][i][j] = coef[t, i, j, 0] * x[t][L][j] + coef[t, i, j, 1] * x[t][L + 1][j]
        for j in range(3, n - 3):
            for i in list(range(0, 3)) + list(range(n - 3, n)):
                L = int(loc_j[t, i, j])
                y[t][i][j] = coef[t, i, j, 0] * x[t][i][L] + coef[t, i, j, 1] * x[t][i][L + 1]
        for c in range(3):
            for k in range(3):
                y[t][k][c] = y[t][3][c]
                y[t][k][n - 3 + c] = y[t][3][n - 3 + c]
                y[t][n - 3 + k][c] = y[t][n - 4][c]
                y[t][n - 3 + k][n - 3 + c] = y[t][n - 4][n - 3 + c]
    return y


def stage(h, hp, u, v, up, vp, consts, dt, beta, last):
    """One RK stage (barotr_pure.py:100-168) by per-point loops.  Arrays in
    and out are numpy (6, n, n); returns the same 13-tuple as barotr_np."""
    (dzph_x, dzph_y, rdx, rdy, a_f, coef, loc_i, loc_j, c2l, l2c, inner, outer,
     gct, sina, rda, c_dy, d_dx) = consts[:17]
    n = h.shape[-1]
    H_, U, V = _arr(h), _arr(u), _arr(v)
    rng = range(n)

    cx, cy = _zeros(n), _zeros(n)
    uct, vct = _zeros(n), _zeros(n)
    for t in range(6):
        for i in rng:
            for j in rng:
                a = _centred(H_, t, i, j, n, 0)
                b = _centred(H_, t, i, j, n, 1)
                cx[t][i][j] = max(np.sqrt(9.8 * (a + dzph_x[t, i, j])), 10.0)
                cy[t][i][j] = max(np.sqrt(9.8 * (b + dzph_y[t, i, j])), 10.0)
                g = gct[t, i, j]
                uct[t][i][j] = (g[0, 0] * U[t][i][j] + g[0, 1] * V[t][i][j]) * sina[t, i, j]
                vct[t][i][j] = (g[1, 0] * U[t][i][j] + g[1, 1] * V[t][i][j]) * sina[t, i, j]

    # C-grid normal components
    ucx, vcy = _zeros(n), _zeros(n)
    for t in range(6):
        for i in rng:
            for j in rng:
                if 3 <= i < n - 2:
                    ucx[t][i][j] = 0.5 * (_e_face(uct, t, i - 1, j, n, 0) + _w_face(uct, t, i, j, n, 0))
                if 3 <= j < n - 2:
                    vcy[t][i][j] = 0.5 * (_e_face(vct, t, i, j - 1, n, 1) + _w_face(vct, t, i, j, n, 1))
    ucx, vcy = _edge_average(ucx, vcy, n)

    # D-grid upwind on the covariant u, v
    ucy, vcx = _zeros(n), _zeros(n)
    for t in range(6):
        for i in rng:
            for j in rng:
                if 3 <= i < n - 2:
                    vcx[t][i][j] = (_e_face(V, t, i - 1, j, n, 0) if ucx[t][i][j] > 0.0
                                    else _w_face(V, t, i, j, n, 0))
                if 3 <= j < n - 2:
                    ucy[t][i][j] = (_e_face(U, t, i, j - 1, n, 1) if vcy[t][i][j] > 0.0
                                    else _w_face(U, t, i, j, n, 1))

    # LMARS velocity correction
    for t in range(6):
        for i in rng:
            for j in rng:
                if 3 <= i < n - 2:
                    ucx[t][i][j] = ucx[t][i][j] + 4.9 / cx[t][i][j] * (
                        _e_face(H_, t, i - 1, j, n, 0) - _w_face(H_, t, i, j, n, 0))
                if 3 <= j < n - 2:
                    vcy[t][i][j] = vcy[t][i][j] + 4.9 / cy[t][i][j] * (
                        _e_face(H_, t, i, j - 1, n, 1) - _w_face(H_, t, i, j, n, 1))

    # upwind mass flux
    fu, fv = _zeros(n), _zeros(n)
    for t in range(6):
        for i in rng:
            for j in rng:
                a = dzph_x[t, i, j] * ucx[t][i][j]
                if i >= 1:
                    up_h = (_e_face(H_, t, i - 1, j, n, 0) if ucx[t][i][j] > 0.0
                            else _w_face(H_, t, i, j, n, 0))
                    a = a + up_h * ucx[t][i][j]
                fu[t][i][j] = a
                b = dzph_y[t, i, j] * vcy[t][i][j]
                if j >= 1:
                    up_h = (_e_face(H_, t, i, j - 1, n, 1) if vcy[t][i][j] > 0.0
                            else _w_face(H_, t, i, j, n, 1))
                    b = b + up_h * vcy[t][i][j]
                fv[t][i][j] = b
    if last:
        fu, fv = _edge_average(fu, fv, n)

    hn = _zeros(n)
    for t in range(6):
        for i in rng:
            for j in rng:
                dv = 0.0
                if i < n - 1 and j < n - 1:
                    dv = rda[t, i, j] * (((fu[t][i + 1][j] * c_dy[t, i + 1, j] - fu[t][i][j] * c_dy[t, i, j])
                                          + fv[t][i][j + 1] * d_dx[t, i, j + 1]) - fv[t][i][j] * d_dx[t, i, j])
                hn[t][i][j] = hp[t, i, j] - dv * dt
    hn = ext_scalar(hn, coef, loc_i, loc_j, n)

    # advection
    zeta = _zeros(n)
    for t in range(6):
        for i in range(n - 1):
            for j in range(n - 1):
                zeta[t][i][j] = rda[t, i, j] * (((ucy[t][i][j] * d_dx[t, i, j] + vcx[t][i + 1][j] * c_dy[t, i + 1, j])
                                                 - ucy[t][i][j + 1] * d_dx[t, i, j + 1]) - vcx[t][i][j] * c_dy[t, i, j])
    advx, advy = _zeros(n), _zeros(n)
    for t in range(6):
        for i in rng:
            for j in rng:
                if i < n - 1:
                    ku1 = 0.5 * (ucx[t][i + 1][j] * ucx[t][i + 1][j] + vcx[t][i + 1][j] * vcx[t][i + 1][j])
                    ku0 = 0.5 * (ucx[t][i][j] * ucx[t][i][j] + vcx[t][i][j] * vcx[t][i][j])
                    advx[t][i][j] = zeta[t][i][j] * vct[t][i][j] - (ku1 - ku0) * rdx[t, i, j]
                if j < n - 1:
                    kv1 = 0.5 * (vcy[t][i][j + 1] * vcy[t][i][j + 1] + ucy[t][i][j + 1] * ucy[t][i][j + 1])
                    kv0 = 0.5 * (vcy[t][i][j] * vcy[t][i][j] + ucy[t][i][j] * ucy[t][i][j])
                    advy[t][i][j] = -zeta[t][i][j] * uct[t][i][j] - (kv1 - kv0) * rdy[t, i, j]

    # LMARS momentum dissipation, pressure gradient, update
    def cxx(t, i, j):
        return 0.5 * cx[t][i][j] * (_e_face(uct, t, i - 1, j, n, 0) - _w_face(uct, t, i, j, n, 0))

    def cyy(t, i, j):
        return 0.5 * cy[t][i][j] * (_e_face(vct, t, i, j - 1, n, 1) - _w_face(vct, t, i, j, n, 1))

    htil = [[[(1.0 - beta) * hn[t][i][j] + beta * H_[t][i][j] for j in rng] for i in rng]
            for t in range(6)]
    un, vn = _zeros(n), _zeros(n)
    for t in range(6):
        for i in rng:
            for j in rng:
                vis_x = vis_y = pgf_u = pgf_v = 0.0
                if 3 <= i < n - 3 and 3 <= j < n - 3:
                    vis_x = -(rdx[t, i, j] * (cxx(t, i + 1, j) - cxx(t, i, j)))
                    vis_y = -(rdy[t, i, j] * (cyy(t, i, j + 1) - cyy(t, i, j)))
                    pgf_u = -9.80 * (rdx[t, i, j] * (_centred(htil, t, i + 1, j, n, 0) - _centred(htil, t, i, j, n, 0)))
                    pgf_v = -9.80 * (rdy[t, i, j] * (_centred(htil, t, i, j + 1, n, 1) - _centred(htil, t, i, j, n, 1)))
                un[t][i][j] = up[t, i, j] + (((pgf_u + advx[t][i][j]) + a_f[t, i, j] * vct[t][i][j]) + vis_x) * dt
                vn[t][i][j] = vp[t, i, j] + (((pgf_v + advy[t][i][j]) - a_f[t, i, j] * uct[t][i][j]) + vis_y) * dt

    # vector halo: rotate to lon/lat, fill, rotate back
    ull, vll = _zeros(n), _zeros(n)
    for t in range(6):
        for i in rng:
            for j in rng:
                c = c2l[t, i, j]
                ull[t][i][j] = (c[0, 0] * un[t][i][j] + c[0, 1] * vn[t][i][j]) * inner[t, i, j]
                vll[t][i][j] = (c[1, 0] * un[t][i][j] + c[1, 1] * vn[t][i][j]) * inner[t, i, j]
    P0 = ext_scalar(ull, coef, loc_i, loc_j, n)
    P1 = ext_scalar(vll, coef, loc_i, loc_j, n)
    uo, vo = _zeros(n), _zeros(n)
    for t in range(6):
        for i in rng:
            for j in rng:
                q = l2c[t, i, j]
                uo[t][i][j] = (q[0, 0] * P0[t][i][j] + q[0, 1] * P1[t][i][j]) * outer[t, i, j] + un[t][i][j] * inner[t, i, j]
                vo[t][i][j] = (q[1, 0] * P0[t][i][j] + q[1, 1] * P1[t][i][j]) * outer[t, i, j] + vn[t][i][j] * inner[t, i, j]

    return tuple(_np(a) for a in (hn, uo, vo, cx, cy, uct, vct, ucx, ucy, vcx, vcy, advx, advy))
