"""The cube-edge frame kernel's block decomposition, modelled on the CPU.

k_frame_edge (csrc/barotr.hip: frame_edge_body, eside) gives block b the
cube edge b / nseg and seg along-edge positions of its first face edge (42 in
single-context launches, 28 in batched ones); the
face edge across the joint owns the mirrored positions.  Each side evaluates
its frame cells over the owned positions widened by 3 (its tile), fills the
velocity and height halo of the owned positions (and, at positions 0 / nx-1
of a W / E face edge, the 3 x 3 halo corners) from the other side's tile, and
writes the next stage's edge values of the owned positions.

This model restates that index logic and the halo plan's entries
(k_plan_build: the strip / corner cells and their two interpolation sources
at along-edge indices L-3, L-2 from k2e_loc) with the library's routing
table, and checks, at every workload resolution:
  * every frame cell, every halo cell (strips and corners) and every edge
    entry (24 nx) is written by exactly one block;
  * both sources of every halo cell a block fills lie in the other side's
    tile, so the kernel reads them from LDS;
  * the six stencil points of every edge value lie in the block's tile or
    among the halo cells it filled.
"""

from __future__ import annotations

import pytest

from licom4py_gpu_amd import _hip, duogrid

DE, DW, DN, DS = 0, 1, 2, 3
HALO = 3
SEGS = (42, 28)  # owned positions per block: kSeg, kBatchSeg


def _halo_entry(nx, N, t, row, col):
    if col >= N - HALO:
        e = 3 * (row - HALO) + (col - (N - HALO))
    elif col < HALO:
        e = 3 * nx + 3 * (row - HALO) + (HALO - 1 - col)
    elif row >= N - HALO:
        e = 6 * nx + (row - (N - HALO)) * nx + (col - HALO)
    else:
        e = 9 * nx + (HALO - 1 - row) * nx + (col - HALO)
    return t * (12 * nx + 36) + e


def _interior_cell(N, d, layer, r):
    if d == DE:
        return 3 + r, N - 3 - layer
    if d == DW:
        return 3 + r, 2 + layer
    if d == DN:
        return N - 3 - layer, 3 + r
    return 2 + layer, 3 + r


def _plan(nx, g, R):
    """Plan entry q -> (halo cell (t, i, j), source 0, source 1)."""
    N = nx + 6
    li, lj = g["k2e_loc_i"], g["k2e_loc_j"]
    per = 12 * nx + 36
    out = {}
    for t in range(6):
        for q0 in range(per):
            qf = q0
            if qf < 6 * nx:  # W/E strips: entry order (d, r, layer)
                d = qf // (3 * nx)
                rem = qf - d * 3 * nx
                qf = d * 3 * nx + (rem % 3) * nx + rem // 3
            if qf < 12 * nx:
                d = qf // (3 * nx)
                rem = qf - d * 3 * nx
                layer, r = 1 + rem // nx, rem % nx
                i, j = {DE: (3 + r, N - 4 + layer), DW: (3 + r, 3 - layer),
                        DN: (N - 4 + layer, 3 + r), DS: (3 - layer, 3 + r)}[d]
                si, sj = i, j
            else:  # corners copy row 3 / N-4 of the W/E strips
                c = qf - 12 * nx
                corner, w = c // 9, c % 9
                north, east = corner >= 2, corner & 1
                i = N - 3 + w // 3 if north else w // 3
                j = N - 3 + w % 3 if east else w % 3
                si, sj = (N - 4 if north else 3), j
            if sj < 3 or sj >= N - 3:
                d, layer, L = (DW, 3 - sj, int(li[t, si, sj])) if sj < 3 else (DE, sj - (N - 4), int(li[t, si, sj]))
            else:
                d, layer, L = (DS, 3 - si, int(lj[t, si, sj])) if si < 3 else (DN, si - (N - 4), int(lj[t, si, sj]))
            tb, db, flip = R[t][d]
            srcs = []
            for x in (L - 3, L - 2):
                rr = nx - 1 - x if flip else x
                srcs.append((tb,) + _interior_cell(N, db, layer, rr))
            out[t * per + q0] = ((t, i, j), srcs[0], srcs[1])
    return out


def _side(nx, N, t, d, lo, hi):
    pa = max(lo - 3, 0)
    m = min(hi + 3, nx) - pa
    if d in (DS, DN):
        tile = (3 if d == DS else N - 6, 3, 3 + pa, m)
    else:
        tile = (3 + pa, m, 3 if d == DW else N - 6, 3)
    cr0, crows, cc0, ccols = tile
    cells = {(t, cr0 + q // ccols, cc0 + q % ccols) for q in range(crows * ccols)}
    return {"t": t, "d": d, "lo": lo, "hi": hi, "cells": cells, "tile": tile}


@pytest.mark.parametrize("nx,seg", [(nx, SEGS[0]) for nx in (7, 8, 15, 16, 22, 48, 96, 192, 384, 768, 1536)]
                         + [(nx, SEGS[1]) for nx in (7, 8, 15, 16, 22, 29, 48, 96, 192, 384, 768, 1536)])
def test_frame_edge_blocks_cover_every_cell_once(nx, seg):
    N = nx + 6
    R = _hip.routing_table()
    plan = _plan(nx, duogrid.grid(nx), R)
    ue = [t * 4 + d for t in range(6) for d in range(4) if t * 4 + d < R[t][d][0] * 4 + R[t][d][1]]
    assert len(ue) == 12
    nseg = (nx + seg - 1) // seg
    frame, halo, edge = {}, {}, {}
    for blk in range(12 * nseg):
        e, p0 = blk // nseg, (blk % nseg) * seg
        p1 = min(p0 + seg, nx)
        ta, da = ue[e] >> 2, ue[e] & 3
        tb, db, flip = (int(x) for x in R[ta][da])
        S = (_side(nx, N, ta, da, p0, p1),
             _side(nx, N, tb, db, nx - p1 if flip else p0, nx - p0 if flip else p1))
        for X in S:
            cr0, crows, cc0, ccols = X["tile"]
            assert crows * ccols <= 3 * (seg + 6) and (crows + 6) * (ccols + 6) <= (seg + 12) * 9  # LDS sizes
            ns = X["d"] in (DS, DN)
            for (t, i, j) in X["cells"]:
                pos = j - 3 if ns else i - 3
                if X["lo"] <= pos < X["hi"] and (ns or 3 <= pos < nx - 3):
                    frame[(t, i, j)] = frame.get((t, i, j), 0) + 1
        filled = [{}, {}]
        for hs in range(2):
            H, Y = S[hs], S[1 - hs]
            nown = H["hi"] - H["lo"]
            we = H["d"] in (DW, DE)
            c_lo, c_hi = int(we and H["lo"] == 0), int(we and H["hi"] == nx)
            assert 3 * nown + 9 * (c_lo + c_hi) <= 3 * seg + 18  # halo entries per side (LDS rotation slots)
            for k in range(3 * nown + 9 * (c_lo + c_hi)):
                if k < 3 * nown:
                    layer, pos = k // nown + 1, H["lo"] + k % nown
                    d = H["d"]
                    row = 3 - layer if d == DS else N - 4 + layer if d == DN else 3 + pos
                    col = 3 - layer if d == DW else N - 4 + layer if d == DE else 3 + pos
                    q = _halo_entry(nx, N, H["t"], row, col)
                else:
                    kc = k - 3 * nown
                    north = kc >= 9 or not c_lo
                    corner = (2 if north else 0) + (1 if H["d"] == DE else 0)
                    q = H["t"] * (12 * nx + 36) + 12 * nx + corner * 9 + kc % 9
                cell, s0, s1 = plan[q]
                assert cell[0] == H["t"]
                assert s0 in Y["cells"] and s1 in Y["cells"], (blk, hs, k)
                halo[cell] = halo.get(cell, 0) + 1
                filled[hs][cell] = True
        for k in range(p1 - p0):
            pa_ = p0 + k
            for X, p in ((0, pa_), (1, nx - 1 - pa_ if flip else pa_)):
                Sx = S[X]
                ns = Sx["d"] in (DS, DN)
                line = 3 if Sx["d"] in (DS, DW) else N - 3
                for kk in range(6):
                    i, j = (line - 3 + kk, 3 + p) if ns else (3 + p, line - 3 + kk)
                    pt = (Sx["t"], i, j)
                    assert pt in Sx["cells"] or pt in filled[X]
                key = (Sx["t"], Sx["d"], p)
                edge[key] = edge.get(key, 0) + 1
    frame_all = {(t, i, j) for t in range(6) for i in range(3, N - 3) for j in range(3, N - 3)
                 if i < 6 or i >= N - 6 or j < 6 or j >= N - 6}
    assert set(frame) == frame_all and max(frame.values()) == 1
    halo_all = {v[0] for v in plan.values()}
    assert len(halo_all) == 6 * (12 * nx + 36)
    assert set(halo) == halo_all and max(halo.values()) == 1
    assert len(edge) == 24 * nx and max(edge.values()) == 1
