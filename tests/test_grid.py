"""Invariants of the generated duo-grid (CPU).

The package's C++ generator restates the reference's Fortran grid library
(chtholly_get_global_grid.F90, global_grid_base.F90, gen_lonlat.F90,
gen_cell.F90, gen_vec.F90, gen_k2e.F90, lib_grid.F90).  No generated field is
shipped with the reference, so the tables are pinned by what the geometry fixes.
"""

from __future__ import annotations

import json
import os

import numpy as np
import pytest

from licom4py_gpu_amd import duogrid

R_EARTH = 6.3712e6          # lib_grid.F90:16
TWO_OMEGA = 2 * 7.292e-5    # global_grid_base.F90:193-202 (alpha = 0)
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _interior(a):
    return a[:, 3:-3, 3:-3]


def _strip_mask(n):
    m = np.zeros((6, n, n), bool)
    m[:, 3:-3, :3] = m[:, 3:-3, -3:] = True
    m[:, :3, 3:-3] = m[:, -3:, 3:-3] = True
    return m


@pytest.mark.parametrize("nx", [8, 24, 48])
def test_cell_areas_tile_the_sphere(nx):
    g = duogrid.grid(nx)
    area = np.sum(1.0 / _interior(g["rda"]))
    assert abs(area / (4 * np.pi * R_EARTH ** 2) - 1) < 1e-11


@pytest.mark.parametrize("nx", [8, 48])
def test_coriolis_is_two_omega_sin_lat(nx):
    g = duogrid.grid(nx)
    lat = g["a_pt"][..., 1]
    np.testing.assert_allclose(_interior(g["a_f"]), TWO_OMEGA * np.sin(_interior(lat)), rtol=1e-13, atol=1e-19)


@pytest.mark.parametrize("nx", [8, 48])
def test_face_symmetries(nx):
    g = duogrid.grid(nx)
    # x <-> y exchange: rdx is rdy transposed, bit for bit
    assert np.array_equal(g["rdx"], np.swapaxes(g["rdy"], 1, 2))
    # the face-independent metric planes are identical on all six faces
    for k in ("rdx", "rdy", "rda", "c_dy", "d_dx"):
        for t in range(1, 6):
            assert np.array_equal(g[k][0], g[k][t]), (k, t)
    # cell area has the 8-fold symmetry of a cube face
    rda = _interior(g["rda"])[0]
    for other in (rda.T, rda[::-1], rda[:, ::-1], rda[::-1, ::-1].T):
        np.testing.assert_allclose(other, rda, rtol=1e-11)
    # a_gct symmetric with equal diagonal (what the kernels' two-plane read uses)
    G = g["a_gct"]
    assert np.array_equal(G[..., 0, 1], G[..., 1, 0])
    assert np.array_equal(G[..., 0, 0], G[..., 1, 1])


@pytest.mark.parametrize("nx", [8, 48])
def test_rotation_pairs_are_inverse(nx):
    g = duogrid.grid(nx)
    I = np.einsum("...ab,...bc->...ac", g["a_c2l"], g["a_l2c"])
    np.testing.assert_allclose(_interior(I[..., 0, 0]), 1.0, atol=1e-12)
    np.testing.assert_allclose(_interior(I[..., 0, 1]), 0.0, atol=1e-12)


@pytest.mark.parametrize("nx", [8, 48])
def test_k2e_tables(nx):
    n = nx + 6
    g = duogrid.grid(nx)
    strips = _strip_mask(n)
    # order-2 Lagrange weights sum to one on the halo strips (lib_grid.F90:213-231)
    np.testing.assert_allclose(g["k2e_coef"].sum(-1)[strips], 1.0, rtol=0, atol=4e-16)
    li, lj = g["k2e_loc_i"], g["k2e_loc_j"]
    ew = li[:, 3:-3, :][:, :, [0, 1, 2, n - 3, n - 2, n - 1]]
    ns = lj[:, [0, 1, 2, n - 3, n - 2, n - 1], :][:, :, 3:-3]
    for loc in (ew, ns.transpose(0, 2, 1)):
        assert loc.min() >= 3 and loc.max() <= nx + 1
        assert np.all(np.diff(loc, axis=1) >= 0)   # monotone along the strip


@pytest.mark.parametrize("nx", [48, 96, 192, 384, 768, 1536])
def test_k2e_has_no_near_tie(nx):
    """The integer k2e tables come from a binary search over coordinates
    computed with libm (gen_k2e.F90:122-132); a searched value within a few
    ulp of a table node could round the other way under another libm.  The
    smallest gap met at every resolution the workloads use (C48 ... C1536) is
    at least 1e-7 rad, nine orders of magnitude above such differences."""
    _, margin = duogrid.generate(nx, with_margin=True)
    assert margin > 1e-7


@pytest.mark.parametrize("nx", [8, 48])
def test_coordinate_line(nx):
    line = duogrid.coordinate_line(nx)
    assert line.shape == (2 * nx + 13,)
    assert np.array_equal(line, -line[::-1])
    assert np.all(np.diff(line) > 0)


def test_npz_round_trip(tmp_path):
    g = duogrid.grid(8)
    p = str(tmp_path / "duogrid_C8.npz")
    duogrid.save_npz(p, g)
    back = duogrid.load_npz(p)
    for k in duogrid.ESSENTIAL_FIELDS:
        assert np.array_equal(back[k], g[k]) and back[k].dtype == g[k].dtype
    with pytest.raises(FileNotFoundError):
        duogrid.load_npz(str(tmp_path / "missing.npz"))
    np.savez(str(tmp_path / "partial.npz"), rdx=g["rdx"])
    with pytest.raises(KeyError):
        duogrid.load_npz(str(tmp_path / "partial.npz"))


def test_grid_matches_golden_digest():
    """Regression pin: the C8 tables hash to the committed digest."""
    with open(os.path.join(GOLDEN, "grid_c8_digest.json")) as f:
        want = json.load(f)
    from tests.golden.make_golden import grid_digest
    assert grid_digest(duogrid.generate(8)) == want
