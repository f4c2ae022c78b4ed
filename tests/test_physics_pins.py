"""Pins of the oracle's velocity physics and RK composition against facts of
the continuous equations and of the stencil mathematics, independent of the
two transcriptions (oracle/barotr_np.py, oracle/loop_ref.py):

* Williamson et al. (1992) case 2 is a steady geostrophic solution of the
  rotating shallow-water equations.  The discrete flow must stay near it, and
  its departure must shrink at second order with the grid spacing.  This
  exercises every velocity term at once: Coriolis, pressure gradient,
  vorticity / kinetic-energy advection (about 4 % of the balance at the
  standard speed) and the LMARS dissipation.  The reference's own case
  (initial/w92_field.py:36-96, ubar = 1 m/s) is run too.
* The LMARS momentum dissipation (barotr_pure.py:63-75) on a Fourier mode
  with unit celerity and metric is the hyperdiffusion -(2/15)(1 - cos t)^3:
  the two mirrored 5-point face values differ by a fifth difference
  (poly.py:50-63), so the operator is a sixth difference.
* Mistakes planted in the oracle on the CPU (a Coriolis sign, a pressure
  gradient sign, advection dropped, dissipation dropped) each break one of
  the W92 bounds, which shows the bounds are tight enough to see them.
* RK3 (momentum/barotr.py:64-87) is three stages at dt/3, dt/2, dt with
  beta = 1, the *p fields held, `last` on the third stage only.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from licom4py_gpu_amd import duogrid, initial
from oracle import barotr_np as B
from tests._util import grid_consts, random_core

DEPTH = 5600.0                             # the reference's resting depth (duogrid.py:213)
U_W92 = 2.0 * math.pi * 6.371e6 / (12 * 86400.0)   # Williamson et al. (1992) case-2 speed
INTERIOR = (slice(None), slice(3, -3), slice(3, -3))


def _w92_departure(nx, ubar, substeps, consts_fix=None, patch=None, monkeypatch=None):
    """Relative L2 departure of h and of the covariant u after `substeps` RK2
    substeps (Courant 0.5) from the steady W92 case-2 state."""
    g = duogrid.grid(nx)
    consts = list(duogrid.barotr_consts(g, DEPTH))
    if consts_fix:
        consts_fix(consts)
    h0, u0, v0 = initial.w92_case2(g, ubar)
    c = math.sqrt(9.8 * DEPTH) + ubar
    dt = 0.5 * float(np.min(np.minimum(1.0 / g["rdx"][INTERIOR], 1.0 / g["rdy"][INTERIOR]))) / c
    if patch:
        for name, wrap in patch.items():
            monkeypatch.setattr(B, name, wrap(getattr(B, name)))
    with np.errstate(invalid="ignore", over="ignore"):
        out = B.barotr_rk2_core(initial.state_tuple(h0, u0, v0), tuple(consts), substeps, dt)

    def rel(a, b):
        d = np.linalg.norm((a - b)[INTERIOR]) / np.linalg.norm(b[INTERIOR])
        return d if np.isfinite(d) else np.inf

    return rel(out[0], h0), rel(out[2], u0), rel(out[3], v0)


def test_reference_w92_case_stays_near_steady():
    """The reference's validation case (ubar = 1 m/s) over about 19 hours."""
    dh, du, dv = _w92_departure(24, 1.0, 100)
    assert dh < 1e-3 and du < 2e-3 and dv < 2e-3


def test_w92_departure_converges_at_second_order():
    """Standard case 2 over the same 17 hours at C24 and C48: small, and about
    4x smaller at twice the resolution (a missing or wrong term leaves an O(1)
    imbalance that does not shrink)."""
    a = _w92_departure(24, U_W92, 100)
    b = _w92_departure(48, U_W92, 200)
    assert b[0] < 2e-4 and b[1] < 2.5e-4 and b[2] < 2.5e-4
    assert a[0] / b[0] > 3.0 and a[1] / b[1] > 3.0


def _neg(f):
    return lambda *a: tuple(-x for x in f(*a))


def _zero(f):
    return lambda *a: tuple(0.0 * x for x in f(*a))


@pytest.mark.parametrize("mistake", ["coriolis_sign", "pgf_sign", "advection_dropped", "dissipation_dropped"])
def test_planted_mistakes_break_the_w92_bounds(mistake, monkeypatch):
    if mistake == "coriolis_sign":
        d = _w92_departure(24, U_W92, 100, consts_fix=lambda c: c.__setitem__(4, -c[4]))
    elif mistake == "pgf_sign":
        d = _w92_departure(24, U_W92, 100, patch={"calculate_pgf": _neg}, monkeypatch=monkeypatch)
    elif mistake == "advection_dropped":
        d = _w92_departure(24, U_W92, 100, patch={"calculate_advection": _zero}, monkeypatch=monkeypatch)
    else:
        d = _w92_departure(48, U_W92, 200, patch={"calculate_pgf_vis_2d": _zero}, monkeypatch=monkeypatch)
    # the bounds of the C48 convergence test, which the correct scheme meets
    assert d[0] >= 2e-4 or d[1] >= 2.5e-4 or d[2] >= 2.5e-4


@pytest.mark.parametrize("theta", [0.1, 0.4, 1.0, 2.0, math.pi])
def test_lmars_dissipation_is_a_sixth_difference(theta):
    n = 40
    i = np.arange(n, dtype=np.float64)[:, None] * np.ones((1, n))
    u = np.broadcast_to(np.cos(theta * i), (6, n, n)).copy()
    v = u.transpose(0, 2, 1).copy()          # the same mode along j
    one = np.ones((6, n, n))
    vis_x, vis_y = B.calculate_pgf_vis_2d(one, one, one, one, u, v)
    lam = (2.0 / 15.0) * (1.0 - math.cos(theta)) ** 3
    np.testing.assert_allclose(vis_x[INTERIOR], -lam * u[INTERIOR], rtol=0, atol=1e-13)
    np.testing.assert_allclose(vis_y[INTERIOR], -lam * v[INTERIOR], rtol=0, atol=1e-13)
    assert np.all(vis_x[:, :3] == 0) and np.all(vis_y[:, :, -3:] == 0)  # zero outside the interior


def test_rk3_is_three_stages():
    nx = 8
    _, consts = grid_consts(nx)
    core = random_core(nx, seed=4, distinct_p=False)
    h, u, v = core[0], core[2], core[3]
    st = initial.state_tuple(h, u, v)
    dt = 45.0
    out = B.barotr_rk3_core(st, consts, 1, dt)
    r1 = B.step_rk_logic(h, h, u, v, u, v, consts, dt / 3.0, 1.0, False)
    r2 = B.step_rk_logic(r1[0], h, r1[1], r1[2], u, v, consts, dt / 2.0, 1.0, False)
    r3 = B.step_rk_logic(r2[0], h, r2[1], r2[2], u, v, consts, dt, 1.0, True)
    assert np.array_equal(out[0], r3[0]) and np.array_equal(out[1], r3[0])
    assert np.array_equal(out[2], r3[1]) and np.array_equal(out[4], r3[1])
    assert np.array_equal(out[3], r3[2]) and np.array_equal(out[5], r3[2])
    for k in range(3, 13):
        assert np.array_equal(out[k + 3], r3[k])
    # two substeps chain through the *p fields of the first substep's output
    two = B.barotr_rk3_core(st, consts, 2, dt)
    again = B.barotr_rk3_core(out, consts, 1, dt)
    for a, b in zip(two, again):
        assert np.array_equal(a, b)
