"""Initial conditions for the barotropic state.

* ``seeded_balanced`` — the benchmark recipe: height anomaly = sum of random
  real spherical harmonics of degree 1..8 (each degree l spanned by 2l+1 zonal
  harmonics P_l(a.r) about random axes a, N(0,1) coefficients), rescaled to
  max|h| = ``amp`` over the interior, with velocity in geostrophic balance at a
  constant f0 = 2*Omega*sin(45 deg).  Gradients are analytic, so the poles of
  faces 3 and 6 need no special case.  Velocity is rotated to cube covariant
  components with ``a_l2c`` as in initial/w92_field.py:68-73, at every padded
  point.
* ``w92_case2`` — the reference's Williamson-92 case-2 field
  (initial/w92_field.py:36-96, ubar = 1 m/s, radius 6.371e6).

``state_tuple`` returns the 23-tuple in ``STATE_KEYS`` order with
h0p = h0, ubp = ub, vbp = vb (w92_field.py:94-96) and zero diagnostics and
coupling fields.
"""

from __future__ import annotations

import concurrent.futures
import math
import os
from typing import Dict, Tuple

import numpy as np

GRAV = 9.8
OMEGA = 7.292e-5
RADIUS = 6.3712e6          # grid radius, lib_grid.F90:16
F0 = 2.0 * OMEGA * math.sin(math.pi / 4.0)
MAX_DEGREE = 8


def _unit_xyz(a_pt: np.ndarray) -> np.ndarray:
    lon, lat = a_pt[..., 0], a_pt[..., 1]
    cl = np.cos(lat)
    return np.stack([cl * np.cos(lon), cl * np.sin(lon), np.sin(lat)], axis=-1)


def _legendre(l: int, x: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """P_l(x) and P_l'(x) by the three-term recurrences."""
    p_prev, p = np.ones_like(x), x.copy()
    d_prev, d = np.zeros_like(x), np.ones_like(x)
    if l == 0:
        return p_prev, d_prev
    for n in range(1, l):
        p_next = ((2 * n + 1) * x * p - n * p_prev) / (n + 1)
        d_next = d_prev + (2 * n + 1) * p
        p_prev, p = p, p_next
        d_prev, d = d, d_next
    return p, d


def harmonic_coefficients(seed: int, perturb: float = 0.0, member_seed: int | None = None):
    """Axes and coefficients of the random harmonic field.

    ``perturb`` > 0 multiplies each coefficient by (1 + perturb * N(0,1)) drawn
    from ``member_seed`` (ensemble members share the base field).
    """
    rng = np.random.default_rng(seed)
    terms = []
    for l in range(1, MAX_DEGREE + 1):
        for _ in range(2 * l + 1):
            a = rng.normal(size=3)
            a /= np.linalg.norm(a)
            terms.append((l, a, rng.normal()))
    if perturb > 0.0:
        prng = np.random.default_rng(member_seed if member_seed is not None else seed + 1)
        terms = [(l, a, c * (1.0 + perturb * prng.normal())) for (l, a, c) in terms]
    return terms


def _harmonic_chunk(r: np.ndarray, terms):
    h = np.zeros(r.shape[:-1])
    grad = np.zeros(r.shape)
    for l, a, c in terms:
        x = r @ a
        p, dp = _legendre(l, x)
        h += c * p
        grad += (c * dp)[..., None] * (a - x[..., None] * r)
    return h, grad


def _harmonic_sum(r: np.ndarray, terms):
    """Sum of the harmonic terms and of their surface gradients at the points
    r (..., 3).  Every operation is elementwise per point, so the points are
    split into chunks evaluated by a thread pool (NumPy releases the GIL in
    its loops); the result does not depend on the split."""
    flat = r.reshape(-1, 3)
    n = flat.shape[0]
    workers = max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 1))
    nchunk = workers if n >= 65536 else 1
    bounds = np.linspace(0, n, nchunk + 1).astype(int)
    h = np.empty(n)
    grad = np.empty((n, 3))

    def run(k):
        lo, hi = bounds[k], bounds[k + 1]
        h[lo:hi], grad[lo:hi] = _harmonic_chunk(flat[lo:hi], terms)

    if nchunk == 1:
        run(0)
    else:
        with concurrent.futures.ThreadPoolExecutor(max_workers=nchunk) as ex:
            list(ex.map(run, range(nchunk)))
    return h.reshape(r.shape[:-1]), grad.reshape(r.shape)


def seeded_balanced(g: Dict[str, np.ndarray], seed: int = 20261203, amp: float = 200.0,
                    perturb: float = 0.0, member_seed: int | None = None,
                    terms=None):
    """(h0, ub, vb, speed) on the padded grid; see module docstring."""
    r = _unit_xyz(g["a_pt"])
    if terms is None:
        terms = harmonic_coefficients(seed, perturb, member_seed)
    h, grad = _harmonic_sum(r, terms)
    interior = h[:, 3:-3, 3:-3]
    scale = amp / np.abs(interior).max()
    h *= scale
    grad *= scale
    # geostrophic velocity u = (g / (f0 R)) r x grad_s h
    uvec = (GRAV / (F0 * RADIUS)) * np.cross(r, grad)
    lon, lat = g["a_pt"][..., 0], g["a_pt"][..., 1]
    elon = np.stack([-np.sin(lon), np.cos(lon), np.zeros_like(lon)], axis=-1)
    elat = np.stack([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)], axis=-1)
    u_lon = np.sum(uvec * elon, axis=-1)
    u_lat = np.sum(uvec * elat, axis=-1)
    l2c = g["a_l2c"]
    ub = l2c[..., 0, 0] * u_lon + l2c[..., 0, 1] * u_lat
    vb = l2c[..., 1, 0] * u_lon + l2c[..., 1, 1] * u_lat
    speed = np.sqrt(u_lon ** 2 + u_lat ** 2)
    return h, ub, vb, speed


def courant_dt(g: Dict[str, np.ndarray], h: np.ndarray, speed: np.ndarray,
               depth: float = 5000.0, courant: float = 0.5) -> float:
    """dt = courant * min over the interior of min(dx, dy) / (sqrt(g (H + h)) + |u|)."""
    s = (slice(None), slice(3, -3), slice(3, -3))
    dx = 1.0 / g["rdx"][s]
    dy = 1.0 / g["rdy"][s]
    c = np.sqrt(GRAV * (depth + h[s])) + speed[s]
    return float(courant * np.min(np.minimum(dx, dy) / c))


def w92_case2(g: Dict[str, np.ndarray], ubar: float = 1.0):
    """Reference W92 case 2 (initial/w92_field.py:53-91): (h0, ub, vb).  The
    reference uses ubar = 1 m/s; Williamson et al. (1992) use
    ubar = 2 pi a / 12 days (about 38.6 m/s)."""
    lon, lat = g["a_pt"][..., 0], g["a_pt"][..., 1]
    alpha, radius, omega, grav = 0.0, 6.371e6, 7.292e-5, 9.80
    u_lon = ubar * (np.cos(alpha) * np.cos(lat) + np.sin(alpha) * np.cos(lon) * np.sin(lat))
    u_lat = -ubar * np.sin(alpha) * np.sin(lon)
    l2c = g["a_l2c"]
    ub = l2c[..., 0, 0] * u_lon + l2c[..., 0, 1] * u_lat
    vb = l2c[..., 1, 0] * u_lon + l2c[..., 1, 1] * u_lat
    h0 = (-(radius * omega * ubar + ubar ** 2 / 2.0)
          * ((-np.cos(lon) * np.cos(lat) * np.sin(alpha) + np.sin(lat) * np.cos(alpha)) ** 2 - 1.0 / 3.0)
          / grav)
    return h0, ub, vb


def state_tuple(h0: np.ndarray, ub: np.ndarray, vb: np.ndarray):
    """23-tuple in STATE_KEYS order (schedule.py:62-92), *p fields = current."""
    z = lambda: np.zeros_like(h0)  # noqa: E731
    return (h0.copy(), h0.copy(), ub.copy(), vb.copy(), ub.copy(), vb.copy(),
            *(z() for _ in range(10)), *(z() for _ in range(7)))
