"""Shared helpers for the test-suite (inputs, comparisons)."""

from __future__ import annotations

import numpy as np

from licom4py_gpu_amd import duogrid, initial

DEPTH = 5000.0


def grid_consts(nx: int, depth: float = DEPTH):
    g = duogrid.grid(nx)
    return g, duogrid.barotr_consts(g, depth)


def random_core(nx: int, seed: int = 0, distinct_p: bool = True):
    """Random (h, hp, u, v, up, vp) with every padded value set (halo too)."""
    n = nx + 6
    rng = np.random.default_rng(seed)
    h = rng.normal(size=(6, n, n)) * 50.0
    u = rng.normal(size=(6, n, n)) * 5.0
    v = rng.normal(size=(6, n, n)) * 5.0
    if distinct_p:
        hp = h + rng.normal(size=(6, n, n))
        up = u + rng.normal(size=(6, n, n)) * 0.1
        vp = v + rng.normal(size=(6, n, n)) * 0.1
    else:
        hp, up, vp = h.copy(), u.copy(), v.copy()
    return h, hp, u, v, up, vp


def seeded_state(nx: int, seed: int = 20261203):
    """23-tuple state of the benchmark recipe and its Courant-0.5 dt."""
    g, consts = grid_consts(nx)
    h, u, v, speed = initial.seeded_balanced(g, seed=seed)
    dt = initial.courant_dt(g, h, speed, depth=DEPTH)
    return initial.state_tuple(h, u, v), consts, dt


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return float(num / den) if den > 0 else float(num)


def max_ulp_diff(a, b) -> int:
    """Largest distance in units of last place between two float64 arrays."""
    a = np.ascontiguousarray(a, np.float64).view(np.int64)
    b = np.ascontiguousarray(b, np.float64).view(np.int64)
    # map to a monotonic integer line
    a = np.where(a < 0, np.int64(-0x8000000000000000) - a, a)
    b = np.where(b < 0, np.int64(-0x8000000000000000) - b, b)
    return int(np.max(np.abs(a - b))) if a.size else 0


def generic_consts(consts, which):
    """Copies of the consts with layout properties the kernels exploit broken:
    "dz" (dzph not uniform), "shared" (rdx differs on face 2), "gsym" (a_gct
    not symmetric on face 3)."""
    c = [np.array(x, copy=True) for x in consts]
    if "dz" in which:
        r = np.random.default_rng(5)
        c[0] = c[0] + r.uniform(-20, 20, size=c[0].shape)
        c[1] = c[1] + r.uniform(-20, 20, size=c[1].shape)
    if "shared" in which:
        c[2][2] *= 1.0 + 1e-3
    if "gsym" in which:
        c[12][3, ..., 0, 1] *= 1.0 + 1e-3
    return tuple(c)
