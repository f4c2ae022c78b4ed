"""Cubed-sphere duo-grid tables: generation, on-disk format and const tuple.

The grid is produced by the package's own C++ restatement of the reference's
Fortran generator (``csrc/duogrid_gen.cpp``) and written in the reference's
``field/duogrid_C{nx}.npz`` layout (initial_field/main.py:30-39, 96-97,
116-124): 14 essential arrays stacked over the face axis, C-contiguous.

``barotr_consts`` assembles the 17-tuple the barotropic step consumes, in the
order of ``CONST_SOURCES`` (mymodule/schedule.py:95-116), including the fields
the reference derives at start-up: ``inner``/``outer`` masks
(duogrid/duogrid.py:160-190), constant depth ``dzph_x``/``dzph_y``
(duogrid.py:193-229) and ``loc_i_local``/``loc_j_local``
(kernel/cube.py:166-207, which for a single device is ``k2e_loc_i/j``).
"""

from __future__ import annotations

import ctypes
import os
from typing import Dict

import numpy as np

from . import build as _build

HALO = 3  # kernel/gpu_mesh.py:27, kernel/cube.py:26

ESSENTIAL_FIELDS = (
    "a_sina", "a_f", "rdx", "rdy", "rda", "a_gct",
    "k2e_coef", "k2e_loc_i", "k2e_loc_j", "a_c2l", "a_l2c",
    "c_dy", "d_dx", "a_pt",
)

_lib = None


def _grid_lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.GRID_SO):
            _build.build_grid()
        lib = ctypes.CDLL(_build.GRID_SO)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        lib.duogrid_generate.restype = ctypes.c_int
        lib.duogrid_generate.argtypes = [ctypes.c_int] + [dp] * 12 + [ip, ip, dp]
        lib.duogrid_line.restype = ctypes.c_int
        lib.duogrid_line.argtypes = [ctypes.c_int, dp]
        _lib = lib
    return _lib


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def generate(nx: int, with_margin: bool = False):
    """Generate the 14 essential duo-grid fields for resolution C{nx}.

    Returns a dict name -> ndarray (float64, int32 for ``k2e_loc_*``); with
    ``with_margin`` also the k2e near-tie audit value (the smallest gap between
    a searched coordinate and a table node, in radians).
    """
    if nx < 4:
        raise ValueError("nx must be >= 4")
    n = nx + 2 * HALO
    f2 = lambda: np.empty((6, n, n), np.float64)  # noqa: E731
    out = {
        "a_pt": np.empty((6, n, n, 2), np.float64),
        "a_sina": f2(), "a_f": f2(), "rdx": f2(), "rdy": f2(), "rda": f2(),
        "a_gct": np.empty((6, n, n, 2, 2), np.float64),
        "a_c2l": np.empty((6, n, n, 2, 2), np.float64),
        "a_l2c": np.empty((6, n, n, 2, 2), np.float64),
        "c_dy": f2(), "d_dx": f2(),
        "k2e_coef": np.empty((6, n, n, 2), np.float64),
        "k2e_loc_i": np.empty((6, n, n), np.int32),
        "k2e_loc_j": np.empty((6, n, n), np.int32),
    }
    margin = np.zeros(1, np.float64)
    d = ctypes.c_double
    rc = _grid_lib().duogrid_generate(
        nx, _ptr(out["a_pt"], d), _ptr(out["a_sina"], d), _ptr(out["a_f"], d),
        _ptr(out["rdx"], d), _ptr(out["rdy"], d), _ptr(out["rda"], d),
        _ptr(out["a_gct"], d), _ptr(out["a_c2l"], d), _ptr(out["a_l2c"], d),
        _ptr(out["c_dy"], d), _ptr(out["d_dx"], d), _ptr(out["k2e_coef"], d),
        _ptr(out["k2e_loc_i"], ctypes.c_int32), _ptr(out["k2e_loc_j"], ctypes.c_int32),
        _ptr(margin, d))
    if rc != 0:
        raise RuntimeError(f"duogrid_generate failed for nx={nx}")
    if with_margin:
        return out, float(margin[0])
    return out


def coordinate_line(nx: int) -> np.ndarray:
    """Super-grid gnomonic coordinate line, indices 1-6 .. 2nx+7."""
    out = np.empty(2 * nx + 13, np.float64)
    if _grid_lib().duogrid_line(nx, _ptr(out, ctypes.c_double)) != 0:
        raise ValueError("bad nx")
    return out


def save_npz(path: str, fields: Dict[str, np.ndarray]) -> None:
    """Write the fields in the reference's ``duogrid_C{nx}.npz`` format."""
    np.savez(path, **{k: fields[k] for k in ESSENTIAL_FIELDS})


def load_npz(path: str) -> Dict[str, np.ndarray]:
    """Read a ``duogrid_C{nx}.npz`` (no pickles are allowed in the file)."""
    if not os.path.exists(path):
        raise FileNotFoundError(f"Duogrid NPZ file not found: {path}")
    with np.load(path, allow_pickle=False) as data:
        out = {k: np.array(data[k]) for k in data.files}
    missing = set(ESSENTIAL_FIELDS) - set(out)
    if missing:
        raise KeyError(f"missing essential fields: {sorted(missing)}")
    out["k2e_loc_i"] = out["k2e_loc_i"].astype(np.int32)
    out["k2e_loc_j"] = out["k2e_loc_j"].astype(np.int32)
    return out


_GRID_CACHE: Dict[int, Dict[str, np.ndarray]] = {}


def grid(nx: int) -> Dict[str, np.ndarray]:
    """Generated grid for C{nx}, memoised for the process."""
    if nx not in _GRID_CACHE:
        _GRID_CACHE[nx] = generate(nx)
    return _GRID_CACHE[nx]


def masks(nx: int):
    """``inner`` = 1 on the interior, 0 on the 3-wide halo; ``outer`` = 1 - inner
    (duogrid/duogrid.py:160-190)."""
    n = nx + 2 * HALO
    inner = np.zeros((6, n, n), np.float64)
    inner[:, HALO:-HALO, HALO:-HALO] = 1.0
    outer = np.ones((6, n, n), np.float64)
    outer[:, HALO:-HALO, HALO:-HALO] = 0.0
    return inner, outer


def barotr_consts(g: Dict[str, np.ndarray], depth: float = 5600.0):
    """The 17 barotropic consts in ``CONST_SOURCES`` order (schedule.py:95-116).

    ``depth`` is the constant resting depth ``dzph`` (duogrid.py:213 uses 5600).
    """
    n = g["a_sina"].shape[-1]
    nx = n - 2 * HALO
    inner, outer = masks(nx)
    dzph = np.full((6, n, n), float(depth), np.float64)
    return (
        dzph, dzph.copy(), g["rdx"], g["rdy"], g["a_f"],
        g["k2e_coef"], g["k2e_loc_i"].astype(np.int32), g["k2e_loc_j"].astype(np.int32),
        g["a_c2l"], g["a_l2c"], inner, outer,
        g["a_gct"], g["a_sina"], g["rda"], g["c_dy"], g["d_dx"],
    )
