"""ctypes binding of ``libbarotr_hip.so`` (the C-ABI in include/licom_sw.h).

The product path has no CPU fallback: if the library is not built, or no GPU
is visible when a context is created, the calls raise ``RuntimeError``.
"""

from __future__ import annotations

import ctypes
import os
from typing import Sequence

from . import build as _build

SW_F64 = 0
SW_F32 = 1
SW_SKIP_DIAG = 1
SW_GRAPH = 2
SW_INPLACE = 4
KERNEL_CLASSES = ("edge", "march_a", "halo_h", "march_b", "halo_uv", "diag", "fused", "frame")
SW_N_CONSTS = 17

# Every symbol declared in include/licom_sw.h.
EXPORTED = (
    "sw_create", "sw_destroy", "sw_set_stream", "sw_rk2", "sw_rk3", "sw_stage",
    "sw_halo_scalar", "sw_halo_vector", "sw_edge_avg", "sw_sync",
    "sw_routing_table", "sw_info", "sw_last_error", "sw_set_tuning", "sw_get_tuning", "sw_layout_flags", "sw_profile", "sw_integrals",
    "sw_create_like", "sw_batch_create", "sw_batch_destroy", "sw_batch_run", "sw_integrals_noalloc",
    "sw_profile_graph", "sw_set_band_split", "sw_get_band_split", "sw_set_frame_form", "sw_get_frame_form",
)


class SWMember(ctypes.Structure):
    """``sw_member`` of include/licom_sw.h."""
    _fields_ = [("ctx", ctypes.c_void_p), ("inp", ctypes.c_void_p * 6), ("out", ctypes.c_void_p * 13),
                ("nbb", ctypes.c_int), ("dtb", ctypes.c_double)]

_lib = None


def lib() -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = _build.HIP_SO
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is not built; run `python -m licom4py_gpu_amd.build` "
            "(or __graft_entry__.build()) first")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    vpp = ctypes.POINTER(ctypes.c_void_p)
    ci = ctypes.c_int
    L.sw_create.argtypes = [ctypes.POINTER(vp), ci, ci, vpp, vp]
    L.sw_create.restype = ci
    L.sw_destroy.argtypes = [vp]
    L.sw_destroy.restype = None
    L.sw_set_stream.argtypes = [vp, vp]
    L.sw_set_stream.restype = ci
    for name in ("sw_rk2", "sw_rk3"):
        f = getattr(L, name)
        f.argtypes = [vp, vpp, vpp, ci, ctypes.c_double, ci]
        f.restype = ci
    L.sw_stage.argtypes = [vp, vpp, vpp, ctypes.c_double, ctypes.c_double, ci]
    L.sw_stage.restype = ci
    L.sw_halo_scalar.argtypes = [vp, vp, ci]
    L.sw_halo_scalar.restype = ci
    L.sw_halo_vector.argtypes = [vp, vp, vp]
    L.sw_halo_vector.restype = ci
    L.sw_edge_avg.argtypes = [vp, vp, vp]
    L.sw_edge_avg.restype = ci
    L.sw_sync.argtypes = [vp]
    L.sw_sync.restype = ci
    L.sw_routing_table.argtypes = [ctypes.POINTER(ctypes.c_int32)]
    L.sw_routing_table.restype = ci
    L.sw_info.argtypes = [vp, ctypes.POINTER(ci), ctypes.POINTER(ci)]
    L.sw_info.restype = ci
    L.sw_profile.argtypes = [vp, vpp, vpp, ci, ctypes.c_double, ci,
                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ci)]
    L.sw_profile.restype = ci
    L.sw_profile_graph.argtypes = [vp, vpp, vpp, ci, ci, ctypes.c_double,
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ci),
                                   ctypes.POINTER(ctypes.c_double)]
    L.sw_profile_graph.restype = ci
    L.sw_integrals.argtypes = [vp, vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
    L.sw_integrals.restype = ci
    L.sw_integrals_noalloc.argtypes = [vp, vp, vp, vp, vp]
    L.sw_integrals_noalloc.restype = ci
    L.sw_set_tuning.argtypes = [vp, ci, ci, ci]
    L.sw_set_tuning.restype = ci
    L.sw_get_tuning.argtypes = [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci)]
    L.sw_get_tuning.restype = ci
    L.sw_set_band_split.argtypes = [vp, ci]
    L.sw_set_band_split.restype = ci
    L.sw_get_band_split.argtypes = [vp, ctypes.POINTER(ci)]
    L.sw_get_band_split.restype = ci
    L.sw_set_frame_form.argtypes = [vp, ci]
    L.sw_set_frame_form.restype = ci
    L.sw_get_frame_form.argtypes = [vp, ctypes.POINTER(ci)]
    L.sw_get_frame_form.restype = ci
    L.sw_layout_flags.argtypes = [vp, ctypes.POINTER(ci)]
    L.sw_layout_flags.restype = ci
    L.sw_create_like.argtypes = [ctypes.POINTER(vp), vp]
    L.sw_create_like.restype = ci
    L.sw_batch_create.argtypes = [ctypes.POINTER(vp)]
    L.sw_batch_create.restype = ci
    L.sw_batch_destroy.argtypes = [vp]
    L.sw_batch_destroy.restype = None
    L.sw_batch_run.argtypes = [vp, ctypes.POINTER(SWMember), ci, ci, ci, vp]
    L.sw_batch_run.restype = ci
    L.sw_last_error.argtypes = []
    L.sw_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def check(rc: int, what: str = "licom_sw") -> None:
    if rc != 0:
        msg = lib().sw_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (code {rc}): {msg}")


def ptr_array(ptrs: Sequence[int]):
    arr = (ctypes.c_void_p * len(ptrs))(*[ctypes.c_void_p(int(p)) for p in ptrs])
    return ctypes.cast(arr, ctypes.POINTER(ctypes.c_void_p)), arr


def routing_table():
    """The C library's cube routing table as a nested list (6, 4, 3)."""
    out = (ctypes.c_int32 * 72)()
    check(lib().sw_routing_table(out), "sw_routing_table")
    flat = list(out)
    return [[flat[(t * 4 + d) * 3:(t * 4 + d) * 3 + 3] for d in range(4)] for t in range(6)]
