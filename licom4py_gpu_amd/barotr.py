"""Drop-in barotropic step: the reference's `_function(state, consts, *static)`
contract over device tensors, executed by the HIP kernels.

Mirrors
  momentum/barotr.py:41-61    `_barotr_rk2_core(state, consts, nbb, dtb, px, py)`
  momentum/barotr.py:64-87    `_barotr_rk3_core(...)`
  momentum/barotr_pure.py:100-168 `_step_rk_logic(...)`
  kernel/cube.py:130-153      `ext_scalar`, `ext_vector`
  kernel/communication.py:566-570 `Communication.boundary_communication`
  mymodule/schedule.py:62-120 `STATE_KEYS`, `CONST_SOURCES`, `N_BAROTR_*`
with the same tuple orders and argument meaning.  `state` is the 23-tuple of
(6, N, N) tensors (float64 for the drop-in, float32 for the FP32 form) on one
GPU; `consts` is the 17-tuple (float64 device tensors, int32 index tables).
Inputs are never modified; new tensors are returned (h0p aliases h0, ubp
aliases ub, vbp aliases vb, and the 7 coupling fields are passed through, as
the reference returns them).  Only a single device holding all six faces is
supported: px = py = 1.
"""

from __future__ import annotations

import ctypes
from collections import OrderedDict
from typing import Sequence, Tuple

import numpy as np
import torch

from . import _hip

STATE_KEYS = (
    "h0", "h0p", "ub", "vb", "ubp", "vbp",
    "celerity_x", "celerity_y", "ub_ct", "vb_ct", "ub_cx", "ub_cy", "vb_cx", "vb_cy",
    "advx", "advy",
    "pax", "pxb", "whx", "pay", "pyb", "why", "wgp",
)

CONST_SOURCES = (
    ("Dg", "dzph_x"), ("Dg", "dzph_y"), ("Dg", "rdx"), ("Dg", "rdy"), ("Dg", "a_f"),
    ("Dg", "k2e_coef"), ("Dg", "loc_i_local"), ("Dg", "loc_j_local"),
    ("Dg", "a_c2l"), ("Dg", "a_l2c"), ("Dg", "inner"), ("Dg", "outer"),
    ("Dg", "a_gct"), ("Dg", "a_sina"), ("Dg", "rda"), ("Dg", "c_dy"), ("Dg", "d_dx"),
)

N_BAROTR_STATE = 23
N_BAROTR_CONSTS = 17
HALO = 3

_INT_CONSTS = (6, 7)


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the barotropic step runs only on the HIP path")
    return torch.device("cuda", torch.cuda.current_device())


def as_device_consts(consts: Sequence, device=None) -> Tuple[torch.Tensor, ...]:
    """Move/convert the 17 consts to contiguous device tensors (float64, the
    two index tables int32).  Accepts numpy arrays or tensors."""
    if len(consts) < N_BAROTR_CONSTS:
        raise ValueError(f"expected {N_BAROTR_CONSTS} consts, got {len(consts)}")
    dev = device or _device()
    out = []
    for k, c in enumerate(consts[:N_BAROTR_CONSTS]):
        dt = torch.int32 if k in _INT_CONSTS else torch.float64
        t = torch.as_tensor(np.asarray(c) if not isinstance(c, torch.Tensor) else c)
        out.append(t.to(device=dev, dtype=dt).contiguous())
    return tuple(out)


def _nx_of(consts) -> int:
    n = consts[13].shape[-1]
    return int(n) - 2 * HALO


class SWContext:
    """Owns one `sw_ctx` (internal const planes, scratch) for a resolution and
    precision.  Launches go to the current torch stream of the call."""

    tuning = (32, 32, 32)  # rows_a, rows_b, rows_f (read from the library after sw_create)
    split = 0              # short band height of the fused kernel's split bands (0: uniform)

    def __init__(self, consts: Sequence, dtype: torch.dtype = torch.float64):
        if consts is None:  # filled in by like()
            return
        if dtype not in (torch.float64, torch.float32):
            raise ValueError("dtype must be torch.float64 or torch.float32")
        dc = as_device_consts(consts)
        self.nx = _nx_of(dc)
        self.N = self.nx + 2 * HALO
        for k, c in enumerate(dc):
            if c.shape[:3] != (6, self.N, self.N):
                raise ValueError(f"const {CONST_SOURCES[k][1]} has shape {tuple(c.shape)}")
        self.dtype = dtype
        self.device = dc[0].device
        arr, _keep = _hip.ptr_array([c.data_ptr() for c in dc])
        handle = ctypes.c_void_p()
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _hip.check(_hip.lib().sw_create(ctypes.byref(handle), self.nx,
                                         _hip.SW_F64 if dtype == torch.float64 else _hip.SW_F32,
                                         arr, ctypes.c_void_p(stream)), "sw_create")
        self._h = handle
        rows = [ctypes.c_int() for _ in range(3)]
        _hip.check(_hip.lib().sw_get_tuning(handle, *(ctypes.byref(r) for r in rows)), "sw_get_tuning")
        self.tuning = tuple(r.value for r in rows)
        self.split = self.band_split()

    def like(self) -> "SWContext":
        """Another context of this resolution and precision with its own
        workspace; the device copy of the consts is shared."""
        other = SWContext(None)
        other.nx, other.N, other.dtype, other.device = self.nx, self.N, self.dtype, self.device
        other.tuning = self.tuning  # sw_create_like copies the band heights
        other.split = self.split
        handle = ctypes.c_void_p()
        _hip.check(_hip.lib().sw_create_like(ctypes.byref(handle), self._h), "sw_create_like")
        other._h = handle
        return other

    def __del__(self):
        h = getattr(self, "_h", None)
        lib = getattr(_hip, "_lib", None) if _hip is not None else None  # None at interpreter exit
        if h is not None and h.value and lib is not None:
            lib.sw_destroy(h)
            self._h = None

    # -- helpers -----------------------------------------------------------
    def _bind_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        _hip.check(_hip.lib().sw_set_stream(self._h, ctypes.c_void_p(s)), "sw_set_stream")

    def _check_field(self, x: torch.Tensor, name: str) -> torch.Tensor:
        if not isinstance(x, torch.Tensor) or x.device != self.device:
            raise ValueError(f"{name} must be a tensor on {self.device}")
        if x.dtype != self.dtype:
            raise ValueError(f"{name} has dtype {x.dtype}, context expects {self.dtype}")
        if tuple(x.shape) != (6, self.N, self.N):
            raise ValueError(f"{name} has shape {tuple(x.shape)}, expected (6, {self.N}, {self.N})")
        return x.contiguous()

    def empty(self, n: int):
        return [torch.empty((6, self.N, self.N), dtype=self.dtype, device=self.device) for _ in range(n)]

    # -- entry points --------------------------------------------------------
    def rk(self, core6: Sequence[torch.Tensor], nbb: int, dtb: float, rk: int = 2,
           diagnostics: bool = True, out=None, graph: bool = True, inplace: bool = False):
        """nbb substeps; returns the output tensors (h, u, v, then the 10
        diagnostics unless ``diagnostics`` is false).  With ``graph`` the
        substep pairs replay recorded hipGraphs; with ``inplace`` out[0..2]
        may be the input tensors h, u, v themselves."""
        ins = [self._check_field(x, STATE_KEYS[k]) for k, x in enumerate(core6)]
        outs = out if out is not None else self.empty(13 if diagnostics else 3)
        if len(outs) < (13 if diagnostics else 3):
            raise ValueError("out needs 13 tensors with diagnostics, 3 without")
        for k, x in enumerate(outs):
            if x.data_ptr() != self._check_field(x, f"out[{k}]").data_ptr():
                raise ValueError(f"out[{k}] must be contiguous")
        self._bind_stream()
        pin, _k1 = _hip.ptr_array([x.data_ptr() for x in ins])
        pout, _k2 = _hip.ptr_array([x.data_ptr() for x in outs] + [0] * (13 - len(outs)))
        fn = _hip.lib().sw_rk2 if rk == 2 else _hip.lib().sw_rk3
        flags = ((0 if diagnostics else _hip.SW_SKIP_DIAG) | (_hip.SW_GRAPH if graph else 0)
                 | (_hip.SW_INPLACE if inplace else 0))
        _hip.check(fn(self._h, pin, pout, int(nbb), float(dtb), flags), f"sw_rk{rk}")
        return outs

    def profile(self, core6: Sequence[torch.Tensor], nbb: int, dtb: float, reps: int = 1):
        """Eager RK2 run, every launch (repeated `reps` times) timed by
        start / stop hipEvents on its own dispatch; returns
        {kernel class: (total ms, launches)}."""
        ins = [self._check_field(x, STATE_KEYS[k]) for k, x in enumerate(core6)]
        outs = self.empty(13)
        self._bind_stream()
        pin, _k1 = _hip.ptr_array([x.data_ptr() for x in ins])
        pout, _k2 = _hip.ptr_array([x.data_ptr() for x in outs])
        ms = (ctypes.c_double * len(_hip.KERNEL_CLASSES))()
        cnt = (ctypes.c_int * len(_hip.KERNEL_CLASSES))()
        _hip.check(_hip.lib().sw_profile(self._h, pin, pout, int(nbb), float(dtb), int(reps), ms, cnt),
                   "sw_profile")
        self.last_profile_out = outs  # the pass computes the same outputs as rk()
        return {name: (ms[k], cnt[k]) for k, name in enumerate(_hip.KERNEL_CLASSES)}

    def profile_graph(self, core6: Sequence[torch.Tensor], pairs: int, replays: int, dtb: float):
        """The fused kernel as it runs in the recorded substep-pair graphs: a
        graph of `pairs` RK2 substep pairs with the kernel's launch stamps on,
        `replays` measured replays, each after 3 unmeasured ones queued back
        to back (after one eager substep); returns {"fused": (total ms,
        launches), "graph": (total ms, replays)} of the measured replays."""
        ins = [self._check_field(x, STATE_KEYS[k]) for k, x in enumerate(core6)]
        outs = self.empty(3)
        self._bind_stream()
        pin, _k1 = _hip.ptr_array([x.data_ptr() for x in ins])
        pout, _k2 = _hip.ptr_array([x.data_ptr() for x in outs] + [0] * 10)
        ms = (ctypes.c_double * len(_hip.KERNEL_CLASSES))()
        cnt = (ctypes.c_int * len(_hip.KERNEL_CLASSES))()
        gms = ctypes.c_double()
        _hip.check(_hip.lib().sw_profile_graph(self._h, pin, pout, int(pairs), int(replays), float(dtb), ms, cnt,
                                               ctypes.byref(gms)), "sw_profile_graph")
        self.last_profile_out = outs
        k = _hip.KERNEL_CLASSES.index("fused")
        return {"fused": (ms[k], cnt[k]), "graph": (gms.value, int(replays))}

    def integrals(self, h: torch.Tensor, u: torch.Tensor, v: torch.Tensor):
        """(mass, energy) in binary64 with compensated summation (see
        include/licom_sw.h); waits for the result."""
        h, u, v = (self._check_field(x, n) for x, n in ((h, "h"), (u, "u"), (v, "v")))
        self._bind_stream()
        out = (ctypes.c_double * 2)()
        _hip.check(_hip.lib().sw_integrals(self._h, ctypes.c_void_p(h.data_ptr()), ctypes.c_void_p(u.data_ptr()),
                                           ctypes.c_void_p(v.data_ptr()), out), "sw_integrals")
        return float(out[0]), float(out[1])

    def set_tuning(self, rows_a: int, rows_b: int, rows_f=None) -> None:
        """Band heights of the separate height / velocity kernels and of the
        fused kernel (rows_f = 0 selects the separate kernels; None keeps the
        current value)."""
        rows_f = self.tuning[2] if rows_f is None else int(rows_f)
        _hip.check(_hip.lib().sw_set_tuning(self._h, int(rows_a), int(rows_b), rows_f), "sw_set_tuning")
        self.tuning = (int(rows_a), int(rows_b), rows_f)

    def set_band_split(self, rows_short: int) -> None:
        """Split bands of the fused kernel: columns alternate bands of rows_f
        rows (the first blocks dispatched) and `rows_short` rows (the later
        ones); 0 restores uniform bands.  Performance only."""
        _hip.check(_hip.lib().sw_set_band_split(self._h, int(rows_short)), "sw_set_band_split")
        self.split = int(rows_short)

    def band_split(self) -> int:
        r = ctypes.c_int()
        _hip.check(_hip.lib().sw_get_band_split(self._h, ctypes.byref(r)), "sw_get_band_split")
        return r.value

    def set_frame_form(self, cube_edge: bool) -> None:
        """The launch after the fused kernel: one cube-edge launch for the
        frame cells, velocity halo and next edge pass (True, the default
        where the grid allows it), or the frame launch and a halo / edge
        launch (False).  Performance only."""
        _hip.check(_hip.lib().sw_set_frame_form(self._h, 1 if cube_edge else 0), "sw_set_frame_form")

    def frame_form(self) -> bool:
        r = ctypes.c_int()
        _hip.check(_hip.lib().sw_get_frame_form(self._h, ctypes.byref(r)), "sw_get_frame_form")
        return bool(r.value)

    def layout_flags(self) -> dict:
        f = ctypes.c_int()
        _hip.check(_hip.lib().sw_layout_flags(self._h, ctypes.byref(f)), "sw_layout_flags")
        return {"dz_uniform": bool(f.value & 1), "shared_metrics": bool(f.value & 2),
                "gct_symmetric": bool(f.value & 4)}

    def stage(self, h, hp, u, v, up, vp, dt: float, beta: float, last: bool):
        ins = [self._check_field(x, n) for x, n in zip((h, hp, u, v, up, vp), STATE_KEYS[:6])]
        outs = self.empty(13)
        self._bind_stream()
        pin, _k1 = _hip.ptr_array([x.data_ptr() for x in ins])
        pout, _k2 = _hip.ptr_array([x.data_ptr() for x in outs])
        _hip.check(_hip.lib().sw_stage(self._h, pin, pout, float(dt), float(beta), int(bool(last))),
                   "sw_stage")
        return tuple(outs)

    def halo_scalar_(self, f: torch.Tensor) -> torch.Tensor:
        """In-place ext_scalar on (6, N, N) or (6, k, N, N)."""
        if f.dtype != self.dtype or not f.is_contiguous() or f.shape[0] != 6 or f.shape[-2:] != (self.N, self.N):
            raise ValueError("field must be a contiguous (6, [k,] N, N) tensor of the context dtype")
        ncomp = 1 if f.dim() == 3 else int(f.shape[1])
        self._bind_stream()
        _hip.check(_hip.lib().sw_halo_scalar(self._h, ctypes.c_void_p(f.data_ptr()), ncomp), "sw_halo_scalar")
        return f

    def halo_vector_(self, u: torch.Tensor, v: torch.Tensor):
        u = self._check_field(u, "u")
        v = self._check_field(v, "v")
        self._bind_stream()
        _hip.check(_hip.lib().sw_halo_vector(self._h, ctypes.c_void_p(u.data_ptr()),
                                             ctypes.c_void_p(v.data_ptr())), "sw_halo_vector")
        return u, v

    def edge_avg_(self, u: torch.Tensor, v: torch.Tensor):
        u = self._check_field(u, "u")
        v = self._check_field(v, "v")
        self._bind_stream()
        _hip.check(_hip.lib().sw_edge_avg(self._h, ctypes.c_void_p(u.data_ptr()),
                                          ctypes.c_void_p(v.data_ptr())), "sw_edge_avg")
        return u, v


# ---------------------------------------------------------------------------
# context cache keyed on the buffers of the consts (the reference keys its
# compiled graph on the static args and the arrays' shardings)
# ---------------------------------------------------------------------------
_CTX_CACHE: "OrderedDict[tuple, tuple]" = OrderedDict()
_CTX_CACHE_MAX = 8
_ctx_builds = 0  # contexts built by context_for (a cache miss each)


def _buffer_key(c) -> tuple:
    """Identity of a const's storage: a new tensor or DLPack view of the same
    buffer (same address, shape, dtype, device, unmodified) gives the same
    key, so a caller that rewraps its arrays on every call still hits.  Host
    arrays carry no version counter: consts are read-only inputs (JAX arrays
    are immutable), and one modified in place needs a new buffer."""
    if isinstance(c, torch.Tensor):
        return ("t", c.data_ptr(), tuple(c.shape), tuple(c.stride()), str(c.dtype), str(c.device),
                c._version)
    a = np.asarray(c)
    return ("n", a.__array_interface__["data"][0], a.shape, a.strides, a.dtype.str)


def context_for(consts: Sequence, dtype: torch.dtype) -> SWContext:
    """The cached context of (consts, dtype), built on first use.  The cache
    holds the const objects, so their buffers cannot be freed and reused while
    an entry refers to them."""
    global _ctx_builds
    key = (str(dtype),) + tuple(_buffer_key(c) for c in consts[:N_BAROTR_CONSTS])
    hit = _CTX_CACHE.get(key)
    if hit is not None:
        _CTX_CACHE.move_to_end(key)
        return hit[0]
    ctx = SWContext(consts, dtype)
    _ctx_builds += 1
    _CTX_CACHE[key] = (ctx, tuple(consts[:N_BAROTR_CONSTS]))
    while len(_CTX_CACHE) > _CTX_CACHE_MAX:
        _CTX_CACHE.popitem(last=False)
    return ctx


def context_builds() -> int:
    """Number of contexts `context_for` has built in this process."""
    return _ctx_builds


def _check_mesh(px: int, py: int) -> None:
    if px != 1 or py != 1:
        raise ValueError(
            f"px={px}, py={py}: this library keeps all six faces on one GPU; "
            "only px = py = 1 is supported")


def _rk_core(state, consts, nbb, dtb, px, py, rk, donate=False):
    _check_mesh(px, py)
    if len(state) < N_BAROTR_STATE:
        raise ValueError(f"state must hold {N_BAROTR_STATE} fields, got {len(state)}")
    if nbb == 0:
        return tuple(state[:N_BAROTR_STATE])
    ctx = context_for(consts, state[0].dtype)
    if donate:
        # in place: h0, ub, vb and the 10 diagnostic slots of `state` receive the
        # results; h0p, ubp, vbp are set equal to them (the reference's return)
        outs = [state[0], state[2], state[3]] + list(state[6:16])
        ctx.rk(state[:6], nbb, dtb, rk=rk, out=outs, inplace=True)
        for dst, src in ((1, 0), (4, 2), (5, 3)):
            if state[dst].data_ptr() != state[src].data_ptr():
                state[dst].copy_(state[src])
        return tuple(state[:16]) + tuple(state[16:23])
    out = ctx.rk(state[:6], nbb, dtb, rk=rk)
    h, u, v = out[0], out[1], out[2]
    coupling = tuple(state[16:23])
    return (h, h, u, v, u, v, *out[3:13]) + coupling


def barotr_rk2_core(state, consts, nbb, dtb, px=1, py=1, donate=False):
    """barotr.py:41-61 — nbb RK2-midpoint substeps (stage 1 at dtb/2, stage 2 at
    dtb from the same *p fields); coupling fields [16:23] pass through.  By
    default returns new tensors; ``donate=True`` updates `state`'s tensors in
    place and returns them."""
    return _rk_core(state, consts, nbb, dtb, px, py, 2, donate)


def barotr_rk3_core(state, consts, nbb, dtb, px=1, py=1, donate=False):
    """barotr.py:64-87 — nbb RK3 substeps (dtb/3, dtb/2, dtb; beta = 1)."""
    return _rk_core(state, consts, nbb, dtb, px, py, 3, donate)


def step_rk_logic(h0, h0p, ub, vb, ubp, vbp, consts, dt, beta_d, is_laststep, px=1, py=1):
    """barotr_pure.py:100-168 — one stage; returns the reference's 13-tuple
    (new_h0, new_ub, new_vb, celerity_x, celerity_y, ub_ct, vb_ct, ub_cx,
    ub_cy, vb_cx, vb_cy, advx, advy)."""
    _check_mesh(px, py)
    ctx = context_for(consts, h0.dtype)
    return ctx.stage(h0, h0p, ub, vb, ubp, vbp, dt, beta_d, is_laststep)


def ext_scalar(var, consts, px=1, py=1):
    """cube.py:130-134 on a (6, N, N) or stacked (6, k, N, N) field; returns a
    new tensor.  (The reference passes coef/loc tables; here they come from the
    context built from `consts`.)"""
    _check_mesh(px, py)
    ctx = context_for(consts, var.dtype)
    return ctx.halo_scalar_(var.clone().contiguous())


def ext_vector(u, v, consts, px=1, py=1):
    """cube.py:137-153; returns new (u, v)."""
    _check_mesh(px, py)
    ctx = context_for(consts, u.dtype)
    return ctx.halo_vector_(u.clone(), v.clone())


def boundary_communication(u, v, consts):
    """communication.py:566-570; returns new (u, v)."""
    ctx = context_for(consts, u.dtype)
    return ctx.edge_avg_(u.clone(), v.clone())
