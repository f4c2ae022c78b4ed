"""Batched stepping of many fields: ensembles of mixed resolutions.

Every member is one field of its own resolution with its own context (a
workspace of stage scratch and ping-pong buffers); members of one resolution
share the device copy of the consts (``SWContext.like``).  ``Batch.run``
advances all members with one launch per kernel class and stage
(``sw_batch_run``); each member's result is bit-identical to running it alone.

    python -m licom4py_gpu_amd.ensemble --members 64 --res 48 96 192 384 \\
        --substeps 200 --dtype f64

runs the seeded ensemble (resolutions drawn uniformly, 1 % noise on the
harmonic coefficients of the initial height), times the batched run against
the same members stepped one after another, checks that both give identical
bits, and prints one JSON line.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import sys
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _hip, barotr, duogrid, initial


@dataclass
class Member:
    """One field of a batched run: its context, the six input fields
    (h0, h0p, ub, vb, ubp, vbp), substep count and length."""
    ctx: "barotr.SWContext"
    core6: Sequence[torch.Tensor]
    nbb: int
    dtb: float
    out: Optional[List[torch.Tensor]] = None
    tag: Dict = field(default_factory=dict)


class Batch:
    """Owns an ``sw_batch`` (device tables reused across runs)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _hip.check(_hip.lib().sw_batch_create(ctypes.byref(h)), "sw_batch_create")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        lib = getattr(_hip, "_lib", None) if _hip is not None else None
        if h is not None and h.value and lib is not None:
            lib.sw_batch_destroy(h)
            self._h = None

    def run(self, members: Sequence[Member], rk: int = 2, diagnostics: bool = True, graph: bool = False):
        """Advance every member by its own nbb substeps; fills and returns each
        member's outputs (h0, ub, vb, then the 10 diagnostics unless
        ``diagnostics`` is false, when 3 output tensors suffice).  With
        ``graph`` the substeps after the first replay recorded hipGraphs of
        substep pairs, one per number of members still stepping."""
        if not members:
            return []
        arr = (_hip.SWMember * len(members))()
        keep = []
        dev = members[0].ctx.device
        for k, m in enumerate(members):
            ins = [m.ctx._check_field(x, barotr.STATE_KEYS[i]) for i, x in enumerate(m.core6)]
            if m.out is None:
                m.out = m.ctx.empty(13 if diagnostics else 3)
            if diagnostics and len(m.out) < 13:
                raise ValueError("diagnostics need 13 output tensors per member")
            keep.append(ins)
            arr[k].ctx = m.ctx._h.value
            for i in range(6):
                arr[k].inp[i] = ins[i].data_ptr()
            for i in range(13):  # diagnostic slots stay null when not written
                arr[k].out[i] = m.out[i].data_ptr() if i < len(m.out) else None
            arr[k].nbb = int(m.nbb)
            arr[k].dtb = float(m.dtb)
        stream = torch.cuda.current_stream(dev).cuda_stream
        flags = (0 if diagnostics else _hip.SW_SKIP_DIAG) | (_hip.SW_GRAPH if graph else 0)
        _hip.check(_hip.lib().sw_batch_run(self._h, arr, len(members), int(rk), flags, ctypes.c_void_p(stream)),
                   "sw_batch_run")
        return [m.out for m in members]


class GridPool:
    """Per (resolution, dtype): the grid, its device consts and a base context
    whose const planes later members share.  With ``tune`` the base context's
    band heights are picked by the autotuner (on the seeded field of that
    resolution, cached per device) and inherited by every member context;
    else ``rows`` = (rows_a, rows_b, rows_f), if given, replaces the library's
    per-resolution defaults (which suit one context alone on the device)."""

    def __init__(self, device=None, tune: bool = False, rows=None):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.tune = tune
        self.rows = tuple(rows) if rows is not None else None
        self.tuning: Dict[tuple, dict] = {}
        self._consts: Dict[int, tuple] = {}
        self._base: Dict[tuple, "barotr.SWContext"] = {}

    def sibling(self, rows=None) -> "GridPool":
        """A pool on the same device sharing this pool's grids and device
        consts, with its own contexts and band heights."""
        other = GridPool(self.device, rows=rows)
        other._consts = self._consts
        return other

    def consts(self, nx: int):
        if nx not in self._consts:
            g = duogrid.grid(nx)
            self._consts[nx] = (g, barotr.as_device_consts(duogrid.barotr_consts(g, 5000.0), self.device))
        return self._consts[nx]

    def context(self, nx: int, dtype: torch.dtype) -> "barotr.SWContext":
        key = (nx, dtype)
        if key not in self._base:
            ctx = barotr.SWContext(self.consts(nx)[1], dtype)
            if self.tune:
                from . import autotune
                g = self.consts(nx)[0]
                h, u, v, sp = initial.seeded_balanced(g)
                core = [torch.as_tensor(x).to(device=self.device, dtype=dtype)
                        for x in initial.state_tuple(h, u, v)[:6]]
                r = autotune.tune(ctx, core, initial.courant_dt(g, h, sp))
                self.tuning[key] = {"rows_a": r["rows_a"], "rows_b": r["rows_b"], "rows_f": r["rows_f"],
                                    "rows_f2": r.get("rows_f2", 0)}
            elif self.rows is not None:
                ctx.set_tuning(*self.rows)
                ctx.set_band_split(0)
            self._base[key] = ctx
            return ctx
        return self._base[key].like()


def seeded_members(pool: GridPool, resolutions: Sequence[int], dtype: torch.dtype, nbb: int,
                   seed: int = 20261203, perturb: float = 0.01) -> List[Member]:
    """Members of the seeded balanced recipe, one per entry of `resolutions`,
    each with 1 % noise on the harmonic coefficients and its own Courant dt."""
    out = []
    for k, nx in enumerate(resolutions):
        g, _ = pool.consts(nx)
        h, u, v, sp = initial.seeded_balanced(g, seed=seed, perturb=perturb, member_seed=k)
        dt = initial.courant_dt(g, h, sp)
        core = [torch.as_tensor(x).to(device=pool.device, dtype=dtype) for x in initial.state_tuple(h, u, v)[:6]]
        out.append(Member(pool.context(nx, dtype), core, nbb, dt, tag={"nx": nx, "member": k}))
    return out


def _cells(members) -> int:
    return sum(6 * m.ctx.nx * m.ctx.nx * m.nbb for m in members)


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--members", type=int, default=64)
    p.add_argument("--res", type=int, nargs="+", default=[48, 96, 192, 384])
    p.add_argument("--substeps", type=int, default=200)
    p.add_argument("--dtype", choices=("f64", "f32"), default="f64")
    p.add_argument("--seed", type=int, default=20261203)
    p.add_argument("--reps", type=int, default=2)
    p.add_argument("--no-autotune", action="store_true")
    a = p.parse_args(argv)
    tdt = torch.float64 if a.dtype == "f64" else torch.float32
    rng = np.random.default_rng(a.seed)
    res = [int(x) for x in rng.choice(a.res, size=a.members)]
    # each resolution tuned alone for the one-after-another comparison ...
    pool = GridPool(tune=not a.no_autotune)
    members = seeded_members(pool, res, tdt, a.substeps, seed=a.seed)
    for m in members:  # like() contexts inherit the tuned base; apply per member explicitly
        t = pool.tuning.get((m.ctx.nx, tdt))
        if t:
            m.ctx.set_tuning(t["rows_a"], t["rows_b"], t["rows_f"])
            m.ctx.set_band_split(t.get("rows_f2", 0))
    for m in members:
        m.ctx.rk(m.core6, 2, m.dtb, diagnostics=False, graph=True)  # record graphs
    torch.cuda.synchronize()
    ts = []
    for _ in range(a.reps):
        t0 = time.perf_counter()
        alone = [m.ctx.rk(m.core6, m.nbb, m.dtb, diagnostics=False, graph=True) for m in members]
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    alone = [[x.clone() for x in o[:3]] for o in alone]
    # ... and the batch tuned as a whole
    batch = Batch()
    batch_tuning = None
    if not a.no_autotune:
        from . import autotune
        batch_tuning = autotune.tune_batch(batch, members)
    batch.run(members, diagnostics=False, graph=True)  # warm-up, records the graph
    torch.cuda.synchronize()
    tb = []
    for _ in range(a.reps):
        t0 = time.perf_counter()
        batch.run(members, diagnostics=False, graph=True)
        torch.cuda.synchronize()
        tb.append(time.perf_counter() - t0)
    got = [[x.clone() for x in m.out[:3]] for m in members]
    same = all(torch.equal(g, s) for gm, am in zip(got, alone) for g, s in zip(gm, am[:3]))
    cells = _cells(members)
    line = {
        "workload": f"ensemble of {a.members} members, C in {sorted(set(res))}, {a.substeps} RK2 substeps, {a.dtype}",
        "members_per_resolution": {str(r): res.count(r) for r in sorted(set(res))},
        "interior_cells": sum(6 * r * r for r in res),
        "batched_s": min(tb), "sequential_s": min(ts),
        "batched_cell_substeps_per_s": cells / min(tb),
        "sequential_cell_substeps_per_s": cells / min(ts),
        "speedup": min(ts) / min(tb),
        "bit_identical_to_alone": bool(same),
        "tuning_alone": {f"C{k[0]}": v for k, v in pool.tuning.items()},
        "tuning_batch": batch_tuning,
    }
    print(json.dumps(line), flush=True)
    return 0 if same else 1


if __name__ == "__main__":
    sys.exit(main())
