"""Project autotuner for the stage kernels' launch shape.

Each wavefront of the two stage kernels marches down a band of R rows of a
63-column strip; R trades parallelism (more, shorter bands fill more SIMDs,
and a launch that just fits the wave slots avoids a second round) against the
6-row stencil warm-up every band repeats.  `tune` times a short eager run
(per-dispatch hipEvents) for every candidate R of each kernel on the
context's own resolution and precision, keeps the fastest, and records it in
a JSON cache keyed by (nx, dtype, device name); `tune_batch` picks one R for
all members of a batched run by timing the whole batch.  Results do not
depend on R (the kernels produce identical bits for any band height), so
tuning is free to change it.
"""

from __future__ import annotations

import json
import os
import time
from typing import Dict, Iterable, Optional, Sequence, Tuple

CANDIDATES: Tuple[int, ...] = (8, 12, 16, 20, 24, 26, 28, 30, 32, 36, 40, 48)
CACHE_ENV = "LICOM_SW_TUNE_CACHE"


def _cache_path() -> str:
    return os.environ.get(CACHE_ENV, os.path.join(os.path.expanduser("~"), ".cache", "licom_sw_tune.json"))


def _key(nx: int, dtype, device_name: str) -> str:
    return f"{device_name}|C{nx}|{str(dtype).replace('torch.', '')}"


def load_cache() -> Dict[str, dict]:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache: Dict[str, dict]) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path + ".tmp", "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(path + ".tmp", path)
    except OSError:
        pass


def tune(ctx, core6: Sequence, dt: float, substeps: int = 2, reps: int = 2,
         candidates: Iterable[int] = CANDIDATES, use_cache: bool = True) -> dict:
    """Pick (rows_a, rows_b) for `ctx` by timing; applies and returns it."""
    import torch

    dev_name = torch.cuda.get_device_name(ctx.device)
    key = _key(ctx.nx, ctx.dtype, dev_name)
    cache = load_cache() if use_cache else {}
    if key in cache:
        best = cache[key]
        ctx.set_tuning(best["rows_a"], best["rows_b"])
        return dict(best, cached=True)
    cands = [c for c in candidates if c <= max(4, ctx.nx)]
    times = {}
    # tune the two kernels independently with the per-class event timer
    for r in cands:
        ctx.set_tuning(r, r)
        ctx.profile(core6, 1, dt)  # warm
        best_a = best_b = float("inf")
        for _ in range(reps):
            prof = ctx.profile(core6, substeps, dt)
            best_a = min(best_a, prof["march_a"][0] / max(prof["march_a"][1], 1))
            best_b = min(best_b, prof["march_b"][0] / max(prof["march_b"][1], 1))
        times[r] = (best_a, best_b)
    ra = min(cands, key=lambda r: times[r][0])
    rb = min(cands, key=lambda r: times[r][1])
    ctx.set_tuning(ra, rb)
    result = {"rows_a": ra, "rows_b": rb,
              "ms_a": times[ra][0], "ms_b": times[rb][1],
              "table": {str(r): {"ms_a": t[0], "ms_b": t[1]} for r, t in times.items()},
              "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    if use_cache:
        cache[key] = result
        save_cache(cache)
    return dict(result, cached=False)


def tune_batch(batch, members, candidates: Iterable[int] = CANDIDATES, substeps: int = 4,
               reps: int = 2) -> dict:
    """Pick one band height for every member of a batched run by timing the
    whole batch (short bands help a resolution alone, but in a batch the
    launches are full anyway and short bands only add warm-up rows).
    Applies and returns it; the members' own substep counts are restored."""
    import torch

    saved = [m.nbb for m in members]
    times = {}
    try:
        for m in members:
            m.nbb = substeps
        for r in candidates:
            for m in members:
                m.ctx.set_tuning(r, r)
            batch.run(members, diagnostics=False)
            torch.cuda.synchronize()
            best = float("inf")
            for _ in range(reps):
                t0 = time.perf_counter()
                batch.run(members, diagnostics=False)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
            times[r] = best
    finally:
        for m, n in zip(members, saved):
            m.nbb = n
    r = min(times, key=times.get)
    for m in members:
        m.ctx.set_tuning(r, r)
    return {"rows": r, "table_ms": {str(k): v * 1e3 for k, v in times.items()}}

