"""Project autotuner for the stage kernels' launch shape.

Each wavefront of the stage kernels marches down a band of R rows of a
column strip; R trades parallelism (more, shorter bands fill more SIMDs, and
a launch that just fits the wave slots avoids a second round) against the
stencil warm-up rows every band repeats.  `tune` times a short eager run
(per-dispatch hipEvents) for every candidate R of the separate height and
velocity kernels and of the fused kernel (with its frame kernel) on the
context's own resolution and precision, keeps the fastest band heights and
the faster of the two paths; for the fused path it then times the fused
kernel's heights again as it runs in graph replays (eager timing does not
reproduce them) and tries split bands around the best one (the first blocks
dispatched march taller bands).  It records the result in a JSON cache keyed
by (nx, dtype, device name); `tune_batch` picks one R for all members of a batched
run by timing the whole batch.  Results do not depend on R or on the path
(every setting produces identical bits), so tuning is free to change them.
"""

from __future__ import annotations

import json
import os
import time
from typing import Dict, Iterable, Optional, Sequence, Tuple

CANDIDATES: Tuple[int, ...] = (8, 12, 16, 20, 24, 26, 28, 30, 32, 36, 40, 48)
CACHE_ENV = "LICOM_SW_TUNE_CACHE"
_KERNEL_SET = "fused4"  # cache entries of other kernel generations are ignored
SPLIT_SHARES = {"float64": (0.62, 0.66, 0.69, 0.72), "float32": (0.55, 0.6, 0.65)}


def one_round_rows(nx: int, fp64: bool, cus: int) -> int:
    """The lowest band height whose fused-kernel waves (6 faces x ceil(nx / 52)
    strips x bands) all fit one round of resident waves (8 per CU in FP64, 12
    in FP32), when it exceeds the fixed candidates (64 rows); else 0.  At C1536
    it is 140 rows FP64 and 91 FP32, the fastest heights measured there."""
    strips = -(-nx // 52)
    bands = (8 if fp64 else 12) * cus // (6 * strips)
    if bands < 1:
        return 0
    rows = -(-nx // bands)
    return rows if rows > 64 else 0


def _cache_path() -> str:
    return os.environ.get(CACHE_ENV, os.path.join(os.path.expanduser("~"), ".cache", "licom_sw_tune.json"))


def _key(nx: int, dtype, device_name: str) -> str:
    return f"{device_name}|C{nx}|{str(dtype).replace('torch.', '')}|{_KERNEL_SET}"


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
    """Pick (rows_a, rows_b, rows_f) for `ctx` by timing; applies and
    returns it (rows_f = 0 when the separate kernels are faster)."""
    import torch

    dev_name = torch.cuda.get_device_name(ctx.device)
    key = _key(ctx.nx, ctx.dtype, dev_name)
    cache = load_cache() if use_cache else {}
    if key in cache:
        best = cache[key]
        ctx.set_tuning(best["rows_a"], best["rows_b"], best["rows_f"])
        ctx.set_band_split(best.get("rows_f2", 0))
        return dict(best, cached=True)
    ctx.set_band_split(0)  # uniform bands while the heights are swept
    cands = [c for c in candidates if c <= max(4, ctx.nx)]
    cus = torch.cuda.get_device_properties(ctx.device).multi_processor_count
    r1 = one_round_rows(ctx.nx, ctx.dtype == torch.float64, cus)
    if r1 and r1 not in cands:
        cands.append(r1)
    times = {}

    def per_launch(prof, *names):
        return sum(prof[k][0] for k in names) / max(prof[names[0]][1], 1)

    # the separate kernels independently, then the fused kernel + frame, with
    # the per-class event timer
    for r in cands:
        ctx.set_tuning(r, r, 0)
        ctx.profile(core6, 1, dt)  # warm
        best_a = best_b = best_f = float("inf")
        for _ in range(reps):
            prof = ctx.profile(core6, substeps, dt)
            best_a = min(best_a, per_launch(prof, "march_a"))
            best_b = min(best_b, per_launch(prof, "march_b"))
        ctx.set_tuning(r, r, r)
        ctx.profile(core6, 1, dt)
        for _ in range(reps):
            prof = ctx.profile(core6, substeps, dt)
            best_f = min(best_f, per_launch(prof, "fused", "frame"))
        times[r] = (best_a, best_b, best_f)
    ra = min(cands, key=lambda r: times[r][0])
    rb = min(cands, key=lambda r: times[r][1])
    rf = min(cands, key=lambda r: times[r][2])
    use_fused = times[rf][2] < times[ra][0] + times[rb][1]
    ctx.set_tuning(ra, rb, rf if use_fused else 0)
    rf_uniform, rf2, split_table = rf, 0, {}
    if use_fused:
        rf, rf2, split_table = _tune_fused_graph(ctx, core6, dt, ra, rb, cands)
    result = {"rows_a": ra, "rows_b": rb, "rows_f": rf if use_fused else 0, "rows_f2": rf2,
              "ms_a": times[ra][0], "ms_b": times[rb][1], "ms_fused_frame": times[rf_uniform][2],
              "table": {str(r): {"ms_a": t[0], "ms_b": t[1], "ms_fused_frame": t[2]} for r, t in times.items()},
              "split_table_us": split_table,
              "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    if use_cache:
        cache[key] = result
        save_cache(cache)
    return dict(result, cached=False)


def _tune_fused_graph(ctx, core6, dt, ra, rb, cands, reps: int = 2):
    """The fused kernel's bands timed as it runs in graph replays (which the
    eager per-dispatch timer does not reproduce): every uniform height, then
    split bands around the best one, pairs of about twice its rows split at a
    few shares for the first blocks dispatched (the split changes only the
    fused kernel).  Applies and returns (long height, short height or 0,
    table of microseconds per launch)."""
    def fused_us(r, r2, n=reps):
        ctx.set_tuning(ra, rb, r)
        ctx.set_band_split(r2)
        best = float("inf")
        for _ in range(n):
            p = ctx.profile_graph(core6, 8, 1, dt)
            best = min(best, p["fused"][0] / p["fused"][1] * 1e3)
        return best

    table = {f"{r}/0": fused_us(r, 0) for r in cands}
    best = min((t, int(k.split("/")[0]), 0) for k, t in table.items())
    # the two best uniform heights: near the optimum the uniform timings are
    # flat (C768 FP32: 24 and 36 rows within 1 %), so the best one alone is
    # partly chance, and the spans searched are around twice either
    top2 = [int(k.split("/")[0]) for k, _ in sorted(table.items(), key=lambda kv: kv[1])[:2]]
    shares = SPLIT_SHARES[str(ctx.dtype).replace("torch.", "")]
    # spans in steps of 2 rows: which span gives as many short-band blocks as
    # long-band ones (one of each per CU) depends on the resolution
    spans = {2 * rf + d for rf in top2 for d in (-4, -2, 0, 2, 4)}
    for span in sorted(spans):
        for share in shares:
            lng = int(round(span * share))
            sh = span - lng
            if lng >= ctx.nx or sh < 4 or f"{lng}/{sh}" in table:
                continue
            t = fused_us(lng, sh)
            table[f"{lng}/{sh}"] = t
            best = min(best, (t, lng, sh))
    # the three fastest timed again with more replays (single timings of
    # neighbouring settings differ by about 1 %)
    final = []
    for k, _ in sorted(table.items(), key=lambda kv: kv[1])[:3]:
        lng, sh = (int(x) for x in k.split("/"))
        final.append((fused_us(lng, sh, 2 * reps), lng, sh))
    best = min(final)
    ctx.set_tuning(ra, rb, best[1])
    ctx.set_band_split(best[2])
    return best[1], best[2], {k: round(v, 2) for k, v in table.items()}


def tune_batch(batch, members, candidates: Iterable[int] = CANDIDATES, substeps: int = 4,
               reps: int = 2) -> dict:
    """Pick one band height, and the fused or the separate stage kernels, for
    every member of a batched run by timing the whole batch (short bands help
    a resolution alone, but in a batch the launches are full anyway and short
    bands only add warm-up rows).  Applies and returns it; the members' own
    substep counts and output tensors are restored."""
    import torch

    saved = [(m.nbb, m.out) for m in members]
    times = {}
    try:
        for m in members:
            m.nbb = substeps
            m.out = None  # the timed runs write their own scratch outputs
            m.ctx.set_band_split(0)  # batched launches use uniform bands
        for r in candidates:
            for fused in (False, True):
                for m in members:
                    m.ctx.set_tuning(r, r, r if fused else 0)
                batch.run(members, diagnostics=False)
                torch.cuda.synchronize()
                best = float("inf")
                for _ in range(reps):
                    t0 = time.perf_counter()
                    batch.run(members, diagnostics=False)
                    torch.cuda.synchronize()
                    best = min(best, time.perf_counter() - t0)
                times[(r, fused)] = best
    finally:
        for m, (n, out) in zip(members, saved):
            m.nbb, m.out = n, out
    r, fused = min(times, key=times.get)
    for m in members:
        m.ctx.set_tuning(r, r, r if fused else 0)
    return {"rows": r, "fused": fused,
            "table_ms": {f"{k}{'f' if f else 's'}": v * 1e3 for (k, f), v in times.items()}}

