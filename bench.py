#!/usr/bin/env python3
"""Benchmark of the barotropic RK2 step on one MI355X.

A *step* is one drop-in call `barotr_rk2_core(state, consts, nbb, dtb)` on the
C768 FP64 duo-grid (6 x 768^2 = 3,538,944 interior cells) with nbb = 2000
substeps, the substep pairs replaying recorded hipGraphs.  Inputs are resident
in HBM before the timed region.  The reported value is interior cells x
substeps / wall time (cell-substeps/s).

The JSON line also carries
  roofline       the dominant kernel (the fused stage kernel) against the
                 path's algorithmic bytes per stage launch, 148 B per interior
                 cell in FP64 (296 B per cell-substep over the two stages of
                 a substep), over its average launch time as it runs in the
                 timed steps: the substep-pair graph the timed steps replay,
                 recorded with the kernel's wave stamps on (lane 0 of every
                 wave writes the device wall clock at entry and exit; a launch
                 lasts from its first entry to its last exit) and replayed
                 after the timed region; plus the kernel's own byte
                 count (face-shared planes counted once for six faces) and
                 its L2-to-fabric bytes per launch from the committed
                 rocprofv3 counter summary,
  roofline_path  the whole step against 296 B per FP64 cell-substep (148 B for
                 FP32) at 8 TB/s,
  fp32           the same workload in the FP32 form: timed steps, its path
                 roofline fraction, its speed-up over FP64, and its height
                 error against FP64 after a fixed number of substeps,
  cpu_baseline   the NumPy FP64 oracle on one host core over a bounded sample
                 of the same workload,
  serving        p50/p99 latency and throughput of served solves (a seeded
                 Poisson trace of mixed resolutions, step counts and
                 precisions under continuous batching, least remaining work
                 first) beside the same trace with every admitted request
                 advancing each tick, and the plain FP64 path serving it one
                 request at a time.
Only one GPU is used: --gpus above 1 exits with a message.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "cell-substeps/s + % HBM roofline at C768; p50/p99 latency of served solves"

# Algorithmic bytes per interior cell per launch of the stage kernels: each
# array the kernel must touch, read or written once, averaged over the two RK2
# stages (stage 1 reads h0p/ubp/vbp at the same addresses as h0/ub/vb).  The
# five face-independent metric planes (rdx, rdy, rda, c_dy, d_dx) count 1/6 per
# cell; a_gct counts two planes (symmetric, equal diagonal).
#   height kernel : h u v [hp] g00 g01 sina + rda c_dy d_dx/6 + write h  = 7.5 / 8.5
#   velocity kernel: h hn u v [up vp] g00 g01 sina a_f + 5/6 + write u v = 10.83 / 12.83
#   fused kernel  : h u v [hp up vp] g00 g01 sina a_f + 5/6 + write h u v = 10.83 / 13.83
# (the fused kernel writes u, v on the deep cells and the frame kernel on the
# 3-cell frame along the face edges; the figure counts every interior cell)
VALUES_PER_CELL = {"march_a": 8.0, "march_b": 11.8333, "fused": 12.3333}
PATH_BYTES = {"f64": 296.0, "f32": 148.0}


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--nx", type=int, default=768)
    p.add_argument("--nbb", type=int, default=2000, help="substeps per step")
    p.add_argument("--dtype", choices=("f64", "f32"), default="f64")
    p.add_argument("--no-graph", action="store_true", help="launch every substep eagerly")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep the default band height instead of timing the candidates")
    p.add_argument("--profile-replays", type=int, default=4,
                   help="measured replays of the wave-stamped substep-pair graph (0 disables)")
    p.add_argument("--fp32-steps", type=int, default=None,
                   help="timed steps of the FP32 leg (default --steps; 0 disables)")
    p.add_argument("--accuracy-substeps", type=int, default=10000,
                   help="substeps after which the FP32 height is compared with FP64")
    p.add_argument("--cpu-substeps", type=int, default=8,
                   help="substeps of the CPU oracle sample (0 disables)")
    p.add_argument("--cpu-nx", type=int, default=None, help="resolution of the CPU sample (default --nx)")
    p.add_argument("--serve-requests", type=int, default=200,
                   help="requests of the served trace (Poisson 200/s, C48-C768, 50-5000 steps, seed 7: the "
                        "200-request served workload; 0 disables)")
    return p.parse_args()


def cpu_baseline(nx: int, substeps: int):
    """Time the NumPy FP64 oracle on one core for `substeps` RK2 substeps."""
    import numpy as np  # noqa: F401
    from licom4py_gpu_amd import duogrid, initial
    from oracle import barotr_np

    g = duogrid.grid(nx)
    h, u, v, sp = initial.seeded_balanced(g)
    dt = initial.courant_dt(g, h, sp)
    consts = duogrid.barotr_consts(g, 5000.0)
    st = initial.state_tuple(h, u, v)
    t0 = time.perf_counter()
    barotr_np.barotr_rk2_core(st, consts, substeps, dt)
    el = time.perf_counter() - t0
    cells = 6 * nx * nx
    return {
        "value": cells * substeps / el,
        "unit": "cell-substeps/s",
        "cores": 1,
        "kind": "port",
        "sample": f"C{nx} FP64 RK2, {substeps} substeps of the same seeded field, NumPy oracle "
                  f"(single-threaded) on the GPU host: {el:.2f} s",
    }


def note(text: str) -> None:
    """A progress line on stderr (stdout carries only the result line)."""
    print(f"bench.py: {text}", file=sys.stderr, flush=True)


def load_traffic(kernel: str, workload: str):
    """HBM bytes per launch measured with rocprofv3 --pmc (committed summary)."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def main() -> int:
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py: the integrator runs on one GPU; --gpus {args.gpus} (WORLD_SIZE {world}) "
              "is not supported", file=sys.stderr)
        return 2

    import torch
    from licom4py_gpu_amd import barotr, duogrid, initial

    if not torch.cuda.is_available():
        print("bench.py: no GPU visible", file=sys.stderr)
        return 1
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    nx, nbb = args.nx, args.nbb
    tdt = torch.float64 if args.dtype == "f64" else torch.float32

    t_setup = time.perf_counter()
    g = duogrid.grid(nx)
    h, u, v, speed = initial.seeded_balanced(g)
    dt = initial.courant_dt(g, h, speed, depth=5000.0)
    consts = barotr.as_device_consts(duogrid.barotr_consts(g, 5000.0), dev)
    state = tuple(torch.as_tensor(x).to(device=dev, dtype=tdt) for x in initial.state_tuple(h, u, v))
    ctx = barotr.context_for(consts, tdt)
    tuning = None
    if not args.no_autotune:
        from licom4py_gpu_amd import autotune
        tuning = autotune.tune(ctx, state[:6], dt, use_cache=False)
        tuning = {k: tuning[k] for k in ("rows_a", "rows_b", "rows_f", "rows_f2", "split_table_us", "table")}
    else:
        tuning = {"rows_a": ctx.tuning[0], "rows_b": ctx.tuning[1], "rows_f": ctx.tuning[2], "rows_f2": ctx.split}
    setup_s = time.perf_counter() - t_setup
    note(f"set up and tuned in {setup_s:.1f} s: {tuning['rows_f']}/{tuning['rows_f2']} rows")

    # served solves first (they run both precisions at many resolutions)
    serving = None
    if args.serve_requests > 0:
        from licom4py_gpu_amd import serving as srv
        rep = srv.serve_report(args.serve_requests, 200.0, [48, 96, 192, 384, 768], (50, 5000), 200.0, 400,
                               7, dev, baselines=("fp64",), policy="srpt", width=2e7)
        # the same trace with every admitted request advancing 100 substeps per tick
        ps = srv.serve_report(args.serve_requests, 200.0, [48, 96, 192, 384, 768], (50, 5000), 200.0, 100,
                              7, dev, baselines=(), policy="all")
        base = rep["one_at_a_time_fp64"]
        serving = {"workload": rep["workload"],
                   "policy": "least remaining work first, ticks of 2e7 cells and up to 400 substeps",
                   "p50_s": round(rep["latency"]["p50_s"], 4), "p99_s": round(rep["latency"]["p99_s"], 4),
                   "throughput_cell_substeps_per_s": rep["throughput_cell_substeps_per_s"],
                   "mean_batch": round(rep["mean_batch"], 2), "peak_hbm_gb": round(rep["peak_hbm_gb"], 3),
                   "every_request_each_tick": {"p50_s": round(ps["latency"]["p50_s"], 4),
                                               "p99_s": round(ps["latency"]["p99_s"], 4),
                                               "throughput_cell_substeps_per_s": ps["throughput_cell_substeps_per_s"]},
                   "plain_fp64_one_at_a_time": {"p50_s": round(base["p50_s"], 4), "p99_s": round(base["p99_s"], 4),
                                                "throughput_cell_substeps_per_s":
                                                    base["throughput_cell_substeps_per_s"]}}
        if not (rep["finite"] and ps["finite"]):
            print("bench.py: non-finite served result", file=sys.stderr)
            return 1
        note(f"served {args.serve_requests} requests: p99 {serving['p99_s']} s")

    # the FP32 form next (its accuracy reference runs FP64 too), so the FP64
    # timed steps and their wave-stamped graph replays are the last FP64
    # stage-kernel launches of the run
    fp32 = None
    fp32_steps = args.steps if args.fp32_steps is None else args.fp32_steps
    if args.dtype == "f64" and fp32_steps > 0 and not args.no_graph:
        fp32 = fp32_leg(consts, state, dt, nbb, fp32_steps, args.warmup, args.accuracy_substeps,
                        args.profile_replays, not args.no_autotune, 6 * nx * nx, f"C{nx}_f32")
        note(f"FP32 leg: {fp32['value']:.4g} cell-substeps/s")

    def step(s):
        return barotr.barotr_rk2_core(s, consts, nbb, dt) if not args.no_graph else \
            _eager(ctx, s, nbb, dt)

    s = state
    for _ in range(args.warmup):
        s = step(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s = step(s)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    hfin = s[0][:, 3:-3, 3:-3]
    if not bool(torch.isfinite(hfin).all()):
        print("bench.py: non-finite height after the timed steps", file=sys.stderr)
        return 1

    cells = 6 * nx * nx
    value = cells * nbb * args.steps / el
    note(f"timed {args.steps} steps: {value:.4g} cell-substeps/s")
    ms_per_step = el / args.steps * 1e3
    path_bytes = PATH_BYTES[args.dtype]
    workload = f"C{nx}_{args.dtype}"

    roofline, kernels = graph_roofline(ctx, state, dt, args.profile_replays, args.dtype, cells, workload) \
        if args.profile_replays > 0 else (None, None)
    path_ach = value * path_bytes / 1e9
    roofline_path = {"bound": "hbm", "achieved": round(path_ach, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(path_ach / PEAK_HBM_GBS, 4),
                     "bytes_per_cell_substep": path_bytes}

    if fp32 is not None:
        fp32["ratio_to_fp64"] = fp32["value"] / value

    cpu = None
    if args.cpu_substeps > 0:
        note("CPU baseline sample")
        cpu = cpu_baseline(args.cpu_nx or nx, args.cpu_substeps)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-substeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic: seeded balanced spherical-harmonic height (degree 1-8, max 200 m) with "
                "geostrophic velocity on a generated equal-edge duo-grid, depth 5000 m, Courant 0.5",
        "config": {"workload": f"C{nx} {args.dtype.upper()} RK2 barotropic step, {nbb} substeps per step"
                               + ("" if args.no_graph else ", hipGraph substep pairs"),
                   "nx": nx, "interior_cells": cells, "substeps_per_step": nbb,
                   "dt_s": dt, "setup_s": round(setup_s, 2), "tuning": tuning},
        "roofline": roofline,
        "roofline_path": roofline_path,
        "kernels": kernels,
        "fp32": fp32,
        "cpu_baseline": cpu,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)
    return 0


KERNEL_TIMING = ("device wall clock at every wave's entry and exit (first entry to last exit per launch) in the "
                 "recorded substep-pair graph the timed steps replay (16 pairs), summed over replays after the "
                 "timed region")


def graph_roofline(ctx, state, dt, replays, dtype, cells, workload):
    """The fused stage kernel against the path's bytes per stage launch, timed
    as it runs in the timed steps' graph replays."""
    if ctx.tuning[2] <= 0:
        return None, {"timing": "the tuned path runs the separate stage kernels; the graph timer covers the fused one"}
    prof = ctx.profile_graph(state[:6], 16, replays, dt)
    (fms, n), (gms, nrep) = prof["fused"], prof["graph"]
    substeps = 32 * nrep
    kernels = {"fused": {"ms_total": round(fms, 4), "launches": n, "us_avg": round(fms / n * 1e3, 3)},
               "graph": {"ms_total": round(gms, 4), "replays": nrep, "us_per_substep": round(gms / substeps * 1e3, 3),
                         "fused_share": round(fms / gms, 4)},
               "timing": KERNEL_TIMING}
    esize = 8 if dtype == "f64" else 4
    avg_s = fms / n * 1e-3
    per_cell = PATH_BYTES[dtype] / 2.0  # the path's bytes per cell and stage launch
    alg = per_cell * cells
    achieved = alg / avg_s / 1e9
    traffic = load_traffic("fused", workload)
    own = VALUES_PER_CELL["fused"] * esize * cells
    roofline = {"bound": "hbm", "kernel": "fused", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
                "traffic_kind": "L2-to-fabric bytes per launch (FETCH_SIZE + WRITE_SIZE, scaled by a calibration "
                                "copy); the 256 MB MALL serves part of them, so they bound HBM bytes from above",
                "algorithmic_bytes_per_launch": int(alg), "bytes_per_cell_per_launch": per_cell,
                "avg_launch_us": round(avg_s * 1e6, 3), "timing": KERNEL_TIMING,
                # the kernel's own count: face-shared planes once for the six faces, a_gct as two planes
                "kernel_bytes_per_launch": int(own), "kernel_frac": round(own / avg_s / 1e9 / PEAK_HBM_GBS, 4),
                "traffic_gbs": round(traffic / avg_s / 1e9, 1) if traffic else None}
    return roofline, kernels


def fp32_leg(consts, state, dt, nbb, steps, warmup, acc_substeps, replays, autotune_on, cells, workload):
    """The same workload in the FP32 form: timed steps, the graph-timed
    roofline of its stage kernel, and its height against FP64."""
    import torch
    from licom4py_gpu_amd import barotr

    st32 = tuple(x.to(torch.float32) for x in state)
    ctx32 = barotr.context_for(consts, torch.float32)
    if autotune_on:
        from licom4py_gpu_amd import autotune
        autotune.tune(ctx32, st32[:6], dt, use_cache=False)
    s = st32
    for _ in range(warmup):
        s = barotr.barotr_rk2_core(s, consts, nbb, dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s = barotr.barotr_rk2_core(s, consts, nbb, dt)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    v32 = cells * nbb * steps / el
    ach = v32 * PATH_BYTES["f32"] / 1e9
    out = {"value": v32, "unit": "cell-substeps/s", "steps": steps, "warmup": warmup, "ms_per_step": el / steps * 1e3,
           "ratio_to_fp64": None,  # filled in once the FP64 steps are timed
           "roofline_path": {"achieved": round(ach, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": round(ach / PEAK_HBM_GBS, 4), "bytes_per_cell_substep": PATH_BYTES["f32"]},
           "tuning": dict(zip(("rows_a", "rows_b", "rows_f", "rows_f2"), (*ctx32.tuning, ctx32.split)))}
    if replays > 0:
        out["roofline"], _ = graph_roofline(ctx32, st32, dt, replays, "f32", cells, workload)
    # accuracy: FP32 against FP64 from the same initial state (binary64
    # compensated mass integrals for the drift)
    ctx64 = barotr.context_for(consts, torch.float64)
    h64 = barotr.barotr_rk2_core(state, consts, acc_substeps, dt)
    h32 = barotr.barotr_rk2_core(st32, consts, acc_substeps, dt)
    inner = (slice(None), slice(3, -3), slice(3, -3))
    rel = [float(torch.linalg.vector_norm(h32[k][inner].double() - h64[k][inner]) /
                 torch.linalg.vector_norm(h64[k][inner])) for k in (0, 2, 3)]
    m0 = ctx32.integrals(st32[0], st32[2], st32[3])[0]
    m1 = ctx32.integrals(h32[0], h32[2], h32[3])[0]
    m64 = ctx64.integrals(h64[0], h64[2], h64[3])[0]
    m64_0 = ctx64.integrals(state[0], state[2], state[3])[0]
    out.update({"substeps_compared": acc_substeps, "height_rel_l2_vs_fp64": rel[0],
                "u_rel_l2_vs_fp64": rel[1], "v_rel_l2_vs_fp64": rel[2],
                "mass_drift_rel_fp32": (m1 - m0) / m0, "mass_drift_rel_fp64": (m64 - m64_0) / m64_0})
    return out


def _eager(ctx, s, nbb, dt):
    out = ctx.rk(s[:6], nbb, dt, graph=False)
    h, u, v = out[0], out[1], out[2]
    return (h, h, u, v, u, v, *out[3:13]) + tuple(s[16:23])


if __name__ == "__main__":
    sys.exit(main())
