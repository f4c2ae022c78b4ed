"""Workloads for rocprofv3 and summaries of its result databases.

    python -m licom4py_gpu_amd.profiling run --nx 768 --dtype f64 --substeps 10
        sets up the benchmark field, runs a plain device copy of a known size
        (the calibration reference for the FETCH_SIZE / WRITE_SIZE counters),
        then `substeps` eager RK2 substeps.  Meant to run under
        `rocprofv3 --kernel-trace --stats` or `rocprofv3 --pmc <counter>`.

    python -m licom4py_gpu_amd.profiling summarize DB [DB ...] --out FILE.md
        writes a short text summary: per-kernel launch counts and average
        durations, and the average value of every collected counter per
        dispatch, for each database given.

    python -m licom4py_gpu_amd.profiling traffic --fetch DB --write DB
        --kernel NAME --copy-mib M --out FILE.json
        HBM bytes per launch of a kernel: FETCH_SIZE and WRITE_SIZE (KiB per
        dispatch) each scaled so that the copy kernel reads and writes exactly
        M MiB.
"""

from __future__ import annotations

import argparse
import json
import os
import sqlite3
import sys
import time


def _run(args) -> int:
    import torch

    from . import barotr, duogrid, initial

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    # calibration copy: exactly copy_mib MiB read and written
    n = args.copy_mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    del src, dst
    g = duogrid.grid(args.nx)
    h, u, v, sp = initial.seeded_balanced(g)
    dt = initial.courant_dt(g, h, sp)
    tdt = torch.float64 if args.dtype == "f64" else torch.float32
    consts = barotr.as_device_consts(duogrid.barotr_consts(g, 5000.0), dev)
    state = [torch.as_tensor(x).to(device=dev, dtype=tdt) for x in initial.state_tuple(h, u, v)]
    ctx = barotr.context_for(consts, tdt)
    if args.rows:
        ctx.set_tuning(*args.rows)
    if args.split is not None:
        ctx.set_band_split(args.split)
    if args.frame_form is not None:
        ctx.set_frame_form(bool(args.frame_form))
    t0 = time.perf_counter()
    out = ctx.rk(state[:6], args.substeps, dt, graph=args.graph)
    torch.cuda.synchronize()
    print(f"profiling run: C{args.nx} {args.dtype} {args.substeps} substeps (bands {ctx.tuning[2]}/{ctx.split}, "
          f"graph {args.graph}) in "
          f"{time.perf_counter() - t0:.3f} s; copy {args.copy_mib} MiB", flush=True)
    del out
    return 0


def _short(name: str) -> str:
    name = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return name.split("(")[0].strip()


def kernel_stats(db_path: str):
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, total_calls, total_duration, average from top_kernels").fetchall()
    return [(_short(n), int(c), float(t), float(a)) for n, c, t, a in rows]


def counter_stats(db_path: str):
    """{(kernel, counter): (dispatches, mean value per dispatch)}."""
    db = sqlite3.connect(db_path)
    cols = [r[1] for r in db.execute("pragma table_info(pmc_events)")]
    if "counter_name" in cols:
        cname = "counter_name"
    elif "name" in cols and "pmc_name" in cols:
        cname = "pmc_name"
    else:
        cname = [c for c in cols if "counter" in c or c.endswith("_name")][0]
    vname = "value" if "value" in cols else [c for c in cols if "value" in c][0]
    rows = db.execute(f"select name, {cname}, dispatch_id, {vname} from pmc_events").fetchall()
    acc = {}
    for kname, counter, disp, val in rows:
        key = (_short(kname), counter)
        per = acc.setdefault(key, {})
        per[disp] = per.get(disp, 0.0) + float(val)  # sum over dimensions of one dispatch
    return {k: (len(v), sum(v.values()) / len(v)) for k, v in acc.items()}


def dispatch_phases(db_path: str, kernel: str, tail: int, prev: int = 0):
    """Average duration (us) of the kernel's dispatches: all of them, the last
    `tail` (bench.py's wave-stamped graph replays run last), the `prev`
    before those (bench.py's timed steps) and the others."""
    db = sqlite3.connect(db_path)
    rows = db.execute("select duration from kernels where name like ? order by start",
                      (f"%{kernel}%",)).fetchall()
    d = [r[0] / 1e3 for r in rows]
    if not d:
        return None
    tail = min(tail, len(d))
    prev = min(prev, len(d) - tail)
    last = d[len(d) - tail:]
    before = d[len(d) - tail - prev:len(d) - tail]
    head = d[:len(d) - tail - prev]
    mean = lambda x: sum(x) / len(x) if x else float("nan")  # noqa: E731
    return {"all": (len(d), mean(d)), "last": (len(last), mean(last)),
            "before": (len(before), mean(before)), "others": (len(head), mean(head))}


def timeline(db_path: str, last: int):
    """The last `last` kernel dispatches in start order, per kernel: their
    count, mean duration and the mean gap from the end of the dispatch before
    each to its start (us); and the wall-clock span of the whole window."""
    db = sqlite3.connect(db_path)
    rows = db.execute("select name, start, end from kernels order by start").fetchall()[-last:]
    per = {}
    for k, (name, start, end) in enumerate(rows):
        n = _short(name)
        e = per.setdefault(n, [0, 0.0, 0.0])
        e[0] += 1
        e[1] += (end - start) / 1e3
        if k:
            e[2] += (start - rows[k - 1][2]) / 1e3
    span = (rows[-1][2] - rows[0][1]) / 1e3 if rows else 0.0
    return {n: (c, d / c, g / c) for n, (c, d, g) in per.items()}, span


def _summarize(args) -> int:
    lines = []
    if args.command_line:
        lines.append(f"Command: `{args.command_line}`")
        lines.append("")
    for path in args.db:
        lines.append(f"## {os.path.basename(os.path.dirname(path)) or path}")
        lines.append("")
        lines.append("| kernel | launches | total us | avg us |")
        lines.append("|---|---|---|---|")
        for name, calls, tot, avg in kernel_stats(path):
            lines.append(f"| {name} | {calls} | {tot:.1f} | {avg:.3f} |")
        try:
            cs = counter_stats(path)
        except (sqlite3.Error, IndexError):
            cs = {}
        if cs:
            lines.append("")
            lines.append("| kernel | counter | dispatches | mean per dispatch |")
            lines.append("|---|---|---|---|")
            for (k, c), (n, m) in sorted(cs.items()):
                lines.append(f"| {k} | {c} | {n} | {m:.1f} |")
        for spec in args.phases or []:
            parts = spec.split(":")
            kname, tail = parts[0], int(parts[1])
            prev = int(parts[2]) if len(parts) > 2 else 0
            ph = dispatch_phases(path, kname, tail, prev)
            if ph:
                text = (f"Dispatches of {kname}: all {ph['all'][0]} avg {ph['all'][1]:.3f} us; "
                        f"last {ph['last'][0]} (the wave-stamped graph replays) avg {ph['last'][1]:.3f} us; ")
                if prev:
                    text += (f"the {ph['before'][0]} before them (graph replays of the timed steps) "
                             f"avg {ph['before'][1]:.3f} us; ")
                text += f"the others (tuner passes, warm-up, served solves) {ph['others'][0]} avg {ph['others'][1]:.3f} us"
                lines.append(text)
        if args.timeline:
            per, span = timeline(path, args.timeline)
            lines.append("")
            lines.append(f"Last {args.timeline} dispatches ({span:.1f} us from first start to last end):")
            lines.append("")
            lines.append("| kernel | dispatches | avg us | avg gap before it, us |")
            lines.append("|---|---|---|---|")
            for n, (c, d, g) in per.items():
                lines.append(f"| {n} | {c} | {d:.3f} | {g:.3f} |")
        lines.append("")
    text = "\n".join(lines)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(text)
    return 0


def _find(cs, kernel, counter):
    """Mean per dispatch of `counter` over every kernel whose name contains
    `kernel` (all template variants, weighted by their dispatch counts)."""
    tot = cnt = 0.0
    for (k, c), (n, m) in cs.items():
        if c == counter and kernel in k:
            tot += n * m
            cnt += n
    if not cnt:
        raise KeyError(f"{counter} for {kernel} not found")
    return tot / cnt


def _traffic(args) -> int:
    fetch = counter_stats(args.fetch)
    write = counter_stats(args.write)
    copy_kib = args.copy_mib * 1024.0
    cf = _find(fetch, args.copy_kernel, "FETCH_SIZE")
    cw = _find(write, args.copy_kernel, "WRITE_SIZE")
    kf = _find(fetch, args.kernel, "FETCH_SIZE")
    kw = _find(write, args.kernel, "WRITE_SIZE")
    sf, sw = copy_kib / cf, copy_kib / cw
    bytes_per_launch = (kf * sf + kw * sw) * 1024.0
    out = {
        "kernel": args.kernel,
        "fetch_kib_per_dispatch": kf, "write_kib_per_dispatch": kw,
        "copy_mib": args.copy_mib, "copy_fetch_kib": cf, "copy_write_kib": cw,
        "fetch_scale": sf, "write_scale": sw,
        "bytes_per_launch": bytes_per_launch,
    }
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))
    return 0


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = p.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--nx", type=int, default=768)
    r.add_argument("--dtype", choices=("f64", "f32"), default="f64")
    r.add_argument("--substeps", type=int, default=10)
    r.add_argument("--copy-mib", type=int, default=1024)
    r.add_argument("--rows", type=int, nargs=3, default=None, help="rows_a rows_b rows_f (0: separate kernels)")
    r.add_argument("--split", type=int, default=None, help="short band height of the fused kernel (0: uniform)")
    r.add_argument("--graph", action="store_true", help="replay recorded substep-pair graphs")
    r.add_argument("--frame-form", type=int, choices=(0, 1), default=None,
                   help="1: cube-edge frame launch, 0: frame launch and halo / edge launch")
    s = sub.add_parser("summarize")
    s.add_argument("db", nargs="+")
    s.add_argument("--out", default=None)
    s.add_argument("--command-line", default="")
    s.add_argument("--timeline", type=int, default=0,
                   help="also summarize the last N dispatches: duration and gap before each, per kernel")
    s.add_argument("--phases", nargs="*",
                   help="KERNEL:N[:M] - split the KERNEL dispatches into the last N, the M before them and the rest")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-kernel", default="MulFunctor")
    t.add_argument("--copy-mib", type=int, default=1024)
    t.add_argument("--out", required=True)
    a = p.parse_args(argv)
    return {"run": _run, "summarize": _summarize, "traffic": _traffic}[a.cmd](a)


if __name__ == "__main__":
    sys.exit(main())
