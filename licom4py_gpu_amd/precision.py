"""FP32 form vs the binary64 step: accuracy and conservation report.

    python -m licom4py_gpu_amd.precision --nx 768 --substeps 10000 [--chunk 1000]

Runs the same seeded balanced field with the binary64 path (bit-identical to
the NumPy FP64 oracle on every case the test-suite compares) and with the FP32
form, in chunks of `chunk` substeps, and prints one JSON line with
  * relative L2 of height anomaly and of u, v (FP32 vs FP64, interior cells),
  * relative drift of mass and energy (binary64 compensated integrals) for
    both precisions,
  * cell-substeps/s of both runs.
"""

from __future__ import annotations

import argparse
import json
import sys
import time


def run(nx: int, substeps: int, chunk: int, seed: int = 20261203):
    import torch

    from . import barotr, duogrid, initial

    dev = torch.device("cuda", 0)
    g = duogrid.grid(nx)
    h, u, v, sp = initial.seeded_balanced(g, seed=seed)
    dt = initial.courant_dt(g, h, sp)
    consts = barotr.as_device_consts(duogrid.barotr_consts(g, 5000.0), dev)
    host = initial.state_tuple(h, u, v)
    res = {"nx": nx, "substeps": substeps, "dt_s": dt, "cells": 6 * nx * nx}
    finals = {}
    for name, tdt in (("f64", torch.float64), ("f32", torch.float32)):
        ctx = barotr.context_for(consts, tdt)
        s = tuple(torch.as_tensor(x).to(device=dev, dtype=tdt) for x in host)
        m0, e0 = ctx.integrals(s[0], s[2], s[3])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < substeps:
            n = min(chunk, substeps - done)
            s = barotr.barotr_rk2_core(s, consts, n, dt)
            done += n
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        m1, e1 = ctx.integrals(s[0], s[2], s[3])
        res[name] = {"seconds": el, "cell_substeps_per_s": 6 * nx * nx * substeps / el,
                     "mass_drift_rel": (m1 - m0) / m0, "energy_drift_rel": (e1 - e0) / e0,
                     "mass0": m0, "energy0": e0}
        finals[name] = [x[:, 3:-3, 3:-3].to(torch.float64) for x in (s[0], s[2], s[3])]
    for k, nm in enumerate(("h", "u", "v")):
        a, b = finals["f32"][k], finals["f64"][k]
        res[f"rel_l2_{nm}"] = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    res["f32_speedup"] = res["f32"]["cell_substeps_per_s"] / res["f64"]["cell_substeps_per_s"]
    return res


def main(argv=None) -> int:
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--nx", type=int, default=768)
    p.add_argument("--substeps", type=int, default=10000)
    p.add_argument("--chunk", type=int, default=1000)
    a = p.parse_args(argv)
    print(json.dumps(run(a.nx, a.substeps, a.chunk)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
