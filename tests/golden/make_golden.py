"""Writes the committed fixtures of tests/golden (run from the repository root):

    python -m tests.golden.make_golden

* grid_c8_digest.json  SHA-256 of every C8 grid table (regression pin of the
                       generator) plus a few scalar summaries;
* routing.json         the cube routing table (the known answer restated from
                       kernel/communication.py:57-87, 277-284 and checked
                       against the Fortran neighbour rule,
                       global_grid_util.F90:151-171);
* stage_c8.npz         one RK stage of the NumPy oracle at C8 on a seeded
                       random padded state (inputs and the 13 outputs), so the
                       GPU parity tests also have a frozen target.

The reference ships no fixtures of its own for this path.
"""

from __future__ import annotations

import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

ROUTING = [  # R[t][d] = (neighbour face, its edge, 0 = ID / 1 = FLIP); E=0, W=1, N=2, S=3
    [[2, 3, 1], [5, 0, 0], [1, 3, 0], [4, 0, 1]],
    [[2, 1, 0], [5, 2, 1], [3, 1, 1], [0, 2, 0]],
    [[4, 3, 1], [1, 0, 0], [3, 3, 0], [0, 0, 1]],
    [[4, 1, 0], [1, 2, 1], [5, 1, 1], [2, 2, 0]],
    [[0, 3, 1], [3, 0, 0], [5, 3, 0], [2, 0, 1]],
    [[0, 1, 0], [3, 2, 1], [1, 1, 1], [4, 2, 0]],
]


def grid_digest(g) -> dict:
    out = {}
    for k in sorted(g):
        a = np.ascontiguousarray(g[k])
        out[k] = hashlib.sha256(a.tobytes()).hexdigest()
    return out


def stage_case(nx: int = 8, seed: int = 8):
    from licom4py_gpu_amd import duogrid
    from oracle import barotr_np as B
    from tests._util import random_core
    g = duogrid.grid(nx)
    consts = duogrid.barotr_consts(g, 5000.0)
    core = random_core(nx, seed=seed)
    outs = B.step_rk_logic(*core, consts, 37.5, 0.0, True)
    return core, outs


def main() -> None:
    from licom4py_gpu_amd import duogrid
    with open(os.path.join(HERE, "grid_c8_digest.json"), "w") as f:
        json.dump(grid_digest(duogrid.generate(8)), f, indent=1, sort_keys=True)
    with open(os.path.join(HERE, "routing.json"), "w") as f:
        json.dump(ROUTING, f)
    core, outs = stage_case()
    names_in = ("h", "hp", "u", "v", "up", "vp")
    names_out = ("h0", "ub", "vb", "celerity_x", "celerity_y", "ub_ct", "vb_ct",
                 "ub_cx", "ub_cy", "vb_cx", "vb_cy", "advx", "advy")
    np.savez_compressed(os.path.join(HERE, "stage_c8.npz"),
                        **{f"in_{k}": v for k, v in zip(names_in, core)},
                        **{f"out_{k}": v for k, v in zip(names_out, outs)},
                        dt=np.float64(37.5), beta=np.float64(0.0), last=np.int32(1))


if __name__ == "__main__":
    main()
