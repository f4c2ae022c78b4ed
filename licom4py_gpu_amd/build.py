"""Build recipes for the package's native libraries (compiled in-tree).

* ``libduogrid_gen.so`` — host C++ generator of the duo-grid tables.
* ``libbarotr_hip.so``  — gfx950 HIP kernels behind the C-ABI of
  ``include/licom_sw.h``.

Both are plain ``g++`` / ``hipcc`` invocations; nothing is JIT-compiled at
import time, so the built ``.so`` files travel with the source tree.
"""

from __future__ import annotations

import os
import re
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")
REPO = os.path.dirname(PKG_DIR)
INCLUDE = os.path.join(REPO, "include")

GRID_SO = os.path.join(PKG_DIR, "libduogrid_gen.so")
HIP_SO = os.path.join(PKG_DIR, "libbarotr_hip.so")

HIP_ARCH = "gfx950"

# -ffp-contract=off keeps every multiply and add separately rounded so the FP64
# path reproduces the reference's operation order bit for bit; packed-FP32
# instruction selection is disabled on this target.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={HIP_ARCH}",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")


def build_grid(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "duogrid_gen.cpp")
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           "-fopenmp", src, "-o", GRID_SO + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(GRID_SO + ".tmp", GRID_SO)
    return GRID_SO


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found")
    return path


def scratch_kernels(remarks: str) -> list:
    """Kernels the compiler's resource report shows with private (scratch)
    memory per lane: register spills or stack objects, which cost every
    launch a scratch allocation and slow memory traffic."""
    out, name = [], None
    for line in remarks.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
        m = re.search(r"ScratchSize \[bytes/lane\]: (\d+)", line)
        if m and int(m.group(1)) > 0 and name:
            out.append((name, int(m.group(1))))
    return out


def build_hip(verbose: bool = False) -> str:
    src = os.path.join(CSRC, "barotr.hip")
    cmd = [hipcc(), *HIP_FLAGS, "-Rpass-analysis=kernel-resource-usage", f"-I{INCLUDE}", f"-I{CSRC}", src,
           "-o", HIP_SO + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    spills = scratch_kernels(proc.stdout)
    if spills:
        os.remove(HIP_SO + ".tmp")
        raise RuntimeError("kernels using scratch memory: " + ", ".join(f"{n} ({b} B/lane)" for n, b in spills))
    os.replace(HIP_SO + ".tmp", HIP_SO)
    return HIP_SO


def build_all(verbose: bool = False) -> None:
    build_grid(verbose)
    build_hip(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
