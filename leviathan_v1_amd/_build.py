"""Build recipe for the native libraries (compiled in-tree, loaded by ctypes).

* ``libleviathan_graph.so`` -- host C++ (g++ -O3 -fopenmp): seeded topology
  builders.
* ``libleviathan_hip.so``   -- the gfx950 HIP kernels and the C-ABI declared in
  ``include/leviathan.h`` and ``include/leviathan_phi.h`` (hipcc --offload-arch=gfx950).

Both land next to this file so they travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

GRAPH_SO = os.path.join(PKG, "libleviathan_graph.so")
HIP_SO = os.path.join(PKG, "libleviathan_hip.so")

HIP_SOURCES = ["leviathan_api.hip", "leviathan_kernels.hip", "leviathan_binned.hip", "plan.cpp", "reservoir.hip",
               "phi.hip"]
HIP_HEADERS = ["leviathan_internal.h", "step_kernel.h", "phi_internal.h", "sincos_bounded.h"]

# Packed FP32 instructions are switched off for every device compile.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-fno-gpu-rdc",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off",  # explicit fmaf only: identical rounding in every kernel
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_graph() -> str:
    src = os.path.join(CSRC, "graph.cpp")
    tmp = GRAPH_SO + ".tmp"
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall", "-o", tmp, src])
    os.replace(tmp, GRAPH_SO)
    return GRAPH_SO


def build_hip() -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = HIP_SO + ".tmp"
    # rocSOLVER (with rocBLAS) factorises the reservoir readout's FP64 Gram on the device
    _run([_hipcc(), *HIPCC_FLAGS, "-I", INCLUDE, "-I", CSRC, "-o", tmp, *srcs, "-lrocsolver", "-lrocblas"])
    os.replace(tmp, HIP_SO)
    return HIP_SO


def build_all() -> None:
    build_graph()
    build_hip()


if __name__ == "__main__":
    which = sys.argv[1:] or ["all"]
    for w in which:
        {"graph": build_graph, "hip": build_hip, "all": build_all}[w]()
