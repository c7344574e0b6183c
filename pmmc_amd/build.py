"""Build recipe for the HIP library (gfx950) and the oracle.

    python -m pmmc_amd.build            # both
    python -m pmmc_amd.build --lib      # libpmmc_hip.so only

The library is built in place (pmmc_amd/libpmmc_hip.so) so that it travels
with the source tree.  Every HIP translation unit is compiled with packed
FP32 instructions disabled (-packed-fp32-ops).  The host-only rows
(tree, densification, expected distances, level drivers) are plain C++
built with g++ and -ffp-contract=off, so their float arithmetic rounds where
the reference's host code does.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpmmc_hip.so")
SOURCES = ["engine.hip", "heatmap_reg.hip", "api.hip", "markarcs.hip", "sweep.hip", "server.hip",
           "spline.hip", "probe.hip", "singleton.hip", "noalloc.hip",
           "arcs_pair.hip", "stress.hip"]
HOST_SOURCES = ["levels.cpp", "drivers.cpp", "metrics.cpp", "formats.cpp"]
CXX = os.environ.get("CXX", "g++")
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra",
              f"-I{os.path.join(ROOT, 'include')}"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-function",
    f"-I{os.path.join(ROOT, 'include')}",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def build_lib(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES + HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD, os.path.splitext(s)[0] + ".o") for s in srcs]

    def one(i):
        if srcs[i].endswith(".cpp"):
            return _run([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, srcs[i]), "-o", objs[i]])
        return _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, srcs[i]), "-o", objs[i]])

    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(srcs))) as ex:
        list(ex.map(one, range(len(srcs))))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> None:
    _run(["make", "-C", os.path.join(ROOT, "oracle"), "all"])


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    a = ap.parse_args(argv)
    both = not (a.lib or a.oracle)
    if a.lib or both:
        print("built", build_lib())
    if a.oracle or both:
        build_oracle()
        print("built oracle")
    return 0


if __name__ == "__main__":
    sys.exit(main())
