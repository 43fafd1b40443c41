"""Build recipe for libcgm_align.so (gfx950 HIP kernels + C ABI + host runtime).

Compiles every source under csrc/ with hipcc for --offload-arch=gfx950 and
links them into ``libcgm_align.so`` next to this file (in-tree, so the built
library travels with the source tree).  Packed-FP32 instruction selection is
disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcgm_align.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["cgm_runtime.hip", "cgm_sw.hip", "cgm_sw_lane.hip", "cgm_wfa.hip", "cgm_sort.hip",
               "cgm_mixed.hip", "cgm_mixed_wave.hip", "cgm_engine.hip", "cgm_encode.hip",
               "cgm_cigar.hip", "cgm_nw.hip", "cgm_sw_start.hip",
               "cgm_noalloc.hip"]
CXX_SOURCES = ["cgm_gen.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    # Keep compiler warnings visible, minus the known host-pass note about
    # the packed-fp32 target feature.
    noise = "is not a recognized feature for this target"
    lines = [l for l in (r.stdout + r.stderr).splitlines() if noise not in l and l.strip()]
    if lines:
        sys.stderr.write("\n".join(lines) + "\n")


def build(verbose: bool = False) -> str:
    """Compile all sources and link libcgm_align.so; returns its path."""
    from concurrent.futures import ThreadPoolExecutor
    with tempfile.TemporaryDirectory(prefix="cgm_build_") as tmp:
        jobs = []
        for src in HIP_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            jobs.append((obj, [HIPCC, *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj]))
        for src in CXX_SOURCES:
            obj = os.path.join(tmp, src + ".o")
            jobs.append((obj, ["g++", *COMMON, "-march=x86-64-v2", "-pthread", "-c",
                               os.path.join(CSRC, src), "-o", obj]))
        with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
            list(ex.map(lambda j: _run(j[1]), jobs))
        objs = [j[0] for j in jobs]
        tmp_lib = LIB + ".tmp"
        _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", *objs, "-o", tmp_lib])
        os.replace(tmp_lib, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
