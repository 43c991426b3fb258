"""Builds the P3 HIP library (scream_amd/libp3x.so) in-tree for gfx950.

Every HIP translation unit is compiled with hipcc for gfx950 only, with
packed-FP32 instruction selection disabled and no floating-point contraction
(the FP64 form must round like the reference).  Translation units compile in
parallel; the library links them with hipcc.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "libp3x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          f"-I{os.path.join(HERE, '..', 'include')}"]

SOURCES = ["p3_kernel_fp64.hip", "p3_kernel_mixed.hip", "p3x_api.hip", "p3x_server.hip",
           "p3x_process.hip", "p3_tables.cpp", "p3_table_gen.hip"]

# Extra flags per translation unit.  The mixed-precision form may use the
# hardware-backed FP32 division, square root and transcendental sequences
# (a few ulp instead of correctly rounded); FP64 arithmetic is not affected by
# these flags, and the FP64 form is a translation unit of its own.
FILE_FLAGS: dict[str, list[str]] = {
    "p3_kernel_mixed.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-fapprox-func"],
}


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + FILE_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-4000:]}")
    return out


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", LIB + ".tmp"] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(LIB + ".tmp", LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
