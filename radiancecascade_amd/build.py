"""Build librcgi.so (the HIP kernels + C ABI) in-tree for gfx950.

Usage: python -m radiancecascade_amd.build   (or __graft_entry__.build()).

Every translation unit is compiled with:
  -ffp-contract=off      each float/double operation rounds as written (the
                         raster and occupancy stages must match bit for bit;
                         the march uses explicit fmaf where it fuses);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                         no packed-FP32 VALU instructions are generated.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librcgi.so")
ARCH = os.environ.get("RC_OFFLOAD_ARCH", "gfx950")
SOURCES = ["rc_scene.hip", "rc_jfa.hip", "rc_edt.hip", "rc_cascade.hip", "rc_norms.hip", "rc_mlp.hip", "rc_surrogate.hip",
           "rc_composite.hip", "rc_dump.cpp", "rc_api.hip", "rc_noalloc.hip", "rc_batch.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
