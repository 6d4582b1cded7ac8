"""Builds libstride_hip.so in-tree (gfx950 device code + host C++).

Usage: python -m meshpress_amd.build [--jobs N]

Device translation units are compiled with hipcc for gfx950 only.  Flags:
  -ffp-contract=off   the FP64 dequantisation must not fuse multiply-adds
                      (bit parity with the reference decoder's numpy chain)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions in device code
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libstride_hip.so"
BENCH = PKG / "stride_bench"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["stride_decode.hip", "stride_search.hip", "stride_capi.hip", "stride_mlp.hip",
               "pdlt_decode.hip"]
CXX_SOURCES = ["stride_host.cpp", "stride_dprb.cpp", "pdlt_sblb.cpp"]

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", "-Wall",
             "-Wno-unused-function"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(jobs: int = 8, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    cmds = []
    objs = []
    for src in HIP_SOURCES:
        p = CSRC / src
        if not p.exists():
            continue
        o = BUILD / (src + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(p), "-o", str(o)])
        objs.append(o)
    cxx = shutil.which("g++") or "g++"
    for src in CXX_SOURCES:
        o = BUILD / (src + ".o")
        cmds.append([cxx, *CXX_FLAGS, "-c", str(CSRC / src), "-o", str(o)])
        objs.append(o)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for r in ex.map(_run, cmds):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout, r.stderr, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", *[str(o) for o in objs], "-o", str(tmp),
          "-pthread"])
    os.replace(tmp, LIB)
    # native decode bench over the C ABI (host code only, links the library)
    cli_tmp = BENCH.with_suffix(".tmp")
    _run([HIPCC, "-O2", "-std=c++17", str(CSRC / "stride_bench_cli.cpp"), "-o", str(cli_tmp),
          f"-L{PKG}", "-lstride_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(cli_tmp, BENCH)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
