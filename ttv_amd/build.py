"""Builds libttvx.so (gfx950) in-tree with hipcc.

Usage: python -m ttv_amd.build [--jobs N]
Every HIP source compiles with the packed-FP32 instructions disabled and for gfx950 only.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
LIBDIR = os.path.join(ROOT, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = [
    "plan.cpp", "selector.cpp", "runtime.hip", "api.hip", "quant.hip",
    "launch_f32.hip", "launch_f64.hip", "launch_bf16.hip", "launch_fp8.hip", "launch_bf16_sub.hip",
    "launch_fp8_sub.hip",
]
OPTIONAL = ["batch.hip", "mlp.hip", "bench_util.hip", "chain.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-result", "-I" + os.path.join(REPO, "include")]


def _compile(src):
    obj = os.path.join(LIBDIR, "obj", os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    return obj


def build(jobs=8, verbose=True):
    os.makedirs(os.path.join(LIBDIR, "obj"), exist_ok=True)
    srcs = SOURCES + [s for s in OPTIONAL if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    out = os.path.join(LIBDIR, "libttvx.so")
    tmp = out + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    a = ap.parse_args()
    build(a.jobs)
