#!/usr/bin/env python3
"""Build libt9hip.so in-tree with hipcc for gfx950.

Host C++ (csrc/host/*.cpp) and HIP kernels (csrc/hip/*.hip) are compiled to
objects under _build/ in parallel and linked into tinker-gpu_amd/libt9hip.so.
Packed-FP32 instructions are disabled for all device code.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libt9hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-std=c++17", "-O3", "-fPIC", "-I" + os.path.join(REPO, "include"), "-I" + os.path.join(HERE, "csrc")]
DEVICE = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src):
    obj = os.path.join(BUILD, os.path.relpath(src, os.path.join(HERE, "csrc")).replace("/", "_") + ".o")
    cmd = [HIPCC] + COMMON + ["-c", src, "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"] + DEVICE
    else:
        cmd[1:1] = ["-x", "c++"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stderr


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "host", "*.cpp")) +
                  glob.glob(os.path.join(HERE, "csrc", "hip", "*.hip")))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for src, obj, rc, err in ex.map(_compile, srcs):
            if rc != 0:
                sys.stderr.write(err)
                raise RuntimeError(f"compile failed: {src}")
            if verbose and err.strip():
                sys.stderr.write(err)
            objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
