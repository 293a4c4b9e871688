"""Build recipe for libxpbd_hip.so (hipcc, gfx950).  No GPU is needed to build.

    python -m bevy_xpbd_amd.build          # or __graft_entry__.build()

Flags that matter for parity and correctness:
  -ffp-contract=off          no fused multiply-add: the kernels restate the
                             reference's f32 arithmetic operation by operation
  -Xclang -target-feature -Xclang -packed-fp32-ops
                             keep the compiler from forming packed FP32 ops
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxpbd_hip.so")
ARCH = os.environ.get("XPBD_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = [
    ("kernels.hip", "hip"),
    ("mlp.hip", "hip"),
    ("topology.hip", "hip"),
    ("xpbd_api.cpp", "hip"),
    ("mesh_prep.cpp", "c++"),
]


def _compile(src: str, lang: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, "-x", lang, *COMMON]
    if lang == "hip":
        cmd += DEVICE
    cmd += ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {src}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [(s, l) for s, l in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda sl: _compile(*sl), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *DEVICE, "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
