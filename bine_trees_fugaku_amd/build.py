"""Build libbine_hip.so in-tree for gfx950.

    python -m bine_trees_fugaku_amd.build          # or build() from Python

Host-only sources (schedule, plan compiler) go through g++; the C-ABI layer
and the kernels through hipcc (--offload-arch=gfx950).  HIP code is compiled
with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libbine_hip.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = os.environ.get("BINE_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-command-line-argument", "-ffp-contract=off"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-ffp-contract=off"]

SOURCES = [
    ("schedule.cpp", "host"),
    ("plan.cpp", "host"),
    ("api.cpp", "hiphost"),
    ("sim.cpp", "hiphost"),
    ("gred.hip", "hip"),
    ("gred_realign.hip", "hip"),
    ("sim.hip", "hip"),
    ("mlp.hip", "hip"),
    ("scratch.hip", "hip"),
]


def _hipcc() -> str:
    p = os.path.join(ROCM, "bin", "hipcc")
    return p if os.path.exists(p) else (shutil.which("hipcc") or "hipcc")


def _compile(src: str, kind: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace(".", "_") + ".o")
    inc = ["-I" + INCLUDE, "-I" + CSRC]
    if kind == "host":
        cmd = ["g++", *HOST_FLAGS, *inc, "-c", path, "-o", obj]
    elif kind == "hiphost":
        cmd = ["g++", *HOST_FLAGS, "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include"), *inc, "-c", path,
               "-o", obj]
    else:
        cmd = [_hipcc(), *HIP_FLAGS, *inc, "-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source and link libbine_hip.so (always recompiles)."""
    os.makedirs(BUILD, exist_ok=True)
    srcs = [(s, k) for s, k in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda sk: _compile(*sk), srcs))
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp, *objs,
           "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
