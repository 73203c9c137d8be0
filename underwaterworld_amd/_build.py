"""Compile the engine (csrc/*.hip) into underwaterworld_amd/libuwmesh.so for gfx950.

    python -m underwaterworld_amd._build [-v]

hipcc cross-compiles without a GPU. Flags:
  --offload-arch=gfx950                     the only target
  -ffp-contract=off                         no implicit a*b+c fusion: the mesher reproduces the
                                            reference's f32 operation order bit for bit; kernels
                                            that want FMA call fmaf() explicitly
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                            no packed FP32 (v_pk_*_f32) code generation
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libuwmesh.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


ROOT = os.path.dirname(HERE)
FACADE_SRC = os.path.join(ROOT, "tests", "cpp", "facade_main.cpp")
FACADE_BIN = os.path.join(ROOT, "tests", "cpp", "facade_main")


def build_facade_test(verbose: bool = False) -> str:
    """The C++ facade driver used by tests/test_gpu_facade.py (links libuwmesh.so by rpath)."""
    cmd = [HIPCC, "-O2", "-std=c++17", "-o", FACADE_BIN, FACADE_SRC, "-L" + HERE, "-luwmesh",
           "-Wl,-rpath,$ORIGIN/../../underwaterworld_amd"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building the C++ facade driver")
    return FACADE_BIN


def build(verbose: bool = False) -> str:
    srcs = sources()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-o", tmp, *srcs]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libuwmesh.so")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    if os.path.exists(FACADE_SRC):
        build_facade_test(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
