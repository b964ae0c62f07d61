"""Build libhprqp.so in-tree with hipcc for gfx950 (no cmake)."""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent.parent
CSRC = PKG.parent / "csrc"
INCLUDE = ROOT / "include"
SOURCES = ["hprqp_solver.hip", "hprqp_capi.hip", "hprqp_qap.hip", "hprqp_lowrank.hip", "hprqp_batch.hip",
           "hprqp_mlp.hip", "hprqp_warm.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build_library(verbose: bool = False) -> Path:
    out = PKG / "libhprqp.so"
    tmp = PKG / "libhprqp.so.tmp"
    cmd = [HIPCC, *FLAGS, f"-I{INCLUDE}", *(str(CSRC / s) for s in SOURCES), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stderr[-4000:])
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, out)
    return out


def build_test_plugin(verbose: bool = False) -> Path:
    """The tests' Q-operator plugin (tests/plugin/user_csr_qop.hip): a HIP
    library that a caller registers through hprqp_qop_create_user."""
    src = ROOT / "tests" / "plugin" / "user_csr_qop.hip"
    out = src.parent / "libuser_csr_qop.so"
    tmp = src.parent / "libuser_csr_qop.so.tmp"
    cmd = [HIPCC, *FLAGS, f"-I{INCLUDE}", str(src), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stderr[-4000:])
    if verbose and r.stderr:
        print(r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
