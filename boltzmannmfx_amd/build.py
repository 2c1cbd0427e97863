"""Builds the gfx950 library libbmx_mi355x.so in-tree with hipcc.

Usage: python -m boltzmannmfx_amd.build   (or build() from Python)

Flags:
  -ffp-contract=off   every floating-point operation is rounded as written,
                      so sphere results match a CPU build of the same
                      expressions bit for bit;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions are generated.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB_NAME = "libbmx_mi355x.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

SOURCES = ["bmx_kernels.hip", "bmx_capi.hip", "bmx_adjoint.hip", "bmx_exchange.hip",
           "bmx_mesh.hip", "bmx_split.hip", "bmx_fusion.hip"]
HOST_SOURCES = ["bmx_container.cpp", "bmx_io.cpp", "bmx_checkpoint.cpp"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def common_flags(arch: str) -> list[str]:
    return [
        f"--offload-arch={arch}",
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
        f"-I{INCLUDE}",
        f"-I{CSRC}",
    ]


def build(arch: str = "gfx950", verbose: bool = False) -> str:
    """Compile every HIP/C++ source and link the shared library; returns its path."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    srcs = [s for s in SOURCES + HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    for src in srcs:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc, *common_flags(arch), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed on {src}")
        objs.append(obj)
    tmp = LIB_PATH + ".tmp"
    cmd = [cc, f"--offload-arch={arch}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB_PATH)
    # deck driver linked against the library (found next to it at run time)
    exe = os.path.join(HERE, "bmx_run")
    cmd = [cc, "-O2", "-std=c++17", f"-I{INCLUDE}", f"-I{CSRC}",
           os.path.join(CSRC, "bmx_run.cpp"), "-o", exe + ".tmp", f"-L{HERE}",
           "-lbmx_mi355x", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("bmx_run link failed")
    os.replace(exe + ".tmp", exe)
    return LIB_PATH


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
