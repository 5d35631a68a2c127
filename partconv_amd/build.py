"""Builds libpartconv_amd.so (HIP kernels + engine + C ABI) for gfx950 in-tree.

    python -m partconv_amd.build
"""
from __future__ import annotations

import os
import pathlib
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
LIB = PKG / "libpartconv_amd.so"
SOURCES = [PKG / "csrc" / "pc_engine.hip", PKG / "csrc" / "pc_io.cpp"]
HEADERS = [PKG / "csrc" / "pc_fft.hpp", PKG / "csrc" / "pc_fftreg.hpp", PKG / "csrc" / "pc_kernels.hpp",
           PKG / "csrc" / "pc_offline.hpp", PKG / "csrc" / "pc_mimo_rt.hpp", PKG / "csrc" / "pc_io.hpp",
           ROOT / "include" / "partconv_amd.h"]
ARCH = os.environ.get("PARTCONV_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-fvisibility=hidden", "-Wall", "-Wno-unused-result", "-Wno-unused-value",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", str(ROOT / "include"),
           "-o", str(LIB) + ".tmp"] + [str(s) for s in SOURCES] + ["-lhsa-runtime64"]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
