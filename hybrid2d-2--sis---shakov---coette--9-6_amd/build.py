"""Build libgsis.so for gfx950 with hipcc (in-tree, so the library travels with
the source tree).  Packed-FP32 instructions are disabled for all device code."""
from __future__ import annotations

import pathlib
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
SRC = [PKG / "csrc" / "gsis_kernels.hip", PKG / "csrc" / "gsis_capi.hip", PKG / "csrc" / "gsis_synthetic.hip"]
OUT = PKG / "libgsis.so"
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument"]


def build(out: pathlib.Path = OUT, extra=()) -> pathlib.Path:
    cmd = [HIPCC, *FLAGS, *extra, "-o", str(out), *map(str, SRC)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed")
    return out


if __name__ == "__main__":
    print(build())
