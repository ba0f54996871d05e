"""Build libgsis.so for gfx950 with hipcc (in-tree, so the library travels with
the source tree).  Packed-FP32 instructions are disabled for all device code.
Also builds the Fortran host side (the ISO_C_BINDING module and the cavity
host program) with flang when flang is present."""
from __future__ import annotations

import pathlib
import shutil
import subprocess
import sys
import tempfile

PKG = pathlib.Path(__file__).resolve().parent
SRC = [PKG / "csrc" / "gsis_kernels.hip", PKG / "csrc" / "gsis_capi.hip", PKG / "csrc" / "gsis_synthetic.hip",
       PKG / "csrc" / "gsis_mlp.hip", PKG / "csrc" / "gsis_closure.hip"]
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


FLANG = "/opt/rocm/lib/llvm/bin/flang"
FORTRAN = [PKG / "fortran" / "gsis_capi.f90", PKG / "fortran" / "gsis_cavity_host.f90"]
FORTRAN_HOST = PKG / "fortran" / "gsis_cavity_host"


def flang() -> str | None:
    return FLANG if pathlib.Path(FLANG).exists() else shutil.which("flang")


def build_fortran(lib: pathlib.Path = OUT) -> pathlib.Path | None:
    """Compile the Fortran interface module and host program against libgsis.so
    (rpath to the package directory).  Returns None when flang is absent."""
    fc = flang()
    if fc is None:
        return None
    with tempfile.TemporaryDirectory() as mod:
        cmd = [fc, "-O2", "-J", mod, *map(str, FORTRAN), f"-L{lib.parent}", "-lgsis",
               "-Wl,-rpath,$ORIGIN/..", "-o", str(FORTRAN_HOST)]
        r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("flang failed")
    return FORTRAN_HOST


if __name__ == "__main__":
    print(build())
    print(build_fortran())
