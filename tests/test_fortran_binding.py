"""The Fortran ISO_C_BINDING module (fortran/gsis_capi.f90) lays out
gsis_params / gsis_bc exactly as the C header does: sizes and field offsets
from flang, gcc and ctypes agree.  No GPU needed."""
import ctypes
import pathlib
import shutil
import subprocess

import pytest

from helpers import load_pkg

ROOT = pathlib.Path(__file__).resolve().parents[1]
PKG = ROOT / "hybrid2d-2--sis---shakov---coette--9-6_amd"
FLANG = "/opt/rocm/lib/llvm/bin/flang"
FIELDS = ("vi", "bc", "cfc", "scheme", "cfl", "ns_tol", "ns_maxit", "walk_tiles")


def _flang():
    return FLANG if pathlib.Path(FLANG).exists() else shutil.which("flang")


@pytest.mark.skipif(_flang() is None, reason="flang not available")
def test_fortran_and_c_layouts_agree(tmp_path):
    f90 = tmp_path / "layout.f90"
    offs = "\n".join(
        f"  print '(a,1x,i0)', '{f}', transfer(c_loc(p%{f}), 0_c_intptr_t) - transfer(c_loc(p), 0_c_intptr_t)"
        for f in FIELDS)
    f90.write_text(f"""program layout
  use iso_c_binding
  use gsis_capi
  implicit none
  type(gsis_params_t), target :: p
  type(gsis_bc_t) :: b
  print '(a,1x,i0)', 'params', c_sizeof(p)
  print '(a,1x,i0)', 'bc', c_sizeof(b)
{offs}
end program
""")
    exe = tmp_path / "layout_f"
    subprocess.run([_flang(), "-J", str(tmp_path), str(PKG / "fortran" / "gsis_capi.f90"), str(f90),
                    f"-L{PKG}", "-lgsis", f"-Wl,-rpath,{PKG}", "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    fort = dict(line.split() for line in subprocess.run([str(exe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    c = tmp_path / "layout.c"
    c.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "gsis_capi.h"\nint main(void) {\n'
                 '  printf("params %zu\\nbc %zu\\n", sizeof(gsis_params), sizeof(gsis_bc));\n'
                 + "".join(f'  printf("{f} %zu\\n", offsetof(gsis_params, {f}));\n' for f in FIELDS)
                 + "  return 0;\n}\n")
    cexe = tmp_path / "layout_c"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(c), "-o", str(cexe)], check=True)
    cres = dict(line.split() for line in subprocess.run([str(cexe)], check=True, capture_output=True,
                                                        text=True).stdout.splitlines())
    capi = load_pkg().capi
    py = {"params": ctypes.sizeof(capi.GsisParams), "bc": ctypes.sizeof(capi.GsisBC)}
    py.update({f: getattr(capi.GsisParams, f).offset for f in FIELDS})
    for k in ("params", "bc") + FIELDS:
        assert int(fort[k]) == int(cres[k]) == py[k], (k, fort[k], cres[k], py[k])
