"""Restart files in the reference's layout, for hosts that drive the library
from Python.

The Fortran host keeps its own backup_Mesco_write / backup_Mesco_read
(src/Backup_write_read.f90:150-209, 56-112) and only needs gsis_download_frd /
gsis_upload_frd around them.  This module writes and reads the same bytes:

* ``restart{1,2}/star00/restart00000.dat``: per block, in block order, one
  Fortran unformatted sequential record holding B%FRD(2,-3:nx+3,-3:ny+3,nvi,nvj)
  as REAL(8) (:201-206, :95-101); the file of MPI rank 0, the only rank of a
  single-GPU run (the file names follow :185-193 with myid = 0);
* ``restart{1,2}/flag.dat``: list-directed 1, Mesh(1)%tt, Mesh(1)%Kstep
  (backup_Macro_write, :114-147).

A record longer than 2147483639 bytes is split into subrecords: each carries a
4-byte length before and after it, the leading length is negative when more
subrecords follow and the trailing length is negative when subrecords came
before (the convention of gfortran and Intel Fortran, which flang follows).
The main loop writes restart1 at Kstep = 500 (mod 1000) and restart2 at
Kstep = 0 (mod 1000) (src/opencfd_ec2d.f90:53-54); a restart takes restart2
when its flag is 1, else restart1 (init_flow_read, :15-46).
"""
from __future__ import annotations

import pathlib
import struct

import numpy as np

MAX_SUBRECORD = 2147483639


def backup_dir(flag: int) -> str:
    """Directory of backup slot 1 or 2."""
    if flag not in (1, 2):
        raise ValueError("backup flag must be 1 or 2")
    return f"restart{flag}"


def write_record(f, payload: bytes | memoryview, max_sub: int = MAX_SUBRECORD) -> None:
    """Append one unformatted sequential record."""
    mv = memoryview(payload).cast("B")
    n = len(mv)
    if n == 0:
        f.write(struct.pack("<ii", 0, 0))
        return
    off = 0
    while off < n:
        k = min(max_sub, n - off)
        more, first = off + k < n, off == 0
        f.write(struct.pack("<i", -k if more else k))
        f.write(mv[off:off + k])
        f.write(struct.pack("<i", k if first else -k))
        off += k


def read_record(f) -> bytes:
    """Read the next unformatted sequential record (all its subrecords)."""
    parts = []
    first = True
    while True:
        head = f.read(4)
        if len(head) < 4:
            raise EOFError("end of file inside or before a record")
        (lead,) = struct.unpack("<i", head)
        k = abs(lead)
        data = f.read(k)
        (trail,) = struct.unpack("<i", f.read(4))
        if len(data) != k or abs(trail) != k or (trail < 0) == first:
            raise ValueError("corrupt record markers")
        parts.append(data)
        first = False
        if lead >= 0:
            return b"".join(parts)


def write_restart(root: str | pathlib.Path, flag: int, frd, tt: float, kstep: int,
                  max_sub: int = MAX_SUBRECORD) -> pathlib.Path:
    """backup_Macro_write(flag): the distribution first, then flag.dat.
    ``frd`` is one block's array or a list of them, one record each."""
    d = pathlib.Path(root) / backup_dir(flag)
    (d / "star00").mkdir(parents=True, exist_ok=True)
    blocks = frd if isinstance(frd, (list, tuple)) else [frd]
    with open(d / "star00" / "restart00000.dat", "wb") as f:
        for one in blocks:
            a = np.asfortranarray(one, dtype="<f8")
            write_record(f, a.ravel(order="F").data, max_sub)
    (d / "flag.dat").write_text(f"{1:12d}\n{float(tt)!r}\n{int(kstep):12d}\n")
    return d


def _flag(path: pathlib.Path):
    """(flag, tt, kstep) of a flag.dat, or None when there is none."""
    if not path.exists():
        return None
    tok = path.read_text().split()
    return int(tok[0]), float(tok[1].replace("D", "E").replace("d", "e")), int(tok[2])


def read_restart(root: str | pathlib.Path, shape) -> tuple:
    """init_flow_read + backup_Mesco_read: (frd, tt, kstep, slot) from the slot
    whose flag is 1, restart2 first.  ``shape`` is FRD's (2, nx+7, ny+7, nvi, nvj),
    or a list of such shapes, one per block, for which frd is a list."""
    root = pathlib.Path(root)
    multi = len(shape) > 0 and isinstance(shape[0], (list, tuple))
    shapes = list(shape) if multi else [shape]
    for slot in (2, 1):
        fl = _flag(root / backup_dir(slot) / "flag.dat")
        if fl is not None and fl[0] == 1:
            frds = []
            with open(root / backup_dir(slot) / "star00" / "restart00000.dat", "rb") as f:
                for m, shp in enumerate(shapes, start=1):
                    raw = read_record(f)
                    n = int(np.prod(shp))
                    if len(raw) != 8 * n:
                        raise ValueError(f"restart record {m} holds {len(raw)} bytes, FRD needs {8 * n}")
                    frds.append(np.frombuffer(raw, dtype="<f8").reshape(shp, order="F").copy())
            return (frds if multi else frds[0]), fl[1], fl[2], slot
    raise FileNotFoundError("No backup data is available")
