"""Restart files in the reference's layout (src/Backup_write_read.f90): the
unformatted sequential record framing, including records split into
subrecords, the flag.dat contents and the slot preference of init_flow_read."""
import io
import struct

import numpy as np
import pytest

from helpers import load_pkg


def _rs():
    return load_pkg().restart


def test_record_framing_matches_fortran_sequential_layout():
    """A record of n bytes is <n> payload <n> with 4-byte little-endian markers."""
    rs = _rs()
    a = np.arange(5, dtype="<f8")
    f = io.BytesIO()
    rs.write_record(f, a.data)
    raw = f.getvalue()
    assert raw[:4] == struct.pack("<i", 40) and raw[-4:] == struct.pack("<i", 40)
    assert raw[4:44] == a.tobytes()
    f.seek(0)
    assert rs.read_record(f) == a.tobytes()


@pytest.mark.parametrize("max_sub", [7, 16, 40, 1000])
def test_subrecords_round_trip(max_sub):
    """Records longer than the subrecord limit: the leading marker is negative
    on every subrecord but the last, the trailing one on every subrecord but
    the first; reading joins them back, and two records in a row stay apart."""
    rs = _rs()
    a = np.random.default_rng(3).standard_normal(37).astype("<f8")
    b = np.arange(3, dtype="<f8")
    f = io.BytesIO()
    rs.write_record(f, a.data, max_sub)
    rs.write_record(f, b.data, max_sub)
    raw = f.getvalue()
    n = a.nbytes
    k = min(max_sub, n)
    (lead,) = struct.unpack("<i", raw[:4])
    (trail,) = struct.unpack("<i", raw[4 + k:8 + k])
    assert lead == (-k if n > max_sub else k) and trail == k
    f.seek(0)
    assert rs.read_record(f) == a.tobytes()
    assert rs.read_record(f) == b.tobytes()
    with pytest.raises(EOFError):
        rs.read_record(f)


def test_restart_slots_and_flag_file(tmp_path):
    """write_restart writes restart{1,2}/star00/restart00000.dat and flag.dat
    (1, tt, Kstep); read_restart takes restart2 when its flag is 1, else
    restart1, and refuses when neither is valid."""
    rs = _rs()
    shape = (2, 9, 8, 3, 4)
    f1 = np.random.default_rng(1).standard_normal(shape)
    f2 = f1 + 1.0
    with pytest.raises(FileNotFoundError):
        rs.read_restart(tmp_path, shape)
    rs.write_restart(tmp_path, 1, f1, 0.25, 1500)
    frd, tt, k, slot = rs.read_restart(tmp_path, shape)
    assert slot == 1 and k == 1500 and tt == 0.25 and np.array_equal(frd, f1)
    rs.write_restart(tmp_path, 2, f2, 0.5, 1000)
    frd, tt, k, slot = rs.read_restart(tmp_path, shape)
    assert slot == 2 and k == 1000 and tt == 0.5 and np.array_equal(frd, f2)
    # a flag other than 1 in restart2 falls back to restart1
    (tmp_path / "restart2" / "flag.dat").write_text("0\n0.5\n1000\n")
    assert rs.read_restart(tmp_path, shape)[3] == 1
    # Fortran list-directed forms of the same values are read too
    (tmp_path / "restart1" / "flag.dat").write_text(" 1\n  0.250000000000000D+00\n        1500\n")
    assert rs.read_restart(tmp_path, shape)[1:3] == (0.25, 1500)
    with pytest.raises(ValueError):
        rs.read_restart(tmp_path, (2, 9, 8, 3, 5))
