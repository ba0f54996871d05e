"""The reference's Tecplot files written from the GPU path's fields of the
shipped 40x40 cylinder's initial state.  The oracle writes these files byte
for byte as the reference shipped them (tests/test_output_files.py).  Each GPU
file equals the shipped one too, or differs from the oracle's only by the
print resolution: a value whose FP64 rounding lands on the other side of a
10th-decimal boundary, or a tiny value printed as -0.0000000000 on one side
and 0.0000000000 on the other."""
import hashlib

import numpy as np
import pytest

from helpers import cylinder_case, load_oracle, solver_mod
from test_output_files import DIGESTS, write_initial_output

pytestmark = pytest.mark.gpu


def test_gpu_initial_state_writes_the_shipped_output_files(tmp_path):
    case = cylinder_case()
    print("case:", case.name, "output files", flush=True)
    s = solver_mod().GsisSolver(case)
    s.initialize()
    write_initial_output(s, case, tmp_path / "gpu")
    s.close()
    o = load_oracle().OracleSolver(case)
    o.initialize()
    write_initial_output(o, case, tmp_path / "oracle")
    o.close()
    exact = []
    for name, d in DIGESTS.items():
        raw = (tmp_path / "gpu" / name).read_bytes()
        ref = (tmp_path / "oracle" / name).read_bytes()
        assert hashlib.sha256(ref).hexdigest() == d["sha256"], name
        if raw == ref:
            exact.append(name)
            continue
        la, lb = raw.decode().splitlines(), ref.decode().splitlines()
        assert len(la) == len(lb), name
        nd = 0
        for x, y in zip(la, lb):
            if x == y:
                continue
            assert len(x) == len(y), (name, x, y)
            a = np.array(x.split(), dtype=float)
            b = np.array(y.split(), dtype=float)
            assert np.all(np.abs(a - b) <= 1.0000001e-10), (name, x, y)
            nd += 1
        print(name, "differs from the shipped file on", nd, "of", len(la), "records, within the print resolution")
    print("byte-identical to the shipped files:", exact)
