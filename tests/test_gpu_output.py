"""The reference's Tecplot files written from the GPU path's fields: the
shipped 40x40 cylinder's initial state on the GPU writes every file that this
state determines byte for byte as the reference shipped it."""
import pytest

from helpers import cylinder_case, solver_mod
from test_output_files import check_digests, write_initial_output

pytestmark = pytest.mark.gpu


def test_gpu_initial_state_writes_the_shipped_output_files(tmp_path):
    case = cylinder_case()
    print("case:", case.name, "output files", flush=True)
    s = solver_mod().GsisSolver(case)
    s.initialize()
    write_initial_output(s, case, tmp_path)
    check_digests(tmp_path)
    s.close()
