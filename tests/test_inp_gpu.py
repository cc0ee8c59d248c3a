"""The parameter-file driver (python -m ouranos_amd.ouranos) on the GPU: the
reference's integration-test files (tests/rt_parameters.inp with
rt_geometry.inp and rt_xs.inp) solved by each solver the file format
selects, against the 64 moment-0 values RadiativeTransfer_test.cc:84-152
requires (its tolerance 1e-3; the values are printed to 6 digits)."""
import os

import numpy as np
import pytest

import ouranos_amd as ou
from ouranos_amd import inp, ouranos
from ouranos_amd import problems as P
from helpers import load_golden, to_reference_numbering

pytestmark = pytest.mark.gpu

INP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "inp")


@pytest.mark.parametrize("solver", ["GMRES", "BICGSTAB", "SI"])
def test_reference_parameter_files_solve_to_the_golden_values(tmp_path, solver):
    text = open(os.path.join(INP, "rt_parameters.inp")).read()
    text = text.replace("set Solver type = GMRES", f"set Solver type = {solver}")
    text = text.replace("./tests/", INP + os.sep)  # the copy lives elsewhere
    if solver == "SI":  # source iteration needs more sweeps than the file allows Krylov
        text = text.replace("set Max inner iterations = 100", "set Max inner iterations = 500")
    path = tmp_path / "rt_parameters.inp"
    path.write_text(text)
    print(f"[gpu] parameter-file driver, {solver}", flush=True)
    phi, summary = ouranos.run(str(path), output=str(tmp_path / "phi"))
    assert summary["solver"] == solver and summary["converged"], summary
    gold = np.array(load_golden("reference_2d_solution.json")["moment0"])
    mine = to_reference_numbering(phi, 4, 4)
    assert np.abs(mine - gold).max() < 1e-3
    assert np.abs(mine - gold).max() < 1e-5
    saved = np.load(summary["output"])
    assert np.array_equal(saved, phi)
    assert os.path.getsize(summary["vtk"]) > 0 and os.path.exists(str(tmp_path / "phi.vtk"))


def test_parameter_file_problem_equals_the_built_in_problem():
    """The plan made from the files is the golden problem's plan: GMRES to
    the same tolerance gives the same iterate bit-for-bit."""
    _, spec = inp.problem_from_parameters(os.path.join(INP, "rt_parameters.inp"))
    a, ia = spec.operator(ou.FP64).solve(solver=ou.GMRES, tol=1e-9, max_it=100)
    b, ib = P.reference_2d().operator(ou.FP64).solve(solver=ou.GMRES, tol=1e-9, max_it=100)
    assert ia["iterations"] == ib["iterations"]
    assert np.array_equal(a, b)
