"""The parameter-file driver (python -m ouranos_amd.ouranos) on the GPU: the
reference's integration-test files (tests/rt_parameters.inp with
rt_geometry.inp and rt_xs.inp) solved by each solver the file format
selects, against the 64 moment-0 values RadiativeTransfer_test.cc:84-152
requires (its tolerance 1e-3; the values are printed to 6 digits)."""
import dataclasses
import os

import numpy as np
import pytest

import ouranos_amd as ou
from ouranos_amd import inp, ouranos
from ouranos_amd import problems as P
from helpers import load_golden, oracle_problem, to_reference_numbering

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


def _oracle_groups(O, specs, cross, outer_tol=1e-12):
    """Block Gauss-Seidel over the groups from the oracle's pieces: the
    cross-group source of each other group through an oracle problem with
    that group's sigma_s(g' -> g), the right-hand side swept with the group's
    source, the within-group system by the oracle's GMRES."""
    pbs = [oracle_problem(O, s) for s in specs]
    xs = [[oracle_problem(O, dataclasses.replace(specs[g], sigma_s=cross[g][gp]))
           for gp in range(len(specs))] for g in range(len(specs))]
    phi = [np.zeros(s.size) for s in specs]
    for _ in range(200):
        old = [p.copy() for p in phi]
        for g, pb in enumerate(pbs):
            S = sum(xs[g][gp].scatter_source(phi[gp]) for gp in range(len(specs)) if gp != g)
            rhs = pb.sweep(S, with_source=True)
            phi[g], _, _ = pb.gmres(rhs, 1e-13, x0=phi[g])
        num = sum(np.sum((o - p) ** 2) for o, p in zip(old, phi))
        if np.sqrt(num / sum(np.sum(p ** 2) for p in phi)) < outer_tol:
            break
    return phi


def test_two_groups_match_the_oracle_group_iteration(tmp_path, oracle):
    from test_inp import multigroup_files
    path = multigroup_files(tmp_path)
    print("[gpu] two-group driver vs oracle", flush=True)
    phis, summary = ouranos.run(path, output=str(tmp_path / "mg"))
    assert summary["converged"] and summary["groups"] == 2, summary
    _, specs, cross = inp.group_problems_from_parameters(path)
    ref = _oracle_groups(oracle, specs, cross)
    for g in range(2):
        err = np.linalg.norm(phis[g] - ref[g]) / np.linalg.norm(ref[g])
        assert err < 1e-9, (g, err)
    # the coupled system holds: A_g phi_g = D L^-1 (q_g + inflow + S_g'->g)
    for g in range(2):
        S = sum(oracle_problem(oracle, dataclasses.replace(specs[g], sigma_s=cross[g][gp]))
                .scatter_source(phis[gp]) for gp in range(2) if gp != g)
        pb = oracle_problem(oracle, specs[g])
        res = pb.apply(phis[g]) - pb.sweep(S, with_source=True)
        assert np.linalg.norm(res) / np.linalg.norm(phis[g]) < 1e-9
    # the VTK piece holds both groups
    text = open(summary["vtk"]).read()
    assert "flux_000g_000m" in text and "flux_001g_000m" in text


def test_uncoupled_groups_equal_single_group_solves(tmp_path):
    from test_inp import multigroup_files
    path = multigroup_files(tmp_path, coupling=False)
    phis, summary = ouranos.run(path, output="")
    assert summary["converged"]
    _, specs, _ = inp.group_problems_from_parameters(path)
    for g in range(2):
        ref, _ = specs[g].operator(ou.FP64).solve(solver=ou.GMRES, tol=1e-12, max_it=400)
        assert np.linalg.norm(phis[g] - ref) / np.linalg.norm(ref) < 1e-11


def test_two_groups_by_source_iteration(tmp_path):
    from test_inp import multigroup_files
    gm, _ = ouranos.run(multigroup_files(tmp_path), output="")
    si, summary = ouranos.run(multigroup_files(tmp_path, solver="SI"), output="")
    assert summary["converged"], summary
    for g in range(2):
        assert np.linalg.norm(si[g] - gm[g]) / np.linalg.norm(gm[g]) < 1e-9
