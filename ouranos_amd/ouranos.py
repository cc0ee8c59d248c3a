"""The reference's executable, driven by its own parameter files.

    python -m ouranos_amd.ouranos [--precision fp64|fp32] [--verbose] parameters.inp

Reads the parameter file and the geometry and cross-section files it names
(ouranos.cc:255-330), builds one plan per energy group and solves with the
solver the file selects, as ouranos.cc:43-209 does:
* SI: source iteration until ||phi_new - phi|| / ||phi_new|| < the inner
  tolerance, at most "Max inner iterations" sweeps (ouranos.cc:62-92);
* GMRES / BICGSTAB: the Krylov solve of (I - D L^-1 M Sigma) phi = rhs with
  "Max inner iterations" and the inner tolerance (ouranos.cc:131-171; GMRES
  restarts at 30, deal.II's default).
With several groups, the outer iteration (ouranos.cc:62-205) sweeps the
groups in order; group g's right-hand side is D L^-1 (sum over the other
groups g' of sigma_s(g'->g) M phi_g' + M q_g / ws + inflow_g), built from the
latest fluxes (RadiativeTransfer.cc:845-876, with the stale-source defect
A.11 fixed), and the iteration stops when sqrt(sum_g |phi_g - phi_g_old|^2 /
sum_g |phi_g|^2) falls below the outer tolerance or after "Max outer
iterations".  Each group's solve starts from its own latest flux.
Writes the flux moments as the reference's output_results does (ouranos.cc:
211-253): `<Output file>0000.vtk` with the point fields flux_000g_<mmm>m and
the record `<Output file>.vtk`.  Also writes them, moment-major [m][dof] with
cells numbered x fastest and each cell's dofs lexicographically, to
`<Output file>.npy`, and prints one JSON line with the solve's outcome.
"""
import argparse
import ctypes
import json
import sys
import time

import numpy as np

import ouranos_amd as ou
from ouranos_amd.inp import Geometry, group_problems_from_parameters, resolve

SOLVER_NAMES = {ou.SI: "SI", ou.GMRES: "GMRES", ou.BICGSTAB: "BICGSTAB"}


def output_results(filename, spec, group_flux, origin=None):
    """Legacy VTK of the flux moments (ou_write_vtk): group_flux is one
    moment vector per group ([m][dof] each) on the mesh of `spec`."""
    L = ou.lib()
    dp, ip = ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.ou_write_vtk.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ip, dp, dp,
                               ctypes.c_int, ctypes.c_int, dp]
    flux = np.ascontiguousarray(np.stack([np.asarray(g, dtype=np.float64)
                                          for g in group_flux]))
    n = (ctypes.c_int * 3)(*(list(spec.n) + [1] * (3 - len(spec.n))))
    h = (ctypes.c_double * 3)(*(list(spec.h) + [1.0] * (3 - len(spec.h))))
    o = (ctypes.c_double * 3)(*((list(origin) if origin is not None else [0.0] * spec.dim)
                                + [0.0] * (3 - spec.dim)))
    ou._check(L.ou_write_vtk(str(filename).encode(), spec.dim, spec.p, n, h, o, len(flux),
                             spec.quad.n_mom, flux.ctypes.data_as(dp)))
    return str(filename) + "0000.vtk"


def solve_groups(prm, specs, cross, precision=ou.FP64, verbose=False):
    """The outer group iteration (ouranos.cc:62-205) on one plan per group.
    Returns ([phi_g], summary)."""
    ops = [s.operator(precision) for s in specs]
    ng = len(ops)
    phi = [np.zeros(op.n) for op in ops]
    solver, tol, max_it = prm.get_solver_type(), prm.get_inner_tolerance(), prm.get_max_inner_it()
    inner, outer_its, change, converged = [], 0, 0.0, True
    for out_it in range(prm.get_max_outer_it() if ng > 1 else 1):
        old = [p.copy() for p in phi]
        for g, op in enumerate(ops):
            if ng == 1:
                phi[g], info = op.solve(solver=solver, tol=tol, max_it=max_it, restart=30,
                                        phi0=phi[g], verbose=verbose, raise_on_fail=False)
            else:
                first = True
                for gp in range(ng):
                    if gp != g:
                        op.scatter_source_with(phi[gp], cross[g][gp], accumulate=not first)
                        first = False
                rhs = op.sweep(None, with_source=True)
                phi[g], info = op.solve_rhs(rhs, solver=solver, tol=tol, max_it=max_it,
                                            restart=30, phi0=phi[g], verbose=verbose,
                                            raise_on_fail=False)
            inner.append(info)
            converged &= info["converged"]
        outer_its = out_it + 1
        if ng > 1:
            num = sum(float(np.sum((o - p) ** 2)) for o, p in zip(old, phi))
            den = sum(float(np.sum(p ** 2)) for p in phi)
            change = float(np.sqrt(num / den)) if den > 0 else 0.0
            if verbose:
                print(f"[ouranos] outer iteration {out_it}: change {change:.3e}", flush=True)
            if change < prm.get_outer_tolerance():
                break
    if ng > 1:
        converged = converged and change < prm.get_outer_tolerance()
    summary = {"groups": ng, "outer_iterations": outer_its, "outer_change": change,
               "inner_iterations": [i["iterations"] for i in inner],
               "rel_residual": max(i["rel_residual"] for i in inner), "converged": converged}
    return phi, summary


def run(parameters_path, precision=ou.FP64, verbose=False, output=None):
    """Solves the problem of one parameter file.  Returns (phi, summary):
    phi is the moment vector of a one-group problem, else one per group."""
    prm, specs, cross = group_problems_from_parameters(parameters_path)
    spec = specs[0]
    geo = Geometry(resolve(prm.get_geometry_filename(), parameters_path), spec.dim)
    t0 = time.perf_counter()
    phis, s = solve_groups(prm, specs, cross, precision, verbose)
    seconds = time.perf_counter() - t0
    summary = {"parameters": str(parameters_path),
               "solver": SOLVER_NAMES.get(prm.get_solver_type(), "?"),
               "dim": spec.dim, "fe_order": spec.p, "n_cells": list(spec.n),
               "n_dir": spec.quad.n_dir, "n_mom": spec.quad.n_mom, "unknowns": spec.size,
               "iterations": sum(s["inner_iterations"]), **s, "seconds": seconds}
    out = output if output is not None else prm.get_output_filename()
    if out:
        np.save(out + ".npy", np.stack(phis) if len(phis) > 1 else phis[0])
        summary["output"] = out + ".npy"
        summary["vtk"] = output_results(out, spec, phis, origin=geo.get_bottom_left())
    return (phis if len(phis) > 1 else phis[0]), summary


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("parameters")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    ap.add_argument("--output", default=None, help="overrides the file's Output file")
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    _, summary = run(a.parameters, ou.FP32 if a.precision == "fp32" else ou.FP64, a.verbose,
                     a.output)
    print(json.dumps(summary), flush=True)
    return 0 if summary["converged"] else 1


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
