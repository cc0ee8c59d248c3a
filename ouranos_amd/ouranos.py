"""The reference's executable, driven by its own parameter files.

    python -m ouranos_amd.ouranos [--precision fp64|fp32] [--verbose] parameters.inp

Reads the parameter file and the geometry and cross-section files it names
(ouranos.cc:255-330), builds the plan and solves with the solver the file
selects, as ouranos.cc:43-209 does for one energy group:
* SI: source iteration until ||phi_new - phi|| / ||phi_new|| < the inner
  tolerance, at most "Max inner iterations" sweeps (ouranos.cc:62-92);
* GMRES / BICGSTAB: the Krylov solve of (I - D L^-1 M Sigma) phi = D L^-1 q
  with "Max inner iterations" and the inner tolerance (ouranos.cc:131-171;
  GMRES restarts at 30, deal.II's default).
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
from ouranos_amd.inp import Geometry, problem_from_parameters, resolve

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


def run(parameters_path, precision=ou.FP64, verbose=False, output=None):
    """Solves the problem of one parameter file.  Returns (phi, summary)."""
    prm, spec = problem_from_parameters(parameters_path)
    geo = Geometry(resolve(prm.get_geometry_filename(), parameters_path), spec.dim)
    op = spec.operator(precision)
    t0 = time.perf_counter()
    phi, info = op.solve(solver=prm.get_solver_type(), tol=prm.get_inner_tolerance(),
                         max_it=prm.get_max_inner_it(), restart=30, verbose=verbose,
                         raise_on_fail=False)
    seconds = time.perf_counter() - t0
    summary = {"parameters": str(parameters_path),
               "solver": SOLVER_NAMES.get(prm.get_solver_type(), "?"),
               "dim": spec.dim, "fe_order": spec.p, "n_cells": list(spec.n),
               "n_dir": spec.quad.n_dir, "n_mom": spec.quad.n_mom, "unknowns": spec.size,
               "iterations": info["iterations"], "rel_residual": info["rel_residual"],
               "converged": info["converged"], "seconds": seconds}
    out = output if output is not None else prm.get_output_filename()
    if out:
        np.save(out + ".npy", phi)
        summary["output"] = out + ".npy"
        summary["vtk"] = output_results(out, spec, [phi], origin=geo.get_bottom_left())
    return phi, summary


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
