"""Time to a relative residual of 1e-8 on a transport problem, three ways:

  plain     FP64 sweep, GMRES(30) (the plain GPU method)
  refined   FP32 sweep with FP64 moment accumulation inside FGMRES(30),
            restarted from FP64 residuals of the FP64 operator
            (iterative refinement), no preconditioner
  learned   as refined, right-preconditioned by the trained U-Net, with
            FGMRES(24) and inner tolerance 7e-5: two refinement cycles reach
            1e-8 with a margin of about 2.5 on this workload and on held-out
            instances of its generator (1e-4 also takes two cycles, at
            a margin of 1.2-1.5; 2e-4 needs a third)

Every residual test is FP64 on the FP64 operator; the FP32 runs report the
relative L2 difference of their scalar flux (moment 0) from the plain FP64
solution.  Prints one JSON object.

    python -m ouranos_amd.headline_solve [--workload headline_3d] [--skip plain]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import ouranos_amd as ou
from ouranos_amd import learned as Lm
from ouranos_amd import problems as P

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "weights", "unet_c5_h128.npy")


def log(msg):
    print(msg, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="headline_3d")
    ap.add_argument("--n", type=int, default=0, help="mesh size override (cells per axis)")
    ap.add_argument("--block", type=int, default=0, help="material block size override")
    ap.add_argument("--seed", type=int, default=-1, help="material seed override")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--inner-tol", type=float, default=3e-5,
                    help="inner tolerance of the refined solve without a network")
    ap.add_argument("--learned-inner-tol", type=float, default=7e-5)
    ap.add_argument("--max-it", type=int, default=3000)
    ap.add_argument("--restart", type=int, default=30,
                    help="restart length of plain GMRES and the refined solve")
    ap.add_argument("--learned-restart", type=int, default=24)
    ap.add_argument("--weights", default=WEIGHTS)
    ap.add_argument("--grid", type=int, default=0,
                    help="coarse grid of the network (default: the one it was trained on)")
    ap.add_argument("--skip", default="", help="comma list of plain,refined,learned")
    ap.add_argument("--reference", default=None,
                    help="saved plain FP64 solution (.npy) to compare against")
    ap.add_argument("--save-plain", default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args(argv)
    skip = set(s for s in a.skip.split(",") if s)

    builder = getattr(P, a.workload)
    kw = {}
    if a.n:
        kw["n"] = a.n
    if a.block:
        kw["block"] = a.block
    if a.seed >= 0:
        kw["seed"] = a.seed
    spec = builder(**kw)
    log(f"[solve] {spec.name}: n={spec.n}, p={spec.p}, {spec.quad.n_dir} directions, "
        f"{spec.quad.n_mom} moments, {spec.size} unknowns")
    t0 = time.perf_counter()
    op64 = spec.operator(ou.FP64)
    op32 = spec.operator(ou.FP32)
    log(f"[solve] plans built in {time.perf_counter() - t0:.1f} s")
    nd = spec.n_dofs
    out = {"workload": spec.name, "tol": a.tol, "restart": a.restart,
           "learned_restart": a.learned_restart, "learned_inner_tol": a.learned_inner_tol,
           "unknowns": spec.size,
           "device": ou.device_name()}
    sols = {}

    if "plain" not in skip:
        phi, info = op64.solve(solver=ou.GMRES, tol=a.tol, max_it=a.max_it, restart=a.restart,
                               verbose=a.verbose, raise_on_fail=False)
        out["plain"] = info
        sols["plain"] = phi[:nd].copy()
        log(f"[solve] plain FP64 GMRES: {json.dumps(info)}")
        if a.save_plain:
            np.save(a.save_plain, sols["plain"])
        del phi
    elif a.reference:
        sols["plain"] = np.load(a.reference, allow_pickle=False)

    net = None
    if "learned" not in skip:
        net = Lm.load(a.weights, grid=a.grid or None)
        log(f"[solve] network on a {net.grid}^3 coarse grid")
    for name, n in (("learned", net), ("refined", None)):
        if name in skip:
            continue
        phi, info = ou.solve_refined(op64, op32, n, tol=a.tol, max_it=a.max_it,
                                     restart=a.learned_restart if n is not None else a.restart,
                                     inner_tol=a.learned_inner_tol if n is not None else a.inner_tol,
                                     verbose=a.verbose, raise_on_fail=False)
        if "plain" in sols:
            ref = sols["plain"]
            info["scalar_flux_rel_diff_vs_plain"] = float(
                np.linalg.norm(phi[:nd] - ref) / np.linalg.norm(ref))
        out[name] = info
        log(f"[solve] {name}: {json.dumps(info)}")
        del phi
    for name in ("learned", "refined"):
        if name in out and "plain" in out:
            out[f"speedup_{name}_vs_plain"] = out["plain"]["seconds"] / out[name]["seconds"]
    print(json.dumps(out), flush=True)
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:])
