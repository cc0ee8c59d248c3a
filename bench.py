"""Benchmark of the transport sweep (the operator application of
RadiativeTransfer::Apply) on one MI355X.

One step = one application y = x - D2M L^-1 M2D S(x) of the full problem:
scattering source, every sweep level of all octants, epilogue.  Inputs are
resident in HBM before the timed region.  Default workload: 3D 128^3 hex
mesh, DG Q2 (27 dofs/cell), Gauss-Legendre-Chebyshev 12x24 = 288 directions,
L=1 anisotropic scattering, seeded 16^3-cell material blocks; FP32 angular
flux with FP64 moment accumulation.

The same line carries the rest of the metric, measured in the same run:
  time_to_tol   time to a 1e-8 relative residual (FP64 residual of the FP64
                operator) on the same workload: plain FP64 GMRES(30) against
                the learned right preconditioner in FGMRES(24) on the FP32
                sweep with FP64 refinement; seconds, iterations, residuals,
                scalar-flux difference and the ratio
  fp64_sweep    the FP64 operator application (the reference's precision):
                ms, DoF-angle updates/s and its HBM roofline by FP64 bytes
  cpu_baseline  the FP64 oracle sweep on the host's cores: one thread and the
                box's CPU share, with the core counts seen

Prints ONE JSON line on stdout (progress goes to stderr).  `--gpus N` with
N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "sweep DoF-angle updates/s; time-to-1e-8 learned-FGMRES vs plain GMRES"
HBM_PEAK_GBS = 8000.0
# host threads that one GPU's share of the box may use
CPU_SHARE = 16
WEIGHTS = os.path.join(ROOT, "ouranos_amd", "weights", "unet_c5_h128.npy")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def workload(name):
    from ouranos_amd import problems as P
    if name == "headline_3d":
        return P.headline_3d(), ("3D 128^3 hex mesh, DG Q2 (27 dofs/cell), Gauss-Legendre-"
                                 "Chebyshev 12 polar x 24 azimuthal = 288 directions, 1 group, "
                                 "L=1 Henyey-Greenstein scattering, seeded 16^3-cell blocks "
                                 "(sigma_t log-U[1,500], c U[0.99,0.9999]), vacuum")
    if name == "hetero_2d":
        return P.hetero_2d(), ("2D 256^2 quads, DG Q2, GLC 16x32 (256 directions, mu>0), L=3, "
                               "seeded 16^2-cell blocks, vacuum")
    if name == "uniform_2d":
        return P.uniform_2d(), "2D 32^2 quads, DG Q1, LS-S8 (40 directions), c=0.5, q=1, vacuum"
    if name == "thick_3d":
        return P.thick_3d(), "3D 32^3 hex, DG Q1, LS-S8 (80 directions), 4^3-cell blocks, c~1"
    if name == "batch_3d":
        return P.batch_3d(n=64), "3D 64^3 hex, DG Q1, LS-S12 (168 directions), 8^3-cell blocks"
    raise SystemExit(f"unknown workload {name}")


def measured_traffic(workload_name, precision):
    """HBM traffic of one sweep (bytes, all level launches) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload, if present."""
    path = os.path.join(ROOT, "profiles", f"pmc_k_sweep_level_{workload_name}_{precision}.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if line.startswith("traffic_bytes_per_sweep:"):
            return int(line.split(":")[1]), os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(spec, seconds_budget, threads):
    """Times the FP64 oracle sweep (plain C restatement of
    RadiativeTransfer::sweep, one cell per task, Crout LU per cell) on a slab
    of the same mesh: the first z-layers (all of x, y) with the same
    materials, one direction per thread, each direction sized to take about
    `seconds_budget`.  Returns updates/s over all threads."""
    from oracle import oracle as O
    O.lib()
    n = list(spec.n)
    per_cell_s = 5e-6 * (spec.n_loc / 27.0) ** 2.2  # measured 5.2 us per Q2 cell-direction
    layers = max(1, min(n[-1], int(seconds_budget / per_cell_s / (n[0] * n[1]))))
    sub_n = n[:-1] + [layers]
    nc = int(np.prod(sub_n))
    q = spec.quad
    pb = O.Problem(spec.dim, spec.p, sub_n, spec.h, q.omega, q.M2D, q.D2M, q.mom_l,
                   spec.sigma_t, spec.sigma_s, spec.mat[:nc], q=spec.q[:nc], ws=spec.ws)
    S = np.random.default_rng(0).random(pb.size)
    dirs = [int(d) for d in np.linspace(0, q.n_dir - 1, threads)]
    out = [None] * threads

    def run(k):
        out[k] = pb.sweep(S, d_begin=dirs[k], d_end=dirs[k] + 1)

    ths = [threading.Thread(target=run, args=(k,)) for k in range(threads)]
    t0 = time.perf_counter()
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    dt = time.perf_counter() - t0
    updates = nc * spec.n_loc * threads
    return {"value": updates / dt, "unit": "DoF-angle updates/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle FP64 sweep of {threads} direction(s) (one per thread) over a "
                       f"{'x'.join(map(str, sub_n))} slab of the same mesh and materials: "
                       f"{updates:.3e} updates in {dt:.2f} s")}


def cpu_legs(spec, seconds_budget):
    """The oracle timed on one thread and on the host threads of one GPU's
    share of the box (at most CPU_SHARE, and no more than the process may run
    on).  The multi-thread leg is the reported baseline; both legs and the
    core counts the host shows are kept."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    threads = max(1, min(CPU_SHARE, affinity))
    one = cpu_baseline(spec, seconds_budget / 2, 1)
    many = cpu_baseline(spec, seconds_budget / 2, threads)
    out = dict(many)
    out["single_thread"] = {"value": one["value"], "unit": one["unit"], "cores": 1,
                            "sample": one["sample"]}
    out["parallel_speedup"] = many["value"] / one["value"]
    out["host_cpus"] = {"os_cpu_count": os.cpu_count(), "affinity": affinity,
                        "used": threads, "share_per_gpu": CPU_SHARE}
    return out


def time_to_tolerance(spec, op32, tol):
    """Time to a `tol` relative residual, every residual test FP64 on the FP64
    operator: plain FP64 GMRES(30), the reference's Krylov method with the
    identity preconditioner (ouranos.cc:131-208), against FGMRES(24) on the
    FP32 operator with the learned right preconditioner and FP64 refinement."""
    import ouranos_amd as ou
    from ouranos_amd import learned as Lm
    op64 = spec.operator(ou.FP64)
    nd = spec.n_dofs
    log("[bench] time to tolerance: plain FP64 GMRES(30)")
    t0 = time.perf_counter()
    phi_p, info_p = op64.solve(solver=ou.GMRES, tol=tol, max_it=3000, restart=30,
                               raise_on_fail=False)
    wall_p = time.perf_counter() - t0
    ref = phi_p[:nd].copy()
    del phi_p
    log(f"[bench]   {info_p['iterations']} iterations, {info_p['seconds']:.2f} s")
    net = Lm.load(WEIGHTS)
    log("[bench] time to tolerance: learned FGMRES(24), FP32 sweep, FP64 refinement")
    t0 = time.perf_counter()
    phi_l, info_l = ou.solve_refined(op64, op32, net, tol=tol, max_it=3000, restart=24,
                                     inner_tol=7e-5, raise_on_fail=False)
    wall_l = time.perf_counter() - t0
    diff = float(np.linalg.norm(phi_l[:nd] - ref) / np.linalg.norm(ref))
    del phi_l
    log(f"[bench]   {info_l['iterations']} iterations, {info_l['seconds']:.2f} s")
    op64.close()
    return {
        "tol": tol,
        "plain": {"method": "FP64 GMRES(30), FP64 sweep", "seconds": info_p["seconds"],
                  "wall_seconds": wall_p, "iterations": info_p["iterations"],
                  "fp64_rel_residual": info_p["rel_residual"],
                  "converged": bool(info_p["converged"])},
        "learned": {"method": ("FGMRES(24), inner tol 7e-5, on the FP32 sweep (FP64 moments), "
                               "FP32 Krylov basis "
                               "(FP64 inner products), learned right preconditioner (U-Net, "
                               "BF16, 64^3 coarse grid), FP64 refinement"),
                    "weights": os.path.relpath(WEIGHTS, ROOT),
                    "seconds": info_l["seconds"], "wall_seconds": wall_l,
                    "iterations": info_l["iterations"],
                    "fp64_applications": info_l.get("fp64_applications"),
                    "fp64_rel_residual": info_l["rel_residual"],
                    "converged": bool(info_l["converged"]),
                    "scalar_flux_rel_diff_vs_plain": diff},
        "ratio": info_p["seconds"] / info_l["seconds"],
    }


def fp64_leg(spec, d_x, d_y, warmup, steps):
    """The operator application in FP64 throughout (the reference's
    precision) on the same vectors: ms per application, updates/s, and the
    sweep kernels' HBM roofline by the FP64 form of the traffic formula."""
    import ouranos_amd as ou
    log("[bench] FP64 sweep leg")
    op64 = spec.operator(ou.FP64)
    if warmup > 0:
        op64.time_apply(d_x, d_y, warmup, 1)
    ou.lib().ou_device_synchronize()
    t0 = time.perf_counter()
    tm = op64.time_apply(d_x, d_y, 0, steps)
    ou.lib().ou_device_synchronize()
    wall = time.perf_counter() - t0
    op64.close()
    sweep_ms = tm["ms_sweep_kernel"] / steps
    alg = spec.algorithmic_bytes_per_sweep(real_bytes=8)
    achieved = alg / (sweep_ms / 1000.0) / 1e9
    return {"dtype": "fp64", "steps": steps, "ms_per_step": 1000.0 * wall / steps,
            "value": spec.updates_per_sweep() * steps / wall, "unit": "DoF-angle updates/s",
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_sweep": alg,
                         "sweep_kernel_ms_per_step": sweep_ms}}


def copy_bandwidth(n_bytes=1 << 30, reps=10):
    """Device-to-device copy rate of the library's plain copy kernel
    (ou_memcpy_dtod: n_bytes read + n_bytes written per copy), GB/s: the
    streaming rate HBM actually delivers on this GPU, beside its 8 TB/s peak."""
    import ctypes
    import ouranos_amd as ou
    lib = ou.lib()
    lib.ou_memcpy_dtod.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    a, b = ou.DeviceBuffer(n_bytes), ou.DeviceBuffer(n_bytes)
    lib.ou_memcpy_dtod(b.ptr, a.ptr, n_bytes)
    lib.ou_device_synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        lib.ou_memcpy_dtod(b.ptr, a.ptr, n_bytes)
    lib.ou_device_synchronize()
    dt = time.perf_counter() - t0
    a.free()
    b.free()
    return 2.0 * n_bytes * reps / dt / 1e9


def config_leg(name, warmup, steps):
    """Another configuration's operator application in its plan's default sweep
    mode (the tile sweep for C1, C2 and C4), FP32 and FP64: ms per application
    (HIP events) and the sweep kernels' share."""
    import ouranos_amd as ou
    from ouranos_amd import problems as P
    spec, desc = workload(name)
    log(f"[bench] {name} leg")
    x = P.seeded_vector(spec.size, 1)
    d_x = ou.DeviceBuffer(x.nbytes)
    d_y = ou.DeviceBuffer(x.nbytes)
    d_x.upload(x)
    out = {"workload": desc, "n_cells": list(spec.n), "n_dir": spec.quad.n_dir,
           "updates_per_sweep": spec.updates_per_sweep()}
    modes = {ou.SWEEP_LEVELS: "levels", ou.SWEEP_PERSISTENT: "persistent", ou.SWEEP_TILES: "tiles"}
    ops = {}
    for prec, pname in ((ou.FP32, "fp32"), (ou.FP64, "fp64")):
        op = spec.operator(prec)
        if warmup > 0:
            op.time_apply(d_x, d_y, warmup, 1)
        tm = op.time_apply(d_x, d_y, 0, steps)
        out[pname] = {"mode": modes[op.sweep_mode], "shape": list(op.sweep_shape),
                      "ms_per_apply": tm["ms_total"] / steps,
                      "sweep_kernel_ms": tm["ms_sweep_kernel"] / steps,
                      "launches_per_sweep": tm["n_sweep_launches"],
                      "updates_per_s": spec.updates_per_sweep() * steps / (tm["ms_total"] / 1000.0)}
        ops[pname] = op
    d_x.free()
    d_y.free()
    if name == "uniform_2d":
        # the configuration's own solve: FP64 source iteration to a 1e-10
        # relative scalar-flux change (ouranos.cc:52-130)
        t0 = time.perf_counter()
        _, info = ops["fp64"].solve(solver=ou.SI, tol=1e-10, max_it=2000)
        out["si_1e-10_fp64"] = {"seconds": time.perf_counter() - t0,
                                "iterations": info["iterations"],
                                "converged": bool(info["converged"])}
    elif name == "hetero_2d":
        # FP64 GMRES(30) against FGMRES(30) on the FP32 operator with FP64
        # refinement, both to a 1e-8 FP64 relative residual (ouranos.cc:131-208)
        nd = spec.n_dofs
        t0 = time.perf_counter()
        phi_p, info_p = ops["fp64"].solve(solver=ou.GMRES, tol=1e-8, max_it=5000, restart=30,
                                          raise_on_fail=False)
        wall_p = time.perf_counter() - t0
        t0 = time.perf_counter()
        phi_r, info_r = ou.solve_refined(ops["fp64"], ops["fp32"], None, tol=1e-8, max_it=5000,
                                         restart=30, inner_tol=3e-5, raise_on_fail=False)
        wall_r = time.perf_counter() - t0
        out["gmres_1e-8"] = {
            "plain_fp64": {"seconds": wall_p, "iterations": info_p["iterations"],
                           "rel_residual": info_p["rel_residual"],
                           "converged": bool(info_p["converged"])},
            "fp32_sweep_fp64_refined": {"seconds": wall_r, "iterations": info_r["iterations"],
                                        "rel_residual": info_r["rel_residual"],
                                        "converged": bool(info_r["converged"])},
            "scalar_flux_rel_diff": float(np.linalg.norm(phi_r[:nd] - phi_p[:nd]) /
                                          np.linalg.norm(phi_p[:nd])),
            "ratio": wall_p / wall_r}
    for op in ops.values():
        op.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="headline_3d")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true",
                    help="skip the time-to-tolerance comparison")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 sweep leg")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the C1 / C2 operator timings")
    ap.add_argument("--tol", type=float, default=1e-8)
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path "
              f"(asked for --gpus {a.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import ouranos_amd as ou
    from ouranos_amd import problems as P

    spec, desc = workload(a.workload)
    prec = ou.FP32 if a.precision == "fp32" else ou.FP64
    log(f"[bench] {spec.name}: building the plan")
    t0 = time.perf_counter()
    op = spec.operator(prec)
    t_plan = time.perf_counter() - t0
    x = P.seeded_vector(spec.size, 1)
    d_x = ou.DeviceBuffer(x.nbytes)
    d_y = ou.DeviceBuffer(x.nbytes)
    d_x.upload(x)
    del x
    # warmup + timed region, device synchronised on both sides
    log(f"[bench] {a.warmup} warmup + {a.steps} timed applications")
    op.time_apply(d_x, d_y, a.warmup, 1) if a.warmup > 0 else None
    ou.lib().ou_device_synchronize()
    t0 = time.perf_counter()
    tm = op.time_apply(d_x, d_y, 0, a.steps)
    ou.lib().ou_device_synchronize()
    wall = time.perf_counter() - t0
    y = np.empty(spec.size)
    d_y.download(y)
    if not np.isfinite(y).all():
        raise SystemExit("non-finite output")
    del y

    updates = spec.updates_per_sweep()
    ms_step = 1000.0 * wall / a.steps
    value = updates * a.steps / wall
    sweep_ms = tm["ms_sweep_kernel"] / a.steps
    alg_bytes = spec.algorithmic_bytes_per_sweep(real_bytes=4 if prec == ou.FP32 else 8)
    achieved = alg_bytes / (sweep_ms / 1000.0) / 1e9
    traffic, traffic_src = measured_traffic(spec.name, a.precision)
    line = {
        "metric": METRIC, "value": value, "unit": "DoF-angle updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32" if prec == ou.FP32 else "fp64", "data": "synthetic",
        "config": {"workload": f"{spec.name}: {desc}", "n_cells": list(spec.n),
                   "fe_order": spec.p, "n_dir": spec.quad.n_dir, "n_mom": spec.quad.n_mom,
                   "dofs": spec.n_dofs, "updates_per_sweep": updates,
                   "accumulation": "fp64 moments", "seed": spec.meta.get("seed")},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": tm.get("sweep_kernel", "k_sweep_level (one sweep)"),
                     "algorithmic_bytes_per_sweep": alg_bytes,
                     "sweep_kernel_ms_per_step": sweep_ms,
                     "launches_per_step": tm["n_sweep_launches"]},
        "event_ms_per_step": tm["ms_total"] / a.steps,
        "plan_seconds": t_plan,
    }
    copy = copy_bandwidth()
    line["copy_GBs"] = {"value": copy, "frac_of_copy": achieved / copy,
                        "what": "ou_memcpy_dtod 1 GiB read + 1 GiB written, 10 copies; the "
                                "sweep's algorithmic GB/s as a share of this rate"}
    if not a.no_fp64 and prec == ou.FP32:
        line["fp64_sweep"] = fp64_leg(spec, d_x, d_y, a.warmup, a.steps)
    d_x.free()
    d_y.free()
    if not a.no_configs and spec.name == "headline_3d":
        line["configs"] = {"C1": config_leg("uniform_2d", a.warmup, a.steps),
                           "C2": config_leg("hetero_2d", a.warmup, a.steps),
                           "C4": config_leg("thick_3d", a.warmup, a.steps)}
    if not a.no_solve and spec.name == "headline_3d":
        try:
            line["time_to_tol"] = time_to_tolerance(spec, op, a.tol)
        except Exception as e:  # reported, not hidden: the line still prints
            line["time_to_tol"] = {"error": f"{type(e).__name__}: {e}"}
    op.close()
    if not a.no_cpu:
        log("[bench] CPU baseline")
        line["cpu_baseline"] = cpu_legs(spec, a.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
