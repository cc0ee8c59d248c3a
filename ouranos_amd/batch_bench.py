"""Throughput of the batched multi-instance launch on the uneven-instance
workload: 32 unit cubes of n^3 Q1 cells (n uniform over {16,24,32,48,64}),
LS S12 (168 directions), block materials, 4 right-hand sides each = 128
operator applications per batch, FP32 sweep with FP64 moment accumulation.

Measured with HIP events on the batch's stream, inputs resident in HBM:
  graph       the batch captured as one HIP graph and replayed (autotuned shape)
  eager       the same batched kernels launched one by one
  sequential  the 128 members applied one after another through their own
              plans at their default shape (a host loop of ou_apply_device)
and the autotuner's table.  Two members (the smallest and the largest) are
checked bit-for-bit against their single-plan Apply.  Prints one JSON object.

    python -m ouranos_amd.batch_bench [--instances 32] [--rhs 4] [--out file.json]
"""
import argparse
import json
import time

import numpy as np

import ouranos_amd as ou
from ouranos_amd import problems as P

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--instances", type=int, default=32)
    ap.add_argument("--rhs", type=int, default=4)
    ap.add_argument("--sizes", default="16,24,32,48,64")
    ap.add_argument("--seed", type=int, default=3)
    ap.add_argument("--block", type=int, default=8, help="material block size (cells per axis)")
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tune-reps", type=int, default=3)
    ap.add_argument("--skip-sequential", action="store_true")
    ap.add_argument("--shape", default=None,
                    help="waves,dir_groups: fixed shape instead of the autotuner")
    ap.add_argument("--mode", default="levels", choices=["levels", "tiles"],
                    help="sweep mode of a fixed --shape")
    ap.add_argument("--si-tol", type=float, default=0.0,
                    help="also solve every member by batched source iteration to this change")
    ap.add_argument("--si-max-it", type=int, default=1000)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    sizes = tuple(int(s) for s in a.sizes.split(","))

    specs = P.batch_instances(n_instances=a.instances, n_rhs=a.rhs, sizes=sizes, seed=a.seed,
                              block=a.block)
    cells = sum(s.n_cells for s in specs)
    updates = sum(s.updates_per_sweep() for s in specs)
    alg_bytes = sum(s.algorithmic_bytes_per_sweep() for s in specs)
    log(f"[batch] {len(specs)} members, {cells} cells, {updates / 1e9:.2f} G updates and "
        f"{alg_bytes / 1e9:.1f} GB algorithmic traffic per batched Apply")
    t0 = time.perf_counter()
    ops = [s.operator(ou.FP32) for s in specs]
    log(f"[batch] plans built in {time.perf_counter() - t0:.1f} s")
    xs = [ou.DeviceBuffer(s.size * 8) for s in specs]
    ys = [ou.DeviceBuffer(s.size * 8) for s in specs]
    for i, (b, s) in enumerate(zip(xs, specs)):
        b.upload(P.seeded_vector(s.size, 1000 + i))

    # the members one by one at their own default shape
    seq_ms = None
    if not a.skip_sequential:
        t0 = time.perf_counter()
        seq_ms = 0.0
        for op, x, y in zip(ops, xs, ys):
            t = op.time_apply(x, y, 1, 3)
            seq_ms += t["ms_total"] / t["applications"]
        log(f"[batch] sequential single-plan applies: {seq_ms:.2f} ms "
            f"(measured in {time.perf_counter() - t0:.1f} s)")

    t0 = time.perf_counter()
    fixed = tuple(int(v) for v in a.shape.split(",")) if a.shape else (4, 1)
    # a tile shape is set after the batch exists (its constructor's shape is the
    # chunk sweep's)
    batch = ou.Batch(ops, waves=fixed[0] if a.mode == "levels" else 4,
                     dir_groups=fixed[1] if a.mode == "levels" else 1)
    if a.shape and a.mode == "tiles":
        batch.set_sweep_mode(ou.SWEEP_TILES)
        batch.set_shape(*fixed)
    log(f"[batch] batch built in {time.perf_counter() - t0:.1f} s: {batch.info()}")
    mode_name = {ou.SWEEP_LEVELS: "levels", ou.SWEEP_PERSISTENT: "persistent",
                 ou.SWEEP_TILES: "tiles"}
    table = None
    if not a.shape:
        table = batch.tune(reps=a.tune_reps)
        for e in sorted(table, key=lambda e: e["ms"]):
            log(f"[batch]   {mode_name[e['mode']]:6s} waves={e['waves']:2d} "
                f"dir_groups={e['dir_groups']}  {e['ms']:.2f} ms")
    info = batch.info()
    mode = batch.sweep_mode
    log(f"[batch] {mode_name[mode]}, shape: waves={info['waves']} dir_groups={info['dir_groups']} "
        f"({'fixed' if a.shape else 'autotuned'})")

    graph_ms = batch.time_apply(xs, ys, a.warmup, a.steps)
    log(f"[batch] graph replay: {graph_ms:.2f} ms per batched Apply")
    batch.set_graph(False)
    eager_ms = batch.time_apply(xs, ys, a.warmup, a.steps)
    batch.set_graph(True)
    log(f"[batch] eager batched launches: {eager_ms:.2f} ms")
    batch.set_streams(1)
    one_stream_ms = batch.time_apply(xs, ys, a.warmup, a.steps)
    batch.set_streams(2)
    log(f"[batch] graph replay, one launch per level (no octant-half side stream): "
        f"{one_stream_ms:.2f} ms")
    # the level launches at their best shape, and every member's levels in one
    # persistent launch inside the graph at that shape
    lev = [e for e in (table or []) if e["mode"] == ou.SWEEP_LEVELS]
    lev_shape = min(lev, key=lambda e: e["ms"]) if lev else None
    levels_ms = persistent_ms = None
    if mode != ou.SWEEP_LEVELS and lev_shape:
        batch.set_sweep_mode(ou.SWEEP_LEVELS)
        batch.set_shape(lev_shape["waves"], lev_shape["dir_groups"])
        levels_ms = batch.time_apply(xs, ys, a.warmup, a.steps)
        log(f"[batch] graph replay, merged level launches at ({lev_shape['waves']},"
            f"{lev_shape['dir_groups']}): {levels_ms:.2f} ms")
    batch.set_sweep_mode(ou.SWEEP_PERSISTENT)
    persistent_ms = batch.time_apply(xs, ys, a.warmup, a.steps)
    log(f"[batch] graph replay, one persistent sweep launch: {persistent_ms:.2f} ms")
    batch.set_sweep_mode(mode)
    batch.set_shape(info["waves"], info["dir_groups"])

    # bit-for-bit spot check of two members against their own Apply in the
    # batch's mode and shape
    for op in ops:
        op.sweep_mode = ou.SWEEP_TILES if mode == ou.SWEEP_TILES else ou.SWEEP_LEVELS
    batch.apply_device(xs, ys)
    ou.lib().ou_device_synchronize()
    order = np.argsort([s.n_cells for s in specs])
    exact = True
    for i in (int(order[0]), int(order[-1])):
        y = np.empty(specs[i].size)
        ys[i].download(y)
        x = np.empty(specs[i].size)
        xs[i].download(x)
        exact &= bool(np.array_equal(y, ops[i].Apply(x)))
    log(f"[batch] members equal their single-plan Apply bit-for-bit: {exact}")

    si = None
    if a.si_tol > 0:
        t0 = time.perf_counter()
        _, sinfo = batch.solve_si(tol=a.si_tol, max_it=a.si_max_it, verbose=True)
        si_s = time.perf_counter() - t0
        its = sinfo["iterations"]
        # the same solves one member at a time (ou_solve SI) for every 8th member
        # in size order, scaled by 8 to the whole batch
        order = np.argsort([s.n_cells for s in specs])
        sample = [int(i) for i in order[::8]]
        t0 = time.perf_counter()
        seq_its = []
        for i in sample:
            _, info_i = ops[i].solve(solver=ou.SI, tol=a.si_tol, max_it=a.si_max_it,
                                     raise_on_fail=False)
            seq_its.append(info_i["iterations"])
            log(f"[batch]   member {i} ({specs[i].n[0]}^3) alone: {info_i['iterations']} "
                f"iterations, {info_i['seconds']:.2f} s")
        seq_s = time.perf_counter() - t0
        share = len(sample) / len(specs)
        si = {"tol": a.si_tol, "batched_seconds": si_s, "batched_iterations": sinfo["n_iterations"],
              "member_iterations_min": int(min(its)), "member_iterations_max": int(max(its)),
              "all_converged": all(it > 0 for it in its),
              "sequential_sample_members": len(sample), "sequential_sample_seconds": seq_s,
              "sequential_sample_iterations": seq_its,
              "sequential_projected_seconds": seq_s / share if share > 0 else None}
        log(f"[batch] batched SI to {a.si_tol:g}: {si_s:.2f} s, {sinfo['n_iterations']} "
            f"iterations (members {min(its)}..{max(its)}); one at a time, {len(sample)} "
            f"sampled members: {seq_s:.2f} s -> {si['sequential_projected_seconds']:.1f} s "
            f"projected for the batch")

    gbs = alg_bytes / (graph_ms * 1e-3) / 1e9
    out = {
        "workload": "batch_instances",
        "members": len(specs), "instances": a.instances, "rhs_per_instance": a.rhs,
        "sizes": list(sizes), "block": a.block, "cells": cells, "directions": specs[0].quad.n_dir,
        "precision": "fp32 sweep, fp64 moments",
        "shape": {"waves": info["waves"], "dir_groups": info["dir_groups"]},
        "mode": mode_name[mode],
        "ms_graph_one_stream": one_stream_ms,
        "ms_graph_levels_best": levels_ms,
        "ms_graph_persistent": persistent_ms,
        "levels_per_batch": info["n_levels"], "chunks_per_batch": info["n_chunks"],
        "ms_graph": graph_ms, "ms_eager": eager_ms, "ms_sequential": seq_ms,
        "updates_per_s": updates / (graph_ms * 1e-3),
        "algorithmic_bytes": alg_bytes, "achieved_GBs": gbs, "roofline_frac": gbs / HBM_PEAK_GBS,
        "speedup_vs_sequential": (seq_ms / graph_ms) if seq_ms else None,
        "tune_table": table, "bitwise_equal_spot_check": exact, "source_iteration": si,
        "device": ou.device_name(),
    }
    print(json.dumps(out))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    batch.close()
    for b in xs + ys:
        b.free()
    return 0 if exact else 1


if __name__ == "__main__":
    raise SystemExit(main())
