"""GPU parity at the benchmark workloads' own angular sets and sizes, through
the C ABI, against the FP64 CPU oracle on the same seeded inputs.

* The production angular sets on small meshes: GLC 12x24 (36 directions per
  octant, Q2, L=1), GLC 16x32 in 2D (64 per quadrant, Q2, L=3) and LS S12
  (21 per octant, Q1), at every workgroup shape the plans and the autotuner
  use.  With one direction group a workgroup takes its octant's directions
  in several rounds of W (9 rounds at 4 waves for GLC 12x24), so the
  prefetch of the next round's traces and inverses runs.
* The 2D 32x32 Q1 LS-S8 problem (c = 0.5, q = 1, vacuum) at its size: one
  Apply, and source iteration to 1e-10.
* The 3D 32^3 Q1 LS-S8 optically thick family at its size.
* The batched workload at its size: 32 instances x 4 right-hand sides,
  n^3 with n in {16, 24, 32, 48, 64}, LS S12.  Every member's batched Apply
  equals its own single-plan Apply bit-for-bit; the 16^3 members match the
  oracle; every member is linear.

Bars: FP64 <= 1e-12 relative, FP32 sweep part <= 2e-5 relative (FP32 angular
flux and local solves, FP64 moment accumulation), printed per case."""
import numpy as np
import pytest

import ouranos_amd as ou
from ouranos_amd import problems as P
from helpers import oracle_problem, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-5


def _production_specs():
    return {
        # GLC 12 polar x 24 azimuthal = 288 directions, L=1, Q2 hexahedra
        "glc12x24_q2_L1": lambda: P.headline_3d(n=6, block=2),
        # GLC 16x32, mu > 0: 256 directions, 64 per quadrant, L=3, Q2 quads
        "glc16x32_2d_q2_L3": lambda: P.hetero_2d(n=64, block=4),
        # LS S12: 168 directions, 21 per octant, Q1 hexahedra
        "ls12_q1": lambda: P.batch_3d(n=6, block=2),
    }


SPECS = _production_specs()
# (waves, direction groups): None = the plan's default
SHAPES = {
    "glc12x24_q2_L1": {ou.FP64: [(4, 1), (8, 1), None], ou.FP32: [(12, 1), (8, 1), (6, 1), None]},
    "glc16x32_2d_q2_L3": {ou.FP64: [(4, 1), None], ou.FP32: [(4, 1), None]},
    "ls12_q1": {ou.FP64: [(2, 1), (4, 1), (8, 1), None], ou.FP32: [(2, 1), (4, 1), (8, 1), None]},
}


@pytest.mark.parametrize("precision", [ou.FP64, ou.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", sorted(SPECS))
def test_production_angular_sets_match_oracle(oracle, name, precision):
    spec = SPECS[name]()
    pb = oracle_problem(oracle, spec)
    xs = [P.seeded_vector(spec.size, s) for s in (11, 12)]
    refs = [pb.apply(x) for x in xs]
    op = spec.operator(precision)
    default = op.sweep_shape
    n_oct = 4 if spec.dim == 2 else 8
    per_oct = spec.quad.n_dir // n_oct
    for shape in SHAPES[name][precision]:
        if shape is not None:
            op.sweep_shape = shape
        else:
            op.sweep_shape = default
        w, g = op.sweep_shape
        rounds = -(-(-(-per_oct // g)) // w)  # ceil(ceil(per_oct / g) / w)
        print(f"[gpu] {name} {'fp64' if precision == ou.FP64 else 'fp32'} shape=({w},{g}) "
              f"{per_oct} directions per octant, {rounds} round(s) per workgroup", flush=True)
        for x, ref in zip(xs, refs):
            y = op.Apply(x)
            if precision == ou.FP64:
                err = rel_err(y, ref)
                assert err < FP64_TOL, (name, op.sweep_shape, err)
            err = rel_err(x - y, x - ref)
            print(f"[gpu]   sweep part relative error vs FP64 oracle: {err:.2e}", flush=True)
            assert err < (FP64_TOL if precision == ou.FP64 else FP32_TOL), (name, op.sweep_shape)
    op.close()


def test_c1_uniform_2d_apply_and_source_iteration(oracle):
    """2D unit square 32x32 Q1, LS-S8 (40 directions), sigma_t = 1, c = 0.5,
    q = 1, vacuum: Apply at 1e-12, and source iteration to a 1e-10 relative
    change lands on the oracle's source-iteration fixed point (ouranos.cc:52-130)."""
    spec = P.uniform_2d()
    assert spec.n == (32, 32) and spec.quad.n_dir == 40
    pb = oracle_problem(oracle, spec)
    op = spec.operator(ou.FP64)
    x = P.seeded_vector(spec.size, 21)
    assert rel_err(op.Apply(x), pb.apply(x)) < FP64_TOL
    phi, info = op.solve(solver=ou.SI, tol=1e-10, max_it=1000)
    assert info["converged"]
    phi_ref, it_ref = pb.si(1e-10, 1000)
    err = rel_err(phi, phi_ref)
    print(f"[gpu] C1 SI: {info['iterations']} iterations (oracle {it_ref}), "
          f"difference {err:.2e}", flush=True)
    assert abs(info["iterations"] - it_ref) <= 1
    assert err < 1e-9
    # the same fixed point by GMRES, and the FP32 sweep's error stated
    phi_g, info_g = op.solve(solver=ou.GMRES, tol=1e-12, max_it=500)
    assert info_g["converged"] and rel_err(phi_g, phi_ref) < 1e-9
    # FP32 sweeps: the change of source iteration stalls at the FP32 rounding
    # level (~1e-7), so the FP32 solve is run to 1e-6
    op32 = spec.operator(ou.FP32)
    err32 = rel_err(x - op32.Apply(x), x - pb.apply(x))
    phi32, info32 = op32.solve(solver=ou.SI, tol=1e-6, max_it=1000)
    err_phi32 = rel_err(phi32, phi_ref)
    print(f"[gpu] C1 FP32: sweep part error {err32:.2e}; SI to 1e-6 in "
          f"{info32['iterations']} iterations, scalar flux error {err_phi32:.2e}", flush=True)
    assert err32 < FP32_TOL and err_phi32 < FP32_TOL


def test_c4_thick_3d_full_size_matches_oracle(oracle):
    """3D 32^3 Q1, LS-S8 (80 directions), 4^3-cell blocks with sigma_t
    log-U[1,1000] and c in [0.99, 0.9999]: one instance of the training
    family at its size."""
    spec = P.thick_3d()
    assert spec.n == (32, 32, 32) and spec.quad.n_dir == 80
    pb = oracle_problem(oracle, spec)
    x = P.seeded_vector(spec.size, 31)
    ref = pb.apply(x)
    assert rel_err(spec.operator(ou.FP64).Apply(x), ref) < FP64_TOL
    err = rel_err(x - spec.operator(ou.FP32).Apply(x), x - ref)
    print(f"[gpu] C4 32^3 FP32 sweep part error vs FP64 oracle: {err:.2e}", flush=True)
    assert err < FP32_TOL


@pytest.mark.parametrize("precision", [ou.FP32, ou.FP64], ids=["fp32", "fp64"])
def test_c3_batched_instances_full_size(oracle, precision):
    """32 instances x 4 right-hand sides (128 members), n^3 Q1 with n drawn
    from {16, 24, 32, 48, 64}, LS S12: the batched Apply of every member
    equals its single-plan Apply bit-for-bit; the 16^3 members match the
    oracle (FP64 at 1e-12; FP32 sweep part at 2e-5); every member's batched
    Apply is linear."""
    specs = P.batch_instances()
    assert len(specs) == 128 and specs[0].quad.n_dir == 168
    sizes = sorted({s.n[0] for s in specs})
    print(f"[gpu] C3 batch: {len(specs)} members, sizes {sizes}, "
          f"{sum(s.n_cells for s in specs)} cells", flush=True)
    ops = [s.operator(precision) for s in specs]
    batch = ou.Batch(ops)
    info = batch.info()
    print(f"[gpu] C3 batch shape waves={info['waves']} groups={info['dir_groups']}, "
          f"{info['n_levels']} merged levels", flush=True)
    xs = [P.seeded_vector(s.size, 500 + i) for i, s in enumerate(specs)]
    ys = batch.apply(xs)
    # The batch sets every member's shape for its mode; a member's own Apply in
    # that mode runs the same arithmetic (the persistent launch equals the
    # level launches).
    mode = batch.sweep_mode
    print(f"[gpu] C3 batch mode {mode}", flush=True)
    for op in ops:
        op.sweep_mode = ou.SWEEP_TILES if mode == ou.SWEEP_TILES else ou.SWEEP_LEVELS
    for i, (op, x, y) in enumerate(zip(ops, xs, ys)):
        assert np.array_equal(y, op.Apply(x)), i
    print("[gpu] C3 batch: every member equals its single-plan Apply bit-for-bit", flush=True)
    # linearity of every member through the batched launch
    zs = [P.seeded_vector(s.size, 900 + i) for i, s in enumerate(specs)]
    yz = batch.apply(zs)
    ylin = batch.apply([2.0 * x - 0.5 * z for x, z in zip(xs, zs)])
    tol_lin = 1e-12 if precision == ou.FP64 else 1e-5
    worst = max(rel_err(yl, 2.0 * y - 0.5 * w) for yl, y, w in zip(ylin, ys, yz))
    print(f"[gpu] C3 batch linearity, worst member: {worst:.2e}", flush=True)
    assert worst < tol_lin
    small = [i for i, s in enumerate(specs) if s.n[0] == 16]
    assert small
    for i in small:
        ref = oracle_problem(oracle, specs[i]).apply(xs[i])
        if precision == ou.FP64:
            assert rel_err(ys[i], ref) < FP64_TOL, i
        else:
            err = rel_err(xs[i] - ys[i], xs[i] - ref)
            print(f"[gpu] C3 member {i} (16^3) FP32 sweep part error: {err:.2e}", flush=True)
            assert err < FP32_TOL
    batch.close()
    for op in ops:
        op.close()
