"""The persistent sweep (ou_set_sweep_mode OU_SWEEP_PERSISTENT): one launch
per sweep whose workgroups take (chunk, direction group) items in level order
and wait for the upwind level of their own octant and group.

Bar: it runs the same per-chunk arithmetic as the level launches, so every
result equals the level-launch result bit-for-bit (FP64 and FP32, Apply, the
source sweep, solvers, batches); FP64 matches the oracle to 1e-12.  The
watchdog: with a zero wait limit the sweep gives up, every kernel exits, the
call returns OU_ERR_TIMEOUT, and the plan works again afterwards."""
import numpy as np
import pytest

import ouranos_amd as ou
from ouranos_amd import problems as P
from helpers import oracle_problem, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _cases():
    q4 = ou.Quadrature(ou.LS, 3, sn=4, L_max=0)
    return {
        "uniform_2d_q1": lambda: P.uniform_2d(),
        "hetero_2d_q2_L3": lambda: P.hetero_2d(n=32, block=4, L=3),
        "hetero_2d_q3_L1": lambda: P.hetero_2d(n=8, block=4, n_polar=4, n_azim=8, L=1, p=3),
        "batch_3d_q1_s12": lambda: P.batch_3d(n=8, block=2),
        "thick_3d_q1": lambda: P.thick_3d(n=8, block=2, quad=q4),
        "headline_3d_q2_glc12x24": lambda: P.headline_3d(n=6, block=2),
        "headline_3d_q1_L3": lambda: P.headline_3d(n=5, block=2, n_polar=4, n_azim=8, L=3, p=1),
    }


CASES = _cases()


@pytest.mark.parametrize("precision", [ou.FP64, ou.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_persistent_equals_levels_and_oracle(oracle, name, precision):
    spec = CASES[name]()
    op = spec.operator(precision)
    xs = [P.seeded_vector(spec.size, 70 + r) for r in range(2)]
    op.sweep_mode = ou.SWEEP_LEVELS
    ref = [op.Apply(x) for x in xs]
    op.compute_scattering_source(xs[0])
    ref_src = op.sweep(None, with_source=True)
    op.sweep_mode = ou.SWEEP_PERSISTENT
    assert op.sweep_mode == ou.SWEEP_PERSISTENT
    w, g = op.sweep_shape
    print(f"[gpu] persistent {name} {'fp64' if precision == ou.FP64 else 'fp32'} "
          f"shape=({w},{g})", flush=True)
    for _ in range(2):  # counters are cleared before every launch
        for x, y0 in zip(xs, ref):
            assert np.array_equal(op.Apply(x), y0)
    op.compute_scattering_source(xs[0])
    assert np.array_equal(op.sweep(None, with_source=True), ref_src)
    # other shapes: still the level launches' results at that shape
    for shape in [(w, 1), (w, min(4, 8))]:
        op.sweep_shape = shape
        op.sweep_mode = ou.SWEEP_LEVELS
        y_lev = op.Apply(xs[0])
        op.sweep_mode = ou.SWEEP_PERSISTENT
        assert np.array_equal(op.Apply(xs[0]), y_lev), shape
    if precision == ou.FP64:
        assert rel_err(op.Apply(xs[1]), oracle_problem(oracle, spec).apply(xs[1])) < FP64_TOL
    op.close()


def test_persistent_solvers_equal_level_solvers():
    """GMRES, BiCGStab and SI in persistent mode retrace the level-launch
    iterations exactly."""
    spec = P.hetero_2d(n=16, block=4, n_polar=4, n_azim=8, L=1)
    op = spec.operator(ou.FP64)
    for solver in (ou.GMRES, ou.BICGSTAB, ou.SI):
        op.sweep_mode = ou.SWEEP_LEVELS
        a, ia = op.solve(solver=solver, tol=1e-10, max_it=300)
        op.sweep_mode = ou.SWEEP_PERSISTENT
        b, ib = op.solve(solver=solver, tol=1e-10, max_it=300)
        assert ia["iterations"] == ib["iterations"] and np.array_equal(a, b), solver
    op.close()


@pytest.mark.parametrize("precision", [ou.FP64, ou.FP32], ids=["fp64", "fp32"])
def test_batch_persistent_launch_equals_levels(precision):
    """Every member's levels in one persistent launch inside the batch's
    captured graph: equal to the merged level launches and to each member's
    own Apply, over several replays with new inputs."""
    q = ou.Quadrature(ou.LS, 3, sn=6, L_max=0)
    specs = P.batch_instances(n_instances=4, n_rhs=2, sizes=(3, 5, 8), block=2, seed=12,
                              quad=q)
    ops = [s.operator(precision) for s in specs]
    batch = ou.Batch(ops, waves=4, dir_groups=2)
    for rep in range(3):
        xs = [P.seeded_vector(s.size, 300 + 10 * rep + i) for i, s in enumerate(specs)]
        batch.set_sweep_mode(ou.SWEEP_LEVELS)
        ys_lev = batch.apply(xs)
        batch.set_sweep_mode(ou.SWEEP_PERSISTENT)
        ys = batch.apply(xs)
        ys2 = batch.apply(xs)  # graph replay
        for i, (a, b, c) in enumerate(zip(ys, ys2, ys_lev)):
            assert np.array_equal(a, c) and np.array_equal(b, c), (rep, i)
        for op, x, y in zip(ops, xs, ys):
            assert np.array_equal(op.Apply(x), y)
    batch.set_sweep_mode(ou.SWEEP_PERSISTENT)
    phis, info = batch.solve_si(tol=1e-8, max_it=200)
    batch.set_sweep_mode(ou.SWEEP_LEVELS)
    phis_lev, info_lev = batch.solve_si(tol=1e-8, max_it=200)
    assert info["iterations"] == info_lev["iterations"]
    for a, b in zip(phis, phis_lev):
        assert np.array_equal(a, b)
    batch.close()


def test_watchdog_gives_up_and_reports_timeout():
    """A zero wait limit: the first workgroup that finds its upwind level
    unfinished gives up, sets the error word, and every workgroup leaves at
    its next ticket or poll.  The call returns OU_ERR_TIMEOUT; the output is
    NaN, never a silent wrong result; the plan then works again."""
    spec = P.hetero_2d(n=64, block=4, L=3)  # 127 levels, 64 directions per quadrant
    op = spec.operator(ou.FP32)
    op.sweep_shape = (4, 1)  # each level-0 item takes 16 rounds of directions
    x = P.seeded_vector(spec.size, 5)
    y_ref = op.Apply(x)
    op.sweep_mode = ou.SWEEP_PERSISTENT
    op.set_sweep_wait_limit(0.0)
    with pytest.raises(ou.OuError) as e:
        op.Apply(x)
    assert e.value.code == -8, e.value
    assert b"gave up" in ou.lib().ou_last_error()
    # a solve reports the watchdog, not the NaN it left behind
    with pytest.raises(ou.OuError) as e:
        op.solve(solver=ou.GMRES, tol=1e-8, max_it=50)
    assert e.value.code == -8
    op.sweep_status()  # cleared by the report
    op.set_sweep_wait_limit(10.0)
    assert np.array_equal(op.Apply(x), y_ref)
    op.sweep_status()
    op.close()
