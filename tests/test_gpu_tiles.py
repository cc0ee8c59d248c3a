"""The tile sweep (OU_SWEEP_TILES): one launch per tile level of 8x8-cell
(2D) or 4x4x4-cell (3D Q1) tiles, the traces of a tile's 15 (10) cell levels
kept in registers.

Bar: the per-cell arithmetic and the D2M accumulation order are the level
sweep's, so at equal direction groups Apply and the source sweep equal the
level launches bit-for-bit (FP64 and FP32); FP64 matches the oracle to 1e-12.
Covers: partial tiles at the mesh edge, one tile, a quadrant without
directions, every tile shape, the C2 workload at full size, and solvers."""
import numpy as np
import pytest

import ouranos_amd as ou
from ouranos_amd import problems as P
from helpers import oracle_problem, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
PRECISIONS = [(ou.FP64, "fp64"), (ou.FP32, "fp32")]


class _Subset:
    """Some directions of a quadrature (the operator takes any omega / M2D / D2M)."""

    def __init__(self, q, keep):
        self.omega = np.ascontiguousarray(q.omega[keep])
        self.M2D = np.ascontiguousarray(q.M2D[keep])
        self.D2M = np.ascontiguousarray(q.D2M[:, keep])
        self.mom_l = q.mom_l
        self.most_normal = np.ascontiguousarray(q.most_normal[:, keep])
        self.weight_sum = q.weight_sum
        self.n_dir = int(self.omega.shape[0])
        self.n_mom = q.n_mom


def _cases():
    return {
        # 32^2 Q1 LS-S8 (10 directions per quadrant), 4 x 4 tiles
        "c1_uniform_2d_q1": lambda: P.uniform_2d(),
        # partial tiles on both axes (13 x 13 cells: 8 + 5)
        "uniform_2d_13": lambda: P.uniform_2d(n=13),
        # one tile only: the reference problem, 4 x 4 Q1 GLC-S4
        "reference_2d": lambda: P.reference_2d(),
        "hetero_2d_q2_L3": lambda: P.hetero_2d(n=32, block=8, n_polar=4, n_azim=8, L=3),
        "hetero_2d_q2_L1_64dirs": lambda: P.hetero_2d(n=24, block=8, n_polar=16, n_azim=32, L=1),
        "hetero_2d_q1_L2": lambda: P.hetero_2d(n=16, block=8, n_polar=4, n_azim=8, L=2, p=1),
        # 3D Q1: LS-S8 (10 per octant) on 4^3 material blocks, LS-S12 (21 per
        # octant) on 8^3 blocks, one material with partial tiles (4 + 4 + 2)
        "thick_3d_q1_s8": lambda: P.thick_3d(n=8, block=4),
        "batch_3d_q1_s12": lambda: P.batch_3d(n=12, block=4),
        "batch_3d_10_one_block": lambda: P.batch_3d(n=10, block=16),
        "headline_3d_q1_L1": lambda: P.headline_3d(n=8, block=4, n_polar=4, n_azim=8, L=1, p=1),
        # 16 moments (L=3) in 3D
        "headline_3d_q1_L3": lambda: P.headline_3d(n=8, block=4, n_polar=4, n_azim=8, L=3, p=1),
    }


CASES = _cases()


def _same_groups_levels(op, x, src_x):
    """Apply and the source sweep in level launches at the tile mode's groups."""
    w, g = op.sweep_shape
    op.sweep_mode = ou.SWEEP_LEVELS
    op.sweep_shape = (4, g)
    y = op.Apply(x)
    op.compute_scattering_source(src_x)
    s = op.sweep(None, with_source=True)
    op.sweep_mode = ou.SWEEP_TILES
    op.sweep_shape = (w, g)
    return y, s


@pytest.mark.parametrize("precision,pname", PRECISIONS, ids=["fp64", "fp32"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_tiles_equal_levels_and_oracle(oracle, name, precision, pname):
    spec = CASES[name]()
    op = spec.operator(precision)
    assert op.sweep_mode == ou.SWEEP_TILES  # the default where a plan has tiles
    w, g = op.sweep_shape
    print(f"[gpu] tiles {name} {pname} shape=({w},{g})", flush=True)
    xs = [P.seeded_vector(spec.size, 80 + r) for r in range(2)]
    ys = [op.Apply(x) for x in xs]
    op.compute_scattering_source(xs[1])
    src = op.sweep(None, with_source=True)
    y_lev, src_lev = _same_groups_levels(op, xs[0], xs[1])
    assert np.array_equal(ys[0], y_lev)
    assert np.array_equal(src, src_lev)
    assert np.array_equal(op.Apply(xs[1]), ys[1])  # back in tile mode
    if precision == ou.FP64:
        pb = oracle_problem(oracle, spec)
        for x, y in zip(xs, ys):
            assert rel_err(y, pb.apply(x)) < FP64_TOL
    op.close()


@pytest.mark.parametrize("precision,pname", PRECISIONS, ids=["fp64", "fp32"])
def test_every_tile_shape_equals_levels(precision, pname):
    """Waves 1, 2, 4 (8 directions each) x direction groups 1..8, several
    rounds per workgroup where the waves hold fewer than the group's
    directions: each equals the level launches at the same groups."""
    spec = P.hetero_2d(n=16, block=8, n_polar=8, n_azim=16, L=3)  # 16 per quadrant
    op = spec.operator(precision)
    x = P.seeded_vector(spec.size, 91)
    lev = {}
    op.sweep_mode = ou.SWEEP_LEVELS
    for g in (1, 2, 3, 8):
        op.sweep_shape = (4, g)
        lev[g] = op.Apply(x)
    op.sweep_mode = ou.SWEEP_TILES
    for w in (1, 2, 4):
        for g in (1, 2, 3, 8):
            try:
                op.sweep_shape = (w, g)
            except ou.OuError as e:  # this many waves does not fit the LDS
                assert e.code == -4 and w == 4, (w, g, e)
                continue
            print(f"[gpu] tile shape ({w},{g}) {pname}", flush=True)
            assert np.array_equal(op.Apply(x), lev[g]), (w, g)
    op.close()


@pytest.mark.parametrize("precision,pname", PRECISIONS, ids=["fp64", "fp32"])
def test_every_3d_tile_shape_equals_levels(precision, pname):
    """3D Q1 tiles at waves 1, 2, 4 (4 directions each) x groups 1, 2, 5:
    several rounds where the waves hold fewer than the group's directions."""
    spec = P.batch_3d(n=8, block=4)  # LS-S12: 21 directions per octant
    op = spec.operator(precision)
    assert op.sweep_mode == ou.SWEEP_TILES
    x = P.seeded_vector(spec.size, 93)
    lev = {}
    op.sweep_mode = ou.SWEEP_LEVELS
    for g in (1, 2, 5):
        op.sweep_shape = (4, g)
        lev[g] = op.Apply(x)
    op.sweep_mode = ou.SWEEP_TILES
    for w in (1, 2, 4):
        for g in (1, 2, 5):
            print(f"[gpu] 3D tile shape ({w},{g}) {pname}", flush=True)
            op.sweep_shape = (w, g)
            assert np.array_equal(op.Apply(x), lev[g]), (w, g)
    op.close()


def test_c4_full_size_tiles(oracle):
    """C4 (32^3 Q1, LS-S8, 4^3 blocks) in the 3D tile sweep: FP64 at 1e-12
    against the oracle, and the level launches' bits at the same groups."""
    spec = P.thick_3d()
    x = P.seeded_vector(spec.size, 41)
    ref = oracle_problem(oracle, spec).apply(x)
    for precision in (ou.FP64, ou.FP32):
        op = spec.operator(precision)
        assert op.sweep_mode == ou.SWEEP_TILES
        print(f"[gpu] C4 tiles shape={op.sweep_shape}", flush=True)
        y = op.Apply(x)
        tol = FP64_TOL if precision == ou.FP64 else 2e-5
        assert rel_err(x - y, x - ref) < tol
        y_lev, _ = _same_groups_levels(op, x, x)
        assert np.array_equal(y, y_lev)
        op.close()


def test_quadrant_without_directions(oracle):
    """No direction in the (-,-) quadrant: no tiles for it, and its
    partial-moment slots are skipped, whatever a previous layout left there."""
    q = ou.Quadrature(ou.PRODUCT_GLC, 2, n_polar=4, n_azim=8, L_max=1)
    keep = ~((q.omega[:, 0] < 0) & (q.omega[:, 1] < 0))
    spec = P.hetero_2d(n=16, block=8, n_polar=4, n_azim=8, L=1)
    spec.quad = _Subset(q, keep)
    x = P.seeded_vector(spec.size, 17)
    ref = oracle_problem(oracle, spec).apply(x)
    for precision in (ou.FP64, ou.FP32):
        op = spec.operator(precision)
        assert op.sweep_mode == ou.SWEEP_TILES
        for mode, shape in [(ou.SWEEP_LEVELS, (4, 3)), (ou.SWEEP_TILES, (1, 1)),
                            (ou.SWEEP_LEVELS, (4, 1)), (ou.SWEEP_TILES, (2, 2))]:
            op.sweep_mode = mode
            op.sweep_shape = shape
            err = rel_err(op.Apply(x), ref)
            assert err < (FP64_TOL if precision == ou.FP64 else 2e-5), (precision, mode, shape, err)
        op.close()


def test_plans_without_tiles_refuse_the_mode():
    """3D Q2 in FP64, 2D Q3+ and tiles that mix materials have no tile sweep:
    they start in level launches and refuse OU_SWEEP_TILES."""
    for spec in (P.hetero_2d(n=16, block=4, n_polar=4, n_azim=8, L=1),   # 4x4 material blocks
                 P.hetero_2d(n=8, block=8, n_polar=4, n_azim=8, L=1, p=3),
                 P.batch_3d(n=4, block=2, quad=ou.Quadrature(ou.LS, 3, sn=4, L_max=0)),
                 P.headline_3d(n=4, block=4, n_polar=4, n_azim=8, L=1)):  # 3D Q2, FP64
        op = spec.operator(ou.FP64)
        assert op.sweep_mode == ou.SWEEP_LEVELS
        with pytest.raises(ou.OuError) as e:
            op.sweep_mode = ou.SWEEP_TILES
        assert e.value.code == -4
        op.close()


@pytest.mark.parametrize("precision,pname", PRECISIONS, ids=["fp64", "fp32"])
def test_c2_full_size_tiles_equal_levels(precision, pname):
    """C2 (256^2 Q2, GLC 16x32 = 64 directions per quadrant, L=3): the tile
    sweep at its default shape equals the level launches at the same groups,
    and the operator is linear."""
    spec = P.hetero_2d()
    op = spec.operator(precision)
    w, g = op.sweep_shape
    print(f"[gpu] C2 tiles {pname} shape=({w},{g})", flush=True)
    x1 = P.seeded_vector(spec.size, 3)
    x2 = P.seeded_vector(spec.size, 4)
    y1 = op.Apply(x1)
    y_lev, _ = _same_groups_levels(op, x1, x2)
    assert np.array_equal(y1, y_lev)
    y12 = op.Apply(x1 + 2.0 * x2)
    y2 = op.Apply(x2)
    tol = 1e-12 if precision == ou.FP64 else 1e-5
    assert rel_err(y12, y1 + 2.0 * y2) < tol
    op.close()


def test_solvers_in_tile_mode_retrace_level_mode():
    """GMRES and SI in tile mode take the level mode's iterations and reach the
    same bits (same direction groups)."""
    spec = P.hetero_2d(n=32, block=8, n_polar=4, n_azim=8, L=1)
    op = spec.operator(ou.FP64)
    w, g = op.sweep_shape
    for solver in (ou.GMRES, ou.SI):
        op.sweep_mode = ou.SWEEP_TILES
        op.sweep_shape = (w, g)
        a, ia = op.solve(solver=solver, tol=1e-10, max_it=300)
        op.sweep_mode = ou.SWEEP_LEVELS
        op.sweep_shape = (4, g)
        b, ib = op.solve(solver=solver, tol=1e-10, max_it=300)
        assert ia["iterations"] == ib["iterations"] and np.array_equal(a, b), solver
    op.close()


@pytest.mark.parametrize("precision,pname", PRECISIONS, ids=["fp64", "fp32"])
def test_batch_tile_mode(precision, pname):
    """A batch of uneven 3D Q1 members (4^3 material blocks) in the merged tile
    levels: every member equals its own tile-mode Apply and the batch's level
    launches at the same groups bit-for-bit, over graph replays; batched
    source iteration retraces the level mode; the autotuner times tile
    shapes too."""
    specs = P.batch_instances(n_instances=3, n_rhs=2, sizes=(4, 8, 12), block=4, seed=5)
    ops = [s.operator(precision) for s in specs]
    assert all(op.sweep_mode == ou.SWEEP_TILES for op in ops)
    batch = ou.Batch(ops, waves=4, dir_groups=2)
    assert batch.sweep_mode == ou.SWEEP_LEVELS
    xs = [P.seeded_vector(s.size, 700 + i) for i, s in enumerate(specs)]
    lev = batch.apply(xs)
    batch.set_sweep_mode(ou.SWEEP_TILES)
    for w in (4, 2):
        batch.set_shape(w, 2)
        assert batch.info()["waves"] == w and batch.info()["dir_groups"] == 2
        print(f"[gpu] batch tiles shape ({w},2) {pname}", flush=True)
        for rep in range(2):  # eager capture, then a graph replay
            ys = batch.apply(xs)
            for i, (y, y_lev) in enumerate(zip(ys, lev)):
                assert np.array_equal(y, y_lev), (w, rep, i)
    for op, x, y in zip(ops, xs, lev):
        assert op.sweep_shape[1] == 2
        assert np.array_equal(op.Apply(x), y)
    phis, info = batch.solve_si(tol=1e-8, max_it=200)
    batch.set_sweep_mode(ou.SWEEP_LEVELS)
    batch.set_shape(4, 2)
    phis_lev, info_lev = batch.solve_si(tol=1e-8, max_it=200)
    assert info["iterations"] == info_lev["iterations"]
    for a, b in zip(phis, phis_lev):
        assert np.array_equal(a, b)
    table = batch.tune(reps=1)
    assert {e["mode"] for e in table} == {ou.SWEEP_LEVELS, ou.SWEEP_TILES}
    best = min(table, key=lambda e: e["ms"])
    assert batch.sweep_mode == best["mode"]
    assert (batch.info()["waves"], batch.info()["dir_groups"]) == (best["waves"], best["dir_groups"])
    batch.close()



@pytest.mark.parametrize("bc", ["isotropic", "most_normal", "mixed"])
def test_tile_source_sweep_with_inflow_matches_oracle(oracle, bc):
    """The right-hand-side sweep in the tile sweep, 2D and 3D: volumetric
    source and boundary inflow on the tiles' upwind faces
    (RadiativeTransfer.cc:914-925, 957-975), FP64 at 1e-12."""
    for spec in (P.hetero_2d(n=16, block=8, n_polar=4, n_azim=8, L=1),
                 P.thick_3d(n=8, block=4, quad=ou.Quadrature(ou.LS, 3, sn=4, L_max=0))):
        nf = 2 * spec.dim
        if bc == "isotropic":
            spec.bc, spec.inc = [1] * nf, [0.5 + f for f in range(nf)]
        elif bc == "most_normal":
            spec.bc, spec.inc = [2] * nf, [1.0 + f for f in range(nf)]
        else:
            spec.bc, spec.inc = [0, 1, 2, 0, 1, 2][:nf], [0.0, 2.0, 3.0, 0.0, 1.0, 0.5][:nf]
        op = spec.operator(ou.FP64)
        assert op.sweep_mode == ou.SWEEP_TILES
        pb = oracle_problem(oracle, spec)
        x = P.seeded_vector(spec.size, 5)
        op.compute_scattering_source(x)
        y0 = P.seeded_vector(spec.size, 6)
        y = op.sweep(y0.copy(), with_source=True)
        S = pb.scatter_source(x)
        assert rel_err(y, y0 + pb.sweep(S, with_source=True)) < FP64_TOL, (spec.dim, bc)
        assert rel_err(op.sweep(None, with_source=False), pb.sweep(S)) < FP64_TOL
        op.close()
