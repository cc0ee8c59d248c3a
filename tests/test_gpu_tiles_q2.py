"""The 3D Q2 tile sweep in FP32 (OU_SWEEP_TILES for 27-dof hexahedra): 4x4x4
tiles, four directions per wave solved together on the matrix cores, ψ
exchanged through LDS once per tile level.  It is opt-in: Q2 plans start in
level launches, which are faster on the headline workload.

Bar: the FP32 sweep is within the FP32 bound (2e-5 relative) of the FP64
oracle, and within 1e-5 of the FP32 level launches. The two FP32 sweeps round
the same arithmetic in different orders, so they are not compared bit for bit.
Covers:
* one, two and three rounds of twelve directions, including partial rounds;
* 1 and 4 moments;
* partial tiles at the mesh edge;
* every tile shape;
* boundary inflow;
* solvers;
* the full-size headline workload."""
import numpy as np
import pytest

import ouranos_amd as ou
from ouranos_amd import problems as P
from helpers import oracle_problem, rel_err

pytestmark = pytest.mark.gpu

FP32_TOL = 2e-5   # FP32 sweep against the FP64 oracle
PAIR_TOL = 1e-5   # FP32 tile sweep against the FP32 level sweep


def _cases():
    return {
        # GLC 4x8: 4 directions per octant, one round with 8 idle slots
        "q2_4dirs_L1": lambda: P.headline_3d(n=8, block=4, n_polar=4, n_azim=8, L=1),
        # GLC 8x16: 16 per octant, a full round and a round of 4
        "q2_16dirs_L1": lambda: P.headline_3d(n=8, block=4, n_polar=8, n_azim=16, L=1),
        # GLC 12x24 as in the headline: 36 per octant, three full rounds
        "q2_36dirs_L1": lambda: P.headline_3d(n=8, block=8, n_polar=12, n_azim=24, L=1),
        # isotropic scattering: one moment
        "q2_16dirs_L0": lambda: P.headline_3d(n=8, block=4, n_polar=8, n_azim=16, L=0),
        # 10^3 cells in one material: partial tiles (4 + 4 + 2) on every axis
        "q2_partial_tiles": lambda: P.headline_3d(n=10, block=16, n_polar=4, n_azim=8, L=1),
    }


CASES = _cases()


def _tiles(spec):
    """The FP32 operator of spec, switched to the tile sweep."""
    op = spec.operator(ou.FP32)
    assert op.sweep_mode == ou.SWEEP_LEVELS  # the Q2 default
    op.sweep_mode = ou.SWEEP_TILES
    return op


def _levels(op, x, src_x):
    """Apply and the source sweep in level launches at the tile mode's groups."""
    w, g = op.sweep_shape
    op.sweep_mode = ou.SWEEP_LEVELS
    op.sweep_shape = (12, g)
    y = op.Apply(x)
    op.compute_scattering_source(src_x)
    s = op.sweep(None, with_source=True)
    op.sweep_mode = ou.SWEEP_TILES
    op.sweep_shape = (w, g)
    return y, s


@pytest.mark.parametrize("name", sorted(CASES))
def test_q2_tiles_match_oracle_and_levels(oracle, name):
    spec = CASES[name]()
    op = _tiles(spec)
    w, g = op.sweep_shape
    print(f"[gpu] q2 tiles {name} shape=({w},{g})", flush=True)
    pb = oracle_problem(oracle, spec)
    xs = [P.seeded_vector(spec.size, 30 + r) for r in range(2)]
    ys = [op.Apply(x) for x in xs]
    op.compute_scattering_source(xs[1])
    src = op.sweep(None, with_source=True)
    y_lev, src_lev = _levels(op, xs[0], xs[1])
    e_lev = rel_err(ys[0], y_lev)
    e_src = rel_err(src, src_lev)
    print(f"[gpu]   vs FP32 levels: apply {e_lev:.2e}, source sweep {e_src:.2e}", flush=True)
    assert e_lev < PAIR_TOL and e_src < PAIR_TOL
    for x, y in zip(xs, ys):
        ref = pb.apply(x)
        err = rel_err(x - y, x - ref)  # the sweep part of Apply
        print(f"[gpu]   sweep part vs FP64 oracle: {err:.2e}", flush=True)
        assert err < FP32_TOL
    S = pb.scatter_source(xs[1])
    assert rel_err(src, pb.sweep(S, with_source=True)) < FP32_TOL
    assert np.array_equal(op.Apply(xs[1]), ys[1])  # back in tile mode, same bits
    op.close()


def test_q2_every_tile_shape():
    """3 and 4 waves (12 and 16 directions per round) x 1, 2, 3 direction
    groups on 16 directions per octant: each within PAIR_TOL of the level
    launches at the same groups."""
    spec = P.headline_3d(n=8, block=4, n_polar=8, n_azim=16, L=1)
    op = _tiles(spec)
    x = P.seeded_vector(spec.size, 7)
    lev = {}
    op.sweep_mode = ou.SWEEP_LEVELS
    for g in (1, 2, 3):
        op.sweep_shape = (12, g)
        lev[g] = op.Apply(x)
    op.sweep_mode = ou.SWEEP_TILES
    for w in (3, 4):
        for g in (1, 2, 3):
            op.sweep_shape = (w, g)
            err = rel_err(op.Apply(x), lev[g])
            print(f"[gpu] q2 tile shape ({w},{g}): {err:.2e}", flush=True)
            assert err < PAIR_TOL, (w, g)
    for w in (1, 2):  # the matrix-core tile kernel has 3 or 4 waves only
        with pytest.raises(ou.OuError) as e:
            op.sweep_shape = (w, 1)
        assert e.value.code == -4
    op.close()


def test_q2_plans_without_the_tile_sweep():
    """FP64, more than 4 moments, or tiles that mix materials: level launches."""
    for spec, prec in ((P.headline_3d(n=8, block=4, n_polar=4, n_azim=8, L=1), ou.FP64),
                       (P.headline_3d(n=8, block=4, n_polar=4, n_azim=8, L=2), ou.FP32),
                       (P.headline_3d(n=8, block=2, n_polar=4, n_azim=8, L=1), ou.FP32)):
        op = spec.operator(prec)
        assert op.sweep_mode == ou.SWEEP_LEVELS
        with pytest.raises(ou.OuError) as e:
            op.sweep_mode = ou.SWEEP_TILES
        assert e.value.code == -4
        op.close()


@pytest.mark.parametrize("bc", ["isotropic", "most_normal", "mixed"])
def test_q2_tile_source_sweep_with_inflow(oracle, bc):
    """Boundary inflow on the tiles' upwind faces (RadiativeTransfer.cc:957-975):
    the source sweep within the FP32 bound of the oracle."""
    spec = P.headline_3d(n=8, block=4, n_polar=4, n_azim=8, L=1)
    if bc == "isotropic":
        spec.bc, spec.inc = [1] * 6, [0.5 + f for f in range(6)]
    elif bc == "most_normal":
        spec.bc, spec.inc = [2] * 6, [1.0 + f for f in range(6)]
    else:
        spec.bc, spec.inc = [0, 1, 2, 0, 1, 2], [0.0, 2.0, 3.0, 0.0, 1.0, 0.5]
    op = _tiles(spec)
    pb = oracle_problem(oracle, spec)
    x = P.seeded_vector(spec.size, 5)
    op.compute_scattering_source(x)
    S = pb.scatter_source(x)
    err = rel_err(op.sweep(None, with_source=True), pb.sweep(S, with_source=True))
    print(f"[gpu] q2 tiles inflow {bc}: {err:.2e}", flush=True)
    assert err < FP32_TOL
    op.close()


def test_q2_solvers_in_tile_mode():
    """GMRES and SI on the FP32 operator in tile mode reach the level mode's
    solution to the FP32 bound."""
    spec = P.headline_3d(n=8, block=4, n_polar=4, n_azim=8, L=1)
    spec.sigma_s = spec.sigma_s * 0.5  # moderate scattering: SI converges in tens of steps
    op = _tiles(spec)
    for solver in (ou.GMRES, ou.SI):
        op.sweep_mode = ou.SWEEP_TILES
        op.sweep_shape = (3, 1)
        a, ia = op.solve(solver=solver, tol=1e-6, max_it=300)
        op.sweep_mode = ou.SWEEP_LEVELS
        op.sweep_shape = (12, 1)
        b, ib = op.solve(solver=solver, tol=1e-6, max_it=300)
        print(f"[gpu] q2 tiles {solver}: {ia['iterations']} vs {ib['iterations']} iterations",
              flush=True)
        assert ia["converged"] and ib["converged"]
        assert rel_err(a, b) < 1e-4, solver
    op.close()


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


def test_q2_octant_without_directions(oracle):
    """No direction in the (-,-,-) octant: no tiles for it, its partial-moment
    slots are skipped whatever a previous shape left there, and the tile sweep
    stays within the FP32 bound of the oracle across shape changes."""
    spec = P.headline_3d(n=8, block=4, n_polar=8, n_azim=16, L=1)
    q = spec.quad
    keep = ~((q.omega[:, 0] < 0) & (q.omega[:, 1] < 0) & (q.omega[:, 2] < 0))
    spec.quad = _Subset(q, keep)
    x = P.seeded_vector(spec.size, 19)
    ref = oracle_problem(oracle, spec).apply(x)
    op = _tiles(spec)
    for mode, shape in [(ou.SWEEP_TILES, (3, 1)), (ou.SWEEP_TILES, (4, 3)),
                        (ou.SWEEP_LEVELS, (12, 2)), (ou.SWEEP_TILES, (3, 1))]:
        op.sweep_mode = mode
        op.sweep_shape = shape
        err = rel_err(x - op.Apply(x), x - ref)
        print(f"[gpu] q2 octant without directions {mode} {shape}: {err:.2e}", flush=True)
        assert err < FP32_TOL, (mode, shape, err)
    op.close()


def test_q2_batch_tile_mode():
    """A batch of uneven Q2 members in the merged tile levels: within PAIR_TOL
    of the batch's level launches, the same bits over graph replays and as each
    member's own tile-mode Apply; the autotuner times the Q2 tile shapes."""
    specs = [P.headline_3d(n=n, block=4, seed=s, n_polar=4, n_azim=8, L=1)
             for n, s in ((4, 1), (8, 2), (12, 3))]
    ops = [s.operator(ou.FP32) for s in specs]
    batch = ou.Batch(ops, waves=12, dir_groups=1)
    xs = [P.seeded_vector(s.size, 900 + i) for i, s in enumerate(specs)]
    lev = batch.apply(xs)
    batch.set_sweep_mode(ou.SWEEP_TILES)
    batch.set_shape(3, 1)
    first = None
    for rep in range(3):  # eager capture, then graph replays
        ys = batch.apply(xs)
        for i, (y, y_lev) in enumerate(zip(ys, lev)):
            assert rel_err(xs[i] - y, xs[i] - y_lev) < PAIR_TOL, (rep, i)
        if first is None:
            first = ys
        else:
            assert all(np.array_equal(a, b) for a, b in zip(ys, first)), rep
    for op, x, y in zip(ops, xs, first):
        op.sweep_mode = ou.SWEEP_TILES
        assert op.sweep_shape == (3, 1)
        assert np.array_equal(op.Apply(x), y)
    table = batch.tune(reps=1)
    assert {(e["mode"], e["waves"]) for e in table if e["mode"] == ou.SWEEP_TILES} == \
        {(ou.SWEEP_TILES, 3), (ou.SWEEP_TILES, 4)}
    batch.close()


def test_q2_full_headline_size_tiles():
    """128^3 Q2, GLC 12x24 (36 directions per octant, three rounds), L=1 in
    the tile sweep: within PAIR_TOL of the FP32 level launches, and linear."""
    spec = P.headline_3d()
    op = _tiles(spec)
    assert op.sweep_shape == (3, 1)  # three full rounds of 12 directions, one group
    print(f"[gpu] full-size q2 tiles shape={op.sweep_shape}", flush=True)
    x1 = P.seeded_vector(spec.size, 3)
    x2 = P.seeded_vector(spec.size, 4)
    y1 = op.Apply(x1)
    y2 = op.Apply(x2)
    y12 = op.Apply(x1 + 2.0 * x2)
    lin = rel_err(x1 + 2.0 * x2 - y12, (x1 - y1) + 2.0 * (x2 - y2))
    op.sweep_mode = ou.SWEEP_LEVELS
    op.sweep_shape = (12, 1)
    err = rel_err(x1 - y1, x1 - op.Apply(x1))
    print(f"[gpu] full-size q2 tiles vs levels {err:.2e}, linearity {lin:.2e}", flush=True)
    assert err < PAIR_TOL
    assert lin < 1e-5
    op.close()
