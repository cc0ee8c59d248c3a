"""The batched multi-instance launch (ou_batch_*) on the GPU.

Bar: every member's batched result equals its own single-plan result
bit-for-bit at the same sweep shape (FP64 and FP32 sweeps: the batched
kernels run the same per-element arithmetic in the same order), matches the
FP64 CPU oracle to 1e-12, and different shapes differ only by FP64 summation
order.  The captured graph is replayed several times with other GPU work
and new vectors in between, and every replay is checked."""
import numpy as np
import pytest

import ouranos_amd as ou
from ouranos_amd import problems as P
from helpers import oracle_problem, rel_err

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12


def _specs(n_instances=3, n_rhs=2, sizes=(3, 4, 5, 6), sn=4, seed=11):
    q = ou.Quadrature(ou.LS, 3, sn=sn, L_max=0)
    return P.batch_instances(n_instances=n_instances, n_rhs=n_rhs, sizes=sizes, block=2,
                             seed=seed, quad=q)


def _uneven_specs():
    """Members of different sizes, aspect ratios and quadratures."""
    specs = _specs()
    q6 = ou.Quadrature(ou.LS, 3, sn=6, L_max=0)
    extra = P.batch_3d(n=4, block=2, seed=21, quad=q6)
    extra.n, extra.h = (7, 3, 5), (1 / 7, 1 / 3, 1 / 5)
    extra.mat = (np.arange(105) % len(extra.sigma_t)).astype(np.int32)
    extra.q = np.linspace(0.0, 1.0, 105)
    return specs + [extra]


def _vectors(specs, seed0):
    return [P.seeded_vector(s.size, seed0 + i) for i, s in enumerate(specs)]


def _upload(vecs):
    bufs = []
    for v in vecs:
        b = ou.DeviceBuffer(v.size * 8)
        b.upload(v)
        bufs.append(b)
    return bufs


def _download(bufs, specs):
    out = []
    for b, s in zip(bufs, specs):
        y = np.empty(s.size)
        b.download(y)
        out.append(y)
    return out


@pytest.mark.parametrize("precision", [ou.FP64, ou.FP32])
@pytest.mark.parametrize("shape", [(4, 1), (2, 3), (8, 2)])
def test_batch_apply_equals_single_bitwise(precision, shape):
    specs = _uneven_specs()
    print(f"[gpu] batch apply vs single, precision={precision} shape={shape} "
          f"members={len(specs)}", flush=True)
    ops = [s.operator(precision) for s in specs]
    batch = ou.Batch(ops, waves=shape[0], dir_groups=shape[1])
    info = batch.info()
    assert (info["waves"], info["dir_groups"]) == shape
    assert info["n_levels"] == max(sum(s.n) - 2 for s in specs)
    for op in ops:
        assert op.sweep_shape == shape
    xs = _vectors(specs, 100)
    ys = batch.apply(xs)
    for i, (op, x, y) in enumerate(zip(ops, xs, ys)):
        y1 = op.Apply(x)
        assert np.array_equal(y, y1), (i, rel_err(y, y1))
    # one launch per merged level instead of two (octant halves on two streams)
    batch.set_streams(1)
    for y, y1 in zip(batch.apply(xs), ys):
        assert np.array_equal(y, y1)
    batch.set_streams(2)
    for y, y1 in zip(batch.apply(xs), ys):
        assert np.array_equal(y, y1)
    with pytest.raises(ou.OuError):
        batch.set_streams(4)
    batch.close()


def test_batch_apply_matches_oracle_and_shapes_agree(oracle):
    specs = _uneven_specs()
    ops = [s.operator(ou.FP64) for s in specs]
    xs = _vectors(specs, 200)
    refs = [oracle_problem(oracle, s).apply(x) for s, x in zip(specs, xs)]
    batch = ou.Batch(ops, waves=4, dir_groups=1)
    first = None
    for shape in [(4, 1), (2, 2), (8, 3), (4, 4)]:
        print(f"[gpu] batch apply vs oracle, shape={shape}", flush=True)
        batch.set_shape(*shape)
        ys = batch.apply(xs)
        for i, (y, ref) in enumerate(zip(ys, refs)):
            assert rel_err(y, ref) < FP64_TOL, (shape, i, rel_err(y, ref))
        if first is None:
            first = ys
        else:  # only the order of the FP64 partial sums differs
            for y, y0 in zip(ys, first):
                assert rel_err(y, y0) < 1e-14
    batch.close()


@pytest.mark.parametrize("bc", ["vacuum", "mixed"])
def test_batch_source_sweep_equals_single(oracle, bc):
    specs = _uneven_specs()
    for s in specs:
        if bc == "mixed":
            s.bc, s.inc = [0, 1, 2, 0, 1, 2], [0.0, 2.0, 3.0, 0.0, 1.0, 0.5]
    print(f"[gpu] batch source sweep, bc={bc}", flush=True)
    ops = [s.operator(ou.FP64) for s in specs]
    batch = ou.Batch(ops, waves=2, dir_groups=2)
    xs = _vectors(specs, 300)
    y0s = _vectors(specs, 400)
    singles = []
    for op, x, y0 in zip(ops, xs, y0s):
        op.compute_scattering_source(x)
        singles.append(op.sweep(y0.copy(), with_source=True))
    bufs = _upload(y0s)
    batch.sweep_device(bufs, with_source=True)
    ou.lib().ou_device_synchronize()
    got = _download(bufs, specs)
    for i, (g, s1, spec, x, y0) in enumerate(zip(got, singles, specs, xs, y0s)):
        assert np.array_equal(g, s1), (i, rel_err(g, s1))
        pb = oracle_problem(oracle, spec)
        ref = y0 + pb.sweep(pb.scatter_source(x), with_source=True)
        assert rel_err(g, ref) < FP64_TOL
    # without the source: the stored scattering sources alone
    for b, y0 in zip(bufs, y0s):
        b.upload(np.zeros_like(y0))
    batch.sweep_device(bufs, with_source=False)
    ou.lib().ou_device_synchronize()
    got = _download(bufs, specs)
    for g, op in zip(got, ops):
        assert np.array_equal(g, op.sweep(None, with_source=False))
    for b in bufs:
        b.free()
    batch.close()


def test_batch_graph_replays_with_work_between():
    """Each replay of the captured graph against the eager launch sequence,
    with single-plan applies, copies and new vectors between replays."""
    specs = _specs(n_instances=4, n_rhs=2, sizes=(4, 6, 8))
    ops = [s.operator(ou.FP32) for s in specs]
    batch = ou.Batch(ops, waves=4, dir_groups=2)
    rounds = [_vectors(specs, 500 + 50 * r) for r in range(4)]
    batch.set_graph(False)
    eager = [batch.apply(xs) for xs in rounds]
    batch.set_graph(True)
    dx = _upload(rounds[0])
    dy = [ou.DeviceBuffer(s.size * 8) for s in specs]
    other = P.seeded_vector(specs[0].size, 999)
    for rep in range(6):
        r = rep % len(rounds)
        print(f"[gpu] graph replay {rep} (input round {r})", flush=True)
        for b, x in zip(dx, rounds[r]):
            b.upload(x)
        batch.apply_device(dx, dy)
        ou.lib().ou_device_synchronize()
        got = _download(dy, specs)
        for i, (g, e) in enumerate(zip(got, eager[r])):
            assert np.array_equal(g, e), (rep, i, rel_err(g, e))
        # eager work on the same plans between replays
        ops[0].Apply(other)
        ops[-1].Apply(P.seeded_vector(specs[-1].size, rep))
    # a second output set: the pointer table, not the graph, carries vectors
    dy2 = [ou.DeviceBuffer(s.size * 8) for s in specs]
    batch.apply_device(dx, dy2)
    ou.lib().ou_device_synchronize()
    for g, e in zip(_download(dy2, specs), eager[(6 - 1) % len(rounds)]):
        assert np.array_equal(g, e)
    for b in dx + dy + dy2:
        b.free()
    batch.close()


def test_batch_autotuner_picks_fastest_and_stays_exact():
    specs = _specs(n_instances=4, n_rhs=2, sizes=(6, 8, 10), sn=6)
    ops = [s.operator(ou.FP32) for s in specs]
    print("[gpu] autotuned batch", flush=True)
    batch = ou.Batch(ops)  # autotuned
    info = batch.info()
    entries = batch.tune(reps=2)
    print("[gpu] tune:", entries, flush=True)
    assert len(entries) >= 4
    assert {(e["waves"], e["dir_groups"]) for e in entries} >= {(4, 1), (2, 1), (8, 1)}
    best = min(entries, key=lambda e: e["ms"])
    info = batch.info()
    assert (info["waves"], info["dir_groups"]) == (best["waves"], best["dir_groups"])
    xs = _vectors(specs, 700)
    ys = batch.apply(xs)
    for op, x, y in zip(ops, xs, ys):
        assert op.sweep_shape == (best["waves"], best["dir_groups"])
        assert np.array_equal(y, op.Apply(x))
    batch.close()


@pytest.mark.parametrize("precision", [ou.FP64, ou.FP32])
def test_batched_source_iteration_equals_single(precision):
    """Each member's batched source-iteration iterates equal its own ou_solve
    SI iterates bit-for-bit (fixed iteration count), and the converged run
    meets the tolerance for every member."""
    specs = _uneven_specs()
    ops = [s.operator(precision) for s in specs]
    batch = ou.Batch(ops, waves=2, dir_groups=1)
    k = 6
    phis, info = batch.solve_si(tol=0.0, max_it=k)
    print(f"[gpu] batched SI, {k} iterations: {info['n_iterations']}", flush=True)
    assert info["n_iterations"] == k
    for i, (op, phi) in enumerate(zip(ops, phis)):
        ref, rinfo = op.solve(solver=ou.SI, tol=1e-300, max_it=k, raise_on_fail=False)
        assert rinfo["iterations"] == k
        assert np.array_equal(phi, ref), (i, rel_err(phi, ref))
    phis, info = batch.solve_si(tol=1e-9, max_it=2000)
    print(f"[gpu] batched SI to 1e-9: {info['n_iterations']} iterations, "
          f"per member {info['iterations']}", flush=True)
    assert all(0 < it <= info["n_iterations"] for it in info["iterations"])
    assert max(info["iterations"]) == info["n_iterations"]
    assert len(set(info["iterations"])) > 1  # members converge at different steps
    # a converged member stops: its result is its own SI after exactly that
    # many iterations, and single-plan SI at the same tolerance stops there too
    # (the two compute the change norm in different summation orders, so a
    # borderline step may differ by one)
    for i, (op, phi) in enumerate(zip(ops, phis)):
        its = info["iterations"][i]
        ref, _ = op.solve(solver=ou.SI, tol=1e-300, max_it=its, raise_on_fail=False)
        assert np.array_equal(phi, ref), (i, its, rel_err(phi, ref))
        _, rinfo = op.solve(solver=ou.SI, tol=1e-9, max_it=2000)
        assert abs(rinfo["iterations"] - its) <= 1, (i, its, rinfo["iterations"])
    # GMRES on the same operator (FP32 operators reach about 1e-7)
    gtol, tol = (1e-11, 1e-6) if precision == ou.FP64 else (1e-6, 1e-5)
    for op, phi in zip(ops, phis):
        ref, _ = op.solve(solver=ou.GMRES, tol=gtol, max_it=2000)
        assert rel_err(phi, ref) < tol
    # every member is active again afterwards: a batched Apply covers all
    xs = _vectors(specs, 11)
    ys = batch.apply(xs)
    for i, (op, x, y) in enumerate(zip(ops, xs, ys)):
        assert np.array_equal(y, op.Apply(x)), i
    batch.close()


def test_batch_errors():
    specs = _specs(n_instances=1, n_rhs=2, sizes=(4,))
    ops = [s.operator(ou.FP64) for s in specs]
    with pytest.raises(ou.OuError):
        ou.Batch([ops[0], ops[0]], waves=4, dir_groups=1)  # a plan twice
    q2 = ou.Quadrature(ou.LS, 2, sn=4, L_max=0)
    two_d = P.uniform_2d(n=4)
    two_d.quad = q2
    op2 = two_d.operator(ou.FP64)
    with pytest.raises(ou.OuError):
        ou.Batch([ops[0], op2], waves=4, dir_groups=1)  # mixed dimensions
    with pytest.raises(ou.OuError):
        ou.Batch(ops, waves=3, dir_groups=1)  # no kernel with 3 waves
    with pytest.raises(ou.OuError):
        ou.Batch(ops, waves=4, dir_groups=9)
    batch = ou.Batch(ops, waves=4, dir_groups=1)
    dx = _upload(_vectors(specs, 1))
    with pytest.raises(ou.OuError):
        batch.apply_device(dx, dx)  # y aliases x
    dy = [ou.DeviceBuffer(specs[0].size * 8)] * 2
    with pytest.raises(ou.OuError):
        batch.apply_device(dx, dy)  # two members share an output
    for b in dx + dy[:1]:
        b.free()
    batch.close()
