"""GPU parity: the HIP operator against the FP64 CPU oracle on the same
seeded inputs, through the C ABI.  Bar: FP64 Apply / sweep agree to 1e-12
relative; FP32 sweeps report their error against the FP64 oracle and stay
within the stated bound."""
import numpy as np
import pytest

import ouranos_amd as ou
from ouranos_amd import problems as P
from helpers import load_golden, oracle_problem, rel_err, to_reference_numbering

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
FP32_TOL = 2e-5  # FP32 angular flux, FP64 moment accumulation; one operator application


def small_cases():
    q3 = ou.Quadrature(ou.LS, 3, sn=4, L_max=0)
    return {
        "reference_2d": lambda: P.reference_2d(),
        "uniform_2d_q1": lambda: P.uniform_2d(n=8),
        "hetero_2d_q2_L3": lambda: P.hetero_2d(n=16, block=4, n_polar=4, n_azim=8, L=3),
        "hetero_2d_q3_L1": lambda: P.hetero_2d(n=8, block=4, n_polar=4, n_azim=8, L=1, p=3),
        "hetero_2d_q1_L2": lambda: P.hetero_2d(n=12, block=3, n_polar=4, n_azim=8, L=2, p=1),
        "batch_3d_q1": lambda: P.batch_3d(n=6, block=2, quad=q3),
        "thick_3d_q1": lambda: P.thick_3d(n=6, block=2, quad=q3),
        "headline_3d_q2_L1": lambda: P.headline_3d(n=5, block=2, n_polar=4, n_azim=8, L=1),
        "headline_3d_q2_L2": lambda: P.headline_3d(n=4, block=2, n_polar=4, n_azim=8, L=2),
        "headline_3d_q1_L3": lambda: P.headline_3d(n=5, block=2, n_polar=4, n_azim=8, L=3, p=1),
        # the reference's higher element orders (RadiativeTransfer.cc:1108-1117)
        "hetero_2d_q4_L3": lambda: P.hetero_2d(n=6, block=3, n_polar=4, n_azim=8, L=3, p=4),
        "hetero_2d_q5_L2": lambda: P.hetero_2d(n=5, block=2, n_polar=4, n_azim=8, L=2, p=5),
        "headline_3d_q3_L1": lambda: P.headline_3d(n=4, block=2, n_polar=4, n_azim=8, L=1, p=3),
        # 128 directions per octant: two direction groups of 64 at least
        "batch_3d_q1_glc16x64": lambda: P.batch_3d(
            n=4, block=2, quad=ou.Quadrature(ou.PRODUCT_GLC, 3, n_polar=16, n_azim=64, L_max=0)),
    }


CASES = small_cases()


def _non_cubic(spec):
    return spec


@pytest.mark.parametrize("name", sorted(CASES))
def test_apply_fp64_matches_oracle(oracle, name):
    spec = CASES[name]()
    print(f"[gpu] apply parity {name} n={spec.n} p={spec.p} dirs={spec.quad.n_dir} "
          f"moms={spec.quad.n_mom}", flush=True)
    op = spec.operator(ou.FP64)
    pb = oracle_problem(oracle, spec)
    for seed in (1, 2):
        x = P.seeded_vector(spec.size, seed)
        y = op.Apply(x)
        y_ref = pb.apply(x)
        assert rel_err(y, y_ref) < FP64_TOL, (name, seed, rel_err(y, y_ref))
        # the sweep part alone (x - y), which carries all the transport work
        assert rel_err(x - y, x - y_ref) < FP64_TOL


@pytest.mark.parametrize("name", sorted(CASES))
def test_apply_fp32_error_vs_fp64_oracle(oracle, name):
    spec = CASES[name]()
    op = spec.operator(ou.FP32)
    pb = oracle_problem(oracle, spec)
    x = P.seeded_vector(spec.size, 3)
    sw = x - op.Apply(x)
    sw_ref = x - pb.apply(x)
    err = rel_err(sw, sw_ref)
    print(f"[gpu] FP32 sweep relative error vs FP64 oracle, {name}: {err:.2e}", flush=True)
    assert err < FP32_TOL


def test_non_cubic_meshes_and_unit_aspect(oracle):
    q = ou.Quadrature(ou.LS, 3, sn=6, L_max=1)
    spec = P.batch_3d(n=4, block=2, quad=q)
    spec.n = (5, 3, 4)
    spec.h = (0.2, 0.5, 0.3)
    spec.mat = (np.arange(60) % 3).astype(np.int32)
    spec.sigma_t = np.array([0.5, 4.0, 40.0])
    spec.sigma_s = np.array([[0.2, 0.05], [3.0, 0.4], [39.0, 1.0]])
    spec.q = np.linspace(0, 1, 60)
    op = spec.operator(ou.FP64)
    pb = oracle_problem(oracle, spec)
    x = P.seeded_vector(spec.size, 9)
    assert rel_err(op.Apply(x), pb.apply(x)) < FP64_TOL


@pytest.mark.parametrize("bc", ["isotropic", "most_normal", "mixed"])
def test_source_sweep_with_inflow_matches_oracle(oracle, bc):
    """The right-hand-side sweep: volumetric source and boundary inflow
    (RadiativeTransfer.cc:914-925, 957-975)."""
    for spec in (P.hetero_2d(n=8, block=4, n_polar=4, n_azim=8, L=1),
                 P.headline_3d(n=4, block=2, n_polar=4, n_azim=8, L=1)):
        nf = 2 * spec.dim
        if bc == "isotropic":
            spec.bc, spec.inc = [1] * nf, [0.5 + f for f in range(nf)]
        elif bc == "most_normal":
            spec.bc, spec.inc = [2] * nf, [1.0 + f for f in range(nf)]
        else:
            spec.bc, spec.inc = [0, 1, 2, 0, 1, 2][:nf], [0.0, 2.0, 3.0, 0.0, 1.0, 0.5][:nf]
        op = spec.operator(ou.FP64)
        pb = oracle_problem(oracle, spec)
        x = P.seeded_vector(spec.size, 5)
        op.compute_scattering_source(x)
        y0 = P.seeded_vector(spec.size, 6)
        y = op.sweep(y0.copy(), with_source=True)
        S = pb.scatter_source(x)
        y_ref = y0 + pb.sweep(S, with_source=True)
        assert rel_err(y, y_ref) < FP64_TOL, bc
        # without the source flag the stored scattering source alone is swept
        y = op.sweep(None, with_source=False)
        assert rel_err(y, pb.sweep(S)) < FP64_TOL


def test_reference_golden_through_gpu_solvers():
    """tests/RadiativeTransfer_test.cc:84-152 solved with the GPU operator."""
    gold = np.array(load_golden("reference_2d_solution.json")["moment0"])
    spec = P.reference_2d()
    op = spec.operator(ou.FP64)
    for solver in (ou.BICGSTAB, ou.GMRES, ou.SI):
        phi, info = op.solve(solver=solver, tol=1e-10, max_it=500)
        assert info["converged"], (solver, info)
        mine = to_reference_numbering(phi, 4, 4)
        assert np.abs(mine - gold).max() < 1e-3
        assert np.abs(mine - gold).max() < 1e-5


def test_gmres_matches_oracle_solution(oracle):
    spec = P.hetero_2d(n=8, block=4, n_polar=4, n_azim=8, L=3)
    op = spec.operator(ou.FP64)
    phi, info = op.solve(solver=ou.GMRES, tol=1e-11, max_it=400)
    assert info["converged"]
    pb = oracle_problem(oracle, spec)
    x, it, rr = pb.gmres(pb.rhs(), 1e-11, max_it=400)
    assert rel_err(phi, x) < 1e-9


def test_operator_surface_and_errors():
    spec = P.reference_2d()
    op = spec.operator(ou.FP64)
    assert op.Label() == "radiative_transfer"
    assert op.domain_size() == op.range_size() == 64
    assert not op.UseTranspose() and not op.HasNormInf() and op.NormInf() == 0.0
    with pytest.raises(ou.OuError) as e:
        op.SetUseTranspose(True)
    assert e.value.code == -4
    with pytest.raises(ou.OuError) as e:
        op.ApplyInverse(np.zeros(64))
    assert e.value.code == -4
    with pytest.raises(ou.OuError) as e:
        op.Apply(np.zeros(63))
    assert e.value.code == -2
    x = np.zeros(64)
    x[5] = np.nan
    with pytest.raises(ou.OuError) as e:
        op.Apply(x)
    assert e.value.code == -5
    with pytest.raises(ou.OuError) as e:
        op.set_group(1)
    assert e.value.code == -4
    op.set_group(0)
    # 3D Q4 and 3D Q3 beyond L = 1 have no GPU sweep kernel: reported, not
    # silently run elsewhere
    q = ou.Quadrature(ou.LS, 3, sn=4)
    with pytest.raises(ou.OuError) as e:
        ou.RadiativeTransfer(3, 4, (2, 2, 2), (1, 1, 1), q, [1.0], [[0.5]], np.zeros(8))
    assert e.value.code == -4
    q3 = ou.Quadrature(ou.LS, 3, sn=4, L_max=2)
    with pytest.raises(ou.OuError) as e:
        ou.RadiativeTransfer(3, 3, (2, 2, 2), (1, 1, 1), q3, [1.0], [[0.5, 0.1, 0.0]], np.zeros(8))
    assert e.value.code == -4
    # non-finite cross section rejected at plan creation
    with pytest.raises(ou.OuError) as e:
        ou.RadiativeTransfer(2, 1, (2, 2), (1, 1), spec.quad, [np.inf], [[0.5]], np.zeros(4))
    assert e.value.code == -5
    # forced non-convergence returns its own status
    with pytest.raises(ou.OuError) as e:
        op.solve(solver=ou.GMRES, tol=1e-14, max_it=2)
    assert e.value.code == -6


def test_multi_column_apply_equals_columns():
    """ou_apply_multi: every column of [n_rhs][size] gets the operator, each
    equal to its own single Apply bit-for-bit."""
    spec = P.hetero_2d(n=8, block=4, n_polar=4, n_azim=8, L=1)
    op = spec.operator(ou.FP32)
    X = np.stack([P.seeded_vector(spec.size, 40 + r) for r in range(3)])
    Y = op.Apply(X)
    assert Y.shape == X.shape
    for r in range(3):
        assert np.array_equal(Y[r], op.Apply(X[r]))
    with pytest.raises(ou.OuError):
        op.Apply(np.zeros((2, spec.size - 1)))


@pytest.mark.parametrize("precision", [ou.FP64, ou.FP32])
@pytest.mark.parametrize("name", ["hetero_2d_q2_L3", "batch_3d_q1", "headline_3d_q2_L1"])
def test_octant_parts_on_several_streams_equal_one(name, precision):
    """Each level as one launch, two (octant halves, two streams) or four
    (quarters): the same chunks run the same arithmetic, so Apply is
    identical bit-for-bit; the source sweep too."""
    spec = CASES[name]()
    op = spec.operator(precision)
    assert op.sweep_streams == 2  # the default
    xs = [P.seeded_vector(spec.size, 60 + r) for r in range(3)]
    two = [op.Apply(x) for x in xs]
    op.compute_scattering_source(xs[0])
    two_src = op.sweep(None, with_source=True)
    op.sweep_streams = 1
    assert op.sweep_streams == 1
    for x, y2 in zip(xs, two):
        assert np.array_equal(op.Apply(x), y2)
    op.compute_scattering_source(xs[0])
    assert np.array_equal(op.sweep(None, with_source=True), two_src)
    op.sweep_streams = 4
    assert np.array_equal(op.Apply(xs[1]), two[1])
    op.sweep_streams = 2
    assert np.array_equal(op.Apply(xs[2]), two[2])
    with pytest.raises(ou.OuError):
        op.sweep_streams = 3


def test_apply_on_a_callers_stream_equals_the_plans():
    """ou_apply_device on a stream the caller created: the octant parts run on
    the plan's own streams, forked from and joined to the caller's, and the
    result equals Apply on the plan's stream bit-for-bit, at 2 and 4 parts."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        spec = CASES["headline_3d_q2_L1"]()
        op = spec.operator(ou.FP32)
        x = P.seeded_vector(spec.size, 71)
        d_x, d_y = ou.DeviceBuffer(x.nbytes), ou.DeviceBuffer(x.nbytes)
        d_x.upload(x)
        for parts in (2, 4):
            op.sweep_streams = parts
            ref = op.Apply(x)
            assert ou.lib().ou_apply_device(op._plan, d_x.ptr, d_y.ptr, stream) == 0
            assert hip.hipStreamSynchronize(stream) == 0
            y = np.empty(spec.size)
            d_y.download(y)
            assert np.array_equal(y, ref), parts
        d_x.free()
        d_y.free()
        op.close()
    finally:
        hip.hipStreamDestroy(stream)


def test_apply_properties_at_full_2d_size():
    """256x256 Q2, 256 directions, L=3: linearity and sigma_s = 0 => identity."""
    spec = P.hetero_2d()
    print(f"[gpu] full-size 2D properties: {spec.size} unknowns", flush=True)
    op = spec.operator(ou.FP64)
    x = P.seeded_vector(spec.size, 1)
    y = P.seeded_vector(spec.size, 2)
    ax, ay = op.Apply(x), op.Apply(y)
    assert rel_err(op.Apply(2.0 * x - 0.5 * y), 2.0 * ax - 0.5 * ay) < 1e-12
    assert np.isfinite(ax).all()
    del op
    spec.sigma_s = np.zeros_like(spec.sigma_s)
    op0 = spec.operator(ou.FP32)
    assert np.array_equal(op0.Apply(x), x)


def test_apply_properties_at_full_headline_size():
    """128^3 Q2, GLC 12x24 (288 directions), L=1, the size no CPU
    restatement finishes in seconds: FP64 linearity, the FP32 sweep part
    within the FP32 bound of the FP64 one, and sigma_s = 0 => identity."""
    spec = P.headline_3d()
    print(f"[gpu] full-size 3D properties: {spec.size} unknowns", flush=True)
    x = P.seeded_vector(spec.size, 1)
    y = P.seeded_vector(spec.size, 2)
    op = spec.operator(ou.FP64)
    ax, ay = op.Apply(x), op.Apply(y)
    assert np.isfinite(ax).all()
    z = 2.0 * x - 0.5 * y
    err = rel_err(z - op.Apply(z), z - (2.0 * ax - 0.5 * ay))
    print(f"[gpu] full-size FP64 linearity (sweep part): {err:.2e}", flush=True)
    assert err < 1e-12
    del op, y, ay, z
    op32 = spec.operator(ou.FP32)
    err = rel_err(x - op32.Apply(x), x - ax)
    print(f"[gpu] full-size FP32 sweep part vs FP64: {err:.2e}", flush=True)
    assert err < FP32_TOL
    del op32, ax
    spec.sigma_s = np.zeros_like(spec.sigma_s)
    op0 = spec.operator(ou.FP32)
    assert np.array_equal(op0.Apply(x), x)


def test_default_direction_groups_follow_level_parallelism(oracle):
    """A 2D mesh has few chunks per wavefront level: the default shape splits
    each quadrant's directions over several workgroups.  Results differ from
    one group only in the order of the FP64 partial sums, and match the oracle."""
    spec = P.hetero_2d(n=64, block=8, n_polar=8, n_azim=16, L=1)
    op = spec.operator(ou.FP64)
    w, g = op.sweep_shape
    assert g > 1, (w, g)
    x = P.seeded_vector(spec.size, 5)
    y = op.Apply(x)
    assert rel_err(y, oracle_problem(oracle, spec).apply(x)) < FP64_TOL
    op.sweep_shape = (w, 1)
    assert rel_err(op.Apply(x), y) < 1e-14
    # the headline's level launches already fill the GPU: one group
    assert P.headline_3d(n=5, block=2, n_polar=4, n_azim=8, L=1).operator(ou.FP32).sweep_shape[1] == 1


class _Subset:
    """A direction set made of some directions of a quadrature (the operator
    takes any omega / M2D / D2M it is handed)."""

    def __init__(self, q, keep):
        self.omega = np.ascontiguousarray(q.omega[keep])
        self.M2D = np.ascontiguousarray(q.M2D[keep])
        self.D2M = np.ascontiguousarray(q.D2M[:, keep])
        self.mom_l = q.mom_l
        self.most_normal = np.ascontiguousarray(q.most_normal[:, keep])
        self.weight_sum = q.weight_sum
        self.n_dir = int(self.omega.shape[0])
        self.n_mom = q.n_mom


def test_reshape_with_an_empty_octant_matches_oracle(oracle):
    """A direction set with no direction in one octant: that octant has no
    chunks, so nothing writes its partial-moment slots.  Changing the
    direction groups (1 -> 4 -> 1 -> 2) changes the slot layout; every shape
    still matches the oracle."""
    q = ou.Quadrature(ou.LS, 3, sn=6, L_max=1)
    keep = ~np.all(q.omega < 0, axis=1)  # drop the (-,-,-) octant
    sub = _Subset(q, keep)
    spec = P.batch_3d(n=5, block=2, quad=q)
    spec.quad = sub
    spec.sigma_s = np.concatenate([spec.sigma_s, 0.3 * spec.sigma_s], axis=1)
    pb = oracle_problem(oracle, spec)
    x = P.seeded_vector(spec.size, 77)
    ref = pb.apply(x)
    for precision in (ou.FP64, ou.FP32):
        op = spec.operator(precision)
        tol = FP64_TOL if precision == ou.FP64 else FP32_TOL
        for shape in [(4, 1), (4, 4), (4, 1), (2, 2), (8, 3)]:
            op.sweep_shape = shape
            err = rel_err(op.Apply(x), ref)
            assert err < tol, (precision, shape, err)
        op.close()


@pytest.mark.parametrize("solver", [ou.GMRES, ou.BICGSTAB, ou.SI])
def test_device_nonfinite_ends_the_solve(solver):
    """A value that turns non-finite on the device ends the solve with
    OU_ERR_NONFINITE at once, not after max_it iterations.  Forced from the
    host with finite inputs: in the FP32 sweep, material 1's scattering
    source (sigma_s = 1e300 times M x) overflows FP32, and its inverse
    (sigma_t = 1e300) underflows to zero, so psi = 0 x inf there."""
    spec = P.uniform_2d(n=8)
    spec.sigma_t = np.array([1.0, 1e300])
    spec.sigma_s = np.array([[0.5], [1e300]])
    spec.mat = (np.arange(64) % 2).astype(np.int32)
    op = spec.operator(ou.FP32)
    rhs = np.ones(spec.size)
    with pytest.raises(ou.OuError) as e:
        op.solve_rhs(rhs, solver=solver, tol=1e-10, max_it=100000)
    assert e.value.code == -5, e.value
    assert b"non-finite" in ou.lib().ou_last_error()


def test_more_than_64_directions_per_octant_need_groups(oracle):
    """An octant's directions are staged in LDS 64 at a time, one direction
    group per workgroup: 128 per octant need at least two groups.  The plan
    starts with enough; fewer are refused; every allowed shape matches the
    oracle."""
    q = ou.Quadrature(ou.PRODUCT_GLC, 3, n_polar=16, n_azim=64, L_max=1)
    assert q.n_dir // 8 == 128
    spec = P.headline_3d(n=3, block=2, quad=q, p=1)
    op = spec.operator(ou.FP64)
    w, g = op.sweep_shape
    assert g >= 2
    with pytest.raises(ou.OuError) as e:
        op.sweep_shape = (w, 1)
    assert e.value.code == -1
    x = P.seeded_vector(spec.size, 8)
    ref = oracle_problem(oracle, spec).apply(x)
    for shape in [(4, 2), (4, 3), (2, 8)]:
        op.sweep_shape = shape
        assert rel_err(op.Apply(x), ref) < FP64_TOL, shape
    op.close()
