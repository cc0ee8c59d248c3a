"""Host-side checks of the product library that need no GPU: the C ABI loads
and exports what include/ouranos_amd.h declares; quadrature, harmonics and
reference-element matrices agree with the oracle and the reference fixtures;
argument errors come back as status codes."""
import ctypes
import os
import re

import numpy as np
import pytest

import ouranos_amd as ou
from helpers import load_golden

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      "include", "ouranos_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ou_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ou.lib()
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_is_empty_string_initially():
    assert isinstance(ou.lib().ou_last_error(), bytes)


@pytest.mark.parametrize("kind,dim,kw", [
    ("ls", 2, dict(sn=4)), ("ls", 2, dict(sn=8)), ("ls", 3, dict(sn=8)), ("ls", 3, dict(sn=12)),
    ("glc", 2, dict(sn=4)), ("glc", 2, dict(sn=10)), ("glc", 3, dict(sn=6)),
    ("product", 2, dict(n_polar=16, n_azim=32)), ("product", 3, dict(n_polar=12, n_azim=24)),
])
@pytest.mark.parametrize("L", [0, 1, 3])
def test_quadrature_matches_oracle(oracle, kind, dim, kw, L):
    k = {"ls": ou.LS, "glc": ou.GLC, "product": ou.PRODUCT_GLC}[kind]
    ws = 4 * np.pi
    q = ou.Quadrature(k, dim, L_max=L, weight_sum=ws, **kw)
    om, w = oracle.quadrature(kind, dim, **kw)
    M2D, D2M, ml = oracle.harmonics(dim, L, om, w, ws)
    assert np.abs(q.omega - om).max() < 1e-15
    assert np.abs(q.weight - w).max() < 1e-16
    assert abs(q.weight.sum() - 1.0) < 1e-14
    assert np.abs(q.M2D - M2D).max() < 1e-13
    assert np.abs(q.D2M - D2M).max() < 1e-13
    assert list(q.mom_l) == list(ml)
    assert (q.most_normal == oracle.most_normal(dim, om)).all()
    n_mom = (L + 1) * (L + 2) // 2 if dim == 2 else (L + 1) ** 2
    assert q.n_mom == n_mom


@pytest.mark.parametrize("dim,n_polar,n_azim,L", [(3, 12, 24, 1), (3, 12, 24, 3), (2, 16, 32, 3)])
def test_product_quadrature_galerkin_identity(dim, n_polar, n_azim, L):
    """Gauss-Legendre x Chebyshev integrates every harmonic product of degree
    <= 2L exactly here, so D2M.M2D = I (orthonormal real harmonics)."""
    q = ou.Quadrature(ou.PRODUCT_GLC, dim, n_polar=n_polar, n_azim=n_azim, L_max=L)
    assert np.abs(q.D2M @ q.M2D - np.eye(q.n_mom)).max() < 1e-13


def test_real_harmonics_closed_forms():
    """Y_00 = 1/sqrt(4 pi); Y_1 = sqrt(3/(4 pi)) (mu, -Omega_x, -Omega_y) in the
    moment order (l=1: m=1 cosine, m=0, m=1 sine) with the Condon-Shortley
    phase; Y_20 = sqrt(5/(16 pi)) (3 mu^2 - 1)."""
    q = ou.Quadrature(ou.LS, 3, sn=8, L_max=2, weight_sum=4 * np.pi)
    Y = q.M2D  # sqrt(4 pi / ws) = 1
    o = q.omega
    assert np.abs(Y[:, 0] - 1 / np.sqrt(4 * np.pi)).max() < 1e-15
    c1 = np.sqrt(3 / (4 * np.pi))
    assert np.abs(Y[:, 1] + c1 * o[:, 0]).max() < 1e-14
    assert np.abs(Y[:, 2] - c1 * o[:, 2]).max() < 1e-14
    assert np.abs(Y[:, 3] + c1 * o[:, 1]).max() < 1e-14
    assert np.abs(Y[:, 6] - np.sqrt(5 / (16 * np.pi)) * (3 * o[:, 2] ** 2 - 1)).max() < 1e-14


def test_real_harmonics_closed_forms_l3():
    """Every l <= 3 harmonic against its Cartesian closed form (orthonormal,
    Condon-Shortley phase), moment order per l: cos m = l..1, m = 0,
    sin m = 1..l.  Written out with x = Omega_x, y = Omega_y, z = mu."""
    q = ou.Quadrature(ou.PRODUCT_GLC, 3, n_polar=8, n_azim=16, L_max=3, weight_sum=4 * np.pi)
    Y = q.M2D
    x, y, z = q.omega[:, 0], q.omega[:, 1], q.omega[:, 2]
    pi = np.pi
    c1, c2a, c2b = np.sqrt(3 / (4 * pi)), np.sqrt(15 / (16 * pi)), np.sqrt(15 / (4 * pi))
    c3a, c3b, c3c = np.sqrt(35 / (32 * pi)), np.sqrt(105 / (16 * pi)), np.sqrt(21 / (32 * pi))
    closed = [
        np.full_like(z, 1 / np.sqrt(4 * pi)),
        -c1 * x, c1 * z, -c1 * y,
        c2a * (x * x - y * y), -c2b * x * z, np.sqrt(5 / (16 * pi)) * (3 * z * z - 1),
        -c2b * y * z, c2a * 2 * x * y,
        -c3a * (x ** 3 - 3 * x * y * y), c3b * z * (x * x - y * y), -c3c * x * (5 * z * z - 1),
        np.sqrt(7 / (16 * pi)) * (5 * z ** 3 - 3 * z),
        -c3c * y * (5 * z * z - 1), c3b * z * 2 * x * y, -c3a * (3 * x * x * y - y ** 3),
    ]
    assert q.n_mom == 16 and list(q.mom_l) == [0] + [1] * 3 + [2] * 5 + [3] * 7
    for m, ref in enumerate(closed):
        assert np.abs(Y[:, m] - ref).max() < 2e-14, m


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("sn", [2, 4, 6, 8, 10, 12])
def test_ls_tables_match_reference(dim, sn):
    """The library's level-symmetric sets against LS.cc:14-169 as printed
    (tests/golden/ls_tables.json): the first octant holds exactly the
    reference's cosines in LS::compute_omega's triangle and the reference's
    weight(i) up to the normalisation to sum 1 over all octants; every
    direction is a unit vector and the weights integrate mu^2k exactly for
    k <= sn/2."""
    t = load_golden("ls_tables.json")[str(sn)]
    cos = np.array(t["cosines"])
    w_ref = np.array([t["classes"][k] for k in t["weights"]])
    q = ou.Quadrature(ou.LS, dim, sn=sn, L_max=0)
    n_oct = len(w_ref)
    n_q = 4 if dim == 2 else 8
    assert q.n_dir == n_q * n_oct
    om, w = q.omega[:n_oct], q.weight[:n_oct]
    assert (om > 0).all() and np.isin(om, cos).all()
    row = 0
    for j in range(len(cos)):
        assert (om[row:row + len(cos) - j, 2] == cos[j]).all()
        row += len(cos) - j
    assert np.abs(w * n_q * w_ref.sum() - w_ref).max() < 1e-16
    assert np.abs((q.omega ** 2).sum(axis=1) - 1.0).max() < 1e-15
    assert abs(q.weight.sum() - 1.0) < 1e-15
    for k in range(sn // 2 + 1):
        assert abs((q.weight * q.omega[:, 2] ** (2 * k)).sum() - 1.0 / (2 * k + 1)) < 1e-15


def test_reference_quadrature_fixture():
    fx = load_golden("quadrature_s4.json")
    ls = fx["ls_s4_galerkin"]
    q = ou.Quadrature(ou.LS, 2, sn=4, L_max=4, galerkin=True)
    assert q.get_n_dir() == 12 and q.get_n_mom() == 12
    assert [q.get_l(i) for i in range(12)] == ls["l"]
    assert tuple(q.get_omega(0)) == tuple(ls["omega0"])
    assert np.abs(q.D2M @ q.M2D - np.eye(12)).max() < 1e-12
    for f in range(4):
        for d in range(12):
            assert q.is_most_normal_direction(f, d) == ls["most_normal"][f][d]
    g = fx["glc_s4"]
    q = ou.Quadrature(ou.GLC, 2, sn=4, L_max=4)
    assert q.get_n_dir() == 12 and q.get_n_mom() == 15
    assert [q.get_l(i) for i in range(15)] == g["l"]
    assert np.abs(q.get_omega(0) - np.array(g["omega0"])).max() < 1e-12


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2), (3, 3)])
def test_fe_matrices_match_oracle(oracle, dim, p):
    h = [0.31, 0.7, 1.3][:dim]
    A = ou.fe_matrices(dim, p, h)
    B = oracle.fe_cell(dim, p, h)
    for a, b in zip(A, B):
        assert np.abs(a - b).max() < 1e-14


def test_fe_q1_fixture():
    fx = load_golden("fecell_q1.json")
    M, G, E, U, nrm = ou.fe_matrices(2, 1, [1.0, 1.0])
    for k, v in fx["mass"].items():
        assert abs(M[int(k[0]), int(k[1])] - v) < 1e-12
    for a in (0, 1):
        for k, v in fx[f"grad_{a}"].items():
            assert abs(G[a, int(k[0]), int(k[1])] - v) < 1e-12
    for f, d in fx["downwind_isolated"].items():
        for k, v in d.items():
            assert abs(E[int(f), int(k[0]), int(k[1])] - v) < 1e-12
    for f, d in fx["upwind_middle_of_3x3"].items():
        for k, v in d.items():
            assert abs(U[int(f), int(k[0]), int(k[1])] - v) < 1e-12


@pytest.mark.parametrize("args,code", [
    ((ou.LS, 2, 14, 0, 0, 0, 0, 1.0), -4),      # LS order outside S2..S12
    ((ou.GLC, 2, 5, 0, 0, 0, 0, 1.0), -1),      # odd GLC order
    ((ou.PRODUCT_GLC, 3, 0, 12, 10, 0, 0, 1.0), -1),  # azimuths not divisible by 4
    ((ou.LS, 3, 4, 0, 0, 4, 1, 1.0), -4),       # Galerkin in 3D
    ((ou.LS, 2, 4, 0, 0, 0, 0, -1.0), -1),      # weight sum
    ((ou.LS, 4, 4, 0, 0, 0, 0, 1.0), -1),       # dim
])
def test_quadrature_errors_are_status_codes(args, code):
    h = ctypes.c_void_p()
    kind, dim, sn, npol, naz, L, gal, ws = args
    r = ou.lib().ou_quadrature_create(kind, dim, sn, npol, naz, L, gal, ctypes.c_double(ws),
                                      ctypes.byref(h))
    assert r == code
    assert len(ou.lib().ou_last_error()) > 0


def test_fe_errors_are_status_codes():
    with pytest.raises(ou.OuError) as e:
        ou.fe_matrices(2, 0, [1.0, 1.0])
    assert e.value.code == -1
    with pytest.raises(ou.OuError):
        ou.fe_matrices(2, 1, [1.0, -1.0])


def test_problem_builders_are_deterministic():
    from ouranos_amd import problems as P
    a = P.hetero_2d(n=32, block=8, seed=11, n_polar=4, n_azim=8)
    b = P.hetero_2d(n=32, block=8, seed=11, n_polar=4, n_azim=8)
    assert (a.sigma_t == b.sigma_t).all() and (a.sigma_s == b.sigma_s).all()
    assert (a.mat == b.mat).all() and a.sigma_t.shape[0] == 16
    # SplitMix64 reference value (Vigna's splitmix64.c with seed 0)
    assert P.SplitMix64(0).next_u64() == 0xE220A8397B1DCDAF
    h = P.headline_3d(n=128, block=16, n_polar=12, n_azim=24)
    assert h.quad.n_dir == 288 and h.quad.n_mom == 4 and h.sigma_t.shape[0] == 512
    assert h.updates_per_sweep() == 16307453952
    assert abs(h.algorithmic_bytes_per_sweep() / 165.7e9 - 1) < 0.01


def test_missing_library_fails_loudly(tmp_path):
    """No CPU fallback: without the built HIP library every entry point
    raises instead of computing."""
    saved_lib, saved_path = ou._lib, ou.LIB_PATH
    try:
        ou._lib = None
        ou.LIB_PATH = str(tmp_path / "libouranos_amd.so")
        with pytest.raises(ImportError):
            ou.lib()
        with pytest.raises(ImportError):
            ou.Quadrature(ou.GLC, 2, sn=4)
    finally:
        ou._lib, ou.LIB_PATH = saved_lib, saved_path
    assert ou.lib() is not None


@pytest.mark.parametrize("dim,p,n,n_dir", [
    (2, 1, (50000, 50000, 1), 40),     # 4 x 2.5e9 (quadrant, cell) pairs
    (3, 2, (1, 2000, 2000), 288),      # 288 x 9 x 4e6 z-face trace values
])
def test_plan_rejects_meshes_past_32_bit_indices(dim, p, n, n_dir):
    """Meshes whose (octant, cell) list or face-trace set would overflow the
    32-bit indices of the schedule and kernels are refused with
    OU_ERR_UNSUPPORTED before any array is read or any device memory is
    touched (so this runs without a GPU; the arrays here are one entry
    long and would be read past their end otherwise)."""
    one = np.zeros(3 * n_dir)
    ml = np.zeros(1, np.int32)
    st = np.ones(1)
    mat = np.zeros(1, np.int32)
    ds = ou.ProblemDesc()
    ds.dim, ds.fe_order = dim, p
    for a in range(3):
        ds.n_cells[a] = n[a]
        ds.cell_size[a] = 1.0
    ds.n_dir, ds.n_mom, ds.n_mat, ds.L = n_dir, 1, 1, 0
    ds.omega = ds.M2D = ds.D2M = one.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ds.mom_l = ml.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    ds.sigma_t = ds.sigma_s = st.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    ds.material_id = mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    ds.weight_sum = 1.0
    plan = ctypes.c_void_p()
    r = ou.lib().ou_plan_create(ctypes.byref(ds), ctypes.byref(plan))
    assert r == -4, ou.lib().ou_last_error()
    assert b"32-bit" in ou.lib().ou_last_error()
    assert not plan.value


def test_training_families_are_held_out_from_the_headline():
    """The learned preconditioner's training and validation instances come
    from seeds far from the headline problem's own (seed 5), so the headline
    solve is always on an instance the network never saw.  headline128 is the
    headline generator at the headline's own size."""
    from ouranos_amd import learned as Lm
    from ouranos_amd import problems as P
    head = P.headline_3d()
    fam = Lm.Family("headline128")
    s = fam.spec(Lm.TRAIN_SEED0 + 6000)
    assert s.sigma_t.shape == head.sigma_t.shape == (512,)
    assert s.updates_per_sweep() == head.updates_per_sweep()
    assert not np.array_equal(s.sigma_t, head.sigma_t)
    assert fam.spec(Lm.TRAIN_SEED0 + 6000).sigma_t.tolist() == s.sigma_t.tolist()
    assert min(Lm.TRAIN_SEED0, Lm.EVAL_SEED0) > 5
    small = Lm.Family("headline").spec(Lm.EVAL_SEED0)
    assert small.sigma_t.shape == (512,) and 8 * small.updates_per_sweep() == s.updates_per_sweep()
    with pytest.raises(ValueError):
        Lm.Family("nope")
