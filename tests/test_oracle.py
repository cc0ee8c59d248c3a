"""Pins the FP64 CPU oracle against the reference's known answers, closed
forms, invariants and brute force.  CPU only."""
import numpy as np
import pytest

from helpers import load_golden, rel_err, to_reference_numbering


def reference_problem(O):
    om, w = O.quadrature("glc", 2, sn=4)
    M2D, D2M, ml = O.harmonics(2, 0, om, w, 1.0)
    return O.Problem(2, 1, (4, 4), (0.5, 0.5), om, M2D, D2M, ml, [1.0], [0.5],
                     np.zeros(16, np.int32), q=np.ones(16), ws=1.0)


def test_reference_golden_solution(oracle):
    """tests/RadiativeTransfer_test.cc:84-152: 64 moment-0 values (1e-3)."""
    gold = np.array(load_golden("reference_2d_solution.json")["moment0"])
    pb = reference_problem(oracle)
    rhs = pb.rhs()
    for solve in ("bicgstab", "gmres"):
        x, it, rr = getattr(pb, solve)(rhs, 1e-10)
        assert it > 0 and rr < 1e-10
        mine = to_reference_numbering(x, 4, 4)
        assert np.abs(mine - gold).max() < 1e-3
        # the golden values are printed to 6 significant digits
        assert np.abs(mine - gold).max() < 1e-5
    phi, it = pb.si(1e-12, 500)
    assert it > 0
    assert np.abs(to_reference_numbering(phi, 4, 4) - gold).max() < 1e-5


def test_fecell_q1_fixture(oracle):
    """tests/FECell_test.cc:45-297: Q1 matrices on the unit square (1e-12)."""
    fx = load_golden("fecell_q1.json")
    M, G, E, U, nrm = oracle.fe_cell(2, 1, [1.0, 1.0])
    for k, v in fx["mass"].items():
        assert abs(M[int(k[0]), int(k[1])] - v) < 1e-12
    for a in (0, 1):
        for k, v in fx[f"grad_{a}"].items():
            assert abs(G[a, int(k[0]), int(k[1])] - v) < 1e-12
    for f, d in fx["downwind_isolated"].items():
        for k, v in d.items():
            assert abs(E[int(f), int(k[0]), int(k[1])] - v) < 1e-12
    # middle cell of a 3x3 mesh on [0,3]^2: unit cells, neighbours on all faces
    for f, d in fx["upwind_middle_of_3x3"].items():
        for k, v in d.items():
            assert abs(U[int(f), int(k[0]), int(k[1])] - v) < 1e-12
    assert np.allclose(nrm[:, :2], np.array(fx["normals"]), atol=1e-12)


def test_quadrature_s4_fixture(oracle):
    """tests/RTQuadrature_test.cc:21-165."""
    fx = load_golden("quadrature_s4.json")
    ls = fx["ls_s4_galerkin"]
    om, w = oracle.quadrature("ls", 2, sn=4)
    assert om.shape[0] == ls["n_dir"]
    assert tuple(om[0]) == tuple(ls["omega0"])  # exact equality in the reference test
    M2D, D2M, ml = oracle.harmonics(2, 4, om, w, 4 * np.pi, galerkin=True, sn=4)
    assert M2D.shape[1] == ls["n_mom"]
    assert list(ml) == ls["l"]
    assert np.abs(D2M @ M2D - np.eye(12)).max() < 1e-12
    mn = oracle.most_normal(2, om)
    assert (mn.astype(bool) == np.array(ls["most_normal"])).all()
    g = fx["glc_s4"]
    om, w = oracle.quadrature("glc", 2, sn=4)
    M2D, D2M, ml = oracle.harmonics(2, 4, om, w, 4 * np.pi)
    assert om.shape[0] == g["n_dir"] and M2D.shape[1] == g["n_mom"]
    assert list(ml) == g["l"]
    assert np.abs(om[0] - np.array(g["omega0"])).max() < 1e-12


LS_ORDERS = [2, 4, 6, 8, 10, 12]


def _ls_fixture(sn):
    t = load_golden("ls_tables.json")[str(sn)]
    return np.array(t["cosines"]), np.array([t["classes"][k] for k in t["weights"]])


@pytest.mark.parametrize("sn", LS_ORDERS)
def test_ls_octant_against_reference_tables(oracle, sn):
    """LS::build_octant (LS.cc:14-169), text-extracted: every cosine of the
    oracle's octant is one of the printed direction[k] and every weight is
    the printed weight(i), exactly and in the reference's order; the layout
    is LS::compute_omega's triangle (omega_z constant over a row, omega_y
    ascending along it)."""
    cos, w_ref = _ls_fixture(sn)
    om, w = oracle.ls_octant(sn)
    n = len(cos)
    assert om.shape[0] == n * (n + 1) // 2 == len(w_ref)
    assert np.array_equal(w, w_ref)
    assert np.isin(om, cos).all()
    row = 0
    for j in range(n):
        blk = om[row:row + n - j]
        assert (blk[:, 2] == cos[j]).all()
        assert (np.diff(blk[:, 1]) > 0).all() and (np.diff(blk[:, 0]) < 0).all()
        row += n - j


@pytest.mark.parametrize("sn", LS_ORDERS)
def test_ls_octant_mathematics(oracle, sn):
    """What the level-symmetric construction fixes, independently of the
    printed tables: unit directions (so a wrong cosine or a wrong index in
    the triangle shows), the three axes alike, and the even polar moments
    sum_d w_d mu_d^2k = sum_d w_d / (2k+1) exact for k <= sn/2 (a wrong
    weight class or a class attached to the wrong direction shows)."""
    om, w = oracle.ls_octant(sn)
    assert np.abs((om ** 2).sum(axis=1) - 1.0).max() < 1e-15
    W = w.sum()
    assert abs(W - 1.0) < 1e-15
    for a in range(3):
        assert abs((w * om[:, a] ** 2).sum() - W / 3.0) < 1e-15
        for k in range(sn // 2 + 1):
            assert abs((w * om[:, a] ** (2 * k)).sum() - W / (2 * k + 1)) < 2e-16 * (k + 1)
    # a plausible slip -- one weight class attached to its neighbour's
    # direction -- breaks the moment conditions
    if sn >= 6:
        bad = w.copy()
        i = int(np.argmax(w != w[0]))
        bad[0], bad[i] = bad[i], bad[0]
        errs = [abs((bad * om[:, a] ** (2 * k)).sum() - bad.sum() / (2 * k + 1))
                for a in range(3) for k in range(sn // 2 + 1)]
        assert max(errs) > 1e-6


def test_gauss_legendre_against_reference_tables(oracle):
    """Newton Gauss-Legendre vs the tables of GLC.cc:60-175.  The S6 table is
    printed unsorted (0.661, 0.239, 0.932); compare as sets."""
    tab = load_golden("gauss_legendre_tables.json")
    for sn in range(2, 17, 2):
        x, w = oracle.gauss_legendre(sn)
        t = tab[str(sn)]
        order = np.argsort(t["nodes"])
        assert np.abs(x[sn // 2:] - np.array(t["nodes"])[order]).max() < 2e-15
        assert np.abs(w[sn // 2:] - np.array(t["weights"])[order]).max() < 2e-15


def test_lu_crout_matches_dense_solve(oracle):
    rng = np.random.default_rng(7)
    for n in (4, 8, 9, 27):
        for trial in range(5):
            A = rng.standard_normal((n, n))
            if trial == 0:
                A[0, 0] = 0.0  # forces a row interchange at step 0
            b = rng.standard_normal(n)
            LU, piv = oracle.lu_decompose(A)
            x = oracle.lu_solve(LU, b, piv)
            xr = np.linalg.solve(A, b)
            assert np.abs(x - xr).max() < 1e-10 * max(1.0, np.abs(xr).max())
            assert np.abs(A @ x - b).max() < 1e-10 * np.abs(b).max() * np.linalg.cond(A)
            if trial == 0:
                assert piv[0] != 0
            # L (with diagonal) times unit-upper U reproduces the row-permuted A
            L = np.tril(LU)
            U = np.triu(LU, 1) + np.eye(n)
            P = np.arange(n)
            for k in range(n):
                P[[k, piv[k]]] = P[[piv[k], k]]
            assert np.abs(L @ U - A[P]).max() < 1e-10 * np.abs(A).max() * n


@pytest.mark.parametrize("dim,p", [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (3, 1), (3, 2), (3, 3)])
def test_fe_closed_forms(oracle, dim, p):
    h = [0.5, 0.25, 2.0][:dim]
    M, G, E, U, nrm = oracle.fe_cell(dim, p, h)
    # row sums of M = integral of each basis function: tensor product of the
    # 1D closed Newton-Cotes weights (trapezoid for Q1, Simpson for Q2, 3/8 for
    # Q3, Boole for Q4, the six-point rule for Q5)
    nc = {1: [1 / 2, 1 / 2], 2: [1 / 6, 2 / 3, 1 / 6], 3: [1 / 8, 3 / 8, 3 / 8, 1 / 8],
          4: [7 / 90, 32 / 90, 12 / 90, 32 / 90, 7 / 90],
          5: [19 / 288, 75 / 288, 50 / 288, 50 / 288, 75 / 288, 19 / 288]}[p]
    w = np.array([1.0])
    for a in range(dim):
        w = np.kron(np.array(nc) * h[a], w)
    assert np.abs(M.sum(axis=1) - w).max() < 1e-14
    # integration by parts: G_d + G_d^T = sum_f n_{f,d} E_f
    for d in range(dim):
        rhs = sum(nrm[f, d] * E[f] for f in range(2 * dim))
        assert np.abs(G[d] + G[d].T - rhs).max() < 1e-13
    # M symmetric positive definite; E_f, U_f supported on face dofs only
    assert np.abs(M - M.T).max() < 1e-15 and np.linalg.eigvalsh(M).min() > 0


def _global_transport_solve(O, dim, p, n, h, om, M2D, D2M, sigt, S):
    """Brute force: assemble the global upwind DG system of every direction
    densely and solve it; returns phi = D2M psi."""
    M, G, E, U, nrm = O.fe_cell(dim, p, h)
    N = M.shape[0]
    nn = list(n) + [1] * (3 - len(n))
    nc = int(np.prod(n))
    strides = [1, nn[0], nn[0] * nn[1]]
    phi = np.zeros(S.shape[0])
    n_dofs = nc * N
    for d in range(om.shape[0]):
        Lg = np.zeros((n_dofs, n_dofs))
        b = np.zeros(n_dofs)
        for c in range(nc):
            ijk = [c % nn[0], (c // nn[0]) % nn[1], c // (nn[0] * nn[1])]
            A = sigt * M - sum(om[d, a] * G[a] for a in range(dim))
            for f in range(2 * dim):
                ndo = om[d] @ nrm[f]
                a, side = f // 2, f % 2
                if ndo >= 0:
                    A = A + ndo * E[f]
                else:
                    nb = ijk[a] + (1 if side else -1)
                    if 0 <= nb < nn[a]:
                        cn = c + (1 if side else -1) * strides[a]
                        Lg[c * N:(c + 1) * N, cn * N:(cn + 1) * N] -= -ndo * U[f]
            Lg[c * N:(c + 1) * N, c * N:(c + 1) * N] += A
            for m in range(M2D.shape[1]):
                b[c * N:(c + 1) * N] += M2D[d, m] * S[m * n_dofs + c * N:m * n_dofs + (c + 1) * N]
        psi = np.linalg.solve(Lg, b)
        for m in range(M2D.shape[1]):
            phi[m * n_dofs:(m + 1) * n_dofs] += D2M[m, d] * psi
    return phi


@pytest.mark.parametrize("dim,p,n,L", [(2, 1, (3, 3), 0), (2, 2, (2, 3), 1),
                                       (3, 1, (2, 2, 2), 1), (3, 2, (2, 1, 2), 0),
                                       (2, 4, (2, 2), 1), (2, 5, (2, 1), 0),
                                       (3, 3, (2, 1, 2), 1)])
def test_sweep_equals_global_dense_solve(oracle, dim, p, n, L):
    """One sweep = exact inverse of the block-lower-triangular transport
    operator: compare with a dense solve of the global system."""
    om, w = oracle.quadrature("ls", dim, sn=4)
    M2D, D2M, ml = oracle.harmonics(dim, L, om, w, 4 * np.pi)
    h = [0.3, 0.5, 0.7][:dim]
    nc = int(np.prod(n))
    pb = oracle.Problem(dim, p, n, h, om, M2D, D2M, ml, [2.0], [[0.0] * (L + 1)],
                        np.zeros(nc, np.int32))
    S = np.random.default_rng(1).random(pb.size)
    phi = pb.sweep(S)
    ref = _global_transport_solve(oracle, dim, p, n, h, om, M2D, D2M, 2.0, S)
    assert rel_err(phi, ref) < 1e-12


def test_full_solve_equals_dense_operator(oracle):
    """(I - D L^-1 M Sigma) phi = rhs solved densely equals GMRES / SI."""
    om, w = oracle.quadrature("ls", 2, sn=4)
    M2D, D2M, ml = oracle.harmonics(2, 1, om, w, 1.0)
    n = (3, 2)
    mat = np.array([0, 1, 1, 0, 1, 0], np.int32)
    pb = oracle.Problem(2, 1, n, (0.4, 0.6), om, M2D, D2M, ml, [1.0, 3.0],
                        [[0.6, 0.2], [2.0, 0.5]], mat, q=np.arange(6) * 0.3 + 0.1, ws=1.0)
    A = np.column_stack([pb.apply(e) for e in np.eye(pb.size)])
    rhs = pb.rhs()
    x_dense = np.linalg.solve(A, rhs)
    x, it, rr = pb.gmres(rhs, 1e-13)
    assert rel_err(x, x_dense) < 1e-11
    phi, it = pb.si(1e-14, 2000)
    assert rel_err(phi, x_dense) < 1e-11


def test_infinite_medium_closed_form(oracle):
    """Isotropic inflow equal to the infinite-medium flux makes the solution
    exactly uniform: phi_0 = sqrt(ws) q / (ws (sigma_t - sigma_s)), also with
    anisotropic scattering (L=1) in 3D and at Q2."""
    for dim, p, L, ws in ((2, 1, 0, 1.0), (2, 2, 1, 4 * np.pi), (3, 1, 1, 4 * np.pi),
                          (3, 2, 0, 2.0)):
        om, w = oracle.quadrature("ls", dim, sn=6)
        M2D, D2M, ml = oracle.harmonics(dim, L, om, w, ws)
        n = (3,) * dim
        nc = int(np.prod(n))
        st, ss, q = 2.0, [1.5, 0.4][:L + 1], 0.7
        inc = q / (st - ss[0])
        pb = oracle.Problem(dim, p, n, [0.3] * dim, om, M2D, D2M, ml, [st], [ss],
                            np.zeros(nc, np.int32), q=np.full(nc, q), ws=ws,
                            bc=[1] * 6, inc=[inc] * 6)
        phi, it = pb.si(1e-14, 5000)
        nd = pb.n_dofs
        expect = np.sqrt(ws) * q / (ws * (st - ss[0]))
        assert np.abs(phi[:nd] - expect).max() < 1e-11 * expect
        if L > 0:
            assert np.abs(phi[nd:]).max() < 1e-11 * expect


def test_apply_is_identity_without_scattering_and_linear(oracle):
    om, w = oracle.quadrature("ls", 3, sn=4)
    M2D, D2M, ml = oracle.harmonics(3, 1, om, w, 4 * np.pi)
    n = (2, 3, 2)
    nc = 12
    pb0 = oracle.Problem(3, 1, n, (0.2, 0.3, 0.4), om, M2D, D2M, ml, [1.0], [[0.0, 0.0]],
                         np.zeros(nc, np.int32))
    x = np.random.default_rng(3).random(pb0.size)
    assert np.abs(pb0.apply(x) - x).max() == 0.0
    pb = oracle.Problem(3, 1, n, (0.2, 0.3, 0.4), om, M2D, D2M, ml, [1.0, 2.0],
                        [[0.5, 0.1], [1.5, 0.3]], np.arange(nc, dtype=np.int32) % 2)
    y = np.random.default_rng(4).random(pb.size)
    lhs = pb.apply(2.0 * x - 3.0 * y)
    assert rel_err(lhs, 2.0 * pb.apply(x) - 3.0 * pb.apply(y)) < 1e-13
