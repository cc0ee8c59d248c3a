"""ctypes binding of the FP64 CPU oracle (oracle/ou_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker.  The product
(ouranos_amd/) never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

BC_VACUUM, BC_ISOTROPIC, BC_MOST_NORMAL = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_ubyte)


class _Problem(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("p", ctypes.c_int), ("n", ctypes.c_int * 3),
        ("h", ctypes.c_double * 3), ("n_dir", ctypes.c_int), ("omega", _dp),
        ("n_mom", ctypes.c_int), ("M2D", _dp), ("D2M", _dp), ("mom_l", _ip),
        ("n_mat", ctypes.c_int), ("L", ctypes.c_int), ("sigma_t", _dp),
        ("sigma_s", _dp), ("mat", _ip), ("q", _dp), ("ws", ctypes.c_double),
        ("bc", ctypes.c_int * 6), ("inc", ctypes.c_double * 6),
        ("most_normal", _up),
    ]


def build():
    """Compile the oracle with its Makefile (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.orc_n_dofs.restype = ctypes.c_long
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, _d(x), _d(w))
    return x, w


def ls_octant(sn):
    """One octant of the level-symmetric set as LS::build_octant leaves it:
    (omega [n, 3], raw weights [n]) before deployment and normalisation."""
    om = np.zeros(3 * 64)
    w = np.zeros(64)
    n = lib().orc_ls_octant(sn, _d(om), _d(w))
    if n < 0:
        raise ValueError("unsupported quadrature order")
    return om[: 3 * n].reshape(n, 3).copy(), w[:n].copy()


def quadrature(kind, dim, sn=None, n_polar=None, n_azim=None):
    """Deployed direction set: kind 'ls' | 'glc' (triangular) | 'product'."""
    cap = 8 * 4096
    om = np.zeros(3 * cap)
    w = np.zeros(cap)
    if kind == "ls":
        n_oct = lib().orc_ls_octant(sn, _d(om), _d(w))
    elif kind == "glc":
        n_oct = lib().orc_glc_octant(sn, _d(om), _d(w))
    elif kind == "product":
        n_oct = lib().orc_product_octant(n_polar, n_azim, _d(om), _d(w))
    else:
        raise ValueError(kind)
    if n_oct < 0:
        raise ValueError("unsupported quadrature order")
    n_dir = lib().orc_deploy(dim, n_oct, _d(om), _d(w))
    return om[: 3 * n_dir].reshape(n_dir, 3).copy(), w[:n_dir].copy()


def harmonics(dim, L, omega, w, ws, galerkin=False, sn=0):
    n_dir = omega.shape[0]
    cap = max(n_dir, (L + 1) ** 2)
    M2D = np.zeros(n_dir * cap)
    D2M = np.zeros(cap * n_dir)
    ml = np.zeros(cap, dtype=np.int32)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    ww = np.ascontiguousarray(w, dtype=np.float64)
    n_mom = lib().orc_harmonics(dim, L, int(galerkin), sn, n_dir, _d(om), _d(ww),
                                ctypes.c_double(ws), _d(M2D), _d(D2M), _i(ml))
    if n_mom < 0:
        raise ValueError("harmonics failed")
    return (M2D[: n_dir * n_mom].reshape(n_dir, n_mom).copy(),
            D2M[: n_mom * n_dir].reshape(n_mom, n_dir).copy(), ml[:n_mom].copy())


def most_normal(dim, omega):
    n_dir = omega.shape[0]
    f = np.zeros(2 * dim * n_dir, dtype=np.uint8)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    lib().orc_most_normal(dim, n_dir, _d(om), f.ctypes.data_as(_up))
    return f.reshape(2 * dim, n_dir)


def fe_cell(dim, p, h):
    N = (p + 1) ** dim
    M = np.zeros(N * N)
    G = np.zeros(dim * N * N)
    E = np.zeros(2 * dim * N * N)
    U = np.zeros(2 * dim * N * N)
    nrm = np.zeros(18)
    hh = np.ascontiguousarray(list(h) + [1.0] * (3 - len(h)), dtype=np.float64)
    lib().orc_fe_cell(dim, p, _d(hh), _d(M), _d(G), _d(E), _d(U), _d(nrm))
    return (M.reshape(N, N), G.reshape(dim, N, N), E.reshape(2 * dim, N, N),
            U.reshape(2 * dim, N, N), nrm.reshape(6, 3)[: 2 * dim])


def lu_decompose(A):
    n = A.shape[0]
    a = np.ascontiguousarray(A, dtype=np.float64).copy()
    piv = np.zeros(n, dtype=np.int32)
    lib().orc_lu_decompose(n, _d(a), _i(piv))
    return a, piv


def lu_solve(LU, b, piv):
    n = LU.shape[0]
    bb = np.ascontiguousarray(b, dtype=np.float64).copy()
    x = np.zeros(n)
    lib().orc_lu_solve(n, _d(np.ascontiguousarray(LU)), _d(bb), _d(x), _i(piv))
    return x


class Problem:
    """A structured-mesh DG-Sn problem held for the oracle's C routines."""

    def __init__(self, dim, p, n, h, omega, M2D, D2M, mom_l, sigma_t, sigma_s,
                 mat, q=None, ws=1.0, bc=None, inc=None, most_normal_flags=None):
        self.keep = []

        def keep(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self.keep.append(a)
            return a

        s = _Problem()
        s.dim, s.p = dim, p
        nn = list(n) + [1] * (3 - len(n))
        hh = list(h) + [1.0] * (3 - len(h))
        for a in range(3):
            s.n[a] = nn[a]
            s.h[a] = hh[a]
        om = keep(omega, np.float64)
        s.n_dir = om.shape[0]
        s.omega = _d(om)
        m2d = keep(M2D, np.float64)
        d2m = keep(D2M, np.float64)
        s.n_mom = m2d.shape[1]
        s.M2D, s.D2M = _d(m2d), _d(d2m)
        s.mom_l = _i(keep(mom_l, np.int32))
        st = keep(sigma_t, np.float64)
        ss = keep(sigma_s, np.float64).reshape(st.shape[0], -1)
        self.keep.append(ss)
        s.n_mat = st.shape[0]
        s.L = ss.shape[1] - 1
        s.sigma_t, s.sigma_s = _d(st), _d(ss)
        s.mat = _i(keep(mat, np.int32))
        s.q = _d(keep(q, np.float64)) if q is not None else None
        s.ws = ws
        for f in range(6):
            s.bc[f] = (bc[f] if bc is not None and f < len(bc) else BC_VACUUM)
            s.inc[f] = (inc[f] if inc is not None and f < len(inc) else 0.0)
        if most_normal_flags is not None:
            s.most_normal = keep(most_normal_flags, np.uint8).ctypes.data_as(_up)
        self.s = s
        self.n_dofs = int(lib().orc_n_dofs(ctypes.byref(s)))
        self.n_mom = s.n_mom
        self.size = self.n_dofs * self.n_mom

    def _ref(self):
        return ctypes.byref(self.s)

    def scatter_source(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        S = np.zeros(self.size)
        lib().orc_scatter_source(self._ref(), _d(x), _d(S))
        return S

    def sweep(self, S, with_source=False, d_begin=0, d_end=None):
        S = np.ascontiguousarray(S, dtype=np.float64)
        phi = np.zeros(self.size)
        d_end = self.s.n_dir if d_end is None else d_end
        lib().orc_sweep_dirs(self._ref(), _d(S), int(with_source), d_begin, d_end,
                             _d(phi))
        return phi

    def apply(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.size)
        lib().orc_apply(self._ref(), _d(x), _d(y))
        return y

    def rhs(self):
        r = np.zeros(self.size)
        lib().orc_rhs(self._ref(), _d(r))
        return r

    def si(self, tol, max_it, phi0=None):
        phi = np.zeros(self.size) if phi0 is None else np.array(phi0, dtype=np.float64)
        it = lib().orc_si(self._ref(), ctypes.c_double(tol), max_it, _d(phi))
        return phi, it

    def gmres(self, b, tol, restart=30, max_it=1000, x0=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.size) if x0 is None else np.array(x0, dtype=np.float64)
        rr = ctypes.c_double(0.0)
        it = lib().orc_gmres(self._ref(), _d(b), _d(x), restart, ctypes.c_double(tol),
                             max_it, ctypes.byref(rr))
        return x, it, rr.value

    def bicgstab(self, b, tol, max_it=1000, x0=None):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.zeros(self.size) if x0 is None else np.array(x0, dtype=np.float64)
        rr = ctypes.c_double(0.0)
        it = lib().orc_bicgstab(self._ref(), _d(b), _d(x), ctypes.c_double(tol),
                                max_it, ctypes.byref(rr))
        return x, it, rr.value
