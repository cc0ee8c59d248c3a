"""MI355X DG-Sn transport-sweep engine: Python binding of the C ABI.

The engine replaces the CPU sweep behind Ouranos' RadiativeTransfer operator
(src/RadiativeTransfer.hh:46-127 in Rombur/ouranos).  This module mirrors the
reference's classes for Python callers:

  Quadrature          GLC / LS (RTQuadrature.hh:26-98) -> ou_quadrature_*
  fe_matrices         FECell (FECell.hh:25-79)         -> ou_fe_matrices
  RadiativeTransfer   the Epetra_Operator (Apply, Label, maps, sweep
                      primitives, RadiativeTransfer.hh:46-127) -> ou_plan_*
  solve               the drivers of ouranos.cc:43-209  -> ou_solve

Every call goes through libouranos_amd.so (HIP kernels for gfx950).  There
is no CPU fallback: if the library is missing, importing the binding fails.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "Quadrature", "RadiativeTransfer", "fe_matrices", "OuError", "lib", "LIB_PATH",
    "FP64", "FP32", "LS", "GLC", "PRODUCT_GLC", "SI", "GMRES", "BICGSTAB", "FGMRES",
    "BC_VACUUM", "BC_ISOTROPIC", "BC_MOST_NORMAL", "DeviceBuffer", "device_name", "UNet",
    "solve_refined", "Batch",
]

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libouranos_amd.so")

FP64, FP32 = 0, 1
LS, GLC, PRODUCT_GLC = 0, 1, 2
SI, GMRES, BICGSTAB, FGMRES = 0, 1, 2, 3
SWEEP_LEVELS, SWEEP_PERSISTENT, SWEEP_TILES = 0, 1, 2
BC_VACUUM, BC_ISOTROPIC, BC_MOST_NORMAL = 0, 1, 2

STATUS = {0: "OK", -1: "ARG", -2: "SIZE", -3: "HIP", -4: "UNSUPPORTED", -5: "NONFINITE",
          -6: "NOT_CONVERGED", -7: "NOMEM", -8: "TIMEOUT"}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_ubyte)
_vp = ctypes.c_void_p


class OuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ouranos_amd error {code} ({STATUS.get(code, '?')}): {msg}")
        self.code = code


class ProblemDesc(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("fe_order", ctypes.c_int), ("n_cells", ctypes.c_int * 3),
        ("cell_size", ctypes.c_double * 3), ("n_dir", ctypes.c_int), ("omega", _dp),
        ("n_mom", ctypes.c_int), ("M2D", _dp), ("D2M", _dp), ("mom_l", _ip),
        ("n_mat", ctypes.c_int), ("L", ctypes.c_int), ("sigma_t", _dp), ("sigma_s", _dp),
        ("material_id", _ip), ("source", _dp), ("weight_sum", ctypes.c_double),
        ("bc", ctypes.c_int * 6), ("inc_flux", ctypes.c_double * 6), ("most_normal", _up),
        ("precision", ctypes.c_int),
    ]


class SolverOpts(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int), ("tol", ctypes.c_double), ("max_it", ctypes.c_int),
                ("restart", ctypes.c_int), ("verbose", ctypes.c_int)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("rel_residual", ctypes.c_double),
                ("seconds", ctypes.c_double), ("converged", ctypes.c_int),
                ("fp64_applications", ctypes.c_int),
                ("preconditioner_applications", ctypes.c_int)]

    def as_dict(self):
        return {"iterations": self.iterations, "rel_residual": self.rel_residual,
                "seconds": self.seconds, "converged": bool(self.converged),
                "fp64_applications": self.fp64_applications,
                "preconditioner_applications": self.preconditioner_applications}


class TuneEntry(ctypes.Structure):
    _fields_ = [("waves", ctypes.c_int), ("dir_groups", ctypes.c_int), ("ms", ctypes.c_double),
                ("mode", ctypes.c_int)]


class Timing(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_sweep_kernel", ctypes.c_double),
                ("ms_other", ctypes.c_double), ("n_sweep_launches", ctypes.c_int),
                ("applications", ctypes.c_int)]


_lib = None


def lib():
    """Loads libouranos_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m ouranos_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        L.ou_last_error.restype = ctypes.c_char_p
        L.ou_label.restype = ctypes.c_char_p
        L.ou_domain_size.restype = ctypes.c_long
        L.ou_range_size.restype = ctypes.c_long
        L.ou_n_dofs.restype = ctypes.c_long
        L.ou_norm_inf.restype = ctypes.c_double
        L.ou_plan_destroy.restype = None
        L.ou_quadrature_destroy.restype = None
        L.ou_device_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(_vp)]
        L.ou_device_free.argtypes = [_vp]
        L.ou_memcpy_htod.argtypes = [_vp, _vp, ctypes.c_size_t]
        L.ou_memcpy_dtoh.argtypes = [_vp, _vp, ctypes.c_size_t]
        L.ou_apply_device.argtypes = [_vp, _vp, _vp, _vp]
        L.ou_sweep_device.argtypes = [_vp, _vp, ctypes.c_int, _vp]
        L.ou_scatter_source_device.argtypes = [_vp, _vp, _vp]
        L.ou_time_apply.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(Timing)]
        _fp = ctypes.POINTER(ctypes.c_float)
        L.ou_unet_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong,
                                     ctypes.POINTER(_vp)]
        L.ou_unet_destroy.argtypes = [_vp]
        L.ou_unet_destroy.restype = None
        L.ou_unet_n_params.argtypes = [_vp]
        L.ou_unet_n_params.restype = ctypes.c_long
        L.ou_unet_grid.argtypes = [_vp]
        L.ou_unet_set_mode.argtypes = [_vp, ctypes.c_int]
        L.ou_unet_mode.argtypes = [_vp]
        L.ou_unet_get_params.argtypes = [_vp, _fp]
        L.ou_unet_set_params.argtypes = [_vp, _fp]
        L.ou_unet_predict.argtypes = [_vp, _fp, ctypes.c_int, _fp]
        L.ou_unet_loss_grad.argtypes = [_vp, _fp, _fp, ctypes.c_int, _dp, _fp]
        L.ou_unet_train_step.argtypes = [_vp, _fp, _fp, ctypes.c_int, ctypes.c_float, _dp]
        L.ou_coarse_features.argtypes = [_vp, ctypes.c_int, _dp, _dp, _fp]
        L.ou_coarse_prolong.argtypes = [_vp, ctypes.c_int, _fp, _dp]
        L.ou_learned_pairs.argtypes = [_vp, ctypes.c_int, _dp, ctypes.c_int,
                                       ctypes.c_ulonglong, _dp, _dp]
        L.ou_learned_pairs_krylov.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.ou_learned_pairs_fgmres.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.ou_solve_learned.argtypes = [_vp, _vp, ctypes.POINTER(SolverOpts), _dp,
                                       ctypes.POINTER(SolveInfo)]
        L.ou_solve_refined.argtypes = [_vp, _vp, _vp, ctypes.POINTER(SolverOpts),
                                       ctypes.c_double, _dp, ctypes.POINTER(SolveInfo)]
        L.ou_set_sweep_shape.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.ou_sweep_shape.argtypes = [_vp, _ip, _ip]
        L.ou_set_sweep_streams.argtypes = [_vp, ctypes.c_int]
        L.ou_set_sweep_mode.argtypes = [_vp, ctypes.c_int]
        L.ou_batch_sweep_mode.argtypes = [_vp, _ip]
        L.ou_sweep_mode.argtypes = [_vp, _ip]
        L.ou_set_sweep_wait_limit.argtypes = [_vp, ctypes.c_double]
        L.ou_sweep_status.argtypes = [_vp]
        L.ou_batch_set_sweep_mode.argtypes = [_vp, ctypes.c_int]
        L.ou_sweep_streams.argtypes = [_vp, _ip]
        L.ou_batch_create.argtypes = [ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.POINTER(_vp)]
        L.ou_batch_destroy.argtypes = [_vp]
        L.ou_batch_destroy.restype = None
        L.ou_batch_info.argtypes = [_vp, _ip, _ip, _ip, ctypes.POINTER(ctypes.c_long)]
        L.ou_batch_set_shape.argtypes = [_vp, ctypes.c_int, ctypes.c_int]
        L.ou_batch_set_graph.argtypes = [_vp, ctypes.c_int]
        L.ou_batch_set_streams.argtypes = [_vp, ctypes.c_int]
        L.ou_batch_apply_device.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]
        L.ou_batch_sweep_device.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_int, _vp]
        L.ou_batch_tune.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(TuneEntry), ctypes.c_int,
                                    _ip]
        L.ou_batch_solve_si.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_double, ctypes.c_int,
                                        _ip, _dp, _ip, ctypes.c_int]
        L.ou_batch_time_apply.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                          ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _check(code):
    if code != 0:
        raise OuError(code, lib().ou_last_error().decode())
    return code


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def device_name():
    buf = ctypes.create_string_buffer(256)
    _check(lib().ou_device_name(buf, 256))
    return buf.value.decode()


class Quadrature:
    """Angular quadrature and moment operators (RTQuadrature + GLC/LS).

    kind LS/GLC take `sn`; PRODUCT_GLC takes n_polar x n_azim.  dim=3
    deploys eight octants.  Arrays: omega (n_dir,3), weight (n_dir),
    M2D (n_dir,n_mom), D2M (n_mom,n_dir), mom_l (n_mom),
    most_normal (2*dim, n_dir).
    """

    def __init__(self, kind, dim, sn=0, n_polar=0, n_azim=0, L_max=0, galerkin=False,
                 weight_sum=4.0 * np.pi):
        h = ctypes.c_void_p()
        _check(lib().ou_quadrature_create(int(kind), int(dim), int(sn), int(n_polar),
                                          int(n_azim), int(L_max), int(bool(galerkin)),
                                          ctypes.c_double(weight_sum), ctypes.byref(h)))
        try:
            nd = lib().ou_quadrature_n_dir(h)
            nm = lib().ou_quadrature_n_mom(h)
            self.omega = np.zeros((nd, 3))
            self.weight = np.zeros(nd)
            self.M2D = np.zeros((nd, nm))
            self.D2M = np.zeros((nm, nd))
            self.mom_l = np.zeros(nm, dtype=np.int32)
            self.most_normal = np.zeros((2 * dim, nd), dtype=np.uint8)
            _check(lib().ou_quadrature_copy(h, _d(self.omega), _d(self.weight), _d(self.M2D),
                                            _d(self.D2M), _i(self.mom_l),
                                            self.most_normal.ctypes.data_as(_up)))
        finally:
            lib().ou_quadrature_destroy(h)
        self.dim, self.weight_sum = dim, weight_sum
        self.n_dir, self.n_mom = nd, nm

    def get_n_dir(self):
        return self.n_dir

    def get_n_mom(self):
        return self.n_mom

    def get_l(self, i):
        return int(self.mom_l[i])

    def get_omega(self, idir):
        return self.omega[idir]

    def is_most_normal_direction(self, face, idir):
        return bool(self.most_normal[face, idir])


def fe_matrices(dim, p, h):
    """M, G (dim,N,N), E (2dim,N,N), U (2dim,N,N), normals (2dim,3) of one cell."""
    N = (p + 1) ** dim
    M = np.zeros((N, N))
    G = np.zeros((dim, N, N))
    E = np.zeros((2 * dim, N, N))
    U = np.zeros((2 * dim, N, N))
    nrm = np.zeros((2 * dim, 3))
    hh = np.ascontiguousarray(list(h) + [1.0] * (3 - len(h)), dtype=np.float64)
    _check(lib().ou_fe_matrices(dim, p, _d(hh), _d(M), _d(G), _d(E), _d(U), _d(nrm)))
    return M, G, E, U, nrm


class DeviceBuffer:
    """Device memory owned through the library (no second HIP runtime needed)."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        _check(lib().ou_device_alloc(ctypes.c_size_t(self.nbytes), ctypes.byref(self.ptr)))

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(lib().ou_memcpy_htod(self.ptr, arr.ctypes.data_as(_vp), arr.nbytes))

    def download(self, arr):
        assert arr.nbytes <= self.nbytes and arr.flags.c_contiguous
        _check(lib().ou_memcpy_dtoh(arr.ctypes.data_as(_vp), self.ptr, arr.nbytes))
        return arr

    def free(self):
        if self.ptr:
            lib().ou_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class RadiativeTransfer:
    """The transport operator y = x - D2M L^-1 M2D S(x) on the GPU.

    Mirrors RadiativeTransfer<dim,tensor_dim> (src/RadiativeTransfer.hh:46-127):
    Apply, Label, OperatorDomainMap/RangeMap sizes, set_group,
    compute_scattering_source and the sweep.  Vectors are FP64 numpy arrays
    of domain_size() entries in the moment-major layout x[m*n_dofs + dof].
    """

    def __init__(self, dim, fe_order, n_cells, cell_size, quad, sigma_t, sigma_s,
                 material_id, source=None, weight_sum=None, bc=None, inc_flux=None,
                 precision=FP64):
        self._keep = []

        def keep(a, dt):
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a

        ds = ProblemDesc()
        ds.dim, ds.fe_order = int(dim), int(fe_order)
        nc = list(n_cells) + [1] * (3 - len(n_cells))
        hs = list(cell_size) + [1.0] * (3 - len(cell_size))
        for a in range(3):
            ds.n_cells[a] = int(nc[a])
            ds.cell_size[a] = float(hs[a])
        ds.n_dir = quad.n_dir
        ds.omega = _d(keep(quad.omega, np.float64))
        ds.n_mom = quad.n_mom
        ds.M2D = _d(keep(quad.M2D, np.float64))
        ds.D2M = _d(keep(quad.D2M, np.float64))
        ds.mom_l = _i(keep(quad.mom_l, np.int32))
        st = keep(sigma_t, np.float64).reshape(-1)
        ss = keep(sigma_s, np.float64).reshape(st.shape[0], -1)
        self._keep.append(ss)
        ds.n_mat, ds.L = st.shape[0], ss.shape[1] - 1
        ds.sigma_t, ds.sigma_s = _d(st), _d(ss)
        ds.material_id = _i(keep(material_id, np.int32))
        ds.source = _d(keep(source, np.float64)) if source is not None else None
        ds.weight_sum = float(quad.weight_sum if weight_sum is None else weight_sum)
        for f in range(6):
            ds.bc[f] = int(bc[f]) if bc is not None and f < len(bc) else BC_VACUUM
            ds.inc_flux[f] = float(inc_flux[f]) if inc_flux is not None and f < len(inc_flux) else 0.0
        ds.most_normal = keep(quad.most_normal, np.uint8).ctypes.data_as(_up)
        ds.precision = int(precision)
        self._desc = ds
        self._plan = ctypes.c_void_p()
        _check(lib().ou_plan_create(ctypes.byref(ds), ctypes.byref(self._plan)))
        self._keep = []
        self.n = int(lib().ou_domain_size(self._plan))
        self.n_dofs = int(lib().ou_n_dofs(self._plan))
        self.n_mom = quad.n_mom

    # --- Epetra_Operator surface -------------------------------------------
    def Label(self):
        return lib().ou_label(self._plan).decode()

    def domain_size(self):
        return self.n

    def range_size(self):
        return int(lib().ou_range_size(self._plan))

    def Apply(self, x):
        """y = x - D2M L^-1 M2D S(x).  x is one vector of domain_size() entries
        or an array [n_rhs][domain_size()] of columns (each column applied)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.ndim == 2:
            if x.shape[1] != self.n:
                raise OuError(-2, f"x has columns of {x.shape[1]} entries, operator expects {self.n}")
            y = np.empty_like(x)
            _check(lib().ou_apply_multi(self._plan, _d(x), _d(y), x.shape[0]))
            return y
        if x.size != self.n:
            raise OuError(-2, f"x has {x.size} entries, operator expects {self.n}")
        y = np.empty(self.n)
        _check(lib().ou_apply(self._plan, _d(x), _d(y)))
        return y

    def ApplyInverse(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(lib().ou_apply_inverse(self._plan, _d(x), _d(y)))
        return y

    def SetUseTranspose(self, flag):
        return _check(lib().ou_set_use_transpose(self._plan, int(bool(flag))))

    def UseTranspose(self):
        return bool(lib().ou_use_transpose(self._plan))

    def HasNormInf(self):
        return bool(lib().ou_has_norm_inf(self._plan))

    def NormInf(self):
        return float(lib().ou_norm_inf(self._plan))

    # --- sweep primitives ------------------------------------------------------
    def set_group(self, g):
        _check(lib().ou_set_group(self._plan, int(g)))

    def compute_scattering_source(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        _check(lib().ou_scatter_source(self._plan, _d(x)))

    def sweep(self, y=None, with_source=False):
        """y += D2M L^-1 (M2D S [+ source + inflow]); returns y."""
        y = np.zeros(self.n) if y is None else np.ascontiguousarray(y, dtype=np.float64)
        _check(lib().ou_sweep(self._plan, _d(y), int(bool(with_source))))
        return y

    def scatter_source_with(self, x, sigma_s, accumulate=False):
        """Stored scattering source (+)= sigma_s[mat][l(m)] M x_m with another
        group's table sigma_s (n_mat x (L+1)): ou_scatter_source_with."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        sig = np.ascontiguousarray(sigma_s, dtype=np.float64)
        _check(lib().ou_scatter_source_with(self._plan, _d(x), _d(sig), int(bool(accumulate))))

    def solve_rhs(self, rhs, solver=GMRES, tol=1e-8, max_it=1000, restart=30, phi0=None,
                  verbose=False, raise_on_fail=True):
        """Solves A phi = rhs (ou_solve_rhs) from phi0 (zero by default)."""
        rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        phi = np.zeros(self.n) if phi0 is None else np.array(phi0, dtype=np.float64)
        opts = SolverOpts(int(solver), float(tol), int(max_it), int(restart), int(verbose))
        info = SolveInfo()
        code = lib().ou_solve_rhs(self._plan, ctypes.byref(opts), _d(rhs), _d(phi),
                                  ctypes.byref(info))
        if code != 0 and (raise_on_fail or code != -6):
            _check(code)
        return phi, info.as_dict()

    # --- drivers ---------------------------------------------------------------
    def solve(self, solver=GMRES, tol=1e-8, max_it=1000, restart=30, phi0=None,
              verbose=False, raise_on_fail=True):
        phi = np.zeros(self.n) if phi0 is None else np.array(phi0, dtype=np.float64)
        opts = SolverOpts(int(solver), float(tol), int(max_it), int(restart), int(verbose))
        info = SolveInfo()
        code = lib().ou_solve(self._plan, ctypes.byref(opts), _d(phi), ctypes.byref(info))
        if code != 0 and (raise_on_fail or code != -6):
            _check(code)
        return phi, info.as_dict()

    def solve_learned(self, net=None, tol=1e-8, max_it=1000, restart=30, phi0=None,
                      verbose=False, raise_on_fail=True):
        """FGMRES(restart), right-preconditioned by the U-Net `net` (None: plain)."""
        phi = np.zeros(self.n) if phi0 is None else np.array(phi0, dtype=np.float64)
        opts = SolverOpts(FGMRES, float(tol), int(max_it), int(restart), int(verbose))
        info = SolveInfo()
        code = lib().ou_solve_learned(self._plan, net._net if net is not None else None,
                                      ctypes.byref(opts), _d(phi), ctypes.byref(info))
        if code != 0 and (raise_on_fail or code != -6):
            _check(code)
        return phi, info.as_dict()

    # --- learned-preconditioner data -----------------------------------------
    def coarse_features(self, grid, v=None):
        """(octant means [G^3][8] of v's moment 0 or None, material features [G^3][3])."""
        P0 = grid ** 3
        mat = np.empty(P0 * 3, dtype=np.float32)
        coarse = None
        if v is not None:
            v = np.ascontiguousarray(v, dtype=np.float64)
            if v.size != self.n:
                raise OuError(-2, f"v has {v.size} entries, operator expects {self.n}")
            coarse = np.empty((P0, 8))
            _check(lib().ou_coarse_features(self._plan, int(grid), _d(v), _d(coarse),
                                            mat.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        else:
            _check(lib().ou_coarse_features(self._plan, int(grid), None, None,
                                            mat.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return coarse, mat.reshape(P0, 3)

    def coarse_prolong(self, grid, corners, z=None):
        """z + P(corners): corner values [G^3][8] -> continuous moment-0 field."""
        z = np.zeros(self.n) if z is None else np.array(z, dtype=np.float64)
        c = np.ascontiguousarray(corners, dtype=np.float32)
        if c.size != 8 * grid ** 3:
            raise OuError(-2, f"corners must hold 8 * grid^3 values, got {c.size}")
        _check(lib().ou_coarse_prolong(self._plan, int(grid),
                                       c.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), _d(z)))
        return z

    def learned_pairs_krylov(self, grid, n_steps, stride=1, max_records=64, restart=30,
                             net=None):
        """Coarse pairs (R A v_j, R v_j) of plain GMRES's Arnoldi vectors, or with
        `net` the pairs (R A z_j, R z_j) of FGMRES preconditioned by it."""
        P0 = grid ** 3
        r = np.empty((max_records, P0, 8))
        e = np.empty((max_records, P0, 8))
        if net is not None:
            got = lib().ou_learned_pairs_fgmres(self._plan, net._net, int(grid), int(n_steps),
                                                int(restart), int(stride), int(max_records),
                                                _d(r), _d(e))
        else:
            got = lib().ou_learned_pairs_krylov(self._plan, int(grid), int(n_steps),
                                                int(restart), int(stride), int(max_records),
                                                _d(r), _d(e))
        if got < 0:
            _check(got)
        return r[:got], e[:got]

    def learned_pairs(self, grid, n_steps, record, e0=None):
        """Coarse pairs (R A e_k, R e_k) of e_{k+1} = e_k - A e_k for k in `record`."""
        mask = 0
        for k in record:
            if not 0 <= k < n_steps:
                raise ValueError("record steps must lie in [0, n_steps)")
            mask |= 1 << int(k)
        n_rec = bin(mask).count("1")
        P0 = grid ** 3
        r = np.empty((n_rec, P0, 8))
        e = np.empty((n_rec, P0, 8))
        if e0 is not None:
            e0 = np.ascontiguousarray(e0, dtype=np.float64)
            if e0.size != self.n:
                raise OuError(-2, f"e0 has {e0.size} entries, operator expects {self.n}")
        got = lib().ou_learned_pairs(self._plan, int(grid), _d(e0) if e0 is not None else None,
                                     int(n_steps), ctypes.c_ulonglong(mask), _d(r), _d(e))
        if got < 0:
            _check(got)
        return r[:got], e[:got]

    # --- sweep workgroup shape ---------------------------------------------------
    @property
    def sweep_shape(self):
        """(waves per workgroup, direction groups per octant)."""
        w, g = ctypes.c_int(), ctypes.c_int()
        _check(lib().ou_sweep_shape(self._plan, ctypes.byref(w), ctypes.byref(g)))
        return w.value, g.value

    @sweep_shape.setter
    def sweep_shape(self, shape):
        _check(lib().ou_set_sweep_shape(self._plan, int(shape[0]), int(shape[1])))

    @property
    def sweep_streams(self):
        """Launches per wavefront level, each on its own stream: 1, 2 (octant
        halves) or 4 (octant quarters)."""
        n = ctypes.c_int()
        _check(lib().ou_sweep_streams(self._plan, ctypes.byref(n)))
        return n.value

    @sweep_streams.setter
    def sweep_streams(self, n):
        _check(lib().ou_set_sweep_streams(self._plan, int(n)))

    @property
    def sweep_mode(self):
        """SWEEP_LEVELS (one launch per wavefront level), SWEEP_PERSISTENT
        (one persistent launch per sweep) or SWEEP_TILES (one launch per tile
        level: 2D Q1-Q2 with one material per 8x8 tile, 3D Q1 with one per
        4x4x4 tile, and, opt-in, 3D Q2 in FP32 with at most 4 moments).
        sweep_shape reads and sets the shape of the mode in use."""
        m = ctypes.c_int()
        _check(lib().ou_sweep_mode(self._plan, ctypes.byref(m)))
        return m.value

    @sweep_mode.setter
    def sweep_mode(self, mode):
        _check(lib().ou_set_sweep_mode(self._plan, int(mode)))

    def set_sweep_wait_limit(self, seconds):
        """Watchdog of the persistent sweep: a wait for an upwind level gives
        up after `seconds` (0: at the first unfinished check)."""
        _check(lib().ou_set_sweep_wait_limit(self._plan, ctypes.c_double(seconds)))

    def sweep_status(self):
        """Raises OuError(OU_ERR_TIMEOUT) if a persistent sweep gave up since
        the last check."""
        _check(lib().ou_sweep_status(self._plan))

    # --- device API --------------------------------------------------------------
    def apply_device(self, d_x, d_y):
        _check(lib().ou_apply_device(self._plan, d_x.ptr, d_y.ptr, None))

    def time_apply(self, d_x, d_y, warmup, steps):
        t = Timing()
        _check(lib().ou_time_apply(self._plan, d_x.ptr, d_y.ptr, int(warmup), int(steps),
                                   ctypes.byref(t)))
        return {"ms_total": t.ms_total, "ms_sweep_kernel": t.ms_sweep_kernel,
                "ms_other": t.ms_other, "n_sweep_launches": t.n_sweep_launches,
                "applications": t.applications}

    def close(self):
        if getattr(self, "_plan", None):
            lib().ou_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr_table(bufs):
    arr = (_vp * len(bufs))()
    for i, b in enumerate(bufs):
        arr[i] = b.ptr if isinstance(b, DeviceBuffer) else b
    return arr


class Batch:
    """Many RadiativeTransfer operators swept as one (ou_batch_*): merged
    wavefront levels, one launch per element-wise stage for all members, the
    whole sequence captured as a HIP graph and replayed.  waves/dir_groups
    None = autotuned.  The operators must outlive the batch."""

    def __init__(self, ops, waves=None, dir_groups=None):
        self.ops = list(ops)
        plans = (_vp * len(self.ops))(*[op._plan for op in self.ops])
        self._b = ctypes.c_void_p()
        _check(lib().ou_batch_create(plans, len(self.ops), int(waves or 0),
                                     int(dir_groups or 0), ctypes.byref(self._b)))

    def info(self):
        w, g, nl, nc = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
        _check(lib().ou_batch_info(self._b, ctypes.byref(w), ctypes.byref(g), ctypes.byref(nl),
                                   ctypes.byref(nc)))
        return {"waves": w.value, "dir_groups": g.value, "n_levels": nl.value,
                "n_chunks": nc.value}

    def set_shape(self, waves, dir_groups):
        _check(lib().ou_batch_set_shape(self._b, int(waves), int(dir_groups)))

    def set_graph(self, on):
        _check(lib().ou_batch_set_graph(self._b, int(bool(on))))

    def set_streams(self, n):
        """Launches per merged level: 2 (octant halves on two streams) or 1."""
        _check(lib().ou_batch_set_streams(self._b, int(n)))

    def set_sweep_mode(self, mode):
        """SWEEP_LEVELS (merged level launches), SWEEP_PERSISTENT (every
        member's levels in one persistent launch) or SWEEP_TILES (merged tile
        levels, when every member has a tile sweep)."""
        _check(lib().ou_batch_set_sweep_mode(self._b, int(mode)))

    @property
    def sweep_mode(self):
        m = ctypes.c_int()
        _check(lib().ou_batch_sweep_mode(self._b, ctypes.byref(m)))
        return m.value

    def apply_device(self, xs, ys, stream=None):
        _check(lib().ou_batch_apply_device(self._b, _ptr_table(xs), _ptr_table(ys), stream))

    def sweep_device(self, ys, with_source=False, stream=None):
        _check(lib().ou_batch_sweep_device(self._b, _ptr_table(ys), int(bool(with_source)),
                                           stream))

    def apply(self, xs):
        """Host convenience: list of numpy vectors in, list out."""
        dx = [DeviceBuffer(x.size * 8) for x in xs]
        dy = [DeviceBuffer(x.size * 8) for x in xs]
        try:
            for b, x in zip(dx, xs):
                b.upload(np.ascontiguousarray(x, dtype=np.float64))
            self.apply_device(dx, dy)
            _check(lib().ou_device_synchronize())
            out = []
            for b, x in zip(dy, xs):
                y = np.empty(x.size)
                b.download(y)
                out.append(y)
            return out
        finally:
            for b in dx + dy:
                b.free()

    def solve_si(self, tol=1e-8, max_it=1000, phis=None, verbose=False):
        """Source iteration of every member at once (zero initial guesses unless
        `phis` gives host vectors).  Returns (list of solutions, info) with
        per-member first-convergence iterations (-1: not converged) and last
        relative changes."""
        n = len(self.ops)
        bufs = []
        try:
            for i, op in enumerate(self.ops):
                b = DeviceBuffer(op.n * 8)
                x = np.zeros(op.n) if phis is None else np.ascontiguousarray(phis[i], np.float64)
                b.upload(x)
                bufs.append(b)
            its = np.zeros(n, dtype=np.int32)
            ch = np.zeros(n)
            total = ctypes.c_int()
            _check(lib().ou_batch_solve_si(self._b, _ptr_table(bufs), float(tol), int(max_it),
                                           _i(its), _d(ch), ctypes.byref(total),
                                           int(bool(verbose))))
            out = []
            for b, op in zip(bufs, self.ops):
                y = np.empty(op.n)
                b.download(y)
                out.append(y)
            return out, {"iterations": its.tolist(), "change": ch.tolist(),
                         "n_iterations": total.value}
        finally:
            for b in bufs:
                b.free()

    def tune(self, reps=3):
        entries = (TuneEntry * 64)()
        n = ctypes.c_int()
        _check(lib().ou_batch_tune(self._b, int(reps), entries, 64, ctypes.byref(n)))
        return [{"waves": e.waves, "dir_groups": e.dir_groups, "ms": e.ms, "mode": e.mode}
                for e in entries[:min(n.value, 64)]]

    def time_apply(self, xs, ys, warmup, steps):
        ms = ctypes.c_double()
        _check(lib().ou_batch_time_apply(self._b, _ptr_table(xs), _ptr_table(ys), int(warmup),
                                         int(steps), ctypes.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "_b", None):
            lib().ou_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class UNet:
    """The learned right preconditioner's network (ou_unet_*): a 3-level
    convolutional U-Net with 32 base channels on a grid^3 coarse cell grid,
    BF16 weights/activations with FP32 accumulation on MFMA.  Features are
    float32 [batch][grid^3][12] (r_hat[8], tau_hat, kappa_hat, g, unused),
    targets and predictions [batch][grid^3][8] (corner values)."""

    N_FEATURES = 12
    N_OUT = 8

    def __init__(self, grid=32, max_batch=8, seed=0):
        self._net = _vp()
        _check(lib().ou_unet_create(int(grid), int(max_batch), ctypes.c_ulonglong(seed),
                                    ctypes.byref(self._net)))
        self.grid = int(grid)
        self.max_batch = int(max_batch)
        self.n_params = int(lib().ou_unet_n_params(self._net))

    @staticmethod
    def _f(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def _feat(self, feat):
        feat = np.ascontiguousarray(feat, dtype=np.float32).reshape(
            -1, self.grid ** 3, self.N_FEATURES)
        return feat, feat.shape[0]

    @property
    def mode(self):
        """Prediction form: bit 0 adds the skip r_hat, bit 1 scales by g."""
        return int(lib().ou_unet_mode(self._net))

    @mode.setter
    def mode(self, m):
        _check(lib().ou_unet_set_mode(self._net, int(m)))

    def get_params(self):
        p = np.empty(self.n_params, dtype=np.float32)
        _check(lib().ou_unet_get_params(self._net, self._f(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise OuError(-2, f"expected {self.n_params} parameters, got {p.size}")
        _check(lib().ou_unet_set_params(self._net, self._f(p)))

    def predict(self, feat):
        feat, B = self._feat(feat)
        out = np.empty((B, self.grid ** 3, self.N_OUT), dtype=np.float32)
        _check(lib().ou_unet_predict(self._net, self._f(feat), B, self._f(out)))
        return out

    def loss_grad(self, feat, target):
        feat, B = self._feat(feat)
        target = np.ascontiguousarray(target, dtype=np.float32).reshape(B, -1)
        loss = ctypes.c_double()
        grad = np.empty(self.n_params, dtype=np.float32)
        _check(lib().ou_unet_loss_grad(self._net, self._f(feat), self._f(target), B,
                                       ctypes.byref(loss), self._f(grad)))
        return loss.value, grad

    def train_step(self, feat, target, lr):
        feat, B = self._feat(feat)
        target = np.ascontiguousarray(target, dtype=np.float32).reshape(B, -1)
        loss = ctypes.c_double()
        _check(lib().ou_unet_train_step(self._net, self._f(feat), self._f(target), B,
                                        ctypes.c_float(lr), ctypes.byref(loss)))
        return loss.value

    def close(self):
        if getattr(self, "_net", None):
            lib().ou_unet_destroy(self._net)
            self._net = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_refined(op64, op32, net=None, tol=1e-8, max_it=5000, restart=30, inner_tol=1e-3,
                  phi0=None, verbose=False, raise_on_fail=True):
    """FP64-residual refinement around FGMRES on the FP32 operator (ou_solve_refined)."""
    phi = np.zeros(op64.n) if phi0 is None else np.array(phi0, dtype=np.float64)
    opts = SolverOpts(FGMRES, float(tol), int(max_it), int(restart), int(verbose))
    info = SolveInfo()
    code = lib().ou_solve_refined(op64._plan, op32._plan, net._net if net is not None else None,
                                  ctypes.byref(opts), ctypes.c_double(inner_tol), _d(phi),
                                  ctypes.byref(info))
    if code != 0 and (raise_on_fail or code != -6):
        _check(code)
    return phi, info.as_dict()
