"""The reference's input files through the C ABI.

Mirrors of the reference's readers, with its getter names:
* `Parameters(path)`: src/Parameters.hh:40-104 (file format Parameters.cc:28-100);
* `Geometry(path, dim)`: the reading half of src/Geometry.hh:39-52 (Geometry.cc:85-128);
* `RTMaterialProperties(path, n_materials, n_groups)`: src/RTMaterialProperties.hh:31-46.

`problem_from_parameters(path)` assembles what ouranos.cc:255-330 builds from a
parameter file (geometry, cross sections, quadrature) into a ProblemSpec that
makes a plan.  Geometry and cross-section file names are taken as written
(relative to the working directory, as the reference does); when no such
file exists, the same name relative to the parameter file's directory, then
its last component in that directory, are tried.
"""
import ctypes
import os

import numpy as np

from . import (BC_ISOTROPIC, BC_MOST_NORMAL, BC_VACUUM, GLC, LS, OuError, Quadrature, _check,
               lib)
from .problems import ProblemSpec

BC_REFLECTIVE = 3


class ParametersInfo(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("fe_order", ctypes.c_int),
                ("n_refinements", ctypes.c_int), ("refinement_factor", ctypes.c_double),
                ("coarsening_factor", ctypes.c_double), ("solver_type", ctypes.c_int),
                ("max_outer_it", ctypes.c_int), ("max_inner_it", ctypes.c_int),
                ("outer_tol", ctypes.c_double), ("inner_tol", ctypes.c_double),
                ("sn_order", ctypes.c_int), ("quad_type", ctypes.c_int),
                ("weight_sum", ctypes.c_double), ("galerkin", ctypes.c_int),
                ("bc_type", ctypes.c_int * 6), ("n_groups", ctypes.c_int),
                ("n_src", ctypes.c_int)]


class GeometryInfo(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("bottom_left", ctypes.c_double * 3),
                ("upper_right", ctypes.c_double * 3), ("n_subdivisions", ctypes.c_int * 3),
                ("n_global_refinements", ctypes.c_int), ("n_materials", ctypes.c_int),
                ("n_cells", ctypes.c_int * 3), ("cell_size", ctypes.c_double * 3)]


def _bind():
    L = lib()
    if getattr(L, "_inp_bound", False):
        return L
    vp, ip, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    L.ou_parameters_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    L.ou_parameters_destroy.argtypes = [vp]
    L.ou_parameters_destroy.restype = None
    L.ou_parameters_get_info.argtypes = [vp, ctypes.POINTER(ParametersInfo)]
    L.ou_parameters_get_filename.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, ip]
    L.ou_parameters_get_inc_flux.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.ou_parameters_get_src.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.ou_geometry_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(vp)]
    L.ou_geometry_destroy.argtypes = [vp]
    L.ou_geometry_destroy.restype = None
    L.ou_geometry_get_info.argtypes = [vp, ctypes.POINTER(GeometryInfo)]
    L.ou_geometry_get_ids.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip, ip]
    L.ou_geometry_cell_ids.argtypes = [vp, ip, ip]
    L.ou_material_properties_read.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(vp)]
    L.ou_material_properties_destroy.argtypes = [vp]
    L.ou_material_properties_destroy.restype = None
    L.ou_material_properties_get_L_max.argtypes = [vp, ip]
    L.ou_material_properties_get_sigma_t.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.ou_material_properties_get_sigma_s.argtypes = [vp, ctypes.c_int, ctypes.c_int,
                                                     ctypes.c_int, ctypes.c_int, dp]
    L._inp_bound = True
    return L


class Parameters:
    """Parameters(parameters_filename) (src/Parameters.cc:10-26)."""

    def __init__(self, path):
        L = _bind()
        self.path = str(path)
        h = ctypes.c_void_p()
        _check(L.ou_parameters_read(self.path.encode(), ctypes.byref(h)))
        try:
            info = ParametersInfo()
            _check(L.ou_parameters_get_info(h, ctypes.byref(info)))
            self._info = info
            self._names = [self._filename(L, h, w) for w in range(3)]
            d, ng = info.dim, info.n_groups
            self._inc = np.zeros((2 * d, ng))
            v = ctypes.c_double()
            for f in range(2 * d):
                for g in range(ng):
                    _check(L.ou_parameters_get_inc_flux(h, f, g, ctypes.byref(v)))
                    self._inc[f, g] = v.value
            self._src = np.zeros((info.n_src, ng))
            for i in range(info.n_src):
                for g in range(ng):
                    _check(L.ou_parameters_get_src(h, i, g, ctypes.byref(v)))
                    self._src[i, g] = v.value
        finally:
            L.ou_parameters_destroy(h)

    @staticmethod
    def _filename(L, h, which):
        need = ctypes.c_int()
        _check(L.ou_parameters_get_filename(h, which, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        _check(L.ou_parameters_get_filename(h, which, buf, need.value, None))
        return buf.value.decode()

    def get_dimension(self):
        return self._info.dim

    def get_fe_order(self):
        return self._info.fe_order

    def get_geometry_filename(self):
        return self._names[0]

    def get_xs_filename(self):
        return self._names[1]

    def get_output_filename(self):
        return self._names[2]

    def get_n_refinements(self):
        return self._info.n_refinements

    def get_refinement_factor(self):
        return self._info.refinement_factor

    def get_coarsening_factor(self):
        return self._info.coarsening_factor

    def get_solver_type(self):
        """SI, GMRES or BICGSTAB (the package constants)."""
        return self._info.solver_type

    def get_max_outer_it(self):
        return self._info.max_outer_it

    def get_max_inner_it(self):
        return self._info.max_inner_it

    def get_outer_tolerance(self):
        return self._info.outer_tol

    def get_inner_tolerance(self):
        return self._info.inner_tol

    def get_sn_order(self):
        return self._info.sn_order

    def get_quad_type(self):
        """LS or GLC (the package constants)."""
        return self._info.quad_type

    def get_weight_sum(self):
        return self._info.weight_sum

    def get_galerkin(self):
        return bool(self._info.galerkin)

    def get_bc_type(self, face):
        if not 0 <= face < 2 * self._info.dim:
            raise IndexError(face)
        return self._info.bc_type[face]

    def get_inc_flux(self, face, group):
        return float(self._inc[face, group])

    def get_n_src(self):
        return self._info.n_src

    def get_n_groups(self):
        return self._info.n_groups

    def get_src(self, i, g):
        return float(self._src[i, g])


class Geometry:
    """Geometry<dim>(geometry_filename, fe): the file and the mesh it
    describes (src/Geometry.cc:10-128), without deal.II's triangulation."""

    def __init__(self, path, dim):
        L = _bind()
        self.path = str(path)
        h = ctypes.c_void_p()
        _check(L.ou_geometry_read(self.path.encode(), int(dim), ctypes.byref(h)))
        try:
            info = GeometryInfo()
            _check(L.ou_geometry_get_info(h, ctypes.byref(info)))
            self._info = info
            self.dim = info.dim
            ns = [info.n_subdivisions[a] for a in range(self.dim)]
            n_coarse = int(np.prod(ns))
            self._mat = np.zeros(n_coarse, np.int32)
            self._src = np.zeros(n_coarse, np.int32)
            m, s = ctypes.c_int(), ctypes.c_int()
            for c in range(n_coarse):
                i, j = c % ns[0], (c // ns[0]) % ns[1]
                k = c // (ns[0] * ns[1]) if self.dim == 3 else 0
                _check(L.ou_geometry_get_ids(h, i, j, k, ctypes.byref(m), ctypes.byref(s)))
                self._mat[c], self._src[c] = m.value, s.value
            self.n_cells = tuple(info.n_cells[a] for a in range(self.dim))
            self.cell_size = tuple(info.cell_size[a] for a in range(self.dim))
            total = int(np.prod(self.n_cells))
            self.cell_material = np.zeros(total, np.int32)
            self.cell_source = np.zeros(total, np.int32)
            ip = ctypes.POINTER(ctypes.c_int)
            _check(L.ou_geometry_cell_ids(h, self.cell_material.ctypes.data_as(ip),
                                          self.cell_source.ctypes.data_as(ip)))
        finally:
            L.ou_geometry_destroy(h)

    def _coarse(self, i, j, k):
        ns = self._info.n_subdivisions
        return i + ns[0] * (j + ns[1] * k)

    def get_n_subdivisions(self, i):
        return self._info.n_subdivisions[i]

    def get_n_materials(self):
        return self._info.n_materials

    def get_material_ids(self, i, j, k=0):
        return int(self._mat[self._coarse(i, j, k)])

    def get_source_ids(self, i, j, k=0):
        return int(self._src[self._coarse(i, j, k)])

    def get_n_global_refinements(self):
        return self._info.n_global_refinements

    def get_bottom_left(self):
        return tuple(self._info.bottom_left[a] for a in range(self.dim))

    def get_upper_right(self):
        return tuple(self._info.upper_right[a] for a in range(self.dim))


class RTMaterialProperties:
    """RTMaterialProperties(filename, n_materials, n_groups)
    (src/RTMaterialProperties.cc:10-54)."""

    def __init__(self, path, n_materials, n_groups):
        L = _bind()
        self.path = str(path)
        self.n_materials, self.n_groups = int(n_materials), int(n_groups)
        h = ctypes.c_void_p()
        _check(L.ou_material_properties_read(self.path.encode(), self.n_materials,
                                             self.n_groups, ctypes.byref(h)))
        try:
            lm = ctypes.c_int()
            _check(L.ou_material_properties_get_L_max(h, ctypes.byref(lm)))
            self.L_max = lm.value
            nm, ng, nl = self.n_materials, self.n_groups, self.L_max + 1
            self._st = np.zeros((nm, ng))
            self._ss = np.zeros((nm, ng, ng, nl))
            v = ctypes.c_double()
            for m in range(nm):
                for g in range(ng):
                    _check(L.ou_material_properties_get_sigma_t(h, m, g, ctypes.byref(v)))
                    self._st[m, g] = v.value
                    for gp in range(ng):
                        for l in range(nl):
                            _check(L.ou_material_properties_get_sigma_s(h, m, g, gp, l,
                                                                        ctypes.byref(v)))
                            self._ss[m, g, gp, l] = v.value
        finally:
            L.ou_material_properties_destroy(h)

    def get_n_materials(self):
        return self.n_materials

    def get_n_groups(self):
        return self.n_groups

    def get_L_max(self):
        return self.L_max

    def get_sigma_t(self, material_id, g):
        return float(self._st[material_id, g])

    def get_sigma_s(self, material_id, g, gp, mom):
        return float(self._ss[material_id, g, gp, mom])


def resolve(name, parameters_path):
    """A file named in a parameter file: as written, else relative to the
    parameter file's directory, else its last component there."""
    base = os.path.dirname(os.path.abspath(parameters_path))
    for cand in (name, os.path.join(base, name), os.path.join(base, os.path.basename(name))):
        if os.path.exists(cand):
            return cand
    return name


def group_problems_from_parameters(path):
    """The problem a parameter file describes (ouranos.cc:255-330): its
    geometry and cross-section files, the LS or GLC quadrature of its Sn
    order with L_max from the cross sections, one ProblemSpec per energy
    group (that group's sigma_t, sigma_s(g->g), source and inflow), and the
    group-to-group scattering tables: cross[g][gp][mat][l] = sigma_s(gp->g)
    (RTMaterialProperties.hh:92-101 with (gp, g) as in
    RadiativeTransfer.cc:869-870).

    Returns (Parameters, [ProblemSpec per group], cross).  Deviations, each
    explicit: adaptive refinement cycles and reflective boundaries raise
    OuError(UNSUPPORTED); a cell whose source id has no intensity in the
    parameter file gets no source (the reference reads past its table)."""
    prm = Parameters(path)
    dim, ng = prm.get_dimension(), prm.get_n_groups()
    if prm.get_n_refinements() != 0:
        raise OuError(-4, "adaptive refinement cycles are not supported")
    geo = Geometry(resolve(prm.get_geometry_filename(), path), dim)
    mp = RTMaterialProperties(resolve(prm.get_xs_filename(), path), geo.get_n_materials(), ng)
    quad = Quadrature(prm.get_quad_type(), dim, sn=prm.get_sn_order(), L_max=mp.get_L_max(),
                      galerkin=prm.get_galerkin(), weight_sum=prm.get_weight_sum())
    nm, nl = geo.get_n_materials(), mp.get_L_max() + 1
    bc = [BC_VACUUM] * 6
    for f in range(2 * dim):
        bc[f] = prm.get_bc_type(f)
        if bc[f] == BC_REFLECTIVE:
            raise OuError(-4, "reflective boundaries are not implemented (the reference "
                              "parses them but never applies them)")
    sid = geo.cell_source
    specs = []
    for g in range(ng):
        sigma_t = np.array([mp.get_sigma_t(m, g) for m in range(nm)])
        sigma_s = np.array([[mp.get_sigma_s(m, g, g, l) for l in range(nl)] for m in range(nm)])
        intensity = np.array([prm.get_src(i, g) for i in range(prm.get_n_src())])
        q = np.zeros(len(sid))
        known = sid < len(intensity)
        q[known] = intensity[sid[known]]
        inc = [0.0] * 6
        for f in range(2 * dim):
            if bc[f] in (BC_ISOTROPIC, BC_MOST_NORMAL):
                inc[f] = prm.get_inc_flux(f, g)
        name = os.path.basename(str(path)) + (f"[g{g}]" if ng > 1 else "")
        specs.append(ProblemSpec(name, dim, prm.get_fe_order(), geo.n_cells, geo.cell_size,
                                 quad, sigma_t, sigma_s, geo.cell_material.copy(), q,
                                 prm.get_weight_sum(), bc=list(bc), inc=inc))
    cross = [[np.array([[mp.get_sigma_s(m, gp, g, l) for l in range(nl)] for m in range(nm)])
              for gp in range(ng)] for g in range(ng)]
    return prm, specs, cross


def problem_from_parameters(path):
    """group_problems_from_parameters for a one-group file: (Parameters,
    ProblemSpec).  Several groups raise OuError(UNSUPPORTED) here; the driver
    (ouranos_amd.ouranos) solves them with the outer group iteration."""
    prm, specs, _ = group_problems_from_parameters(path)
    if len(specs) != 1:
        raise OuError(-4, "the parameter file has several energy groups: use "
                          "group_problems_from_parameters")
    return prm, specs[0]


__all__ = ["Parameters", "Geometry", "RTMaterialProperties", "problem_from_parameters",
           "group_problems_from_parameters",
           "resolve", "BC_REFLECTIVE", "BC_VACUUM", "BC_ISOTROPIC", "BC_MOST_NORMAL", "LS", "GLC"]
