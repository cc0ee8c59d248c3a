"""The input-file readers (ou_parameters_*, ou_geometry_*, ou_material_properties_*)
on the CPU.

The fixtures under tests/golden/inp/ are the reference's own test inputs
(tests/*.inp); the expected values are those its tests require:
Parameters_test.cc:20-50, Geometry_test.cc:20-76, RTMaterialProperties_test.cc:18-50.
"""
import math
import os

import numpy as np
import pytest

import ouranos_amd as ou
from ouranos_amd import inp
from ouranos_amd import problems as P

HERE = os.path.dirname(os.path.abspath(__file__))
INP = os.path.join(HERE, "golden", "inp")


def fixture(name):
    return os.path.join(INP, name)


def test_parameters_fixture():
    p = inp.Parameters(fixture("parameters.inp"))
    assert p.get_dimension() == 2
    assert p.get_fe_order() == 4
    assert p.get_geometry_filename() == "geometry.inp"
    assert p.get_n_refinements() == 2
    assert p.get_refinement_factor() == 0.3
    assert p.get_coarsening_factor() == 0.1
    assert p.get_output_filename() == "output"
    assert p.get_xs_filename() == "xs.inp"
    assert p.get_solver_type() == ou.BICGSTAB
    assert p.get_max_outer_it() == 2
    assert p.get_max_inner_it() == 10
    assert p.get_outer_tolerance() == 1e-8
    assert p.get_inner_tolerance() == 1e-10
    assert p.get_sn_order() == 6
    assert p.get_quad_type() == ou.GLC
    assert p.get_weight_sum() == 2.0 * math.pi
    assert p.get_galerkin() is False
    assert p.get_bc_type(0) == inp.BC_VACUUM
    assert p.get_bc_type(1) == inp.BC_ISOTROPIC
    assert p.get_bc_type(2) == inp.BC_REFLECTIVE
    assert p.get_bc_type(3) == inp.BC_MOST_NORMAL
    assert p.get_inc_flux(1, 0) == 1.0
    assert p.get_inc_flux(1, 1) == 2.0
    assert p.get_inc_flux(3, 0) == 8.0
    assert p.get_inc_flux(3, 1) == 9.0
    # faces whose boundary takes no values read none
    assert p.get_inc_flux(0, 0) == 0.0 and p.get_inc_flux(2, 1) == 0.0
    assert p.get_n_src() == 1
    assert p.get_n_groups() == 2
    assert p.get_src(0, 0) == 4.5
    assert p.get_src(0, 1) == 8.9


def test_parameter_defaults():
    """An empty file gives every declared default (Parameters.cc:28-100)."""
    p = inp.Parameters(os.devnull)
    assert (p.get_dimension(), p.get_fe_order(), p.get_n_refinements()) == (2, 1, 0)
    assert p.get_geometry_filename() == "geometry.inp"
    assert p.get_xs_filename() == "xs.inp" and p.get_output_filename() == "output"
    assert p.get_solver_type() == ou.SI
    assert (p.get_max_outer_it(), p.get_max_inner_it()) == (1, 1)
    assert (p.get_outer_tolerance(), p.get_inner_tolerance()) == (1e-4, 1e-6)
    assert p.get_sn_order() == 2 and p.get_quad_type() == ou.GLC
    assert p.get_weight_sum() == 4.0 * math.pi and p.get_galerkin() is False
    assert [p.get_bc_type(f) for f in range(4)] == [inp.BC_VACUUM] * 4
    assert (p.get_n_src(), p.get_n_groups()) == (0, 1)


def test_geometry_2d_fixture():
    g = inp.Geometry(fixture("geometry_2D.inp"), 2)
    assert g.get_n_subdivisions(0) == 2
    assert g.get_n_subdivisions(1) == 3
    mats = {(0, 0): 0, (1, 0): 1, (0, 1): 0, (1, 1): 1, (0, 2): 0, (1, 2): 1}
    srcs = {(0, 0): 1, (1, 0): 0, (0, 1): 1, (1, 1): 0, (0, 2): 1, (1, 2): 0}
    for (i, j), m in mats.items():
        assert g.get_material_ids(i, j) == m
        assert g.get_source_ids(i, j) == srcs[(i, j)]
    assert g.get_n_materials() == 2
    assert g.get_bottom_left() == (0.3, 2.0) and g.get_upper_right() == (10.0, 20.0)
    assert g.n_cells == (2, 3)
    assert g.cell_size == pytest.approx(((10.0 - 0.3) / 2, 18.0 / 3))


def test_geometry_3d_fixture():
    g = inp.Geometry(fixture("geometry_3D.inp"), 3)
    assert [g.get_n_subdivisions(a) for a in range(3)] == [2, 3, 2]
    mat0 = {(0, 0): 0, (1, 0): 1, (0, 1): 0, (1, 1): 1, (0, 2): 0, (1, 2): 1}
    src0 = {(0, 0): 1, (1, 0): 0, (0, 1): 1, (1, 1): 0, (0, 2): 1, (1, 2): 0}
    src1 = {(0, 0): 2, (1, 0): 1, (0, 1): 2, (1, 1): 1, (0, 2): 2, (1, 2): 1}
    for (i, j), m in mat0.items():
        assert g.get_material_ids(i, j, 0) == m
        assert g.get_material_ids(i, j, 1) == 1 - m
        assert g.get_source_ids(i, j, 0) == src0[(i, j)]
        assert g.get_source_ids(i, j, 1) == src1[(i, j)]
    assert g.get_n_materials() == 2


def test_material_properties_fixture():
    mp = inp.RTMaterialProperties(fixture("rtmaterial_properties.inp"), 2, 3)
    assert mp.get_n_materials() == 2
    assert mp.get_n_groups() == 3
    assert mp.get_L_max() == 0
    st = {(0, 0): 1.0, (0, 1): 1.5, (0, 2): 2.0, (1, 0): 2.0, (1, 1): 2.5, (1, 2): 3.0}
    for (m, g), v in st.items():
        assert mp.get_sigma_t(m, g) == v
    ss = [1., 0.5, 0.4, 0.3, 2., 0.2, 0.1, 0.0, 3., 11., 10.5, 10.4, 10.3, 12., 10.2, 10.1,
          10.0, 13.]
    k = 0
    for m in range(2):
        for g in range(3):
            for gp in range(3):
                assert mp.get_sigma_s(m, g, gp, 0) == ss[k], (m, g, gp)
                k += 1


def test_reference_problem_from_its_parameter_files():
    """rt_parameters.inp (+ rt_geometry.inp, rt_xs.inp) is the reference's
    integration test problem: the same arrays as problems.reference_2d()."""
    prm, spec = inp.problem_from_parameters(fixture("rt_parameters.inp"))
    ref = P.reference_2d()
    assert prm.get_solver_type() == ou.GMRES
    assert (prm.get_max_inner_it(), prm.get_inner_tolerance()) == (100, 1e-9)
    assert (spec.dim, spec.p, tuple(spec.n)) == (ref.dim, ref.p, tuple(ref.n))
    assert spec.h == pytest.approx(ref.h, abs=0)
    for a in ("omega", "weight", "M2D", "D2M", "mom_l"):
        assert np.array_equal(getattr(spec.quad, a), getattr(ref.quad, a)), a
    assert np.array_equal(spec.sigma_t, ref.sigma_t)
    assert np.array_equal(spec.sigma_s, ref.sigma_s)
    assert np.array_equal(spec.mat, ref.mat)
    assert np.array_equal(spec.q, ref.q)
    assert spec.ws == ref.ws and list(spec.bc) == list(ref.bc)


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_refinement_keeps_parent_ids(tmp_path):
    g = inp.Geometry(_write(tmp_path, "g.inp", """
set Bottom left point = 0.,0.,0.
set Upper right point = 4.,6.,2.
set Number of subdivisions = 2,3,1
set Number of global refinements = 2
set Material IDs = 0,1,2,3,4,5
set Source IDs = 5,4,3,2,1,0
"""), 3)
    assert g.n_cells == (8, 12, 4)
    assert g.cell_size == pytest.approx((0.5, 0.5, 0.5))
    n = g.n_cells
    mat = g.cell_material.reshape(n[2], n[1], n[0])
    src = g.cell_source.reshape(n[2], n[1], n[0])
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                parent = (i >> 2) + 2 * (j >> 2)
                assert mat[k, j, i] == parent and src[k, j, i] == 5 - parent


def test_list_semantics_and_format(tmp_path):
    """Short lists read as zeros past their end (Parameters.cc:304-328);
    comments, continued lines and blank runs inside names are accepted."""
    path = _write(tmp_path, "xs.inp", """# cross sections
set L_max = 1   # two orders
set Sigma_t = 1., \\
              2.5
set   Sigma_s   = 0.5,0.25,0.125
""")
    mp = inp.RTMaterialProperties(path, 3, 1)
    assert [mp.get_sigma_t(m, 0) for m in range(3)] == [1.0, 2.5, 0.0]
    assert [mp.get_sigma_s(m, 0, 0, l) for m in range(3) for l in range(2)] == \
        [0.5, 0.25, 0.125, 0.0, 0.0, 0.0]


@pytest.mark.parametrize("text, where", [
    ("set Dimension = 4\n", "line 1"),                       # outside Integer(2,3)
    ("set FE order = two\n", "line 1"),
    ("\nset No such entry = 1\n", "line 2"),                 # undeclared
    ("subsection Radiative Transfer parameters\n  set Solver type = CG\nend\n", "line 2"),
    ("subsection Radiative Transfer parameters\n  set Sn order = 4\n", "not closed"),
    ("end\n", "without a subsection"),
    ("set Dimension 2\n", "without '='"),
    ("subsection Radiative Transfer parameters\n  set Xmin BC values = 1.,-2.\nend\n", "line 2"),
])
def test_invalid_parameter_files(tmp_path, text, where):
    with pytest.raises(ou.OuError) as e:
        inp.Parameters(_write(tmp_path, "p.inp", text))
    assert e.value.code == -1 and where in str(e.value)


def test_missing_files_and_unsupported_problems(tmp_path):
    with pytest.raises(ou.OuError) as e:
        inp.Parameters(str(tmp_path / "absent.inp"))
    assert "not found" in str(e.value)
    # the reference's parameter fixture: two groups, adaptive cycles, a
    # reflective face
    with pytest.raises(ou.OuError) as e:
        inp.problem_from_parameters(fixture("parameters.inp"))
    assert e.value.code == -4
    base = open(fixture("rt_parameters.inp")).read()
    refl = base.replace("set Ymin BC = vacuum", "set Ymin BC = reflective")
    with pytest.raises(ou.OuError) as e:
        inp.problem_from_parameters(_write(tmp_path, "refl.inp", refl.replace(
            "./tests/", INP + os.sep)))
    assert e.value.code == -4 and "reflective" in str(e.value)


def _read_vtk(path):
    """Sections of a legacy ASCII VTK file: points, cells, cell types and the
    named scalar fields."""
    tok = open(path).read().split()
    out, i = {"fields": {}}, 0
    while i < len(tok):
        t = tok[i]
        if t == "POINTS":
            n = int(tok[i + 1])
            out["points"] = np.array(tok[i + 3:i + 3 + 3 * n], float).reshape(n, 3)
            i += 3 + 3 * n
        elif t == "CELLS":
            n, m = int(tok[i + 1]), int(tok[i + 2])
            out["cells"] = np.array(tok[i + 3:i + 3 + m], int).reshape(n, -1)
            i += 3 + m
        elif t == "CELL_TYPES":
            n = int(tok[i + 1])
            out["types"] = np.array(tok[i + 2:i + 2 + n], int)
            i += 2 + n
        elif t in ("CELL_DATA", "POINT_DATA"):
            out[t] = int(tok[i + 1])
            i += 2
        elif t == "SCALARS":
            name = tok[i + 1]
            n = out["POINT_DATA"] if "POINT_DATA" in out else out["CELL_DATA"]
            out["fields"][name] = np.array(tok[i + 6:i + 6 + n], float)
            i += 6 + n
        else:
            i += 1
    return out


@pytest.mark.parametrize("dim, p", [(2, 1), (3, 2)])
def test_vtk_output(tmp_path, dim, p):
    """output_results (ouranos.cc:211-253): every cell with its own corners,
    the corner dof values per moment, the partitioning field and the record."""
    from ouranos_amd import ouranos
    n = (3, 2) if dim == 2 else (2, 2, 2)
    q = ou.Quadrature(ou.GLC, dim, sn=4, L_max=1, weight_sum=1.0)
    spec = P.ProblemSpec("vtk", dim, p, n, (0.5,) * dim, q, np.ones(1), np.zeros((1, 2)),
                         np.zeros(int(np.prod(n)), np.int32), np.ones(int(np.prod(n))), 1.0)
    phi = np.arange(spec.size, dtype=float) * 0.25
    path = ouranos.output_results(str(tmp_path / "out"), spec, [phi], origin=(1.0,) * dim)
    v = _read_vtk(path)
    nv, cells, N = 2 ** dim, int(np.prod(n)), (p + 1) ** dim
    assert v["points"].shape == (cells * nv, 3)
    assert (v["types"] == (9 if dim == 2 else 12)).all()
    assert (v["cells"][:, 0] == nv).all()
    assert np.array_equal(v["cells"][:, 1:].ravel(), np.arange(cells * nv))
    assert set(v["fields"]) == {"partitioning"} | {f"flux_000g_{m:03d}m" for m in range(q.n_mom)}
    assert (v["fields"]["partitioning"] == 0).all()
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1),
               (0, 1, 1)][:nv]
    for c in range(cells):
        ijk = (c % n[0], (c // n[0]) % n[1], c // (n[0] * n[1]))
        for k, cr in enumerate(corners):
            pt = v["points"][c * nv + k]
            for a in range(dim):
                assert pt[a] == pytest.approx(1.0 + 0.5 * (ijk[a] + cr[a]))
            local = sum(cr[a] * p * (p + 1) ** a for a in range(dim))
            for m in range(q.n_mom):
                want = phi[m * spec.n_dofs + c * N + local]
                assert v["fields"][f"flux_000g_{m:03d}m"][c * nv + k] == want
    assert open(str(tmp_path / "out.vtk")).read().split() == ["!NBLOCKS", "1", "out0000.vtk"]


def multigroup_files(tmp_path, coupling=True, solver="GMRES"):
    """The reference's integration-test problem with two energy groups:
    sigma_t = (1, 1.5), sigma_s(g -> g') in RTMaterialProperties order
    (mat, g, g', l), sources 1 and 0.5."""
    ss = "0.5,0.3,0.1,0.6" if coupling else "0.5,0.,0.,0.6"
    xs = _write(tmp_path, "xs2.inp", f"set L_max = 0\nset Sigma_t = 1.,1.5\nset Sigma_s = {ss}\n")
    params = open(fixture("rt_parameters.inp")).read()
    params = params.replace("./tests/rt_geometry.inp", fixture("rt_geometry.inp"))
    params = params.replace("./tests/rt_xs.inp", xs)
    params = params.replace("set Number of groups = 1", "set Number of groups = 2")
    params = params.replace("set Source intensity = 1.", "set Source intensity = 1.,0.5")
    params = params.replace("set Max outer iterations = 1", "set Max outer iterations = 100")
    params = params.replace("set Tolerance outer solver = 1e-4", "set Tolerance outer solver = 1e-11")
    params = params.replace("set Tolerance inner solver = 1e-9", "set Tolerance inner solver = 1e-12")
    params = params.replace("set Max inner iterations = 100", "set Max inner iterations = 400")
    params = params.replace("set Solver type = GMRES", f"set Solver type = {solver}")
    return _write(tmp_path, "p2.inp", params)


def test_group_problems_from_parameters(tmp_path):
    prm, specs, cross = inp.group_problems_from_parameters(multigroup_files(tmp_path))
    assert prm.get_n_groups() == 2 and len(specs) == 2
    assert [s.sigma_t[0] for s in specs] == [1.0, 1.5]
    assert [s.sigma_s[0, 0] for s in specs] == [0.5, 0.6]      # sigma_s(g -> g)
    assert (specs[0].q == 1.0).all() and (specs[1].q == 0.5).all()
    # cross[g][g'] = sigma_s(g' -> g): the file lists (g, g') = (0,1) as 0.3
    assert cross[1][0][0, 0] == 0.3 and cross[0][1][0, 0] == 0.1
    with pytest.raises(ou.OuError) as e:
        inp.problem_from_parameters(multigroup_files(tmp_path))
    assert e.value.code == -4


def test_refinement_overflow_is_an_error(tmp_path):
    with pytest.raises(ou.OuError) as e:
        inp.Geometry(_write(tmp_path, "g.inp", "set Number of global refinements = 40\n"), 2)
    assert e.value.code == -1 and "too many cells" in str(e.value)
