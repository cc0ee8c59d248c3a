"""Extracts the known-answer data of the reference's own tests into JSON.

Reads Rombur/ouranos sources as text (nothing is compiled, imported or run)
and writes small fixtures next to this script:

  reference_2d_solution.json  64 moment-0 values, tests/RadiativeTransfer_test.cc:84-148
  fecell_q1.json              Q1 matrices/normals, tests/FECell_test.cc:52-297
  quadrature_s4.json          LS/GLC S4 checks, tests/RTQuadrature_test.cc:21-165
  gauss_legendre_tables.json  polar nodes/weights of GLC.cc:60-175 (orders 2..16)
  ls_tables.json              level-symmetric octant tables of LS.cc:14-169 (orders 2..12)

Usage: python tests/golden/make_fixtures.py [/root/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _num(s):
    return float(s.replace("(", "").replace(")", ""))


def reference_solution(ref):
    src = open(os.path.join(ref, "tests/RadiativeTransfer_test.cc")).read()
    vals = {int(i): float(v) for i, v in re.findall(r"solution\[(\d+)\]\s*=\s*([0-9.eE+-]+);", src)}
    assert sorted(vals) == list(range(64))
    return {"source": "tests/RadiativeTransfer_test.cc:84-148", "tolerance": 1e-3,
            "problem": {"domain": [[0, 2], [0, 2]], "cells": [4, 4], "fe_order": 1,
                        "quadrature": "GLC S4", "weight_sum": 1.0, "sigma_t": 1.0,
                        "sigma_s": 0.5, "source": 1.0, "bc": "vacuum"},
            "numbering": "cell = Morton(i,j) with the x bit first; local dofs lexicographic, x fastest",
            "moment0": [vals[i] for i in range(64)]}


def fecell(ref):
    src = open(os.path.join(ref, "tests/FECell_test.cc")).read()
    one, two = src.split('TEST_CASE("FECell/two cells"')
    out = {"source": "tests/FECell_test.cc:45-297"}

    def grab(text, name):
        d = {}
        for i, j, v in re.findall(name + r"\[(\d)\]\[(\d)\]\s*-\s*\(([^;]*?)\)\)\s*<1e-12", text):
            d[f"{i}{j}"] = v
        for i, j in re.findall(name + r"\[(\d)\]\[(\d)\]==0\.", text):
            d[f"{i}{j}"] = "0"
        return d

    def ev(expr):
        expr = expr.replace("grad_ratio", "(1./12.)")
        return float(eval(expr, {"__builtins__": {}}))  # arithmetic literals only

    # mass
    mass = {}
    for i, j, v in re.findall(r"mass_matrix\[(\w)\]\[(\w)\]-\(([^)]*)\)\)", one):
        if i == "i":
            for k in range(4):
                mass[f"{k}{k}"] = ev(v)
        else:
            mass[f"{i}{j}"] = ev(v)
    out["mass"] = mass
    for axis, name in ((0, "x_grad_matrix"), (1, "y_grad_matrix")):
        g = {}
        for i, j, v in re.findall(name + r"\[(\d)\]\[(\d)\]-\(([^;]*?)\)\)<1e-12", one):
            g[f"{i}{j}"] = ev(v)
        out[f"grad_{axis}"] = g
    # downwind matrices of the isolated cell, face by face
    faces = re.split(r"// (Left|Right|Bottom|Top) face", one.split("// Check the downwind matrices")[1].split("// Check the upwind matrices")[0])
    dw = {}
    for k in range(1, len(faces), 2):
        f = ["Left", "Right", "Bottom", "Top"].index(faces[k])
        body = faces[k + 1]
        d = {}
        for i, j, v in re.findall(r"upwind_matrix\[(\d)\]\[(\d)\]-\(([^;]*?)\)\)<1e-12", body):
            d[f"{i}{j}"] = ev(v)
        for i, j in re.findall(r"upwind_matrix\[(\d)\]\[(\d)\]==0\.", body):
            d[f"{i}{j}"] = 0.0
        for r in re.findall(r"REQUIRE\(upwind_matrix\[(\d)\]\[i\]==0\.\)", body):
            for c in range(4):
                d[f"{r}{c}"] = 0.0
        dw[str(f)] = d
    out["downwind_isolated"] = dw
    faces = re.split(r"// (Left|Right|Bottom|Top) face", two.split("// Check the upwind matrices")[1])
    uw = {}
    for k in range(1, len(faces), 2):
        f = ["Left", "Right", "Bottom", "Top"].index(faces[k])
        body = faces[k + 1]
        d = {}
        for i, j, v in re.findall(r"upwind_matrix\[(\d)\]\[(\d)\]-\(([^;]*?)\)\)<1e-12", body):
            d[f"{i}{j}"] = ev(v)
        for i, j in re.findall(r"upwind_matrix\[(\d)\]\[(\d)\]==0\.", body):
            d[f"{i}{j}"] = 0.0
        for r in re.findall(r"REQUIRE\(upwind_matrix\[(\d)\]\[i\]==0\.\)", body):
            for c in range(4):
                d[f"{r}{c}"] = 0.0
        uw[str(f)] = d
    out["upwind_middle_of_3x3"] = uw
    out["normals"] = [[-1, 0], [1, 0], [0, -1], [0, 1]]
    out["upwind_isolated_all_zero"] = True
    return out


def quadrature(ref):
    src = open(os.path.join(ref, "tests/RTQuadrature_test.cc")).read()
    ls, glc = src.split('TEST_CASE("RTQuadrature/GLC"')
    mn = [[False] * 12 for _ in range(4)]
    for f, d, v in re.findall(r"is_most_normal_direction\((\d),(\d+)\)==(true|false)", ls):
        mn[int(f)][int(d)] = v == "true"
    ls_l = [int(v) for _, v in re.findall(r"get_l\((\d+)\)==(\d+)", ls)]
    glc_l = [int(v) for _, v in re.findall(r"get_l\((\d+)\)==(\d+)", glc)]
    ls_om = [float(v) for v in re.findall(r"omega\[\d\] = ([0-9.]+);", ls)]
    glc_om = [float(v) for v in re.findall(r"omega\[\d\] = ([0-9.]+);", glc)]
    return {"source": "tests/RTQuadrature_test.cc:21-165",
            "ls_s4_galerkin": {"sn": 4, "L_max": 4, "weight_sum": "4pi", "n_dir": 12,
                               "n_mom": 12, "most_normal": mn, "l": ls_l, "omega0": ls_om,
                               "galerkin_identity_tol": 1e-12},
            "glc_s4": {"sn": 4, "L_max": 4, "weight_sum": "4pi", "n_dir": 12, "n_mom": 15,
                       "l": glc_l, "omega0": glc_om, "tol": 1e-12}}


def gl_tables(ref):
    src = open(os.path.join(ref, "src/GLC.cc")).read()
    body = src.split("build_gauss_legendre_quadrature")[2]
    out = {"source": "src/GLC.cc:60-175"}
    for case in re.finditer(r"case (\d+):\s*\{(.*?)break;", body, re.S):
        sn = int(case.group(1))
        if sn > 16:
            continue
        nodes = {int(i): v for i, v in re.findall(r"nodes\[(\d+)\]\s*=\s*([0-9.]+)", case.group(2))}
        wts = {int(i): v for i, v in re.findall(r"weight\[(\d+)\]\s*=\s*([0-9.]+)", case.group(2))}
        # keep 20 significant digits: the tables print far more than FP64 holds
        out[str(sn)] = {"nodes": [float(nodes[i][:22]) for i in sorted(nodes)],
                        "weights": [float(wts[i][:22]) for i in sorted(wts)]}
    return out


def ls_tables(ref):
    """The level-symmetric octant data of LS::build_octant (LS.cc:14-169),
    orders 2..12, as printed: the distinct direction cosines direction[k],
    the weight classes w_k, and weight(i) of every octant direction i in the
    order LS::compute_omega lays them out."""
    src = open(os.path.join(ref, "src/LS.cc")).read()
    body = src.split("void LS::build_octant()")[1].split("void LS::compute_omega")[0]
    out = {"source": "src/LS.cc:14-169"}
    for case in re.finditer(r"case (\d+)\s*:\s*\{(.*?)break;", body, re.S):
        sn = int(case.group(1))
        if sn > 12:
            continue
        text = case.group(2)
        if sn == 2:
            d = re.search(r"direction\(([0-9.]+)\)", text).group(1)
            w = re.search(r"const double w\(([0-9./]+)\)", text).group(1)
            out["2"] = {"cosines": [float(d)], "classes": {"w": float(eval(w))},
                        "weights": ["w"]}
            continue
        cos = {int(i): float(v) for i, v in re.findall(r"direction\[(\d+)\]\s*=\s*([0-9.]+)", text)}
        classes = {}
        for name, v in re.findall(r"const double (w(?:_\d+)?)\(([0-9./]+)\)", text):
            classes[name] = float(eval(v, {"__builtins__": {}}))  # decimal or a/b literals only
        wts = {int(i): name for i, name in re.findall(r"weight\((\d+)\)\s*=\s*(w(?:_\d+)?);", text)}
        out[str(sn)] = {"cosines": [cos[i] for i in sorted(cos)], "classes": classes,
                        "weights": [wts[i] for i in sorted(wts)]}
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    fixtures = {
        "reference_2d_solution.json": reference_solution(ref),
        "fecell_q1.json": fecell(ref),
        "quadrature_s4.json": quadrature(ref),
        "gauss_legendre_tables.json": gl_tables(ref),
        "ls_tables.json": ls_tables(ref),
    }
    for name, data in fixtures.items():
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(data, f, indent=1)
        print("wrote", name)


if __name__ == "__main__":
    main()
