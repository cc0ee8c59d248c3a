"""Test helpers: hand a ProblemSpec to the CPU oracle, reference numbering."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_golden(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def oracle_problem(O, spec, quad=None):
    """The oracle problem for the same arrays as `spec` (same quadrature)."""
    q = quad or spec.quad
    return O.Problem(spec.dim, spec.p, spec.n, spec.h, q.omega, q.M2D, q.D2M, q.mom_l,
                     spec.sigma_t, spec.sigma_s, spec.mat, q=spec.q, ws=spec.ws,
                     bc=spec.bc, inc=spec.inc, most_normal_flags=q.most_normal)


def morton_2d(i, j, bits):
    c = 0
    for b in range(bits):
        c |= ((i >> b) & 1) << (2 * b) | ((j >> b) & 1) << (2 * b + 1)
    return c


def to_reference_numbering(x, n, n_loc):
    """Moment-0 block of x (cell = i + n*j) in the reference's deal.II numbering
    (cells in Morton order, x bit first; tests/RadiativeTransfer_test.cc)."""
    bits = int(np.log2(n))
    out = np.zeros(n * n * n_loc)
    for j in range(n):
        for i in range(n):
            c, mc = i + n * j, morton_2d(i, j, bits)
            out[mc * n_loc:(mc + 1) * n_loc] = x[c * n_loc:(c + 1) * n_loc]
    return out


def rel_err(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
