"""Synthetic problem builders for the engine's workloads.

Each builder returns a `ProblemSpec`: the arrays of one DG-Sn problem on a
structured mesh (materials drawn per block of cells from a SplitMix64 stream,
so every host produces the same bits).  `spec.operator(precision)` builds the
GPU operator; tests hand the same arrays to the CPU oracle.

Workloads (sizes as named in the benchmark description; `scale` shrinks the
mesh while keeping the block structure, for tests):
  reference_2d  the 4x4 Q1 GLC-S4 problem of the reference's integration test
                (tests/rt_parameters.inp, rt_geometry.inp, rt_xs.inp)
  uniform_2d    2D unit square 32x32 Q1, LS-S8, sigma_t=1, c=0.5, q=1, vacuum
  hetero_2d     2D 256x256 Q2, product GLC 16x32 (mu>0 half: 256 directions),
                L=3 Henyey-Greenstein, 16x16-cell blocks
  batch_3d      3D n^3 Q1, LS-S12 (168 directions), 8^3-cell blocks
  thick_3d      3D 32^3 Q1, LS-S8 (80 directions), 4^3-cell blocks, c near 1
  headline_3d   3D 128^3 Q2, product GLC 12x24 (288 directions), L=1,
                16^3-cell blocks, c in [0.99, 0.9999]
"""
from dataclasses import dataclass, field

import numpy as np

from . import FP64, GLC, LS, PRODUCT_GLC, Quadrature, RadiativeTransfer

_MASK = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) producing doubles in [0,1)."""

    def __init__(self, seed):
        self.state = seed & _MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo=0.0, hi=1.0):
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def log_uniform(self, lo, hi):
        return float(np.exp(self.uniform(np.log(lo), np.log(hi))))

    def array(self, n, lo=0.0, hi=1.0):
        return np.array([self.uniform(lo, hi) for _ in range(n)])


@dataclass
class ProblemSpec:
    name: str
    dim: int
    p: int
    n: tuple
    h: tuple
    quad: Quadrature
    sigma_t: np.ndarray
    sigma_s: np.ndarray      # (n_mat, L+1)
    mat: np.ndarray          # (n_cells,)
    q: np.ndarray            # (n_cells,)
    ws: float
    bc: list = field(default_factory=lambda: [0] * 6)
    inc: list = field(default_factory=lambda: [0.0] * 6)
    meta: dict = field(default_factory=dict)

    @property
    def n_cells(self):
        return int(np.prod(self.n))

    @property
    def n_loc(self):
        return (self.p + 1) ** self.dim

    @property
    def n_dofs(self):
        return self.n_cells * self.n_loc

    @property
    def size(self):
        return self.n_dofs * self.quad.n_mom

    def operator(self, precision=FP64):
        return RadiativeTransfer(self.dim, self.p, self.n, self.h, self.quad, self.sigma_t,
                                 self.sigma_s, self.mat, source=self.q, weight_sum=self.ws,
                                 bc=self.bc, inc_flux=self.inc, precision=precision)

    def updates_per_sweep(self):
        return self.n_dofs * self.quad.n_dir

    def algorithmic_bytes_per_sweep(self, real_bytes=4):
        """HBM bytes of one sweep by the fixed traffic formula: every interior
        face trace written and read once in the sweep precision (2 x 4 B per
        value in FP32), plus per octant batch a read of the moment source in
        the sweep precision and an FP64 read-modify-write of the moment
        accumulators (4 + 16 B per value in FP32).  real_bytes = 8 gives the
        same formula for the FP64 sweep."""
        n = list(self.n)
        f_int = 0
        for a in range(self.dim):
            prod = 1
            for b in range(self.dim):
                prod *= (n[b] - 1) if b == a else n[b]
            f_int += prod
        f_v = (self.p + 1) ** (self.dim - 1)
        n_batch = 4 if self.dim == 2 else 8
        return (self.quad.n_dir * f_int * f_v * 2 * real_bytes
                + n_batch * self.n_cells * self.n_loc * self.quad.n_mom * (real_bytes + 16))


def _block_map(n, block):
    """Cell -> block index for blocks of `block` cells per axis."""
    dim = len(n)
    nb = [max(1, (n[a] + block[a] - 1) // block[a]) for a in range(dim)]
    idx = np.indices(n[::-1]).reshape(dim, -1)[::-1]  # idx[a] = coordinate along axis a
    b = np.zeros(idx.shape[1], dtype=np.int64)
    stride = 1
    for a in range(dim):
        b += (idx[a] // block[a]) * stride
        stride *= nb[a]
    return b.astype(np.int32), int(np.prod(nb))


def reference_2d():
    """The reference integration test (tests/RadiativeTransfer_test.cc:30-153):
    [0,2]^2, 4x4 Q1, GLC S4, ws=1, sigma_t=1, sigma_s=0.5, q=1, vacuum."""
    quad = Quadrature(GLC, 2, sn=4, L_max=0, weight_sum=1.0)
    return ProblemSpec("reference_2d", 2, 1, (4, 4), (0.5, 0.5), quad, np.array([1.0]),
                       np.array([[0.5]]), np.zeros(16, np.int32), np.ones(16), 1.0)


def uniform_2d(n=32, ws=1.0):
    quad = Quadrature(LS, 2, sn=8, L_max=0, weight_sum=ws)
    return ProblemSpec("uniform_2d", 2, 1, (n, n), (1.0 / n, 1.0 / n), quad,
                       np.array([1.0]), np.array([[0.5]]), np.zeros(n * n, np.int32),
                       np.ones(n * n), ws)


def hetero_2d(n=256, block=16, seed=2, n_polar=16, n_azim=32, L=3, p=2, ws=4.0 * np.pi):
    quad = Quadrature(PRODUCT_GLC, 2, n_polar=n_polar, n_azim=n_azim, L_max=L, weight_sum=ws)
    mat, nb = _block_map((n, n), (block, block))
    rng = SplitMix64(seed)
    st, ss, qb = np.zeros(nb), np.zeros((nb, L + 1)), np.zeros(nb)
    for b in range(nb):
        st[b] = rng.log_uniform(0.1, 100.0)
        c = rng.uniform(0.0, 0.99)
        g = rng.uniform(0.0, 0.9)
        qb[b] = rng.uniform(0.0, 1.0)
        ss[b] = [c * st[b] * g ** l for l in range(L + 1)]
    return ProblemSpec("hetero_2d", 2, p, (n, n), (1.0 / n, 1.0 / n), quad, st, ss, mat,
                       qb[mat], ws, meta={"seed": seed, "block": block})


def batch_3d(n=32, block=8, seed=3, sn=12, ws=4.0 * np.pi, quad=None):
    quad = quad or Quadrature(LS, 3, sn=sn, L_max=0, weight_sum=ws)
    mat, nb = _block_map((n, n, n), (block,) * 3)
    rng = SplitMix64(seed)
    st, ss, qb = np.zeros(nb), np.zeros((nb, 1)), np.zeros(nb)
    for b in range(nb):
        st[b] = rng.log_uniform(0.1, 50.0)
        ss[b, 0] = rng.uniform(0.0, 0.95) * st[b]
        qb[b] = rng.uniform(0.0, 1.0)
    return ProblemSpec("batch_3d", 3, 1, (n, n, n), (1.0 / n,) * 3, quad, st, ss, mat,
                       qb[mat], ws, meta={"seed": seed, "block": block})


def batch_instances(n_instances=32, n_rhs=4, sizes=(16, 24, 32, 48, 64), block=8, seed=3,
                    sn=12, ws=4.0 * np.pi, quad=None):
    """The batched workload: n_instances independent unit cubes of n^3 Q1
    cells, n drawn uniformly from `sizes`, each with its own block materials
    (sigma_t log-U[0.1,50], c ~ U[0,0.95], isotropic scattering) and n_rhs
    block sources q ~ U[0,1].  Returns n_instances * n_rhs specs, instance
    major; the specs of one instance share everything but q.  Draw order per
    instance: n, then (sigma_t, c) per block, then q per block per RHS."""
    quad = quad or Quadrature(LS, 3, sn=sn, L_max=0, weight_sum=ws)
    rng = SplitMix64(seed)
    specs = []
    for i in range(n_instances):
        n = int(sizes[min(int(rng.uniform() * len(sizes)), len(sizes) - 1)])
        mat, nb = _block_map((n, n, n), (block,) * 3)
        st, ss = np.zeros(nb), np.zeros((nb, 1))
        for b in range(nb):
            st[b] = rng.log_uniform(0.1, 50.0)
            ss[b, 0] = rng.uniform(0.0, 0.95) * st[b]
        for r in range(n_rhs):
            qb = np.array([rng.uniform(0.0, 1.0) for _ in range(nb)])
            specs.append(ProblemSpec(f"batch_instance_{i}_rhs_{r}", 3, 1, (n, n, n),
                                     (1.0 / n,) * 3, quad, st, ss, mat, qb[mat], ws,
                                     meta={"seed": seed, "instance": i, "rhs": r,
                                           "block": block}))
    return specs


def thick_3d(n=32, block=4, seed=4, sn=8, ws=4.0 * np.pi, quad=None):
    quad = quad or Quadrature(LS, 3, sn=sn, L_max=0, weight_sum=ws)
    mat, nb = _block_map((n, n, n), (block,) * 3)
    rng = SplitMix64(seed)
    st, ss, qb = np.zeros(nb), np.zeros((nb, 1)), np.zeros(nb)
    for b in range(nb):
        st[b] = rng.log_uniform(1.0, 1000.0)
        ss[b, 0] = rng.uniform(0.99, 0.9999) * st[b]
        qb[b] = rng.uniform(0.0, 1.0)
    return ProblemSpec("thick_3d", 3, 1, (n, n, n), (1.0 / n,) * 3, quad, st, ss, mat,
                       qb[mat], ws, meta={"seed": seed, "block": block})


def headline_3d(n=128, block=16, seed=5, n_polar=12, n_azim=24, L=1, p=2,
                ws=4.0 * np.pi, quad=None):
    quad = quad or Quadrature(PRODUCT_GLC, 3, n_polar=n_polar, n_azim=n_azim, L_max=L,
                              weight_sum=ws)
    mat, nb = _block_map((n, n, n), (block,) * 3)
    rng = SplitMix64(seed)
    st, ss, qb = np.zeros(nb), np.zeros((nb, L + 1)), np.zeros(nb)
    for b in range(nb):
        st[b] = rng.log_uniform(1.0, 500.0)
        c = rng.uniform(0.99, 0.9999)
        g = rng.uniform(0.0, 0.5)
        qb[b] = rng.uniform(0.0, 1.0)
        ss[b] = [c * st[b] * g ** l for l in range(L + 1)]
    return ProblemSpec("headline_3d", 3, p, (n, n, n), (1.0 / n,) * 3, quad, st, ss, mat,
                       qb[mat], ws, meta={"seed": seed, "block": block})


def seeded_vector(size, seed):
    """x ~ U[0,1) from SplitMix64-seeded numpy PCG64 (fast for large sizes)."""
    return np.random.Generator(np.random.PCG64(seed)).random(size)
