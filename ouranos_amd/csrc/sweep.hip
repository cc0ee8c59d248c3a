// HIP kernels of the transport operator and their launch sequences.
//
//   k_build_inverses  assemble A = sigma_t M - sum_a Omega_a G_a
//                     + sum_{f downwind} (n_f.Omega) E_f per (material, direction)
//                     and factor it in LDS with the reference's Crout LU
//                     (RadiativeTransfer.cc:899-905,977-981,1026-1105)
//   k_scatter_source  S_m = sigma_{s,l(m)} M x_m       (RadiativeTransfer.cc:811-843)
//   k_trace_init      boundary inflow traces           (RadiativeTransfer.cc:957-975)
//   k_sweep_level     one wavefront level of every octant: the per-task cell
//                     solve of RadiativeTransfer.cc:878-1012 for 64 cells x
//                     all directions of an octant per workgroup
//   k_epilogue        y = x - phi (Apply, RadiativeTransfer.cc:805-806) or y += phi
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

#include "fe.hpp"
#include "plan.hpp"

namespace ou {

namespace {

// Read-only data that every lane of a wave indexes identically is loaded
// through the constant address space, which lets the compiler use scalar
// loads (one SGPR operand per FMA instead of a vector register).
template <typename T>
using cptr = const __attribute__((address_space(4))) T *;

template <typename T>
__device__ __forceinline__ cptr<T> as_const(const void *p) {
  return (cptr<T>)(p);
}

// ------------------------------------------------------------------------
// Inverse matrices: one 64-lane wave per (material, direction).
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(64) void k_build_inverses(
    int N, int dim, int p, int n_dir, const double *__restrict__ omega,
    const double *__restrict__ sigma_t, const double *__restrict__ Mm,
    const double *__restrict__ Gm, const double *__restrict__ Em,
    T *__restrict__ ainv) {
  extern __shared__ double lds[];
  double *A = lds;           // N x N, row-major
  double *BX = lds + N * N;  // N rows x 64 lanes: b, overwritten by x
  const int lane = threadIdx.x;
  const int md = blockIdx.x, mat = md / n_dir, d = md - mat * n_dir;
  const double om[3] = {omega[3 * d], omega[3 * d + 1], omega[3 * d + 2]};
  const double st = sigma_t[mat];
  const int NN = N * N;
  // Assembly in the order of RadiativeTransfer.cc:899-905 and :977-981.
  for (int e = lane; e < NN; e += 64) {
    double v = Mm[e] * st;
    for (int a = 0; a < dim; ++a) v += -om[a] * Gm[a * NN + e];
    for (int f = 0; f < 2 * dim; ++f) {
      const double ndo = (f & 1) ? om[f >> 1] : -om[f >> 1];
      if (ndo >= 0.0) v += ndo * Em[f * NN + e];
    }
    A[e] = v;
  }
  __syncthreads();
  // Crout LU with row partial pivoting (RadiativeTransfer.cc:1026-1065):
  // the pivot is the first row of largest |A[j][k]|, found by a wave-level
  // arg-max reduction; L keeps the diagonal, U is unit upper triangular.
  int my_pivot = lane;
  for (int k = 0; k < N; ++k) {
    double v = (lane >= k && lane < N) ? fabs(A[lane * N + k]) : -1.0;
    int idx = lane;
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(v, off);
      const int oi = __shfl_xor(idx, off);
      if (ov > v || (ov == v && oi < idx)) {
        v = ov;
        idx = oi;
      }
    }
    if (lane == k) my_pivot = idx;
    if (idx != k && lane < N) {
      const double t = A[k * N + lane];
      A[k * N + lane] = A[idx * N + lane];
      A[idx * N + lane] = t;
    }
    __syncthreads();
    const double akk = A[k * N + k];
    __syncthreads();
    if (lane > k && lane < N) A[k * N + lane] /= akk;
    __syncthreads();
    if (lane > k && lane < N) {
      const double ukj = A[k * N + lane];
      for (int i = k + 1; i < N; ++i) A[i * N + lane] -= A[i * N + k] * ukj;
    }
    __syncthreads();
  }
  // Lane j solves A x = e_j with LU_solve (RadiativeTransfer.cc:1067-1105):
  // column j of the inverse.  Each lane only touches its own column of BX.
  for (int r = 0; r < N; ++r) BX[r * 64 + lane] = r == lane ? 1.0 : 0.0;
  for (int k = 0; k < N; ++k) {
    const int pk = __shfl(my_pivot, k);
    if (pk != k) {
      const double t = BX[k * 64 + lane];
      BX[k * 64 + lane] = BX[pk * 64 + lane];
      BX[pk * 64 + lane] = t;
    }
    double xk = BX[k * 64 + lane];
    for (int i = 0; i < k; ++i) xk -= BX[i * 64 + lane] * A[k * N + i];
    BX[k * 64 + lane] = xk / A[k * N + k];
  }
  for (int k = N - 1; k >= 0; --k) {
    double xk = BX[k * 64 + lane];
    for (int i = k + 1; i < N; ++i) xk -= BX[i * 64 + lane] * A[k * N + i];
    BX[k * 64 + lane] = xk;
  }
  // Stored in the canonical frame of the direction's octant (see the sweep
  // kernel): reflection is an involution, so row r goes to canon(r).
  if (lane < N) {
    int o = 0;
    for (int a = 0; a < dim; ++a)
      if (om[a] < 0.0) o |= 1 << a;
    const int P1 = p + 1;
    auto canon = [&](int i) {
      int ix = i % P1, iy = (i / P1) % P1, iz = i / (P1 * P1);
      if (o & 1) ix = p - ix;
      if (o & 2) iy = p - iy;
      if (o & 4) iz = p - iz;
      return ix + P1 * (iy + P1 * iz);
    };
    T *out = ainv + size_t(md) * NN;
    const int cj = canon(lane);
    for (int r = 0; r < N; ++r) out[canon(r) * N + cj] = T(BX[r * 64 + lane]);
  }
}

// ------------------------------------------------------------------------
// Scattering source S_m|K = sigma_{s,l(m)} (M x_m|K).
// ------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__(256) void k_scatter_source(
    long n_cells, int n_mom, long n_dofs, const double *__restrict__ x,
    const int *__restrict__ mat, const double *__restrict__ sigma_s, int L,
    const int *__restrict__ mom_l, const double *Mm, T *__restrict__ S) {
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n_cells * n_mom) return;
  const int m = int(gid / n_cells);
  const long c = gid - long(m) * n_cells;
  const int l = mom_l[m];
  const double sig = l <= L ? sigma_s[long(mat[c]) * (L + 1) + l] : 0.0;
  const double *xc = x + long(m) * n_dofs + c * N;
  double xv[N];
#pragma unroll
  for (int j = 0; j < N; ++j) xv[j] = xc[j];
  cptr<double> M = as_const<double>(Mm);
  T *sc = S + long(m) * n_dofs + c * N;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < N; ++j) s += M[i * N + j] * xv[j];
    sc[i] = T(s * sig);
  }
}

// Moment 0 of the source sweep: S_0 + (M 1) q / (ws M2D[.,0]), so that
// M2D[d,0] * S0 adds the volumetric source M q/ws of RadiativeTransfer.cc:914-925.
template <typename T, int N>
__global__ __launch_bounds__(256) void k_source_moment0(
    long n_cells, const T *__restrict__ S, const double *__restrict__ q,
    const double *__restrict__ rowsum, double factor, T *__restrict__ S0) {
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n_cells * N) return;
  const long c = gid / N;
  const int i = int(gid - c * N);
  S0[gid] = T(double(S[gid]) + rowsum[i] * (q[c] * factor));
}

// ------------------------------------------------------------------------
// Trace buffers: per direction and axis, one slot per line of cells along
// the axis; a slot holds the angular flux on the face the next cell of the
// line enters through.  Before a sweep every slot holds the boundary inflow.
// ------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_trace_init(T *__restrict__ tr, long per_dir,
                                                    long total,
                                                    const double *__restrict__ inflow,
                                                    int axis, int with_source) {
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const long d = gid / per_dir;
  tr[gid] = with_source ? T(inflow[d * 3 + axis]) : T(0);
}

// ------------------------------------------------------------------------
// One sweep level.
//
// Every direction of an octant is handled in the octant's canonical frame:
// local node indices are reflected along each axis where Omega_a < 0, so the
// upwind face of every cell is node 0 and the downwind face node p along
// each axis.  Register arrays are then indexed by compile-time constants
// only; the reflection is applied to LDS addresses (scattering source in,
// angular flux out) and is built into the stored inverses.  The face mass
// matrices are unchanged by the reflection (equidistant nodes are symmetric).
// ------------------------------------------------------------------------
template <int DIM, int P>
__device__ __forceinline__ int natural_index(int i, int o) {
  constexpr int P1 = P + 1;
  int ix = i % P1, iy = (i / P1) % P1, iz = i / (P1 * P1);
  if (o & 1) ix = P - ix;
  if (o & 2) iy = P - iy;
  if (DIM == 3 && (o & 4)) iz = P - iz;
  return ix + P1 * (iy + P1 * iz);
}

template <int DIM, int P, int A, int SIDE>
__device__ constexpr int face_dof(int fi) {
  constexpr int P1 = P + 1;
  const int u = fi % P1, v = fi / P1;
  if (DIM == 2) return A == 0 ? SIDE + P1 * u : u + P1 * SIDE;
  return A == 0 ? SIDE + P1 * (u + P1 * v)
                : (A == 1 ? u + P1 * (SIDE + P1 * v) : u + P1 * (v + P1 * SIDE));
}

// b[upwind face dofs of axis A] += |Omega_A| Mf_A t: the upwind term
// U_f (|n.Omega| psi_up) of RadiativeTransfer.cc:935-956 restricted to the
// non-zero block of U_f.  t holds the upwind trace (prefetched).
template <typename T, int DIM, int P, int A>
__device__ __forceinline__ void add_upwind(cptr<T> face_mass, T oa, const T *t, T *b) {
  constexpr int P1 = P + 1, F = DIM == 2 ? P1 : P1 * P1;
  cptr<T> Mf = face_mass + A * F * F;
#pragma unroll
  for (int fi = 0; fi < F; ++fi) {
    T s = T(0);
#pragma unroll
    for (int fj = 0; fj < F; ++fj) s += Mf[fi * F + fj] * t[fj];
    b[face_dof<DIM, P, A, 0>(fi)] += oa * s;
  }
}

template <typename T, int DIM, int P, int A>
__device__ __forceinline__ void load_trace(const void *trace, long ps, int d, long slot,
                                           bool active, T *t) {
  constexpr int P1 = P + 1, F = DIM == 2 ? P1 : P1 * P1;
  const T *tr = static_cast<const T *>(trace) + long(d) * F * ps + slot;
#pragma unroll
  for (int f = 0; f < F; ++f) t[f] = active ? tr[f * ps] : T(0);
}

template <typename T, int DIM, int P, int A>
__device__ __forceinline__ void store_trace(void *trace, long ps, int d, long slot,
                                            const T *psi) {
  constexpr int P1 = P + 1, F = DIM == 2 ? P1 : P1 * P1;
  T *tr = static_cast<T *>(trace) + long(d) * F * ps + slot;
#pragma unroll
  for (int f = 0; f < F; ++f) tr[f * ps] = psi[face_dof<DIM, P, A, P>(f)];
}

template <typename T, int DIM, int P, int MM, int W>
struct SweepCfg {
  static constexpr int P1 = P + 1;
  static constexpr int N = DIM == 2 ? P1 * P1 : P1 * P1 * P1;
  static constexpr int F = DIM == 2 ? P1 : P1 * P1;
  static constexpr int NP = N | 1;            // odd LDS stride of one cell
  static constexpr int NC = (N + W - 1) / W;  // moment columns per lane
  // FP32 Q2 hexahedra solve on the matrix cores (see solve_mfma27).
  static constexpr bool mfma = std::is_same<T, float>::value && N == 27;
  static constexpr size_t psi_bytes = size_t(W) * 64 * NP * sizeof(T);
  static constexpr size_t src_bytes = size_t(MM) * 64 * NP * sizeof(T);
  static constexpr bool stage_src = psi_bytes + src_bytes <= 98304;
  static constexpr size_t lds_bytes = psi_bytes + (stage_src ? src_bytes : 0);
};

typedef float v16f __attribute__((ext_vector_type(16)));
typedef float v4f __attribute__((ext_vector_type(4)));

// A-operand fragment of the 27x27 inverse for v_mfma_f32_32x32x2_f32:
// element kk of lane l is Ainv[l % 32][2 kk + l / 32] (zero padded to 32x32).
struct AFrag {
  v4f q[4];
};

__device__ __forceinline__ void load_afrag(const float *base, int lane, AFrag &a) {
  const v4f *p = reinterpret_cast<const v4f *>(base + lane * 16);
#pragma unroll
  for (int k = 0; k < 4; ++k) a.q[k] = p[k];
}

// psi = Ainv b for 64 cells (lane = cell) on the matrix cores.  The 27 x 64
// right-hand sides are split into two 32-column blocks (cells 0-31 and
// 32-63); v_permlane32_swap regroups two rows of b into the B operands of
// both blocks, 14 k-steps of v_mfma_f32_32x32x2_f32 per block accumulate
// 32x32 results, and 16 more swaps return every cell's 27 values to its
// own lane.  Output row i of a block sits in register 4 (i / 8) + i % 4 of
// lane (i / 4) % 2 * 32 + column.
__device__ __forceinline__ void solve_mfma27(const AFrag &a, const float *b, float *psi) {
  v16f accA = {}, accB = {};
#pragma unroll
  for (int kk = 0; kk < 14; ++kk) {
    const float b0 = b[2 * kk];
    const float b1 = 2 * kk + 1 < 27 ? b[2 * kk + 1] : 0.0f;
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(b0), __float_as_uint(b1),
                                                     false, false);
    const float ak = a.q[kk >> 2][kk & 3];
    accA = __builtin_amdgcn_mfma_f32_32x32x2f32(ak, __uint_as_float(sw[0]), accA, 0, 0, 0);
    accB = __builtin_amdgcn_mfma_f32_32x32x2f32(ak, __uint_as_float(sw[1]), accB, 0, 0, 0);
  }
  float lo[16], hi[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(accA[r]),
                                                     __float_as_uint(accB[r]), false, false);
    lo[r] = __uint_as_float(sw[0]);
    hi[r] = __uint_as_float(sw[1]);
  }
#pragma unroll
  for (int i = 0; i < 27; ++i) psi[i] = ((i >> 2) & 1) ? hi[(i >> 3) * 4 + (i & 3)] : lo[(i >> 3) * 4 + (i & 3)];
}

// Workgroup = one chunk (<= 64 cells of one level, one octant, one material)
// x all directions of the octant; wave w takes directions w, w+W, ...  Lane
// = cell.  Per direction a lane builds b from the scattering source and the
// upwind traces, multiplies by the (material, direction) inverse (matrix
// cores for FP32 Q2 hexahedra, scalar-operand FMAs otherwise) and writes its
// downwind traces; the traces and inverse of the wave's next direction are
// loaded while the current one is computed.  The moment accumulation
// phi_m += D2M[m,d] psi_d (RadiativeTransfer.cc:990-996) is FP64: psi is
// exchanged through LDS and, after each round of W directions, every wave
// accumulates a 1/W share of the chunk's (cell, dof) columns.
template <typename T, int DIM, int P, int MM, int W>
__global__ __launch_bounds__(64 * W) void k_sweep_level(SweepView v, int chunk_begin) {
  using C = SweepCfg<T, DIM, P, MM, W>;
  constexpr int N = C::N, NP = C::NP, NC = C::NC, F = C::F;
  extern __shared__ unsigned char lds_raw[];
  T *s_psi = reinterpret_cast<T *>(lds_raw);  // [W][64*NP]
  T *s_src = s_psi + W * 64 * NP;             // [MM][64*NP]
  __shared__ int s_cell[64];

  const Chunk ch = v.chunks[chunk_begin + blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int cnt = ch.count, o = ch.octant;
  const bool active = lane < cnt;
  if (tid < 64) s_cell[tid] = tid < cnt ? v.cells[ch.cell_offset + tid] : 0;
  __syncthreads();

  const T *S = static_cast<const T *>(v.S);
  const T *S0 = static_cast<const T *>(v.S0);
  const long n_dofs = v.n_dofs;
  const int n_mom = v.n_mom, n_dir = v.n_dir;
  if constexpr (C::stage_src) {
    for (int e = tid; e < n_mom * 64 * N; e += 64 * W) {
      const int m = e / (64 * N), r = e - m * 64 * N, cl = r / N, i = r - cl * N;
      T val = T(0);
      if (cl < cnt) {
        const long g = long(s_cell[cl]) * N + i;
        val = (m == 0 && S0) ? S0[g] : S[long(m) * n_dofs + g];
      }
      s_src[m * 64 * NP + cl * NP + i] = val;
    }
    __syncthreads();
  }

  const int c = s_cell[lane];
  const int nx = v.n[0], ny = v.n[1];
  const int ci = c % nx, cj = (c / nx) % ny, ck = DIM == 3 ? c / (nx * ny) : 0;
  const int ii = (o & 1) ? nx - 1 - ci : ci;
  const int jj = (o & 2) ? ny - 1 - cj : cj;
  const int kk = (DIM == 3 && (o & 4)) ? v.n[2] - 1 - ck : ck;
  const long slot0 = jj + long(ny) * kk, slot1 = ii + long(nx) * kk,
             slot2 = long(ii + jj) * ny + jj;

  double acc[MM][NC];
#pragma unroll
  for (int m = 0; m < MM; ++m)
#pragma unroll
    for (int t = 0; t < NC; ++t) acc[m][t] = 0.0;

  const int dbeg = v.oct_off[o], ndir_o = v.oct_off[o + 1] - dbeg;
  cptr<int> dirs = as_const<int>(v.oct_dirs) + dbeg;
  cptr<T> m2d = as_const<T>(v.m2d);
  cptr<double> d2m = as_const<double>(v.d2m);
  cptr<T> omega_abs = as_const<T>(v.omega_abs);
  cptr<T> face_mass = as_const<T>(v.face_mass);
  const int mat = ch.material;
  const T *src_cell = S + long(c) * N;
  const T *src0_cell = (S0 ? S0 : S) + long(c) * N;
  const float *afrag_base = v.ainv_mfma + long(mat) * n_dir * 64 * 16;

  // Prefetched state of the wave's current direction.
  T tx[F], ty[F], tz[F];
  AFrag af;
  auto prefetch = [&](int slot, T *px, T *py, T *pz, AFrag &pa) {
    const int d = dirs[slot];
    load_trace<T, DIM, P, 0>(v.trace[0], v.plane_size[0], d, slot0, active, px);
    load_trace<T, DIM, P, 1>(v.trace[1], v.plane_size[1], d, slot1, active, py);
    if constexpr (DIM == 3) load_trace<T, DIM, P, 2>(v.trace[2], v.plane_size[2], d, slot2, active, pz);
    if constexpr (C::mfma) load_afrag(afrag_base + long(d) * 64 * 16, lane, pa);
  };
  if (w < ndir_o) prefetch(w, tx, ty, tz, af);

  for (int r0 = 0; r0 < ndir_o; r0 += W) {
    const int slot = r0 + w;
    if (slot < ndir_o) {
      const int d = dirs[slot];
      T nx_[F], ny_[F], nz_[F];
      AFrag naf;
      const bool more = slot + W < ndir_o;
      if (more) prefetch(slot + W, nx_, ny_, nz_, naf);

      T b[N];
#pragma unroll
      for (int i = 0; i < N; ++i) b[i] = T(0);
      for (int m = 0; m < MM; ++m) {
        if (m >= n_mom) break;
        const T cm = m2d[d * n_mom + m];
        if constexpr (C::stage_src) {
          const T *src = s_src + m * 64 * NP + lane * NP;
#pragma unroll
          for (int i = 0; i < N; ++i) b[i] += cm * src[natural_index<DIM, P>(i, o)];
        } else {
          const T *src = m == 0 ? src0_cell : src_cell + long(m) * n_dofs;
#pragma unroll
          for (int i = 0; i < N; ++i)
            b[i] += cm * (active ? src[natural_index<DIM, P>(i, o)] : T(0));
        }
      }
      {
        const T o0 = omega_abs[d * 3 + 0], o1 = omega_abs[d * 3 + 1];
        if (o0 != T(0)) add_upwind<T, DIM, P, 0>(face_mass, o0, tx, b);
        if (o1 != T(0)) add_upwind<T, DIM, P, 1>(face_mass, o1, ty, b);
        if constexpr (DIM == 3) {
          const T o2 = omega_abs[d * 3 + 2];
          if (o2 != T(0)) add_upwind<T, DIM, P, 2>(face_mass, o2, tz, b);
        }
      }

      T psi[N];
      if constexpr (C::mfma) {
        solve_mfma27(af, reinterpret_cast<const float *>(b), reinterpret_cast<float *>(psi));
      } else {
        cptr<T> Ai = as_const<T>(v.ainv) + (long(mat) * n_dir + d) * (N * N);
#pragma unroll
        for (int r = 0; r < N; ++r) {
          T s = T(0);
#pragma unroll
          for (int q = 0; q < N; ++q) s += Ai[r * N + q] * b[q];
          psi[r] = s;
        }
      }
      if (active) {
        store_trace<T, DIM, P, 0>(v.trace[0], v.plane_size[0], d, slot0, psi);
        store_trace<T, DIM, P, 1>(v.trace[1], v.plane_size[1], d, slot1, psi);
        if constexpr (DIM == 3) store_trace<T, DIM, P, 2>(v.trace[2], v.plane_size[2], d, slot2, psi);
      }
      T *dst = s_psi + w * 64 * NP + lane * NP;
#pragma unroll
      for (int i = 0; i < N; ++i) dst[natural_index<DIM, P>(i, o)] = psi[i];
      if (more) {
#pragma unroll
        for (int f = 0; f < F; ++f) {
          tx[f] = nx_[f];
          ty[f] = ny_[f];
          if constexpr (DIM == 3) tz[f] = nz_[f];
        }
        if constexpr (C::mfma) af = naf;
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < NC; ++t) {
      const int col = (t * W + w) * 64 + lane;
      if (col < 64 * N) {
        const int cl = col / N, i = col - cl * N;
        for (int ww = 0; ww < W; ++ww) {
          const int sl = r0 + ww;
          if (sl >= ndir_o) break;
          const int dd = dirs[sl];
          const double pv = double(s_psi[ww * 64 * NP + cl * NP + i]);
#pragma unroll
          for (int m = 0; m < MM; ++m)
            if (m < n_mom) acc[m][t] += d2m[m * n_dir + dd] * pv;
        }
      }
    }
    __syncthreads();
  }

#pragma unroll
  for (int t = 0; t < NC; ++t) {
    const int col = (t * W + w) * 64 + lane;
    if (col < 64 * N) {
      const int cl = col / N, i = col - cl * N;
      if (cl < cnt) {
        double *out = v.partial + long(o) * n_mom * n_dofs + long(s_cell[cl]) * N + i;
#pragma unroll
        for (int m = 0; m < MM; ++m)
          if (m < n_mom) out[long(m) * n_dofs] = acc[m][t];
      }
    }
  }
}

// Rearranges the canonical row-major inverses of FP32 Q2 hexahedra into the
// A-operand layout of solve_mfma27: out[md][lane][kk] = A[lane%32][2kk+lane/32].
__global__ __launch_bounds__(64) void k_ainv_to_mfma(const float *__restrict__ ainv,
                                                     float *__restrict__ out) {
  const int lane = threadIdx.x;
  const long md = blockIdx.x;
  const float *A = ainv + md * 27 * 27;
  float *o = out + (md * 64 + lane) * 16;
  const int m = lane & 31;
  for (int kk = 0; kk < 16; ++kk) {
    const int k = 2 * kk + (lane >> 5);
    o[kk] = (m < 27 && k < 27) ? A[m * 27 + k] : 0.0f;
  }
}

// y = x - sum_o partial_o (Apply) or y += sum_o partial_o (sweep).
__global__ __launch_bounds__(256) void k_epilogue(long n, int n_oct, long stride,
                                                  const double *__restrict__ partial,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ y, int apply) {
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  double s = 0.0;
  for (int o = 0; o < n_oct; ++o) s += partial[o * stride + gid];
  y[gid] = apply ? x[gid] - s : y[gid] + s;
}

inline unsigned grid_for(long n, int block) { return unsigned((n + block - 1) / block); }

// ------------------------------------------------------------------------
// Dispatch tables.
// ------------------------------------------------------------------------
using LevelLauncher = void (*)(const Plan &, const SweepView &, hipStream_t, hipEvent_t,
                               hipEvent_t);

template <typename T, int DIM, int P, int MM, int W>
void launch_levels_impl(const Plan &pl, const SweepView &v, hipStream_t s,
                        hipEvent_t before, hipEvent_t after) {
  using C = SweepCfg<T, DIM, P, MM, W>;
  auto kern = k_sweep_level<T, DIM, P, MM, W>;
  static bool attr_set = false;
  if (!attr_set) {
    OU_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               int(C::lds_bytes)));
    attr_set = true;
  }
  if (before) OU_HIP(hipEventRecord(before, s));
  for (int l = 0; l < pl.n_levels; ++l) {
    const int b = pl.level_chunk_off[l], e = pl.level_chunk_off[l + 1];
    if (e > b) hipLaunchKernelGGL(kern, dim3(e - b), dim3(64 * W), C::lds_bytes, s, v, b);
  }
  if (after) OU_HIP(hipEventRecord(after, s));
  OU_HIP(hipGetLastError());
}

template <typename T, int DIM, int P, int W>
LevelLauncher pick_mm(int mm) {
  switch (mm) {
    case 1: return launch_levels_impl<T, DIM, P, 1, W>;
    case 4: return launch_levels_impl<T, DIM, P, 4, W>;
    case 10: if constexpr (W == 4) return launch_levels_impl<T, DIM, P, 10, W>; break;
    case 16: if constexpr (W == 4) return launch_levels_impl<T, DIM, P, 16, W>; break;
  }
  return nullptr;
}

template <typename T>
LevelLauncher pick(int dim, int p, int mm, int waves) {
  if (waves == 4) {
    if (dim == 2 && p == 1) return pick_mm<T, 2, 1, 4>(mm);
    if (dim == 2 && p == 2) return pick_mm<T, 2, 2, 4>(mm);
    if (dim == 2 && p == 3) return pick_mm<T, 2, 3, 4>(mm);
    if (dim == 3 && p == 1) return pick_mm<T, 3, 1, 4>(mm);
    if (dim == 3 && p == 2) return pick_mm<T, 3, 2, 4>(mm);
  } else if (waves == 8) {
    if (dim == 3 && p == 2 && mm <= 4) return pick_mm<T, 3, 2, 8>(mm);
  }
  return nullptr;
}

template <typename T>
size_t lds_need(int dim, int p, int mm, int waves) {
  const int P1 = p + 1, N = dim == 2 ? P1 * P1 : P1 * P1 * P1, NP = N | 1;
  const size_t psi = size_t(waves) * 64 * NP * sizeof(T);
  const size_t src = size_t(mm) * 64 * NP * sizeof(T);
  return psi + (psi + src <= 98304 ? src : 0) + 256;
}

template <typename T>
void scatter_impl(Plan &pl, const double *d_x, hipStream_t s) {
  const long n = pl.n_cells * pl.n_mom;
  T *S = static_cast<T *>(pl.d_S);
#define OU_SCAT(NN)                                                                  \
  case NN:                                                                           \
    hipLaunchKernelGGL((k_scatter_source<T, NN>), dim3(grid_for(n, 256)), dim3(256), \
                       0, s, pl.n_cells, pl.n_mom, pl.n_dofs, d_x, pl.d_mat,         \
                       pl.d_sigma_s, pl.L, pl.d_mom_l, pl.d_mass, S);                \
    break;
  switch (pl.N) {
    OU_SCAT(4) OU_SCAT(9) OU_SCAT(16) OU_SCAT(8) OU_SCAT(27)
    default: fail(OU_ERR_UNSUPPORTED, "scatter source: unsupported element size");
  }
#undef OU_SCAT
  OU_HIP(hipGetLastError());
}

template <typename T>
void source0_impl(Plan &pl, hipStream_t s) {
  const long n = pl.n_cells * pl.N;
  const double factor = 1.0 / (pl.ws * pl.M2D[0]);
#define OU_SRC(NN)                                                                   \
  case NN:                                                                           \
    hipLaunchKernelGGL((k_source_moment0<T, NN>), dim3(grid_for(n, 256)), dim3(256), \
                       0, s, pl.n_cells, static_cast<const T *>(pl.d_S), pl.d_source, \
                       pl.d_rowsum, factor, static_cast<T *>(pl.d_S0));               \
    break;
  switch (pl.N) {
    OU_SRC(4) OU_SRC(9) OU_SRC(16) OU_SRC(8) OU_SRC(27)
    default: fail(OU_ERR_UNSUPPORTED, "source: unsupported element size");
  }
#undef OU_SRC
  OU_HIP(hipGetLastError());
}

template <typename T>
void trace_init_impl(Plan &pl, bool with_source, hipStream_t s) {
  for (int a = 0; a < pl.dim; ++a) {
    const long per_dir = long(pl.F) * pl.plane_size[a];
    const long total = per_dir * pl.n_dir;
    hipLaunchKernelGGL(k_trace_init<T>, dim3(grid_for(total, 256)), dim3(256), 0, s,
                       static_cast<T *>(pl.d_trace[a]), per_dir, total, pl.d_inflow, a,
                       int(with_source));
  }
  OU_HIP(hipGetLastError());
}

}  // namespace

SweepView Plan::view() const {
  SweepView v;
  for (int a = 0; a < 3; ++a) {
    v.n[a] = n[a];
    v.trace[a] = d_trace[a];
    v.plane_size[a] = plane_size[a];
  }
  v.n_dir = n_dir;
  v.n_mom = n_mom;
  v.n_dofs = n_dofs;
  v.ainv = d_ainv;
  v.ainv_mfma = d_ainv_mfma;
  v.S = d_S;
  v.S0 = nullptr;
  v.m2d = d_m2d;
  v.d2m = d_d2m;
  v.omega_abs = d_omega_abs;
  v.dir_oct = d_dir_oct;
  v.oct_dirs = d_oct_dirs;
  v.oct_off = d_oct_off;
  v.face_mass = d_face_mass;
  v.partial = d_partial;
  v.chunks = d_chunks;
  v.cells = d_cells;
  return v;
}

void launch_build_inverses(Plan &pl) {
  const int N = pl.N;
  std::vector<double> M, G, E, tmp;
  RefElement re(pl.dim, pl.p, pl.h);
  re.mass(M);
  G.resize(size_t(pl.dim) * N * N);
  for (int a = 0; a < pl.dim; ++a) {
    re.grad(a, tmp);
    std::copy(tmp.begin(), tmp.end(), G.begin() + size_t(a) * N * N);
  }
  E.resize(size_t(2 * pl.dim) * N * N);
  for (int f = 0; f < 2 * pl.dim; ++f) {
    re.face(f, false, tmp);
    std::copy(tmp.begin(), tmp.end(), E.begin() + size_t(f) * N * N);
  }
  double *dM = nullptr, *dG = nullptr, *dE = nullptr, *dom = nullptr, *dst = nullptr;
  OU_HIP(hipMalloc(&dM, M.size() * 8));
  OU_HIP(hipMalloc(&dG, G.size() * 8));
  OU_HIP(hipMalloc(&dE, E.size() * 8));
  OU_HIP(hipMalloc(&dom, pl.omega.size() * 8));
  OU_HIP(hipMalloc(&dst, pl.sigma_t.size() * 8));
  OU_HIP(hipMemcpy(dM, M.data(), M.size() * 8, hipMemcpyHostToDevice));
  OU_HIP(hipMemcpy(dG, G.data(), G.size() * 8, hipMemcpyHostToDevice));
  OU_HIP(hipMemcpy(dE, E.data(), E.size() * 8, hipMemcpyHostToDevice));
  OU_HIP(hipMemcpy(dom, pl.omega.data(), pl.omega.size() * 8, hipMemcpyHostToDevice));
  OU_HIP(hipMemcpy(dst, pl.sigma_t.data(), pl.sigma_t.size() * 8, hipMemcpyHostToDevice));
  const size_t lds = (size_t(N) * N + size_t(N) * 64) * sizeof(double);
  const unsigned blocks = unsigned(pl.n_mat * pl.n_dir);
  if (pl.precision == OU_FP32)
    hipLaunchKernelGGL(k_build_inverses<float>, dim3(blocks), dim3(64), lds, pl.stream, N,
                       pl.dim, pl.p, pl.n_dir, dom, dst, dM, dG, dE,
                       static_cast<float *>(pl.d_ainv));
  else
    hipLaunchKernelGGL(k_build_inverses<double>, dim3(blocks), dim3(64), lds, pl.stream, N,
                       pl.dim, pl.p, pl.n_dir, dom, dst, dM, dG, dE,
                       static_cast<double *>(pl.d_ainv));
  OU_HIP(hipGetLastError());
  if (pl.use_mfma)
    hipLaunchKernelGGL(k_ainv_to_mfma, dim3(blocks), dim3(64), 0, pl.stream,
                       static_cast<const float *>(pl.d_ainv), pl.d_ainv_mfma);
  OU_HIP(hipGetLastError());
  OU_HIP(hipStreamSynchronize(pl.stream));
  hipFree(dM);
  hipFree(dG);
  hipFree(dE);
  hipFree(dom);
  hipFree(dst);
}

bool sweep_supported(int dim, int p, int mm, int precision, int waves, int max_lds) {
  if (precision == OU_FP32)
    return pick<float>(dim, p, mm, waves) != nullptr &&
           lds_need<float>(dim, p, mm, waves) <= size_t(max_lds);
  return pick<double>(dim, p, mm, waves) != nullptr &&
         lds_need<double>(dim, p, mm, waves) <= size_t(max_lds);
}

void launch_scatter_source(Plan &pl, const double *d_x, hipStream_t s) {
  if (pl.precision == OU_FP32)
    scatter_impl<float>(pl, d_x, s);
  else
    scatter_impl<double>(pl, d_x, s);
  pl.S_is_zero = false;
}

void launch_source_moment0(Plan &pl, hipStream_t s) {
  if (pl.precision == OU_FP32)
    source0_impl<float>(pl, s);
  else
    source0_impl<double>(pl, s);
}

void launch_trace_init(Plan &pl, bool with_source, hipStream_t s) {
  if (pl.precision == OU_FP32)
    trace_init_impl<float>(pl, with_source, s);
  else
    trace_init_impl<double>(pl, with_source, s);
}

void launch_sweep_levels(Plan &pl, const void *S0, hipStream_t s, hipEvent_t before,
                         hipEvent_t after) {
  SweepView v = pl.view();
  v.S0 = S0;
  LevelLauncher f = pl.precision == OU_FP32 ? pick<float>(pl.dim, pl.p, pl.mm, pl.waves)
                                            : pick<double>(pl.dim, pl.p, pl.mm, pl.waves);
  if (!f) fail(OU_ERR_UNSUPPORTED, "no sweep kernel for this element / moment count");
  f(pl, v, s, before, after);
}

void launch_epilogue(Plan &pl, const double *d_x, double *d_y, bool apply, hipStream_t s) {
  const long n = pl.vec_size;
  hipLaunchKernelGGL(k_epilogue, dim3(grid_for(n, 256)), dim3(256), 0, s, n, pl.n_oct,
                     pl.vec_size, pl.d_partial, d_x, d_y, int(apply));
  OU_HIP(hipGetLastError());
}

void apply_device(Plan &pl, const double *d_x, double *d_y, hipStream_t s) {
  launch_scatter_source(pl, d_x, s);
  launch_trace_init(pl, false, s);
  launch_sweep_levels(pl, nullptr, s);
  launch_epilogue(pl, d_x, d_y, true, s);
}

void sweep_device(Plan &pl, double *d_y, bool with_source, hipStream_t s) {
  const void *S0 = nullptr;
  if (with_source) {
    launch_source_moment0(pl, s);
    S0 = pl.d_S0;
  }
  launch_trace_init(pl, with_source, s);
  launch_sweep_levels(pl, S0, s);
  launch_epilogue(pl, nullptr, d_y, false, s);
}

}  // namespace ou
