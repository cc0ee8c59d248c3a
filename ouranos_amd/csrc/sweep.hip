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
#include <cstdint>
#include <type_traits>

#include "fe.hpp"
#include "plan.hpp"
#include "sweep_common.hpp"

namespace ou {

namespace {

// ------------------------------------------------------------------------
// Inverse matrices: one 64-lane wave per (material, direction class).  The
// plan passes one representative |Omega| per class (all components >= 0),
// so the matrix is assembled directly in the canonical frame.
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
__device__ __forceinline__ void scatter_one(long gid, long n_cells, long n_dofs,
                                            const double *__restrict__ x,
                                            const int *__restrict__ mat,
                                            const double *__restrict__ sigma_s, int L,
                                            const int *__restrict__ mom_l, const double *Mm,
                                            T *__restrict__ S, bool accumulate = false) {
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
    sc[i] = accumulate ? T(double(sc[i]) + s * sig) : T(s * sig);
  }
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_scatter_source(
    long n_cells, int n_mom, long n_dofs, const double *__restrict__ x,
    const int *__restrict__ mat, const double *__restrict__ sigma_s, int L,
    const int *__restrict__ mom_l, const double *Mm, T *__restrict__ S, int accumulate) {
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n_cells * n_mom) return;
  scatter_one<T, N>(gid, n_cells, n_dofs, x, mat, sigma_s, L, mom_l, Mm, S, accumulate != 0);
}

// Moment 0 of the source sweep: S_0 + (M 1) q / (ws M2D[.,0]), so that
// M2D[d,0] * S0 adds the volumetric source M q/ws of RadiativeTransfer.cc:914-925.
template <typename T, int N>
__device__ __forceinline__ void source0_one(long gid, const T *__restrict__ S,
                                            const double *__restrict__ q,
                                            const double *__restrict__ rowsum, double factor,
                                            T *__restrict__ S0) {
  const long c = gid / N;
  const int i = int(gid - c * N);
  S0[gid] = T(double(S[gid]) + rowsum[i] * (q[c] * factor));
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_source_moment0(
    long n_cells, const T *__restrict__ S, const double *__restrict__ q,
    const double *__restrict__ rowsum, double factor, T *__restrict__ S0) {
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n_cells * N) return;
  source0_one<T, N>(gid, S, q, rowsum, factor, S0);
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
#include "sweep_level.inc"

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

// The same for FP64 Q2 and v_mfma_f64_16x16x4f64: out[md][lane][7 mt + ks] =
// A[16 mt + lane%16][4 ks + lane/16] (zero padded to 32 x 28), the A operand of
// row tile mt and k-step ks of solve_mfma27_f64.
__global__ __launch_bounds__(64) void k_ainv_to_mfma64(const double *__restrict__ ainv,
                                                       double *__restrict__ out) {
  const int lane = threadIdx.x;
  const long md = blockIdx.x;
  const double *A = ainv + md * 27 * 27;
  double *o = out + (md * 64 + lane) * 14;
  for (int mt = 0; mt < 2; ++mt)
    for (int ks = 0; ks < 7; ++ks) {
      const int row = 16 * mt + (lane & 15), col = 4 * ks + (lane >> 4);
      o[7 * mt + ks] = (row < 27 && col < 27) ? A[row * 27 + col] : 0.0;
    }
}

// y = x - sum_o partial_o (Apply, mode 1), y += sum_o partial_o (sweep, mode
// 0) or y = 0 + sum_o partial_o (a sweep into a fresh vector, mode 2), summed
// over the n_parts = octants x `groups` direction groups partial moment
// vectors of the octants in oct_mask (an octant without directions has no
// workgroups, so its slots hold whatever another layout left there).
// A persistent sweep whose watchdog gave up (error word set) leaves its
// partial moments incomplete: the output is NaN then, so that no caller
// mistakes it for a result.
__device__ __forceinline__ void epilogue_one(long gid, int n_parts, int groups,
                                             unsigned oct_mask, long stride,
                                             const double *__restrict__ partial,
                                             const double *__restrict__ x,
                                             double *__restrict__ y, int apply,
                                             const int *error = nullptr) {
  double s = 0.0;
  for (int k = 0; k < n_parts; ++k)
    if ((oct_mask >> (k / groups)) & 1u) s += partial[k * stride + gid];
  if (error && *error) s = __builtin_nan("");
  y[gid] = apply == 1 ? x[gid] - s : (apply == 2 ? 0.0 + s : y[gid] + s);
}

__global__ __launch_bounds__(256) void k_epilogue(long n, int n_parts, int groups,
                                                  unsigned oct_mask, long stride,
                                                  const double *__restrict__ partial,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ y, int apply,
                                                  const int *error) {
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= n) return;
  epilogue_one(gid, n_parts, groups, oct_mask, stride, partial, x, y, apply, error);
}

// Clears the persistent sweep's completion flags and ticket before its launch
// (a kernel, so that captured graphs hold kernel nodes only).
__global__ __launch_bounds__(256) void k_persist_reset(unsigned *done, long n,
                                                       unsigned *ticket) {
  const long gid = long(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid < n) done[gid] = 0u;
  if (gid == 0) *ticket = 0u;
}

// ------------------------------------------------------------------------
// Batched element-wise stages: one launch covers every member of a batch;
// each block walks the members and strides over each member's elements.
// The per-element arithmetic is the single-plan kernels' own.
// ------------------------------------------------------------------------
template <typename T, int N>
__global__ __launch_bounds__(256) void k_batch_scatter(const MemberView *__restrict__ mv,
                                                       int n_mem, const int *__restrict__ active,
                                                       const double *const *__restrict__ xs) {
  const long stride = long(gridDim.x) * blockDim.x;
  for (int u = 0; u < n_mem; ++u) {
    if (!active[u]) continue;
    const MemberView &m = mv[u];
    const long n = m.n_cells * m.n_mom;
    for (long gid = long(blockIdx.x) * blockDim.x + threadIdx.x; gid < n; gid += stride)
      scatter_one<T, N>(gid, m.n_cells, m.n_dofs, xs[u], m.mat, m.sigma_s, m.L, m.mom_l,
                        m.mass, static_cast<T *>(m.S));
  }
}

template <typename T, int N>
__global__ __launch_bounds__(256) void k_batch_source0(const MemberView *__restrict__ mv,
                                                       int n_mem, const int *__restrict__ active) {
  const long stride = long(gridDim.x) * blockDim.x;
  for (int u = 0; u < n_mem; ++u) {
    if (!active[u]) continue;
    const MemberView &m = mv[u];
    const long n = m.n_cells * N;
    for (long gid = long(blockIdx.x) * blockDim.x + threadIdx.x; gid < n; gid += stride)
      source0_one<T, N>(gid, static_cast<const T *>(m.S), m.source, m.rowsum, m.s0_factor,
                        static_cast<T *>(m.S0));
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_batch_trace_init(const MemberView *__restrict__ mv,
                                                          int n_mem, const int *__restrict__ active,
                                                          int dim, int with_source) {
  const long stride = long(gridDim.x) * blockDim.x;
  for (int u = 0; u < n_mem; ++u) {
    if (!active[u]) continue;
    const MemberView &m = mv[u];
    for (int a = 0; a < dim; ++a) {
      T *tr = static_cast<T *>(m.trace[a]);
      const long per_dir = m.trace_per_dir[a], total = per_dir * m.n_dir;
      for (long gid = long(blockIdx.x) * blockDim.x + threadIdx.x; gid < total; gid += stride)
        tr[gid] = with_source ? T(m.inflow[(gid / per_dir) * 3 + a]) : T(0);
    }
  }
}

__global__ __launch_bounds__(256) void k_batch_epilogue(const MemberView *__restrict__ mv,
                                                        int n_mem, const int *__restrict__ active,
                                                        const double *const *__restrict__ xs,
                                                        double *const *__restrict__ ys,
                                                        int apply, const int *error) {
  const long stride = long(gridDim.x) * blockDim.x;
  for (int u = 0; u < n_mem; ++u) {
    if (!active[u]) continue;
    const MemberView &m = mv[u];
    const double *x = apply == 1 ? xs[u] : nullptr;
    double *y = ys[u];
    for (long gid = long(blockIdx.x) * blockDim.x + threadIdx.x; gid < m.vec_size;
         gid += stride)
      epilogue_one(gid, m.n_parts, m.groups, m.oct_mask, m.vec_size, m.partial, x, y, apply,
                   error);
  }
}

// Per member and block: sum (y - x)^2 and sum y^2 over a grid-strided share
// of the member's vector (blockIdx.y = member), for the source-iteration
// change ||phi_new - phi|| / ||phi_new||.
__global__ __launch_bounds__(256) void k_batch_change_norms(const MemberView *__restrict__ mv,
                                                            const int *__restrict__ active,
                                                            const double *const *__restrict__ xs,
                                                            const double *const *__restrict__ ys,
                                                            double *__restrict__ partial) {
  __shared__ double red[2][4];
  const int u = blockIdx.y;
  if (!active[u]) return;
  const long n = mv[u].vec_size;
  const double *x = xs[u], *y = ys[u];
  double d2 = 0.0, y2 = 0.0;
  for (long k = long(blockIdx.x) * blockDim.x + threadIdx.x; k < n;
       k += long(gridDim.x) * blockDim.x) {
    const double yk = y[k], dk = yk - x[k];
    d2 += dk * dk;
    y2 += yk * yk;
  }
  for (int o = 32; o > 0; o >>= 1) {
    d2 += __shfl_xor(d2, o);
    y2 += __shfl_xor(y2, o);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = d2;
    red[1][threadIdx.x >> 6] = y2;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double t = 0.0;
    for (int q = 0; q < 4; ++q) t += red[threadIdx.x][q];
    partial[(long(u) * gridDim.x + blockIdx.x) * 2 + threadIdx.x] = t;
  }
}

// out[u][0..1] = sums of the nb partials of member u (blockIdx.x = member).
__global__ __launch_bounds__(64) void k_batch_sum_pairs(const double *__restrict__ partial,
                                                        int nb, double *__restrict__ out) {
  const int u = blockIdx.x;
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < nb; k += 64) {
    a += partial[(long(u) * nb + k) * 2];
    b += partial[(long(u) * nb + k) * 2 + 1];
  }
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o);
    b += __shfl_xor(b, o);
  }
  if (threadIdx.x == 0) {
    out[2 * u] = a;
    out[2 * u + 1] = b;
  }
}

inline unsigned grid_for(long n, int block) { return unsigned((n + block - 1) / block); }

// ------------------------------------------------------------------------
// Dispatch tables.
// ------------------------------------------------------------------------
// The level kernels of one instantiation: the single-plan launch sequence
// and the batched one (batched: `launch` false only sets the attributes).
struct LevelKernels {
  void (*single)(const Plan &, const SweepView &, hipStream_t, hipEvent_t, hipEvent_t);
  void (*batch)(const BatchLaunch &, const SweepView *, hipStream_t, bool launch);
  void (*persistent)(const Plan &, const SweepView &, hipStream_t, hipEvent_t, hipEvent_t);
  void (*batch_persistent)(const BatchLaunch &, const SweepView *, hipStream_t, bool launch);
};

template <typename T, int DIM, int P, int MM, int W>
void set_lds_attr(const void *kern) {
  using C = SweepCfg<T, DIM, P, MM, W>;
  OU_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                             int(C::lds_bytes)));
}

template <typename T, int DIM, int P, int MM, int W>
void launch_levels_impl(const Plan &pl, const SweepView &v, hipStream_t s,
                        hipEvent_t before, hipEvent_t after) {
  using C = SweepCfg<T, DIM, P, MM, W>;
  auto kern = k_sweep_level<T, DIM, P, MM, W>;
  static bool attr_set = false;
  if (!attr_set) {
    set_lds_attr<T, DIM, P, MM, W>(reinterpret_cast<const void *>(kern));
    attr_set = true;
  }
  if (before) OU_HIP(hipEventRecord(before, s));
  // Octant parts on separate streams: the side streams' launches wait for
  // what s holds before the sweep, and s waits for them before what follows.
  // The parts run on the plan's own streams, created together (so on
  // distinct hardware queues): on a caller's stream, part 0 moves to the
  // plan's stream too, since the caller's stream may share a hardware queue
  // with a side stream and serialise the parts.
  const int ns = pl.sweep_streams;
  const bool foreign = ns > 1 && s != pl.stream;
  hipStream_t part0 = foreign ? pl.stream : s;
  if (ns > 1) {
    OU_HIP(hipEventRecord(pl.ev_fork, s));
    for (int k = 0; k < ns - 1; ++k) OU_HIP(hipStreamWaitEvent(pl.side[k], pl.ev_fork, 0));
    if (foreign) OU_HIP(hipStreamWaitEvent(pl.stream, pl.ev_fork, 0));
  }
  const unsigned groups = unsigned(pl.dir_groups);
  for (int l = 0; l < pl.n_levels; ++l)
    for (int k = 0; k < ns; ++k) {
      const int b = pl.quarter_begin(l, 4 * k / ns), e = pl.quarter_begin(l, 4 * (k + 1) / ns);
      if (e > b)
        hipLaunchKernelGGL(kern, dim3(e - b, groups), dim3(64 * W), C::lds_bytes,
                           k == 0 ? part0 : pl.side[k - 1], v, b);
    }
  if (ns > 1) {
    for (int k = 0; k < ns - 1; ++k) {
      OU_HIP(hipEventRecord(pl.ev_join[k], pl.side[k]));
      OU_HIP(hipStreamWaitEvent(s, pl.ev_join[k], 0));
    }
    if (foreign) {
      OU_HIP(hipEventRecord(pl.ev_join[3], pl.stream));
      OU_HIP(hipStreamWaitEvent(s, pl.ev_join[3], 0));
    }
  }
  if (after) OU_HIP(hipEventRecord(after, s));
  OU_HIP(hipGetLastError());
}

template <typename T, int DIM, int P, int MM, int W>
void launch_batch_levels_impl(const BatchLaunch &bl, const SweepView *views, hipStream_t s,
                              bool launch) {
  using C = SweepCfg<T, DIM, P, MM, W>;
  auto kern = k_sweep_level_batch<T, DIM, P, MM, W>;
  if (!launch) {
    set_lds_attr<T, DIM, P, MM, W>(reinterpret_cast<const void *>(kern));
    return;
  }
  const unsigned groups = unsigned(bl.dir_groups);
  const int n_levels = int(bl.level_off.size()) - 1;
  // the second octant half of every member on the side stream (see
  // launch_levels_impl)
  const bool split = bl.side != nullptr;
  if (split) {
    OU_HIP(hipEventRecord(bl.ev_fork, s));
    OU_HIP(hipStreamWaitEvent(bl.side, bl.ev_fork, 0));
  }
  for (int l = 0; l < n_levels; ++l) {
    const int b = bl.level_off[l], e = bl.level_off[l + 1];
    const int m = split ? bl.level_split[l] : e;
    if (m > b)
      hipLaunchKernelGGL(kern, dim3(m - b, groups), dim3(64 * W), C::lds_bytes, s, views,
                         bl.chunk_view, bl.chunks, b, bl.active);
    if (e > m)
      hipLaunchKernelGGL(kern, dim3(e - m, groups), dim3(64 * W), C::lds_bytes, bl.side, views,
                         bl.chunk_view, bl.chunks, m, bl.active);
  }
  if (split) {
    OU_HIP(hipEventRecord(bl.ev_join, bl.side));
    OU_HIP(hipStreamWaitEvent(s, bl.ev_join, 0));
  }
  OU_HIP(hipGetLastError());
}

// Resident workgroups of a persistent kernel on the whole device (its LDS
// attribute must be set first).
int resident_blocks(const void *kern, int threads, size_t lds, int cus) {
  int per_cu = 0;
  OU_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds));
  return std::max(1, per_cu) * std::max(1, cus);
}

void launch_persist_reset(const PersistView &pv, hipStream_t s) {
  const long n = long(pv.n_items);  // chunks x groups flags
  hipLaunchKernelGGL(k_persist_reset, dim3(std::max(1u, grid_for(n, 256))), dim3(256), 0, s,
                     pv.done, n, pv.ticket);
}

template <typename T, int DIM, int P, int MM, int W>
void launch_persistent_impl(const Plan &pl, const SweepView &v, hipStream_t s,
                            hipEvent_t before, hipEvent_t after) {
  using C = SweepCfg<T, DIM, P, MM, W>;
  auto kern = k_sweep_persistent<T, DIM, P, MM, W>;
  static int blocks_per_device = 0;
  if (!blocks_per_device) {
    set_lds_attr<T, DIM, P, MM, W>(reinterpret_cast<const void *>(kern));
    blocks_per_device =
        resident_blocks(reinterpret_cast<const void *>(kern), 64 * W, C::lds_bytes, pl.cus);
  }
  const PersistView pv = pl.persist_view();
  if (long(pl.n_chunks) * pl.dir_groups > long(INT32_MAX))
    fail(OU_ERR_UNSUPPORTED, "too many sweep items for one persistent launch");
  if (before) OU_HIP(hipEventRecord(before, s));
  launch_persist_reset(pv, s);
  const int grid = std::max(1, std::min(blocks_per_device, pv.n_items));
  if (pv.n_items > 0)
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * W), C::lds_bytes, s, v, pv);
  if (after) OU_HIP(hipEventRecord(after, s));
  OU_HIP(hipGetLastError());
}

template <typename T, int DIM, int P, int MM, int W>
void launch_batch_persistent_impl(const BatchLaunch &bl, const SweepView *views, hipStream_t s,
                                  bool launch) {
  using C = SweepCfg<T, DIM, P, MM, W>;
  auto kern = k_sweep_persistent_batch<T, DIM, P, MM, W>;
  static int blocks_per_device = 0;
  if (!launch) {
    set_lds_attr<T, DIM, P, MM, W>(reinterpret_cast<const void *>(kern));
    if (!blocks_per_device)
      blocks_per_device =
          resident_blocks(reinterpret_cast<const void *>(kern), 64 * W, C::lds_bytes, bl.cus);
    return;
  }
  PersistView pv = bl.persist;
  pv.groups = bl.dir_groups;
  pv.n_items = int(bl.level_off.back() * long(bl.dir_groups));
  launch_persist_reset(pv, s);
  const int grid = std::max(1, std::min(blocks_per_device, pv.n_items));
  if (pv.n_items > 0)
    hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * W), C::lds_bytes, s, views, bl.chunk_view,
                       bl.chunks, bl.active, pv);
  OU_HIP(hipGetLastError());
}

template <typename T, int DIM, int P, int MM, int W>
const LevelKernels *kernels_of() {
  static const LevelKernels k{launch_levels_impl<T, DIM, P, MM, W>,
                              launch_batch_levels_impl<T, DIM, P, MM, W>,
                              launch_persistent_impl<T, DIM, P, MM, W>,
                              launch_batch_persistent_impl<T, DIM, P, MM, W>};
  return &k;
}

// MAX_MM bounds the moment counts built for an element: the higher orders
// keep their FP64 accumulators (MM x N / W per lane) within the registers.
template <typename T, int DIM, int P, int W, int MAX_MM = 16>
const LevelKernels *pick_mm(int mm) {
  switch (mm) {
    case 1: return kernels_of<T, DIM, P, 1, W>();
    case 4: if constexpr (MAX_MM >= 4) return kernels_of<T, DIM, P, 4, W>(); break;
    case 10: if constexpr (W == 4 && MAX_MM >= 10) return kernels_of<T, DIM, P, 10, W>(); break;
    case 16: if constexpr (W == 4 && MAX_MM >= 16) return kernels_of<T, DIM, P, 16, W>(); break;
  }
  return nullptr;
}

// Instantiated shapes: 4 waves everywhere; the 3D shapes with at most 4
// moments also at the other wave counts the autotuner tries.
template <typename T>
const LevelKernels *pick(int dim, int p, int mm, int waves) {
  if (waves == 4) {
    if (dim == 2 && p == 1) return pick_mm<T, 2, 1, 4>(mm);
    if (dim == 2 && p == 2) return pick_mm<T, 2, 2, 4>(mm);
    if (dim == 2 && p == 3) return pick_mm<T, 2, 3, 4>(mm);
    if (dim == 2 && p == 4) return pick_mm<T, 2, 4, 4, 10>(mm);
    if (dim == 2 && p == 5) return pick_mm<T, 2, 5, 4, 10>(mm);
    if (dim == 3 && p == 1) return pick_mm<T, 3, 1, 4>(mm);
    if (dim == 3 && p == 2) return pick_mm<T, 3, 2, 4>(mm);
    if (dim == 3 && p == 3) return pick_mm<T, 3, 3, 4, 4>(mm);
  } else if (dim == 3 && mm <= 4) {
    if (p == 1) {
      if (waves == 2) return pick_mm<T, 3, 1, 2>(mm);
      if (waves == 8) return pick_mm<T, 3, 1, 8>(mm);
    } else if (p == 2) {
      if (waves == 6) return pick_mm<T, 3, 2, 6>(mm);
      if (waves == 8) return pick_mm<T, 3, 2, 8>(mm);
      if (waves == 12) return pick_mm<T, 3, 2, 12>(mm);
    }
  }
  return nullptr;
}

template <typename T>
size_t lds_need(int dim, int p, int mm, int waves) {
  const int P1 = p + 1, N = dim == 2 ? P1 * P1 : P1 * P1 * P1;
  const int NP = sizeof(T) == 4 ? (N + 3) & ~3 : (N + 1) & ~1;
  const size_t psi = size_t(waves) * 64 * NP * sizeof(T);
  const size_t src = size_t(mm) * 64 * NP * sizeof(T);
  const size_t tab =
      size_t(kMaxSlots) * (2 * sizeof(int) + (mm + 4) * sizeof(T) + mm * sizeof(double));
  return psi + (psi + src + tab <= 120 * 1024 ? src : 0) + tab + 256;
}

template <typename T>
void scatter_impl(Plan &pl, const double *d_x, const double *d_sigma, bool accumulate,
                  hipStream_t s) {
  const long n = pl.n_cells * pl.n_mom;
  T *S = static_cast<T *>(pl.d_S);
#define OU_SCAT(NN)                                                                  \
  case NN:                                                                           \
    hipLaunchKernelGGL((k_scatter_source<T, NN>), dim3(grid_for(n, 256)), dim3(256), \
                       0, s, pl.n_cells, pl.n_mom, pl.n_dofs, d_x, pl.d_mat, d_sigma, \
                       pl.L, pl.d_mom_l, pl.d_mass, S, int(accumulate));              \
    break;
  switch (pl.N) {
    OU_SCAT(4) OU_SCAT(9) OU_SCAT(16) OU_SCAT(25) OU_SCAT(36) OU_SCAT(8) OU_SCAT(27) OU_SCAT(64)
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
    OU_SRC(4) OU_SRC(9) OU_SRC(16) OU_SRC(25) OU_SRC(36) OU_SRC(8) OU_SRC(27) OU_SRC(64)
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
  v.n_cls = n_cls;
  v.slot_cls = d_slot_cls;
  v.ainv = d_ainv;
  v.ainv_mfma = d_ainv_mfma;
  v.ainv_mfma64 = d_ainv_mfma64;
  v.S = d_S;
  v.S0 = nullptr;
  v.slot_m2d = d_slot_m2d;
  v.slot_d2m = d_slot_d2m;
  v.slot_c = d_slot_c;
  for (int k = 0; k < kMaxP1Sq; ++k) {
    v.m1f[k] = float(m1[k]);
    v.m1d[k] = m1[k];
  }
  v.oct_dirs = d_oct_dirs;
  v.oct_off = d_oct_off;
  v.partial = d_partial;
  v.chunks = d_chunks;
  v.cells = d_cells;
  v.dir_groups = dir_groups;
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
  OU_HIP(hipMalloc(&dom, pl.cls_omega.size() * 8));
  OU_HIP(hipMalloc(&dst, pl.sigma_t.size() * 8));
  OU_HIP(hipMemcpy(dM, M.data(), M.size() * 8, hipMemcpyHostToDevice));
  OU_HIP(hipMemcpy(dG, G.data(), G.size() * 8, hipMemcpyHostToDevice));
  OU_HIP(hipMemcpy(dE, E.data(), E.size() * 8, hipMemcpyHostToDevice));
  OU_HIP(hipMemcpy(dom, pl.cls_omega.data(), pl.cls_omega.size() * 8, hipMemcpyHostToDevice));
  OU_HIP(hipMemcpy(dst, pl.sigma_t.data(), pl.sigma_t.size() * 8, hipMemcpyHostToDevice));
  const size_t lds = (size_t(N) * N + size_t(N) * 64) * sizeof(double);
  const unsigned blocks = unsigned(pl.n_mat * pl.n_cls);
  if (pl.precision == OU_FP32)
    hipLaunchKernelGGL(k_build_inverses<float>, dim3(blocks), dim3(64), lds, pl.stream, N,
                       pl.dim, pl.p, pl.n_cls, dom, dst, dM, dG, dE,
                       static_cast<float *>(pl.d_ainv));
  else
    hipLaunchKernelGGL(k_build_inverses<double>, dim3(blocks), dim3(64), lds, pl.stream, N,
                       pl.dim, pl.p, pl.n_cls, dom, dst, dM, dG, dE,
                       static_cast<double *>(pl.d_ainv));
  OU_HIP(hipGetLastError());
  if (pl.use_mfma)
    hipLaunchKernelGGL(k_ainv_to_mfma, dim3(blocks), dim3(64), 0, pl.stream,
                       static_cast<const float *>(pl.d_ainv), pl.d_ainv_mfma);
  if (pl.use_mfma64)
    hipLaunchKernelGGL(k_ainv_to_mfma64, dim3(blocks), dim3(64), 0, pl.stream,
                       static_cast<const double *>(pl.d_ainv), pl.d_ainv_mfma64);
  OU_HIP(hipGetLastError());
  OU_HIP(hipStreamSynchronize(pl.stream));
  (void)hipFree(dM);
  (void)hipFree(dG);
  (void)hipFree(dE);
  (void)hipFree(dom);
  (void)hipFree(dst);
}

bool sweep_supported(int dim, int p, int mm, int precision, int waves, int max_lds) {
  if (precision == OU_FP32)
    return pick<float>(dim, p, mm, waves) != nullptr &&
           lds_need<float>(dim, p, mm, waves) <= size_t(max_lds);
  return pick<double>(dim, p, mm, waves) != nullptr &&
         lds_need<double>(dim, p, mm, waves) <= size_t(max_lds);
}

void launch_scatter_source(Plan &pl, const double *d_x, hipStream_t s) {
  launch_scatter_source_with(pl, d_x, pl.d_sigma_s, false, s);
}

void launch_scatter_source_with(Plan &pl, const double *d_x, const double *d_sigma,
                                bool accumulate, hipStream_t s) {
  if (pl.precision == OU_FP32)
    scatter_impl<float>(pl, d_x, d_sigma, accumulate, s);
  else
    scatter_impl<double>(pl, d_x, d_sigma, accumulate, s);
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
  const LevelKernels *k = pl.precision == OU_FP32 ? pick<float>(pl.dim, pl.p, pl.mm, pl.waves)
                                                  : pick<double>(pl.dim, pl.p, pl.mm, pl.waves);
  if (!k) fail(OU_ERR_UNSUPPORTED, "no sweep kernel for this element / moment count");
  if (pl.sweep_mode == OU_SWEEP_TILES) {
    v.dir_groups = pl.tile_groups;
    launch_sweep_tiles(pl, v, s, before, after);
  } else if (pl.sweep_mode == OU_SWEEP_PERSISTENT) {
    k->persistent(pl, v, s, before, after);
  } else {
    k->single(pl, v, s, before, after);
  }
}

void launch_epilogue(Plan &pl, const double *d_x, double *d_y, bool apply, hipStream_t s) {
  const long n = pl.vec_size;
  const int groups = pl.groups_in_use();
  hipLaunchKernelGGL(k_epilogue, dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                     pl.n_oct * groups, groups, pl.oct_mask, pl.vec_size, pl.d_partial, d_x,
                     d_y, int(apply),
                     pl.sweep_mode == OU_SWEEP_PERSISTENT ? pl.d_error : nullptr);
  OU_HIP(hipGetLastError());
}

void apply_device(Plan &pl, const double *d_x, double *d_y, hipStream_t s) {
  launch_scatter_source(pl, d_x, s);
  launch_trace_init(pl, false, s);
  launch_sweep_levels(pl, nullptr, s);
  launch_epilogue(pl, d_x, d_y, true, s);
}

namespace {

const LevelKernels *batch_kernels(const BatchLaunch &b) {
  const LevelKernels *k = b.precision == OU_FP32 ? pick<float>(b.dim, b.p, b.mm, b.waves)
                                                 : pick<double>(b.dim, b.p, b.mm, b.waves);
  if (!k) fail(OU_ERR_UNSUPPORTED, "no batched sweep kernel for this shape");
  return k;
}

unsigned batch_grid(const BatchLaunch &b) {
  return std::max(1u, std::min(4096u, grid_for(b.max_elems, 256)));
}

template <typename T>
void batch_scatter_impl(const BatchLaunch &b, const double *const *d_xs, hipStream_t s) {
  const unsigned g = batch_grid(b);
#define OU_BSCAT(NN)                                                                   \
  case NN:                                                                             \
    hipLaunchKernelGGL((k_batch_scatter<T, NN>), dim3(g), dim3(256), 0, s, b.members, \
                       b.n_members, b.active, d_xs);                                   \
    break;
  switch (b.N) {
    OU_BSCAT(4) OU_BSCAT(9) OU_BSCAT(16) OU_BSCAT(25) OU_BSCAT(36) OU_BSCAT(8) OU_BSCAT(27) OU_BSCAT(64)
    default: fail(OU_ERR_UNSUPPORTED, "batched scatter: unsupported element size");
  }
#undef OU_BSCAT
}

template <typename T>
void batch_source0_impl(const BatchLaunch &b, hipStream_t s) {
  const unsigned g = batch_grid(b);
#define OU_BSRC(NN)                                                                    \
  case NN:                                                                             \
    hipLaunchKernelGGL((k_batch_source0<T, NN>), dim3(g), dim3(256), 0, s, b.members, \
                       b.n_members, b.active);                                         \
    break;
  switch (b.N) {
    OU_BSRC(4) OU_BSRC(9) OU_BSRC(16) OU_BSRC(25) OU_BSRC(36) OU_BSRC(8) OU_BSRC(27) OU_BSRC(64)
    default: fail(OU_ERR_UNSUPPORTED, "batched source: unsupported element size");
  }
#undef OU_BSRC
}

void batch_trace_init(const BatchLaunch &b, bool with_source, hipStream_t s) {
  const unsigned g = batch_grid(b);
  if (b.precision == OU_FP32)
    hipLaunchKernelGGL(k_batch_trace_init<float>, dim3(g), dim3(256), 0, s, b.members,
                       b.n_members, b.active, b.dim, int(with_source));
  else
    hipLaunchKernelGGL(k_batch_trace_init<double>, dim3(g), dim3(256), 0, s, b.members,
                       b.n_members, b.active, b.dim, int(with_source));
}

void batch_epilogue(const BatchLaunch &b, const double *const *d_xs, double *const *d_ys,
                    int mode, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_epilogue, dim3(batch_grid(b)), dim3(256), 0, s, b.members,
                     b.n_members, b.active, d_xs, d_ys, mode,
                     b.sweep_mode == OU_SWEEP_PERSISTENT ? b.persist.error : nullptr);
}

// The merged levels of every member: one persistent launch or level launches.
void batch_levels(const BatchLaunch &b, const LevelKernels *k, const SweepView *views,
                  hipStream_t s) {
  if (b.sweep_mode == OU_SWEEP_TILES)
    launch_batch_tiles(b, views, s, true);
  else if (b.sweep_mode == OU_SWEEP_PERSISTENT)
    k->batch_persistent(b, views, s, true);
  else
    k->batch(b, views, s, true);
}

}  // namespace

void batch_prepare(const BatchLaunch &b) {
  const LevelKernels *k = batch_kernels(b);
  k->batch(b, nullptr, nullptr, false);
  k->batch_persistent(b, nullptr, nullptr, false);
  if (b.tiles_ok) launch_batch_tiles(b, nullptr, nullptr, false);
}

void batch_apply(const BatchLaunch &b, const double *const *d_xs, double *const *d_ys,
                 hipStream_t s) {
  const LevelKernels *k = batch_kernels(b);
  if (b.precision == OU_FP32)
    batch_scatter_impl<float>(b, d_xs, s);
  else
    batch_scatter_impl<double>(b, d_xs, s);
  batch_trace_init(b, false, s);
  batch_levels(b, k, b.views, s);
  batch_epilogue(b, d_xs, d_ys, 1, s);
  OU_HIP(hipGetLastError());
}

void batch_sweep(const BatchLaunch &b, double *const *d_ys, bool with_source, hipStream_t s) {
  const LevelKernels *k = batch_kernels(b);
  if (with_source) {
    if (b.precision == OU_FP32)
      batch_source0_impl<float>(b, s);
    else
      batch_source0_impl<double>(b, s);
  }
  batch_trace_init(b, with_source, s);
  batch_levels(b, k, with_source ? b.views_src : b.views, s);
  batch_epilogue(b, nullptr, d_ys, 0, s);
  OU_HIP(hipGetLastError());
}

void batch_si_step(const BatchLaunch &b, const double *const *d_xs, double *const *d_ys,
                   double *d_partial, int partial_blocks, double *d_norms, hipStream_t s) {
  const LevelKernels *k = batch_kernels(b);
  if (b.precision == OU_FP32) {
    batch_scatter_impl<float>(b, d_xs, s);
    batch_source0_impl<float>(b, s);
  } else {
    batch_scatter_impl<double>(b, d_xs, s);
    batch_source0_impl<double>(b, s);
  }
  batch_trace_init(b, true, s);
  batch_levels(b, k, b.views_src, s);
  batch_epilogue(b, nullptr, d_ys, 2, s);
  hipLaunchKernelGGL(k_batch_change_norms, dim3(unsigned(partial_blocks), unsigned(b.n_members)),
                     dim3(256), 0, s, b.members, b.active, d_xs,
                     static_cast<const double *const *>(d_ys), d_partial);
  hipLaunchKernelGGL(k_batch_sum_pairs, dim3(unsigned(b.n_members)), dim3(64), 0, s, d_partial,
                     partial_blocks, d_norms);
  OU_HIP(hipGetLastError());
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
