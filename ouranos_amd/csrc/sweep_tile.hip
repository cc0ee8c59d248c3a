// The tile sweep (OU_SWEEP_TILES): one launch per tile level instead of
// one per cell level.
//
// A workgroup sweeps one 8 x 8 tile of one quadrant (all of one material) for
// a group of that quadrant's directions.  Wave w holds 8 directions, lane =
// 8 x direction + a, a the tile column: at tile-internal level t the lane
// solves cell (a, t - a) of its direction.  Within the tile the traces never
// leave the registers: the cell's y-inflow is the y-outflow the same lane
// computed one level earlier, its x-inflow the x-outflow of lane - 1 (a
// row_shr:1 lane shift).  Only the tile's upwind edges read the trace slabs
// (written by the tiles of the previous tile level, the previous launch) and
// only its downwind edges write them.  A 256 x 256 quadrant is then 63
// launches of 15 register-resident levels instead of 511 launches that each
// pass every trace through memory.
//
// The 3D Q1 tiles are 4 x 4 x 4 cells: lane = 16 x direction + 4 b + a,
// cell (a, b, t - a - b) at tile level t.
//
// The per-cell arithmetic is the level sweep's own, in the same order (M2D
// rows of the staged source, upwind terms x then y, the (material, class)
// inverse, FP64 D2M accumulation in slot order per direction group), so at
// the same direction groups both modes give the same bits.
// (RadiativeTransfer.cc:878-1012 per cell; the level sweep in sweep.hip.)
// The 3D Q2 tiles solve on the matrix cores instead (sweep_tile_q2.inc).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "plan.hpp"
#include "sweep_common.hpp"

namespace ou {

namespace {

// Tile geometry per dimension: edge, lanes (positions) per direction, directions
// per wave, cell levels inside a tile.
template <int DIM>
struct TileGeom {
  static constexpr int TE = DIM == 2 ? kTileEdge2 : kTileEdge3;
  static constexpr int PPD = DIM == 2 ? TE : TE * TE;
  static constexpr int DPW = 64 / PPD;
  static constexpr int NL = DIM * (TE - 1) + 1;
};

template <typename T, int DIM, int P, int MM, int W>
struct TileCfg {
  using G = TileGeom<DIM>;
  static constexpr int P1 = P + 1, N = DIM == 2 ? P1 * P1 : P1 * P1 * P1;
  static constexpr int F = DIM == 2 ? P1 : P1 * P1;
  // LDS stride of one cell, padded to 16 bytes for 128-bit LDS accesses
  static constexpr int NP = sizeof(T) == 4 ? (N + 3) & ~3 : (N + 1) & ~1;
  static constexpr int NC = (N + W - 1) / W;  // (cell, dof) columns per lane in D2M
  static constexpr size_t psi_bytes = size_t(W) * G::DPW * 64 * NP * sizeof(T);
  static constexpr size_t src_bytes = size_t(MM) * 64 * NP * sizeof(T);
  static constexpr size_t d2m_bytes = size_t(kMaxSlotsPerGroup) * MM * sizeof(double);
  static constexpr size_t lds_bytes = psi_bytes + src_bytes + d2m_bytes;
};

// x of lane l - S within each row of 16 lanes (DPP row_shr:S); a lane whose
// source would fall before its row's start keeps its own value.
template <int S>
__device__ __forceinline__ int lane_shr_bits(int b) {
  return __builtin_amdgcn_update_dpp(b, b, 0x110 + S, 0xF, 0xF, false);
}

template <int S>
__device__ __forceinline__ float lane_shr(float x) {
  return __int_as_float(lane_shr_bits<S>(__float_as_int(x)));
}

template <int S>
__device__ __forceinline__ double lane_shr(double x) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
  const unsigned lo = unsigned(lane_shr_bits<S>(int(unsigned(b))));
  const unsigned hi = unsigned(lane_shr_bits<S>(int(unsigned(b >> 32))));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// out = cand[k] for the lane's k (0 <= k < K), zeros otherwise: a select
// chain, with no register array indexed by a lane-dependent value.
template <typename T, int K, int F>
__device__ __forceinline__ void select_of(const T (&cand)[K][F], int k, T *out) {
#pragma unroll
  for (int f = 0; f < F; ++f) out[f] = T(0);
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int f = 0; f < F; ++f) out[f] = k == j ? cand[j][f] : out[f];
}

// One tile of one octant for direction group `group` (see the file comment).
// 2D: lane = 8 x direction + a, cell (a, t - a) at tile level t; x-inflow
// from lane - 1, y-inflow the lane's own previous cell.  3D: lane = 16 x
// direction + 4 b + a, cell (a, b, t - a - b); x-inflow from lane - 1,
// y-inflow from lane - 4, z-inflow the lane's own previous cell.
template <typename T, int DIM, int P, int MM, int W>
__device__ __forceinline__ void sweep_tile_item(const SweepView &v, const TileItem tile,
                                                int group) {
  using C = TileCfg<T, DIM, P, MM, W>;
  using G = TileGeom<DIM>;
  constexpr int N = C::N, F = C::F, NP = C::NP, NC = C::NC;
  constexpr int TE = G::TE, PPD = G::PPD, DPW = G::DPW, NL = G::NL;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  T *s_psi = reinterpret_cast<T *>(lds_raw);                      // [W*DPW slots][64 cells][NP]
  T *s_src = reinterpret_cast<T *>(lds_raw + C::psi_bytes);       // [MM][64 cells][NP]
  double *t_d2m = reinterpret_cast<double *>(lds_raw + C::psi_bytes + C::src_bytes);  // [slot][MM]

  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int o = tile.octant;
  const int nx = v.n[0], ny = v.n[1], nz = v.n[2];
  const int n_groups = v.dir_groups, dall = v.oct_off[o + 1] - v.oct_off[o];
  const int glen = (dall + n_groups - 1) / n_groups;
  const int dbeg = v.oct_off[o] + group * glen;
  const int ndir_o = max(0, min(glen, dall - group * glen));
  const int ii0 = tile.ti * TE, jj0 = tile.tj * TE, kk0 = tile.tk * TE;
  const long n_dofs = v.n_dofs;
  const int n_mom = v.n_mom;

  // Tile cell cl = pos + PPD e (pos = a + TE b in 3D): canonical
  // (ii0 + a, jj0 + b[, kk0 + c]), -1 outside the mesh.
  auto cell_of = [&](int cl) {
    const int ii = ii0 + cl % TE;
    const int jj = jj0 + (cl / TE) % TE;
    const int kk = DIM == 3 ? kk0 + cl / (TE * TE) : 0;
    if (ii >= nx || jj >= ny || kk >= nz) return -1;
    const int i = (o & 1) ? nx - 1 - ii : ii;
    const int j = (o & 2) ? ny - 1 - jj : jj;
    const int k = (DIM == 3 && (o & 4)) ? nz - 1 - kk : kk;
    return i + nx * (j + ny * k);
  };

  // ---- stage the group's D2M columns and the tile's source ----------------
  for (int e = tid; e < ndir_o * MM; e += 64 * W) t_d2m[e] = v.slot_d2m[long(dbeg) * MM + e];
  {
    const T *S = static_cast<const T *>(v.S);
    const T *S0 = static_cast<const T *>(v.S0);
    constexpr int Q = 16 / int(sizeof(T)), NQ = (N + Q - 1) / Q;
    constexpr int KS = (MM * 64 * NQ + 64 * W - 1) / (64 * W), KB = KS < 4 ? KS : 4;
#pragma unroll 1
    for (int k0 = 0; k0 < KS; k0 += KB) {
      T val[KB][Q];
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const int e = tid + (k0 + k) * 64 * W;
        const int q = e % NQ, mc = e / NQ, m = mc / 64, cl = mc % 64;
#pragma unroll
        for (int j = 0; j < Q; ++j) val[k][j] = T(0);
        const int c = e < MM * 64 * NQ ? cell_of(cl) : -1;
        if (c >= 0 && m < n_mom) {
          const long g = long(c) * N + q * Q;
          const T *src = (m == 0 && S0) ? S0 + g : S + long(m) * n_dofs + g;
          if (q * Q + Q <= N) {
#pragma unroll
            for (int j = 0; j < Q; ++j) val[k][j] = src[j];
          } else {
#pragma unroll
            for (int j = 0; j < Q; ++j)
              if (q * Q + j < N) val[k][j] = src[j];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const int e = tid + (k0 + k) * 64 * W;
        const int q = e % NQ, mc = e / NQ, m = mc / 64, cl = mc % 64;
        if (e < MM * 64 * NQ) {
#pragma unroll
          for (int j = 0; j < Q; ++j)
            if (q * Q + j < N)
              s_src[m * 64 * NP + cl * NP + natural_index<DIM, P>(q * Q + j, o)] = val[k][j];
        }
      }
    }
  }
  lds_barrier();

  const T *m1 = m1_of<T>(v);
  const int pos = lane % PPD, dsub = lane / PPD;
  const int a = pos % TE, b = DIM == 3 ? pos / TE : 0;  // 2D: b is the level's row
  const int ii = ii0 + a, jj3 = jj0 + b;
  const bool col_in = ii < nx && (DIM == 2 || jj3 < ny);
  const long ps0 = v.plane_size[0], ps1 = v.plane_size[1], ps2 = v.plane_size[2];

  double acc[MM][NC];
#pragma unroll
  for (int m = 0; m < MM; ++m)
#pragma unroll
    for (int t = 0; t < NC; ++t) acc[m][t] = 0.0;

  for (int r0 = 0; r0 < ndir_o; r0 += W * DPW) {
    const int sl = r0 + w * DPW + dsub;  // slot within the group
    const bool has_dir = sl < ndir_o;
    const int gsl = dbeg + (has_dir ? sl : 0);
    const int d = v.oct_dirs[gsl];
    const int cls = v.slot_cls[gsl];
    T m2d[MM];
#pragma unroll
    for (int m = 0; m < MM; ++m)
      m2d[m] = has_dir ? static_cast<const T *>(v.slot_m2d)[long(gsl) * MM + m] : T(0);
    T cf[DIM];
#pragma unroll
    for (int x = 0; x < DIM; ++x) cf[x] = static_cast<const T *>(v.slot_c)[long(gsl) * 4 + x];
    // the (tile material, class) inverse, canonical frame, in registers
    const T *Ai = static_cast<const T *>(v.ainv) + (long(tile.material) * v.n_cls + cls) * (N * N);
    T A[N * N];
#pragma unroll
    for (int k = 0; k < N * N; ++k) A[k] = Ai[k];
    // trace slabs of direction d
    T *const dx = static_cast<T *>(v.trace[0]) + long(d) * F * ps0;
    T *const dy = static_cast<T *>(v.trace[1]) + long(d) * F * ps1;
    T *const dz = DIM == 3 ? static_cast<T *>(v.trace[2]) + long(d) * F * ps2 : nullptr;

    // Upwind edges of the tile.  2D: rows jj0 + b (lanes a = 0, row = level),
    // column ii0 + a.  3D: x-face cells (0, b, c) and y-face cells (a, 0, c)
    // for c = 0..3 (the level picks c), z-face cell (a, b, 0).
    T xin[TE][F], yin[DIM == 2 ? 1 : TE][F], zin[F];
    if constexpr (DIM == 2) {
#pragma unroll
      for (int r = 0; r < TE; ++r)
#pragma unroll
        for (int f = 0; f < F; ++f)
          xin[r][f] = (a == 0 && jj0 + r < ny) ? dx[long(jj0 + r) * F + f] : T(0);
#pragma unroll
      for (int f = 0; f < F; ++f) yin[0][f] = col_in ? dy[long(ii) * F + f] : T(0);
    } else {
#pragma unroll
      for (int c = 0; c < TE; ++c) {
        const int kk = kk0 + c;
#pragma unroll
        for (int f = 0; f < F; ++f) {
          xin[c][f] = (a == 0 && jj3 < ny && kk < nz) ? dx[(long(jj3) + long(ny) * kk) * F + f] : T(0);
          yin[c][f] = (b == 0 && ii < nx && kk < nz) ? dy[(long(ii) + long(nx) * kk) * F + f] : T(0);
        }
      }
#pragma unroll
      for (int f = 0; f < F; ++f)
        zin[f] = col_in ? dz[(long(ii + jj3) * ny + jj3) * F + f] : T(0);
    }

    // M2D source term of the lane's cell at tile level t (clamped to the
    // tile for lanes outside it at that level)
    auto source_term = [&](int t, T *out) {
      const int e = t - (DIM == 2 ? a : a + b);
      const int cl = pos + PPD * min(max(e, 0), TE - 1);
#pragma unroll
      for (int i = 0; i < N; ++i) out[i] = T(0);
#pragma unroll
      for (int m = 0; m < MM; ++m) {
        const T *src = s_src + m * 64 * NP + cl * NP;
#pragma unroll
        for (int i = 0; i < N; ++i) out[i] += m2d[m] * src[i];
      }
    };
    // FP32 with up to 10 moments forms it a level ahead (the registers allow)
    constexpr bool ahead = sizeof(T) == 4 && MM <= 10;
    T tx[F], ty[F], tz[F], bnext[N];
#pragma unroll
    for (int f = 0; f < F; ++f) tx[f] = ty[f] = tz[f] = T(0);
    if constexpr (ahead) source_term(0, bnext);
#pragma unroll
    for (int t = 0; t < NL; ++t) {
      const int e = t - (DIM == 2 ? a : a + b);  // the last coordinate of the lane's cell
      const int ec = min(max(e, 0), TE - 1);
      const int cl = pos + PPD * ec;
      T xi[F], yi[F], zi[F];
      bool active;
      if constexpr (DIM == 2) {
        active = has_dir && e >= 0 && e < TE && col_in && jj0 + e < ny;
#pragma unroll
        for (int f = 0; f < F; ++f) {
          xi[f] = a == 0 ? (t < TE ? xin[t < TE ? t : 0][f] : T(0)) : tx[f];
          yi[f] = e == 0 ? yin[0][f] : ty[f];
        }
      } else {
        active = has_dir && e >= 0 && e < TE && col_in && kk0 + e < nz;
        T xs[F], ys[F];
        select_of<T, TE, F>(xin, e, xs);
        select_of<T, TE, F>(yin, e, ys);
#pragma unroll
        for (int f = 0; f < F; ++f) {
          xi[f] = a == 0 ? xs[f] : tx[f];
          yi[f] = b == 0 ? ys[f] : ty[f];
          zi[f] = e == 0 ? zin[f] : tz[f];
        }
      }
      // the source term depends on no earlier cell: formed one level ahead it
      // stays off the chain of dependent cell solves
      T bv[N];
      if constexpr (ahead) {
#pragma unroll
        for (int i = 0; i < N; ++i) bv[i] = bnext[i];
        if (t + 1 < NL) source_term(t + 1, bnext);
      } else {
        source_term(t, bv);
      }
      add_upwind<T, DIM, P, 0>(m1, cf[0], xi, bv);
      add_upwind<T, DIM, P, 1>(m1, cf[1], yi, bv);
      if constexpr (DIM == 3) add_upwind<T, DIM, P, 2>(m1, cf[2], zi, bv);
      T psi[N];
#pragma unroll
      for (int r = 0; r < N; ++r) {
        T s = T(0);
#pragma unroll
        for (int q = 0; q < N; ++q) s += A[r * N + q] * bv[q];
        psi[r] = s;
      }
      if (active) {
        T *dst = s_psi + ((w * DPW + dsub) * 64 + cl) * NP;
#pragma unroll
        for (int i = 0; i < N; ++i) dst[i] = psi[i];
        // downwind faces of the tile (or of the mesh) go to the trace slabs
        if constexpr (DIM == 2) {
          if (a == TE - 1 || ii == nx - 1) store_trace<T, 2, P, 0>(dx, unsigned(jj0 + e), psi);
          if (e == TE - 1 || jj0 + e == ny - 1) store_trace<T, 2, P, 1>(dy, unsigned(ii), psi);
        } else {
          const int kk = kk0 + e;
          if (a == TE - 1 || ii == nx - 1)
            store_trace<T, 3, P, 0>(dx, unsigned(jj3 + ny * kk), psi);
          if (b == TE - 1 || jj3 == ny - 1)
            store_trace<T, 3, P, 1>(dy, unsigned(ii + nx * kk), psi);
          if (e == TE - 1 || kk == nz - 1)
            store_trace<T, 3, P, 2>(dz, unsigned((ii + jj3) * ny + jj3), psi);
        }
      }
      // next level: x-outflow to the lane of a + 1, 3D y-outflow to the lane
      // of b + 1, the last axis's outflow kept
#pragma unroll
      for (int f = 0; f < F; ++f) {
        tx[f] = lane_shr<1>(psi[face_dof<DIM, P, 0, P>(f)]);
        if constexpr (DIM == 2) {
          ty[f] = psi[face_dof<DIM, P, 1, P>(f)];
        } else {
          ty[f] = lane_shr<TE>(psi[face_dof<DIM, P, 1, P>(f)]);
          tz[f] = psi[face_dof<DIM, P, 2, P>(f)];
        }
      }
    }
    lds_barrier();
    // phi_m += D2M[m, d] psi_d over the round's directions in slot order, for
    // this wave's share of the tile's (cell, dof) columns
    int col_off[NC];
#pragma unroll
    for (int t = 0; t < NC; ++t) {
      const int col = (t * W + w) * 64 + lane;
      const int cl = col / N;
      col_off[t] = col < 64 * N ? cl * NP + (col - cl * N) : 0;
    }
    const int n_round = min(W * DPW, ndir_o - r0);
#pragma unroll 1
    for (int s = 0; s < n_round; ++s) {
      double dm[MM];
#pragma unroll
      for (int m = 0; m < MM; ++m) dm[m] = t_d2m[(r0 + s) * MM + m];
      const T *src = s_psi + s * 64 * NP;
#pragma unroll
      for (int t = 0; t < NC; ++t) {
        const double pv = double(src[col_off[t]]);
#pragma unroll
        for (int m = 0; m < MM; ++m) acc[m][t] += dm[m] * pv;
      }
    }
    lds_barrier();
  }

#pragma unroll
  for (int t = 0; t < NC; ++t) {
    const int col = (t * W + w) * 64 + lane;
    if (col < 64 * N) {
      const int cl = col / N, ic = col - cl * N;
      const int c = cell_of(cl);
      if (c >= 0) {
        double *out = v.partial + (long(o) * n_groups + group) * n_mom * n_dofs + long(c) * N +
                      natural_index<DIM, P>(ic, o);
#pragma unroll
        for (int m = 0; m < MM; ++m)
          if (m < n_mom) out[long(m) * n_dofs] = acc[m][t];
      }
    }
  }
}

template <typename T, int DIM, int P, int MM, int W>
__global__ __launch_bounds__(64 * W) void k_sweep_tile(SweepView v,
                                                       const TileItem *__restrict__ tiles,
                                                       int tile_begin) {
  sweep_tile_item<T, DIM, P, MM, W>(v, tiles[tile_begin + blockIdx.x], blockIdx.y);
}

// The batch's merged tile levels: tile i belongs to the member whose view is
// views[tile_view[i]]; members whose flag is cleared are skipped.
template <typename T, int DIM, int P, int MM, int W>
__global__ __launch_bounds__(64 * W) void k_sweep_tile_batch(const SweepView *__restrict__ views,
                                                             const TileItem *__restrict__ tiles,
                                                             const int *__restrict__ tile_view,
                                                             int tile_begin,
                                                             const int *__restrict__ active) {
  const int i = tile_begin + int(blockIdx.x);
  const int u = tile_view[i];
  if (!active[u]) return;  // the whole workgroup, before any barrier
  sweep_tile_item<T, DIM, P, MM, W>(views[u], tiles[i], blockIdx.y);
}

#include "sweep_tile_q2.inc"

using TileLaunch = void (*)(const Plan &, const SweepView &, hipStream_t);
using BatchTileLaunch = void (*)(const BatchLaunch &, const SweepView *, hipStream_t, bool);

template <typename T, int DIM, int P, int MM, int W>
void launch_batch_tiles_impl(const BatchLaunch &bl, const SweepView *views, hipStream_t s,
                             bool launch) {
  using C = TileCfg<T, DIM, P, MM, W>;
  auto kern = k_sweep_tile_batch<T, DIM, P, MM, W>;
  if (!launch) {
    OU_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(C::lds_bytes)));
    return;
  }
  const unsigned groups = unsigned(bl.tile_groups);
  for (size_t l = 0; l + 1 < bl.tile_level_off.size(); ++l) {
    const int b = bl.tile_level_off[l], e = bl.tile_level_off[l + 1];
    if (e > b)
      hipLaunchKernelGGL(kern, dim3(e - b, groups), dim3(64 * W), C::lds_bytes, s, views,
                         bl.tiles, bl.tile_view, b, bl.active);
  }
}

template <typename T, int DIM, int P, int MM, int W>
void launch_tiles_impl(const Plan &pl, const SweepView &v, hipStream_t s) {
  using C = TileCfg<T, DIM, P, MM, W>;
  auto kern = k_sweep_tile<T, DIM, P, MM, W>;
  static bool attr_set = false;
  if (!attr_set) {
    OU_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(C::lds_bytes)));
    attr_set = true;
  }
  const unsigned groups = unsigned(pl.tile_groups);
  for (size_t l = 0; l + 1 < pl.tile_level_off.size(); ++l) {
    const int b = pl.tile_level_off[l], e = pl.tile_level_off[l + 1];
    if (e > b)
      hipLaunchKernelGGL(kern, dim3(e - b, groups), dim3(64 * W), C::lds_bytes, s, v,
                         pl.d_tiles, b);
  }
}

template <int MM, int W>
void launch_q2_batch_tiles_impl(const BatchLaunch &bl, const SweepView *views, hipStream_t s,
                                bool launch) {
  using C = Q2TileCfg<MM, W>;
  auto kern = k_sweep_tile_q2_batch<MM, W>;
  if (!launch) {
    OU_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(C::lds_bytes)));
    return;
  }
  const unsigned groups = unsigned(bl.tile_groups);
  for (size_t l = 0; l + 1 < bl.tile_level_off.size(); ++l) {
    const int b = bl.tile_level_off[l], e = bl.tile_level_off[l + 1];
    if (e > b)
      hipLaunchKernelGGL(kern, dim3(e - b, groups), dim3(64 * W), C::lds_bytes, s, views,
                         bl.tiles, bl.tile_view, b, bl.active);
  }
}

template <int MM, int W>
void launch_q2_tiles_impl(const Plan &pl, const SweepView &v, hipStream_t s) {
  using C = Q2TileCfg<MM, W>;
  auto kern = k_sweep_tile_q2<MM, W>;
  static bool attr_set = false;
  if (!attr_set) {
    OU_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(C::lds_bytes)));
    attr_set = true;
  }
  const unsigned groups = unsigned(pl.tile_groups);
  for (size_t l = 0; l + 1 < pl.tile_level_off.size(); ++l) {
    const int b = pl.tile_level_off[l], e = pl.tile_level_off[l + 1];
    if (e > b)
      hipLaunchKernelGGL(kern, dim3(e - b, groups), dim3(64 * W), C::lds_bytes, s, v,
                         pl.d_tiles, b);
  }
}

// The single-plan and batched launchers of one instantiation.
struct TileKernels {
  TileLaunch single;
  BatchTileLaunch batch;
};

template <int MM, int W>
const TileKernels *q2_kernels_of() {
  static const TileKernels k{launch_q2_tiles_impl<MM, W>, launch_q2_batch_tiles_impl<MM, W>};
  return &k;
}

// 3D Q2: FP32 on the matrix cores, L <= 1 (1 or 4 moments), 3 or 4 waves
// (the D2M accumulators of 1728 (cell, dof) columns need 192 lanes or more).
template <typename T>
const TileKernels *pick_tile_q2(int mm, int waves) {
  if constexpr (std::is_same<T, float>::value) {
    if (waves == 3) return mm == 1 ? q2_kernels_of<1, 3>() : mm == 4 ? q2_kernels_of<4, 3>() : nullptr;
    if (waves == 4) return mm == 1 ? q2_kernels_of<1, 4>() : mm == 4 ? q2_kernels_of<4, 4>() : nullptr;
  }
  return nullptr;
}

template <typename T, int DIM, int P, int MM, int W>
const TileKernels *tile_kernels_of() {
  static const TileKernels k{launch_tiles_impl<T, DIM, P, MM, W>,
                             launch_batch_tiles_impl<T, DIM, P, MM, W>};
  return &k;
}

template <typename T, int DIM, int P, int W>
const TileKernels *pick_tile_mm(int mm) {
  switch (mm) {
    case 1: return tile_kernels_of<T, DIM, P, 1, W>();
    case 4: return tile_kernels_of<T, DIM, P, 4, W>();
    case 10: return tile_kernels_of<T, DIM, P, 10, W>();
    case 16: return tile_kernels_of<T, DIM, P, 16, W>();
  }
  return nullptr;
}

// Instantiated: 2D Q1-Q2 and 3D Q1, up to 16 moments, at 1, 2 and 4 waves.
template <typename T>
const TileKernels *pick_tile(int dim, int p, int mm, int waves) {
  if (dim == 2 && p == 1) {
    if (waves == 1) return pick_tile_mm<T, 2, 1, 1>(mm);
    if (waves == 2) return pick_tile_mm<T, 2, 1, 2>(mm);
    if (waves == 4) return pick_tile_mm<T, 2, 1, 4>(mm);
  } else if (dim == 2 && p == 2) {
    if (waves == 1) return pick_tile_mm<T, 2, 2, 1>(mm);
    if (waves == 2) return pick_tile_mm<T, 2, 2, 2>(mm);
    if (waves == 4) return pick_tile_mm<T, 2, 2, 4>(mm);
  } else if (dim == 3 && p == 1) {
    if (waves == 1) return pick_tile_mm<T, 3, 1, 1>(mm);
    if (waves == 2) return pick_tile_mm<T, 3, 1, 2>(mm);
    if (waves == 4) return pick_tile_mm<T, 3, 1, 4>(mm);
  } else if (dim == 3 && p == 2) {
    return pick_tile_q2<T>(mm, waves);
  }
  return nullptr;
}

template <typename T>
size_t tile_lds(int dim, int p, int mm, int waves) {
  if (dim == 3 && p == 2)  // Q2TileCfg: psi exchange, staged source, inflow, level tables
    return size_t(4 * waves) * mm * sizeof(double) +
           2 * size_t(4 * waves) * q2::kWidest * q2::kRow * sizeof(float) +
           size_t(64) * q2::kN * mm * sizeof(float) +
           size_t(4 * waves) * 3 * 16 * q2::kF * sizeof(float) +
           size_t(4 * waves) * (mm + 4) * sizeof(float) + 16 * q2::kLevels +
           q2::kWidest * q2::kLevels + 8;
  const int P1 = p + 1, N = dim == 2 ? P1 * P1 : P1 * P1 * P1;
  const int NP = sizeof(T) == 4 ? (N + 3) & ~3 : (N + 1) & ~1;
  const int dpw = dim == 2 ? TileGeom<2>::DPW : TileGeom<3>::DPW;
  return size_t(waves) * dpw * 64 * NP * sizeof(T) + size_t(mm) * 64 * NP * sizeof(T) +
         size_t(kMaxSlotsPerGroup) * mm * sizeof(double);
}

}  // namespace

bool tile_sweep_supported(int dim, int p, int mm, int precision, int waves, int max_lds) {
  if (precision == OU_FP32)
    return pick_tile<float>(dim, p, mm, waves) != nullptr &&
           tile_lds<float>(dim, p, mm, waves) <= size_t(max_lds);
  return pick_tile<double>(dim, p, mm, waves) != nullptr &&
         tile_lds<double>(dim, p, mm, waves) <= size_t(max_lds);
}

int tile_waves_max(int dim, int p, int mm, int precision, int max_lds) {
  for (int w : {4, 3, 2, 1})
    if (tile_sweep_supported(dim, p, mm, precision, w, max_lds)) return w;
  return 0;
}

void launch_sweep_tiles(Plan &pl, const SweepView &v, hipStream_t s, hipEvent_t before,
                        hipEvent_t after) {
  const TileKernels *k = pl.precision == OU_FP32
                             ? pick_tile<float>(pl.dim, pl.p, pl.mm, pl.tile_waves)
                             : pick_tile<double>(pl.dim, pl.p, pl.mm, pl.tile_waves);
  if (!k || !pl.tiles_ok) fail(OU_ERR_UNSUPPORTED, "no tile sweep for this plan");
  if (before) OU_HIP(hipEventRecord(before, s));
  k->single(pl, v, s);
  if (after) OU_HIP(hipEventRecord(after, s));
  OU_HIP(hipGetLastError());
}

void launch_batch_tiles(const BatchLaunch &b, const SweepView *views, hipStream_t s, bool launch) {
  const TileKernels *k = b.precision == OU_FP32
                             ? pick_tile<float>(b.dim, b.p, b.mm, b.tile_waves)
                             : pick_tile<double>(b.dim, b.p, b.mm, b.tile_waves);
  if (!k || !b.tiles_ok) fail(OU_ERR_UNSUPPORTED, "no tile sweep for this batch");
  k->batch(b, views, s, launch);
  if (launch) OU_HIP(hipGetLastError());
}

}  // namespace ou
