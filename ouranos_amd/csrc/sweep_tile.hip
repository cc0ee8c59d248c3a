// The 2D tile sweep (OU_SWEEP_TILES): one launch per tile level instead of
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
// The per-cell arithmetic is the level sweep's own, in the same order (M2D
// rows of the staged source, upwind terms x then y, the (material, class)
// inverse, FP64 D2M accumulation in slot order per direction group), so at
// the same direction groups both modes give the same bits.
// (RadiativeTransfer.cc:878-1012 per cell; the level sweep in sweep.hip.)
#include <hip/hip_runtime.h>

#include <algorithm>

#include "plan.hpp"
#include "sweep_common.hpp"

namespace ou {

namespace {

constexpr int kTE = kTileEdge;
constexpr int kDirsPerWave = 64 / kTE;

template <typename T, int P, int MM, int W>
struct TileCfg {
  static constexpr int P1 = P + 1, N = P1 * P1, F = P1;
  // LDS stride of one cell, padded to 16 bytes for 128-bit LDS accesses
  static constexpr int NP = sizeof(T) == 4 ? (N + 3) & ~3 : (N + 1) & ~1;
  static constexpr int NC = (N + W - 1) / W;  // (cell, dof) columns per lane in D2M
  static constexpr size_t psi_bytes = size_t(W) * kDirsPerWave * 64 * NP * sizeof(T);
  static constexpr size_t src_bytes = size_t(MM) * 64 * NP * sizeof(T);
  static constexpr size_t d2m_bytes = size_t(kMaxSlotsPerGroup) * MM * sizeof(double);
  static constexpr size_t lds_bytes = psi_bytes + src_bytes + d2m_bytes;
};

// x of lane l - 1 within each row of 16 lanes (DPP row_shr:1); a lane at the
// start of a row keeps its own value.
__device__ __forceinline__ float lane_shr1(float x) {
  const int b = __float_as_int(x);
  return __int_as_float(__builtin_amdgcn_update_dpp(b, b, 0x111, 0xF, 0xF, false));
}

__device__ __forceinline__ double lane_shr1(double x) {
  const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
  const int lo = int(unsigned(b)), hi = int(unsigned(b >> 32));
  const unsigned lo2 = unsigned(__builtin_amdgcn_update_dpp(lo, lo, 0x111, 0xF, 0xF, false));
  const unsigned hi2 = unsigned(__builtin_amdgcn_update_dpp(hi, hi, 0x111, 0xF, 0xF, false));
  return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi2) << 32) | lo2));
}

template <typename T, int P, int MM, int W>
__global__ __launch_bounds__(64 * W) void k_sweep_tile(SweepView v,
                                                       const TileItem *__restrict__ tiles,
                                                       int tile_begin) {
  using C = TileCfg<T, P, MM, W>;
  constexpr int N = C::N, F = C::F, NP = C::NP, NC = C::NC;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  T *s_psi = reinterpret_cast<T *>(lds_raw);                      // [W*8 slots][64 cells][NP]
  T *s_src = reinterpret_cast<T *>(lds_raw + C::psi_bytes);       // [MM][64 cells][NP]
  double *t_d2m = reinterpret_cast<double *>(lds_raw + C::psi_bytes + C::src_bytes);  // [slot][MM]

  const TileItem tile = tiles[tile_begin + blockIdx.x];
  const int group = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int o = tile.octant;
  const int nx = v.n[0], ny = v.n[1];
  const int n_groups = v.dir_groups, dall = v.oct_off[o + 1] - v.oct_off[o];
  const int glen = (dall + n_groups - 1) / n_groups;
  const int dbeg = v.oct_off[o] + group * glen;
  const int ndir_o = max(0, min(glen, dall - group * glen));
  const int ii0 = tile.ti * kTE, jj0 = tile.tj * kTE;
  const long n_dofs = v.n_dofs;
  const int n_mom = v.n_mom;

  // Tile cell cl = kTE b + a: canonical (ii0 + a, jj0 + b), -1 outside the mesh.
  auto cell_of = [&](int cl) {
    const int ii = ii0 + cl % kTE, jj = jj0 + cl / kTE;
    if (ii >= nx || jj >= ny) return -1;
    const int i = (o & 1) ? nx - 1 - ii : ii;
    const int j = (o & 2) ? ny - 1 - jj : jj;
    return i + nx * j;
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
              s_src[m * 64 * NP + cl * NP + natural_index<2, P>(q * Q + j, o)] = val[k][j];
        }
      }
    }
  }
  lds_barrier();

  const T *m1 = m1_of<T>(v);
  const int a = lane % kTE, dsub = lane / kTE;
  const bool col_in = ii0 + a < nx;
  T *const tr0 = static_cast<T *>(v.trace[0]);
  T *const tr1 = static_cast<T *>(v.trace[1]);
  const long ps0 = v.plane_size[0], ps1 = v.plane_size[1];

  double acc[MM][NC];
#pragma unroll
  for (int m = 0; m < MM; ++m)
#pragma unroll
    for (int t = 0; t < NC; ++t) acc[m][t] = 0.0;

  for (int r0 = 0; r0 < ndir_o; r0 += W * kDirsPerWave) {
    const int sl = r0 + w * kDirsPerWave + dsub;  // slot within the group
    const bool has_dir = sl < ndir_o;
    const int gsl = dbeg + (has_dir ? sl : 0);
    const int d = v.oct_dirs[gsl];
    const int cls = v.slot_cls[gsl];
    T m2d[MM];
#pragma unroll
    for (int m = 0; m < MM; ++m)
      m2d[m] = has_dir ? static_cast<const T *>(v.slot_m2d)[long(gsl) * MM + m] : T(0);
    const T c0 = static_cast<const T *>(v.slot_c)[long(gsl) * 4 + 0];
    const T c1 = static_cast<const T *>(v.slot_c)[long(gsl) * 4 + 1];
    // the (tile material, class) inverse, canonical frame, in registers
    const T *Ai = static_cast<const T *>(v.ainv) + (long(tile.material) * v.n_cls + cls) * (N * N);
    T A[N * N];
#pragma unroll
    for (int k = 0; k < N * N; ++k) A[k] = Ai[k];
    // Trace slabs of direction d: x slot = canonical row jj, y slot = column ii.
    T *const dx = tr0 + long(d) * F * ps0;
    T *const dy = tr1 + long(d) * F * ps1;
    // upwind edge of the tile: rows jj0 + b (lanes a = 0), column ii0 + a
    T xin[kTE][F], yin[F];
#pragma unroll
    for (int b = 0; b < kTE; ++b)
#pragma unroll
      for (int f = 0; f < F; ++f)
        xin[b][f] = (a == 0 && jj0 + b < ny) ? dx[long(jj0 + b) * F + f] : T(0);
#pragma unroll
    for (int f = 0; f < F; ++f) yin[f] = col_in ? dy[long(ii0 + a) * F + f] : T(0);

    T tx[F], ty[F];
#pragma unroll
    for (int f = 0; f < F; ++f) tx[f] = ty[f] = T(0);
#pragma unroll
    for (int t = 0; t < 2 * kTE - 1; ++t) {
      const int b = t - a;
      const int bc = min(max(b, 0), kTE - 1);
      const int cl = bc * kTE + a;
      const bool active = has_dir && b >= 0 && b < kTE && col_in && jj0 + b < ny;
      T xi[F], yi[F];
#pragma unroll
      for (int f = 0; f < F; ++f) {
        xi[f] = a == 0 ? (t < kTE ? xin[t < kTE ? t : 0][f] : T(0)) : tx[f];
        yi[f] = b == 0 ? yin[f] : ty[f];
      }
      T bv[N];
#pragma unroll
      for (int i = 0; i < N; ++i) bv[i] = T(0);
#pragma unroll
      for (int m = 0; m < MM; ++m) {
        const T *src = s_src + m * 64 * NP + cl * NP;
#pragma unroll
        for (int i = 0; i < N; ++i) bv[i] += m2d[m] * src[i];
      }
      add_upwind<T, 2, P, 0>(m1, c0, xi, bv);
      add_upwind<T, 2, P, 1>(m1, c1, yi, bv);
      T psi[N];
#pragma unroll
      for (int r = 0; r < N; ++r) {
        T s = T(0);
#pragma unroll
        for (int q = 0; q < N; ++q) s += A[r * N + q] * bv[q];
        psi[r] = s;
      }
      if (active) {
        T *dst = s_psi + ((w * kDirsPerWave + dsub) * 64 + cl) * NP;
#pragma unroll
        for (int i = 0; i < N; ++i) dst[i] = psi[i];
        // downwind edges of the tile (or of the mesh) go to the trace slabs
        if (a == kTE - 1 || ii0 + a == nx - 1) store_trace<T, 2, P, 0>(dx, unsigned(jj0 + b), psi);
        if (b == kTE - 1 || jj0 + b == ny - 1) store_trace<T, 2, P, 1>(dy, unsigned(ii0 + a), psi);
      }
      // next level: x-outflow to the lane of column a + 1, y-outflow kept
#pragma unroll
      for (int f = 0; f < F; ++f) {
        tx[f] = lane_shr1(psi[face_dof<2, P, 0, P>(f)]);
        ty[f] = psi[face_dof<2, P, 1, P>(f)];
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
    const int n_round = min(W * kDirsPerWave, ndir_o - r0);
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
                      natural_index<2, P>(ic, o);
#pragma unroll
        for (int m = 0; m < MM; ++m)
          if (m < n_mom) out[long(m) * n_dofs] = acc[m][t];
      }
    }
  }
}

using TileLaunch = void (*)(const Plan &, const SweepView &, hipStream_t);

template <typename T, int P, int MM, int W>
void launch_tiles_impl(const Plan &pl, const SweepView &v, hipStream_t s) {
  using C = TileCfg<T, P, MM, W>;
  auto kern = k_sweep_tile<T, P, MM, W>;
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

template <typename T, int P, int W>
TileLaunch pick_tile_mm(int mm) {
  switch (mm) {
    case 1: return launch_tiles_impl<T, P, 1, W>;
    case 4: return launch_tiles_impl<T, P, 4, W>;
    case 10: return launch_tiles_impl<T, P, 10, W>;
    case 16: return launch_tiles_impl<T, P, 16, W>;
  }
  return nullptr;
}

template <typename T>
TileLaunch pick_tile(int p, int mm, int waves) {
  if (p == 1) {
    if (waves == 1) return pick_tile_mm<T, 1, 1>(mm);
    if (waves == 2) return pick_tile_mm<T, 1, 2>(mm);
    if (waves == 4) return pick_tile_mm<T, 1, 4>(mm);
  } else if (p == 2) {
    if (waves == 1) return pick_tile_mm<T, 2, 1>(mm);
    if (waves == 2) return pick_tile_mm<T, 2, 2>(mm);
    if (waves == 4) return pick_tile_mm<T, 2, 4>(mm);
  }
  return nullptr;
}

template <typename T>
size_t tile_lds(int p, int mm, int waves) {
  const int P1 = p + 1, N = P1 * P1;
  const int NP = sizeof(T) == 4 ? (N + 3) & ~3 : (N + 1) & ~1;
  return size_t(waves) * kDirsPerWave * 64 * NP * sizeof(T) + size_t(mm) * 64 * NP * sizeof(T) +
         size_t(kMaxSlotsPerGroup) * mm * sizeof(double);
}

}  // namespace

bool tile_sweep_supported(int p, int mm, int precision, int waves, int max_lds) {
  if (precision == OU_FP32)
    return pick_tile<float>(p, mm, waves) != nullptr && tile_lds<float>(p, mm, waves) <= size_t(max_lds);
  return pick_tile<double>(p, mm, waves) != nullptr && tile_lds<double>(p, mm, waves) <= size_t(max_lds);
}

int tile_waves_max(int p, int mm, int precision, int max_lds) {
  for (int w : {4, 2, 1})
    if (tile_sweep_supported(p, mm, precision, w, max_lds)) return w;
  return 0;
}

void launch_sweep_tiles(Plan &pl, const SweepView &v, hipStream_t s, hipEvent_t before,
                        hipEvent_t after) {
  const TileLaunch k = pl.precision == OU_FP32 ? pick_tile<float>(pl.p, pl.mm, pl.tile_waves)
                                               : pick_tile<double>(pl.p, pl.mm, pl.tile_waves);
  if (!k || !pl.tiles_ok) fail(OU_ERR_UNSUPPORTED, "no tile sweep for this plan");
  if (before) OU_HIP(hipEventRecord(before, s));
  k(pl, v, s);
  if (after) OU_HIP(hipEventRecord(after, s));
  OU_HIP(hipGetLastError());
}

}  // namespace ou
