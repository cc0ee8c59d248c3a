// FP64 vector kernels for the Krylov drivers (and Gram-Schmidt over an FP32
// basis); see vec.hpp.
//
// All kernels stream HBM: each thread walks a grid-strided range in 16-byte
// (double2) steps when every operand is 16-byte aligned and the leading
// dimension is even, else in 8-byte steps.  Dot products keep one FP64
// accumulator per vector in registers, reduce across the wave with shuffles,
// then across the block's four waves in LDS, and leave one partial per
// block; a second kernel sums the partials in a fixed order, so results are
// deterministic for a given n.
#include "ou_common.hpp"
#include "vec.hpp"

namespace ou {

namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kMaxBlocks = 2048;
constexpr int kMaxDots = 32;

inline int blocks_for(long n) {
  long b = (n + kBlock - 1) / kBlock;
  return int(b < kMaxBlocks ? (b < 1 ? 1 : b) : kMaxBlocks);
}

inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }
inline bool aligned8(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 7u) == 0; }

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Partial sums of m dot products <V_i, w> per block.
template <int M, int VEC>
__global__ __launch_bounds__(kBlock) void k_multi_dot(long n, int m, const double *V,
                                                      long ldv, const double *w,
                                                      double *partial) {
  __shared__ double red[kWaves][M];
  double acc[M];
#pragma unroll
  for (int i = 0; i < M; ++i) acc[i] = 0.0;
  const long stride = long(gridDim.x) * kBlock;
  if (VEC == 2) {
    const long n2 = n / 2;
    const double2 *w2 = reinterpret_cast<const double2 *>(w);
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n2; k += stride) {
      const double2 wk = w2[k];
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < m) {
          const double2 v = reinterpret_cast<const double2 *>(V + i * ldv)[k];
          acc[i] += v.x * wk.x + v.y * wk.y;
        }
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0)
      for (int i = 0; i < M; ++i)
        if (i < m) acc[i] += V[i * ldv + n - 1] * w[n - 1];
  } else {
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += stride) {
      const double wk = w[k];
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < m) acc[i] += V[i * ldv + k] * wk;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < M; ++i)
    if (i < m) {
      const double s = wave_sum(acc[i]);
      if (lane == 0) red[wave][i] = s;
    }
  __syncthreads();
  if (int(threadIdx.x) < m) {
    double s = 0.0;
    for (int q = 0; q < kWaves; ++q) s += red[q][threadIdx.x];
    partial[long(blockIdx.x) * m + threadIdx.x] = s;
  }
}

__global__ __launch_bounds__(kBlock) void k_reduce_partials(int nb, int m,
                                                            const double *partial,
                                                            double *out) {
  __shared__ double red[kWaves];
  for (int i = 0; i < m; ++i) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) s += partial[long(b) * m + i];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < kWaves; ++q) t += red[q];
      out[i] = t;
    }
    __syncthreads();
  }
}

// w -= alpha sum_i h[i] V_i
template <int VEC>
__global__ __launch_bounds__(kBlock) void k_multi_axpy(long n, int m, const double *V,
                                                       long ldv, const double *h,
                                                       double alpha, double *w) {
  __shared__ double hs[64];
  if (int(threadIdx.x) < m) hs[threadIdx.x] = alpha * h[threadIdx.x];
  __syncthreads();
  const long stride = long(gridDim.x) * kBlock;
  if (VEC == 2) {
    const long n2 = n / 2;
    double2 *w2 = reinterpret_cast<double2 *>(w);
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n2; k += stride) {
      double sx = 0.0, sy = 0.0;
      for (int i = 0; i < m; ++i) {
        const double2 v = reinterpret_cast<const double2 *>(V + i * ldv)[k];
        sx += hs[i] * v.x;
        sy += hs[i] * v.y;
      }
      double2 wk = w2[k];
      wk.x -= sx;
      wk.y -= sy;
      w2[k] = wk;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
      double s = 0.0;
      for (int i = 0; i < m; ++i) s += hs[i] * V[i * ldv + n - 1];
      w[n - 1] -= s;
    }
  } else {
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += stride) {
      double s = 0.0;
      for (int i = 0; i < m; ++i) s += hs[i] * V[i * ldv + k];
      w[k] -= s;
    }
  }
}

// w -= sum_{i<m} h[i] V_i, then, in the same pass, the partial sums of
// <V_i, w> for i < ndots (ndots is 0 or m) and of <w, w> (slot ndots) per
// block: the projection of one Gram-Schmidt pass fused with the dot products
// of the next and the norm of the result.  A step's m <= 16 vector values
// stay in registers between the update and the dot products.
template <int M, int VEC>
__global__ __launch_bounds__(kBlock) void k_axpy_dots(long n, int m, const double *V, long ldv,
                                                      const double *h, double *w, int ndots,
                                                      double *partial) {
  __shared__ double hs[M];
  __shared__ double red[kWaves][M + 1];
  if (int(threadIdx.x) < m) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  double acc[M + 1];
#pragma unroll
  for (int i = 0; i <= M; ++i) acc[i] = 0.0;
  const long stride = long(gridDim.x) * kBlock;
  if (VEC == 2) {
    double2 *w2 = reinterpret_cast<double2 *>(w);
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n / 2; k += stride) {
      double2 vv[M];
      double sx = 0.0, sy = 0.0;
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < m) {
          vv[i] = reinterpret_cast<const double2 *>(V + i * ldv)[k];
          sx += hs[i] * vv[i].x;
          sy += hs[i] * vv[i].y;
        }
      double2 wk = w2[k];
      wk.x -= sx;
      wk.y -= sy;
      w2[k] = wk;
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < ndots) acc[i] += vv[i].x * wk.x + vv[i].y * wk.y;
      acc[M] += wk.x * wk.x + wk.y * wk.y;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
      double sum = 0.0;
      for (int i = 0; i < m; ++i) sum += hs[i] * V[i * ldv + n - 1];
      const double wk = w[n - 1] - sum;
      w[n - 1] = wk;
      for (int i = 0; i < M; ++i)
        if (i < ndots) acc[i] += V[i * ldv + n - 1] * wk;
      acc[M] += wk * wk;
    }
  } else {
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += stride) {
      double sum = 0.0;
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < m) sum += hs[i] * V[i * ldv + k];
      const double wk = w[k] - sum;
      w[k] = wk;
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < ndots) acc[i] += V[i * ldv + k] * wk;
      acc[M] += wk * wk;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i <= M; ++i)
    if (i < ndots || i == M) {
      const double t = wave_sum(acc[i]);
      if (lane == 0) red[wave][i] = t;
    }
  __syncthreads();
  if (int(threadIdx.x) <= ndots) {
    const int i = int(threadIdx.x) < ndots ? int(threadIdx.x) : M;
    double t = 0.0;
    for (int q = 0; q < kWaves; ++q) t += red[q][i];
    partial[long(blockIdx.x) * (ndots + 1) + threadIdx.x] = t;
  }
}

// Gram-Schmidt pass over an FP32 basis (the inner solve of the FP32
// operator stores its Krylov vectors in FP32: half the bytes of every pass).
// Basis values widen to FP64 in registers; w, h and every sum are FP64.
// When `update`, w -= sum_{i<m} h[i] V_i first; then the partial sums of
// <V_i, w> for i < ndots (0 or m) and of <w, w> (slot ndots) per block.
__device__ __forceinline__ double2 ld_pair(const float *p, long k) {
  const float2 v = reinterpret_cast<const float2 *>(p)[k];
  return double2{double(v.x), double(v.y)};
}

template <int M, int VEC>
__global__ __launch_bounds__(kBlock) void k_gs_f32(long n, int m, const float *V, long ldv,
                                                   const double *h, double *w, int update,
                                                   int ndots, double *partial) {
  __shared__ double hs[M];
  __shared__ double red[kWaves][M + 1];
  if (int(threadIdx.x) < m) hs[threadIdx.x] = update ? h[threadIdx.x] : 0.0;
  __syncthreads();
  double acc[M + 1];
#pragma unroll
  for (int i = 0; i <= M; ++i) acc[i] = 0.0;
  const long stride = long(gridDim.x) * kBlock;
  if (VEC == 4) {
    // four values per lane: one 16-byte load per basis vector, two for w
    double2 *w2 = reinterpret_cast<double2 *>(w);
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n / 4; k += stride) {
      float4 vv[M];
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < m) {
          vv[i] = reinterpret_cast<const float4 *>(V + i * ldv)[k];
          s0 += hs[i] * double(vv[i].x);
          s1 += hs[i] * double(vv[i].y);
          s2 += hs[i] * double(vv[i].z);
          s3 += hs[i] * double(vv[i].w);
        }
      double2 wa = w2[2 * k], wb = w2[2 * k + 1];
      if (update) {
        wa.x -= s0;
        wa.y -= s1;
        wb.x -= s2;
        wb.y -= s3;
        w2[2 * k] = wa;
        w2[2 * k + 1] = wb;
      }
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < ndots)
          acc[i] += double(vv[i].x) * wa.x + double(vv[i].y) * wa.y + double(vv[i].z) * wb.x +
                    double(vv[i].w) * wb.y;
      acc[M] += wa.x * wa.x + wa.y * wa.y + wb.x * wb.x + wb.y * wb.y;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (long k = n & ~3L; k < n; ++k) {
        double sum = 0.0;
        for (int i = 0; i < m; ++i) sum += hs[i] * double(V[i * ldv + k]);
        const double wk = w[k] - sum;
        if (update) w[k] = wk;
        for (int i = 0; i < M; ++i)
          if (i < ndots) acc[i] += double(V[i * ldv + k]) * wk;
        acc[M] += wk * wk;
      }
  } else if (VEC == 2) {
    double2 *w2 = reinterpret_cast<double2 *>(w);
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n / 2; k += stride) {
      double2 vv[M];
      double sx = 0.0, sy = 0.0;
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < m) {
          vv[i] = ld_pair(V + i * ldv, k);
          sx += hs[i] * vv[i].x;
          sy += hs[i] * vv[i].y;
        }
      double2 wk = w2[k];
      if (update) {
        wk.x -= sx;
        wk.y -= sy;
        w2[k] = wk;
      }
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < ndots) acc[i] += vv[i].x * wk.x + vv[i].y * wk.y;
      acc[M] += wk.x * wk.x + wk.y * wk.y;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
      double sum = 0.0;
      for (int i = 0; i < m; ++i) sum += hs[i] * double(V[i * ldv + n - 1]);
      const double wk = w[n - 1] - sum;
      if (update) w[n - 1] = wk;
      for (int i = 0; i < M; ++i)
        if (i < ndots) acc[i] += double(V[i * ldv + n - 1]) * wk;
      acc[M] += wk * wk;
    }
  } else {
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += stride) {
      double sum = 0.0;
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < m) sum += hs[i] * double(V[i * ldv + k]);
      const double wk = w[k] - sum;
      if (update) w[k] = wk;
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (i < ndots) acc[i] += double(V[i * ldv + k]) * wk;
      acc[M] += wk * wk;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i <= M; ++i)
    if (i < ndots || i == M) {
      const double t = wave_sum(acc[i]);
      if (lane == 0) red[wave][i] = t;
    }
  __syncthreads();
  if (int(threadIdx.x) <= ndots) {
    const int i = int(threadIdx.x) < ndots ? int(threadIdx.x) : M;
    double t = 0.0;
    for (int q = 0; q < kWaves; ++q) t += red[q][i];
    partial[long(blockIdx.x) * (ndots + 1) + threadIdx.x] = t;
  }
}

// x += alpha sum_i h[i] V_i over an FP32 basis (h on the device).
template <int VEC>
__global__ __launch_bounds__(kBlock) void k_multi_axpy_f32(long n, int m, const float *V,
                                                           long ldv, const double *h,
                                                           double alpha, double *x) {
  __shared__ double hs[64];
  if (int(threadIdx.x) < m) hs[threadIdx.x] = alpha * h[threadIdx.x];
  __syncthreads();
  const long stride = long(gridDim.x) * kBlock;
  if (VEC == 4) {
    double2 *x2 = reinterpret_cast<double2 *>(x);
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n / 4; k += stride) {
      double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
      for (int i = 0; i < m; ++i) {
        const float4 v = reinterpret_cast<const float4 *>(V + i * ldv)[k];
        s0 += hs[i] * double(v.x);
        s1 += hs[i] * double(v.y);
        s2 += hs[i] * double(v.z);
        s3 += hs[i] * double(v.w);
      }
      double2 xa = x2[2 * k], xb = x2[2 * k + 1];
      xa.x += s0;
      xa.y += s1;
      xb.x += s2;
      xb.y += s3;
      x2[2 * k] = xa;
      x2[2 * k + 1] = xb;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
      for (long k = n & ~3L; k < n; ++k) {
        double t = 0.0;
        for (int i = 0; i < m; ++i) t += hs[i] * double(V[i * ldv + k]);
        x[k] += t;
      }
  } else if (VEC == 2) {
    double2 *x2 = reinterpret_cast<double2 *>(x);
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n / 2; k += stride) {
      double sx = 0.0, sy = 0.0;
      for (int i = 0; i < m; ++i) {
        const double2 v = ld_pair(V + i * ldv, k);
        sx += hs[i] * v.x;
        sy += hs[i] * v.y;
      }
      double2 xk = x2[k];
      xk.x += sx;
      xk.y += sy;
      x2[k] = xk;
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
      double t = 0.0;
      for (int i = 0; i < m; ++i) t += hs[i] * double(V[i * ldv + n - 1]);
      x[n - 1] += t;
    }
  } else {
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += stride) {
      double t = 0.0;
      for (int i = 0; i < m; ++i) t += hs[i] * double(V[i * ldv + k]);
      x[k] += t;
    }
  }
}

// y = (float) (a x); x itself becomes a x when `scale`.
template <int VEC>
__global__ __launch_bounds__(kBlock) void k_scale_store_f32(long n, double a, double *x,
                                                            int scale, float *y) {
  const long stride = long(gridDim.x) * kBlock;
  if (VEC == 2) {
    double2 *x2 = reinterpret_cast<double2 *>(x);
    float2 *y2 = reinterpret_cast<float2 *>(y);
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n / 2; k += stride) {
      double2 v = x2[k];
      v.x *= a;
      v.y *= a;
      if (scale) x2[k] = v;
      y2[k] = float2{float(v.x), float(v.y)};
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0) {
      const double v = a * x[n - 1];
      if (scale) x[n - 1] = v;
      y[n - 1] = float(v);
    }
  } else {
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += stride) {
      const double v = a * x[k];
      if (scale) x[k] = v;
      y[k] = float(v);
    }
  }
}

// Element-wise kernels: z = a x + b y (x, y may be null when unused), in
// double2 steps when aligned (VEC 2), with a scalar tail for odd n.
// kind: 0 axpby (y = a x + b y), 1 lincomb (z = a x + b y), 2 scale (x *= a),
// 3 fill (x = a), 4 copy (y = x).
__device__ __forceinline__ double elem(int kind, double a, double b, double xv, double yv) {
  switch (kind) {
    case 0:
    case 1: return a * xv + b * yv;
    case 2: return a * xv;
    case 3: return a;
    default: return xv;
  }
}

template <int KIND, int VEC>
__global__ __launch_bounds__(kBlock) void k_elementwise(long n, double a, const double *x,
                                                        double b, const double *y, double *z) {
  const long stride = long(gridDim.x) * kBlock;
  const bool rx = KIND != 3, ry = KIND <= 1;
  if (VEC == 2) {
    const double2 *x2 = reinterpret_cast<const double2 *>(x);
    const double2 *y2 = reinterpret_cast<const double2 *>(y);
    double2 *z2 = reinterpret_cast<double2 *>(z);
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n / 2; k += stride) {
      const double2 xv = rx ? x2[k] : double2{0.0, 0.0};
      const double2 yv = ry ? y2[k] : double2{0.0, 0.0};
      z2[k] = double2{elem(KIND, a, b, xv.x, yv.x), elem(KIND, a, b, xv.y, yv.y)};
    }
    if ((n & 1) && blockIdx.x == 0 && threadIdx.x == 0)
      z[n - 1] = elem(KIND, a, b, rx ? x[n - 1] : 0.0, ry ? y[n - 1] : 0.0);
  } else {
    for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += stride)
      z[k] = elem(KIND, a, b, rx ? x[k] : 0.0, ry ? y[k] : 0.0);
  }
}

__global__ __launch_bounds__(kBlock) void k_zero_words(long n, unsigned *p) {
  for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += long(gridDim.x) * kBlock)
    p[k] = 0u;
}

__global__ __launch_bounds__(kBlock) void k_check_finite(long n, const double *x, int *flag) {
  int bad = 0;
  for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += long(gridDim.x) * kBlock)
    bad |= !isfinite(x[k]);
  if (bad) atomicOr(flag, 1);
}

template <int M>
void launch_dot(bool v2, int nb, hipStream_t s, long n, int m, const double *V, long ldv,
                const double *w, double *scratch) {
  if (v2)
    hipLaunchKernelGGL((k_multi_dot<M, 2>), dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, w,
                       scratch);
  else
    hipLaunchKernelGGL((k_multi_dot<M, 1>), dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, w,
                       scratch);
}

template <int M>
void launch_axpy_dots(bool v2, int nb, hipStream_t s, long n, int m, const double *V, long ldv,
                      const double *h, double *w, int ndots, double *scratch) {
  if (v2)
    hipLaunchKernelGGL((k_axpy_dots<M, 2>), dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, h, w,
                       ndots, scratch);
  else
    hipLaunchKernelGGL((k_axpy_dots<M, 1>), dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, h, w,
                       ndots, scratch);
}

// Values per lane of the FP32-basis kernels: 4 (float4 basis loads) when
// every basis vector is 16-byte aligned, 2 (float2) when 8-byte aligned, else
// 1; the FP64 vector is read as double2 in the first two forms.
int f32_width(const float *V, long ldv, int m, const double *w) {
  if (!aligned16(w)) return 1;
  if (aligned16(V) && (m == 1 || ldv % 4 == 0)) return 4;
  if (aligned8(V) && (m == 1 || ldv % 2 == 0)) return 2;
  return 1;
}

template <int M>
void launch_gs_f32(int vec, int nb, hipStream_t s, long n, int m, const float *V, long ldv,
                   const double *h, double *w, int update, int ndots, double *scratch) {
  if (vec == 4)
    hipLaunchKernelGGL((k_gs_f32<M, 4>), dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, h, w,
                       update, ndots, scratch);
  else if (vec == 2)
    hipLaunchKernelGGL((k_gs_f32<M, 2>), dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, h, w,
                       update, ndots, scratch);
  else
    hipLaunchKernelGGL((k_gs_f32<M, 1>), dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, h, w,
                       update, ndots, scratch);
}

}  // namespace

long reduce_scratch_doubles(int m) { return long(kMaxBlocks) * (m < 1 ? 1 : m); }

void vec_multi_dot(long n, int m, const double *V, long ldv, const double *w,
                   double *scratch, double *out, hipStream_t s) {
  if (m < 1) return;
  if (m > kMaxDots) fail(OU_ERR_ARG, "too many simultaneous dot products");
  const bool v2 = aligned16(V) && aligned16(w) && (m == 1 || ldv % 2 == 0);
  const int nb = blocks_for(v2 ? (n + 1) / 2 : n);
  if (m == 1)
    launch_dot<1>(v2, nb, s, n, m, V, ldv, w, scratch);
  else if (m <= 8)
    launch_dot<8>(v2, nb, s, n, m, V, ldv, w, scratch);
  else if (m <= 16)
    launch_dot<16>(v2, nb, s, n, m, V, ldv, w, scratch);
  else
    launch_dot<kMaxDots>(v2, nb, s, n, m, V, ldv, w, scratch);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s, nb, m, scratch, out);
  OU_HIP(hipGetLastError());
}

void vec_axpy_dots(long n, int m, const double *V, long ldv, const double *h, double *w,
                   bool dots, double *scratch, double *out, hipStream_t s) {
  if (m < 1 || m > kMaxDots) fail(OU_ERR_ARG, "axpy_dots takes 1 to 32 vectors");
  const int ndots = dots ? m : 0;
  if (m > 16) {
    // Beyond 16 vectors the fused kernel holds too many registers: update,
    // then one dot-product pass (which also takes the norm when w follows
    // the m vectors in memory, as the next Arnoldi vector does).
    vec_multi_axpy(n, m, V, ldv, h, 1.0, w, s);
    if (dots && w == V + long(m) * ldv && m + 1 <= kMaxDots) {
      vec_multi_dot(n, m + 1, V, ldv, w, scratch, out, s);
    } else {
      if (dots) vec_multi_dot(n, m, V, ldv, w, scratch, out, s);
      vec_multi_dot(n, 1, w, 0, w, scratch, out + ndots, s);
    }
    return;
  }
  const bool v2 = aligned16(V) && aligned16(w) && (m == 1 || ldv % 2 == 0);
  const int nb = blocks_for(v2 ? (n + 1) / 2 : n);
  if (m <= 8)
    launch_axpy_dots<8>(v2, nb, s, n, m, V, ldv, h, w, ndots, scratch);
  else
    launch_axpy_dots<16>(v2, nb, s, n, m, V, ldv, h, w, ndots, scratch);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s, nb, ndots + 1, scratch,
                     out);
  OU_HIP(hipGetLastError());
}

void vec_multi_axpy(long n, int m, const double *V, long ldv, const double *h, double alpha,
                    double *w, hipStream_t s) {
  if (m < 1) return;
  if (m > 64) fail(OU_ERR_ARG, "multi_axpy supports at most 64 vectors");
  const bool v2 = aligned16(V) && aligned16(w) && (m == 1 || ldv % 2 == 0);
  const int nb = blocks_for(v2 ? (n + 1) / 2 : n);
  if (v2)
    hipLaunchKernelGGL(k_multi_axpy<2>, dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, h, alpha, w);
  else
    hipLaunchKernelGGL(k_multi_axpy<1>, dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, h, alpha, w);
  OU_HIP(hipGetLastError());
}

void vec_gs_f32(long n, int m, const float *V, long ldv, const double *h, double *w,
                bool update, bool dots, double *scratch, double *out, hipStream_t s) {
  if (m < 1 || m > kMaxDots) fail(OU_ERR_ARG, "gs_f32 takes 1 to 32 vectors");
  const int ndots = dots ? m : 0;
  const int vec = f32_width(V, ldv, m, w);
  const int nb = blocks_for((n + vec - 1) / vec);
  const int up = update ? 1 : 0;
  if (m <= 8)
    launch_gs_f32<8>(vec, nb, s, n, m, V, ldv, h, w, up, ndots, scratch);
  else if (m <= 16)
    launch_gs_f32<16>(vec, nb, s, n, m, V, ldv, h, w, up, ndots, scratch);
  else
    launch_gs_f32<kMaxDots>(vec, nb, s, n, m, V, ldv, h, w, up, ndots, scratch);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s, nb, ndots + 1, scratch,
                     out);
  OU_HIP(hipGetLastError());
}

void vec_multi_axpy_f32(long n, int m, const float *V, long ldv, const double *h, double alpha,
                        double *x, hipStream_t s) {
  if (m < 1) return;
  if (m > 64) fail(OU_ERR_ARG, "multi_axpy supports at most 64 vectors");
  const int vec = f32_width(V, ldv, m, x);
  const int nb = blocks_for((n + vec - 1) / vec);
  if (vec == 4)
    hipLaunchKernelGGL(k_multi_axpy_f32<4>, dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, h, alpha,
                       x);
  else if (vec == 2)
    hipLaunchKernelGGL(k_multi_axpy_f32<2>, dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, h, alpha,
                       x);
  else
    hipLaunchKernelGGL(k_multi_axpy_f32<1>, dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, h, alpha,
                       x);
  OU_HIP(hipGetLastError());
}

void vec_scale_store_f32(long n, double a, double *x, bool scale, float *y, hipStream_t s) {
  const bool v2 = aligned16(x) && aligned8(y);
  const int nb = blocks_for(v2 ? (n + 1) / 2 : n);
  if (v2)
    hipLaunchKernelGGL(k_scale_store_f32<2>, dim3(nb), dim3(kBlock), 0, s, n, a, x,
                       scale ? 1 : 0, y);
  else
    hipLaunchKernelGGL(k_scale_store_f32<1>, dim3(nb), dim3(kBlock), 0, s, n, a, x,
                       scale ? 1 : 0, y);
  OU_HIP(hipGetLastError());
}

namespace {

template <int KIND>
void elementwise(long n, double a, const double *x, double b, const double *y, double *z,
                 hipStream_t s) {
  const bool v2 = (!x || aligned16(x)) && (!y || aligned16(y)) && aligned16(z);
  const int nb = blocks_for(v2 ? (n + 1) / 2 : n);
  if (v2)
    hipLaunchKernelGGL((k_elementwise<KIND, 2>), dim3(nb), dim3(kBlock), 0, s, n, a, x, b, y, z);
  else
    hipLaunchKernelGGL((k_elementwise<KIND, 1>), dim3(nb), dim3(kBlock), 0, s, n, a, x, b, y, z);
  OU_HIP(hipGetLastError());
}

}  // namespace

void vec_axpby(long n, double a, const double *x, double b, double *y, hipStream_t s) {
  elementwise<0>(n, a, x, b, y, y, s);
}

void vec_lincomb(long n, double a, const double *x, double b, const double *y, double *z,
                 hipStream_t s) {
  elementwise<1>(n, a, x, b, y, z, s);
}

void vec_scale(long n, double a, double *x, hipStream_t s) {
  elementwise<2>(n, a, x, 0.0, nullptr, x, s);
}

void vec_copy(long n, const double *x, double *y, hipStream_t s) {
  elementwise<4>(n, 0.0, x, 0.0, nullptr, y, s);
}

void vec_fill(long n, double v, double *x, hipStream_t s) {
  elementwise<3>(n, v, nullptr, 0.0, nullptr, x, s);
}

void vec_zero_bytes(void *p, size_t bytes, hipStream_t s) {
  const long n = long(bytes / 4);
  hipLaunchKernelGGL(k_zero_words, dim3(blocks_for(n)), dim3(kBlock), 0, s, n,
                     static_cast<unsigned *>(p));
  OU_HIP(hipGetLastError());
}

void vec_check_finite(long n, const double *x, int *flag, hipStream_t s) {
  hipLaunchKernelGGL(k_check_finite, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, x, flag);
  OU_HIP(hipGetLastError());
}

}  // namespace ou
