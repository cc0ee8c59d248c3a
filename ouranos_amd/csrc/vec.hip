// FP64 vector kernels for the Krylov drivers; see vec.hpp.
#include "ou_common.hpp"
#include "vec.hpp"

namespace ou {

namespace {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 1024;
constexpr int kMaxDots = 32;

inline int blocks_for(long n) {
  long b = (n + kBlock - 1) / kBlock;
  return int(b < kMaxBlocks ? (b < 1 ? 1 : b) : kMaxBlocks);
}

// Partial sums of m dot products per block; each thread walks a
// grid-strided range, then the block reduces in LDS.
template <int M>
__global__ __launch_bounds__(kBlock) void k_multi_dot(long n, int m, const double *V,
                                                      long ldv, const double *w,
                                                      double *partial) {
  __shared__ double red[kBlock];
  double acc[M];
#pragma unroll
  for (int i = 0; i < M; ++i) acc[i] = 0.0;
  for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += long(gridDim.x) * kBlock) {
    const double wk = w[k];
#pragma unroll
    for (int i = 0; i < M; ++i)
      if (i < m) acc[i] += V[i * ldv + k] * wk;
  }
  for (int i = 0; i < m; ++i) {
    red[threadIdx.x] = acc[i < M ? i : 0];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[long(blockIdx.x) * m + i] = red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_reduce_partials(int nb, int m,
                                                            const double *partial,
                                                            double *out) {
  __shared__ double red[kBlock];
  for (int i = 0; i < m; ++i) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += kBlock) s += partial[long(b) * m + i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[i] = red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kBlock) void k_multi_axpy(long n, int m, const double *V,
                                                       long ldv, const double *h,
                                                       double alpha, double *w) {
  __shared__ double hs[64];
  if (threadIdx.x < m) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += long(gridDim.x) * kBlock) {
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += hs[i] * V[i * ldv + k];
    w[k] -= alpha * s;
  }
}

__global__ __launch_bounds__(kBlock) void k_axpby(long n, double a, const double *x,
                                                  double b, double *y) {
  for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += long(gridDim.x) * kBlock)
    y[k] = a * x[k] + b * y[k];
}

__global__ __launch_bounds__(kBlock) void k_lincomb(long n, double a, const double *x,
                                                    double b, const double *y, double *z) {
  for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += long(gridDim.x) * kBlock)
    z[k] = a * x[k] + b * y[k];
}

__global__ __launch_bounds__(kBlock) void k_scale(long n, double a, double *x) {
  for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += long(gridDim.x) * kBlock)
    x[k] *= a;
}

__global__ __launch_bounds__(kBlock) void k_fill(long n, double v, double *x) {
  for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += long(gridDim.x) * kBlock)
    x[k] = v;
}

__global__ __launch_bounds__(kBlock) void k_copy(long n, const double *x, double *y) {
  for (long k = long(blockIdx.x) * kBlock + threadIdx.x; k < n; k += long(gridDim.x) * kBlock)
    y[k] = x[k];
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

}  // namespace

long reduce_scratch_doubles(int m) { return long(kMaxBlocks) * (m < 1 ? 1 : m); }

void vec_multi_dot(long n, int m, const double *V, long ldv, const double *w,
                   double *scratch, double *out, hipStream_t s) {
  if (m < 1) return;
  if (m > kMaxDots) fail(OU_ERR_ARG, "too many simultaneous dot products");
  const int nb = blocks_for(n);
  if (m == 1)
    hipLaunchKernelGGL(k_multi_dot<1>, dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, w, scratch);
  else if (m <= 8)
    hipLaunchKernelGGL(k_multi_dot<8>, dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, w, scratch);
  else
    hipLaunchKernelGGL(k_multi_dot<kMaxDots>, dim3(nb), dim3(kBlock), 0, s, n, m, V, ldv, w,
                       scratch);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, s, nb, m, scratch, out);
  OU_HIP(hipGetLastError());
}

void vec_multi_axpy(long n, int m, const double *V, long ldv, const double *h, double alpha,
                    double *w, hipStream_t s) {
  if (m < 1) return;
  if (m > 64) fail(OU_ERR_ARG, "multi_axpy supports at most 64 vectors");
  hipLaunchKernelGGL(k_multi_axpy, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, m, V, ldv, h,
                     alpha, w);
  OU_HIP(hipGetLastError());
}

void vec_axpby(long n, double a, const double *x, double b, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_axpby, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, a, x, b, y);
  OU_HIP(hipGetLastError());
}

void vec_lincomb(long n, double a, const double *x, double b, const double *y, double *z,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_lincomb, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, a, x, b, y, z);
  OU_HIP(hipGetLastError());
}

void vec_scale(long n, double a, double *x, hipStream_t s) {
  hipLaunchKernelGGL(k_scale, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, a, x);
  OU_HIP(hipGetLastError());
}

void vec_copy(long n, const double *x, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_copy, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, x, y);
  OU_HIP(hipGetLastError());
}

void vec_fill(long n, double v, double *x, hipStream_t s) {
  hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(kBlock), 0, s, n, v, x);
  OU_HIP(hipGetLastError());
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
