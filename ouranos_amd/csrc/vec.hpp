// FP64 vector kernels of the Krylov drivers (the reference leaves these to
// AztecOO / Epetra).  All device pointers; all enqueued on `s`.
#pragma once

#include <hip/hip_runtime.h>

namespace ou {

// Scratch for reductions: at least reduce_scratch_doubles(m) doubles.
long reduce_scratch_doubles(int m);

// out[i] = <V_i, w> for i < m, V_i = V + i*ldv (FP64, deterministic order).
// `out` is device memory; partial sums use `scratch`.
void vec_multi_dot(long n, int m, const double *V, long ldv, const double *w,
                   double *scratch, double *out, hipStream_t s);
// w -= sum_{i<m} h[i] V_i (h on the device, m <= 32), then out[i] =
// <V_i, w_new> for i < m when `dots`, and out[m or 0] = <w_new, w_new>: one
// fused pass for m <= 16, an update and one dot-product pass beyond.
// scratch: reduce_scratch_doubles(m + 1) doubles.
void vec_axpy_dots(long n, int m, const double *V, long ldv, const double *h, double *w,
                   bool dots, double *scratch, double *out, hipStream_t s);
// w -= sum_i h[i] V_i  (h on the device, m <= 64)
void vec_multi_axpy(long n, int m, const double *V, long ldv, const double *h,
                    double alpha, double *w, hipStream_t s);
// The FP32-basis forms, for a Krylov basis stored in FP32 (w, h, x and
// every sum FP64):
// when `update`, w -= sum_{i<m} h[i] V_i; then out[i] = <V_i, w> for i < m
// when `dots`, and out[m or 0] = <w, w>; one pass, m <= 32.
// scratch: reduce_scratch_doubles(m + 1) doubles.
void vec_gs_f32(long n, int m, const float *V, long ldv, const double *h, double *w,
                bool update, bool dots, double *scratch, double *out, hipStream_t s);
// x += alpha sum_i h[i] V_i  (h on the device, m <= 64)
void vec_multi_axpy_f32(long n, int m, const float *V, long ldv, const double *h,
                        double alpha, double *x, hipStream_t s);
// y = (float)(a x), and x = a x when `scale`
void vec_scale_store_f32(long n, double a, double *x, bool scale, float *y, hipStream_t s);
// y = a*x + b*y
void vec_axpby(long n, double a, const double *x, double b, double *y, hipStream_t s);
// z = a*x + b*y
void vec_lincomb(long n, double a, const double *x, double b, const double *y, double *z,
                 hipStream_t s);
void vec_scale(long n, double a, double *x, hipStream_t s);
void vec_copy(long n, const double *x, double *y, hipStream_t s);
void vec_fill(long n, double v, double *x, hipStream_t s);
// zero `bytes` bytes (multiple of 4) with a fill kernel
void vec_zero_bytes(void *p, size_t bytes, hipStream_t s);
// counts non-finite entries into *flag (device int, must be zeroed)
void vec_check_finite(long n, const double *x, int *flag, hipStream_t s);

}  // namespace ou
