// U-Net kernels (BF16 MFMA convolutions, pooling, head, loss, Adam) and the
// host-side UNet that sequences them.  The network and its parameter layout
// are described in unet.hpp.
//
// Convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16 with the
// weights as the A operand (rows = output channels) and the activations as
// the B operand (columns = 32 grid positions), so that one lane's operand is
// 8 consecutive channels of one position: a single 16-byte load from the
// channels-last tensor.  The accumulator then holds, in lane l, position
// l%32 and output channels 8(r/4) + 4(l/32) + r%4 for register r: four
// consecutive channels per register quad, stored as 8-byte BF16 quads.
//
// Weight gradients reduce over grid positions (the K dimension); a wave
// stages 16 positions x 32 channels of the output gradient and of the
// shifted input through LDS and reads them transposed.  Partial sums per
// position split go to a scratch buffer and are reduced in a fixed order,
// so gradients are deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "unet.hpp"

namespace ou {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bf2f(uint32_t h) { return __uint_as_float(h << 16); }

// FP32 -> BF16, round to nearest even (NaN stays NaN).
__device__ __forceinline__ uint32_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ bool bf_pos(uint32_t h) { return !(h & 0x8000u) && (h & 0x7fffu); }

__device__ __forceinline__ u32x4 ld16(const uint16_t *p) {
  return *reinterpret_cast<const u32x4 *>(p);
}

__device__ __forceinline__ void unpack8(u32x4 v, float *f) {
  for (int i = 0; i < 4; ++i) {
    f[2 * i] = bf2f(v[i] & 0xffffu);
    f[2 * i + 1] = bf2f(v[i] >> 16);
  }
}

__device__ __forceinline__ u32x4 pack8(const float *f) {
  u32x4 v;
  for (int i = 0; i < 4; ++i) v[i] = f2bf(f[2 * i]) | (f2bf(f[2 * i + 1]) << 16);
  return v;
}

// ---------------------------------------------------------------------------
// 3x3x3 convolution, zero padding, stride 1: y = act(W * x + b) [* mask]
// ---------------------------------------------------------------------------
struct ConvArgs {
  const uint16_t *x;
  int ldx, xoff, cin;
  const uint16_t *w;  // [27][cout][cin]
  const float *bias;  // [cout] or null
  uint16_t *y;
  int ldy, yoff, cout;
  const uint16_t *mask;  // output zeroed where mask <= 0 (null: no mask)
  int ldm, moff;
  int D;
  long P;  // positions = samples * D^3
};

template <int NT, bool RELU, bool MASK>
__global__ __launch_bounds__(256) void k_conv3(ConvArgs a) {
  const int lane = threadIdx.x & 63;
  const long base = (long(blockIdx.x) * 4 + (threadIdx.x >> 6)) * 32;
  if (base >= a.P) return;  // whole wave
  const long pos = base + (lane & 31);
  const bool live = pos < a.P;
  const int D = a.D;
  const long pc = live ? pos : base;
  const long D2 = long(D) * D;
  const int xq = int(pc % D), yq = int((pc / D) % D), zq = int((pc / D2) % D);
  const long bq = pc / (D2 * D);
  const int kh = (lane >> 5) * 8;
  f32x16 acc[NT];
  for (int t = 0; t < NT; ++t)
    for (int r = 0; r < 16; ++r) acc[t][r] = 0.f;
  for (int tap = 0; tap < 27; ++tap) {
    const int zz = zq + tap / 9 - 1, yy = yq + (tap / 3) % 3 - 1, xx = xq + tap % 3 - 1;
    const bool valid = live && unsigned(zz) < unsigned(D) && unsigned(yy) < unsigned(D) &&
                       unsigned(xx) < unsigned(D);
    const long src = valid ? ((bq * D + zz) * D + yy) * D + xx : 0;
    const uint16_t *xp = a.x + src * a.ldx + a.xoff + kh;
    const uint16_t *wp = a.w + (long(tap) * a.cout + (lane & 31)) * a.cin + kh;
    const uint32_t vm = valid ? 0xffffffffu : 0u;
    for (int c0 = 0; c0 < a.cin; c0 += 16) {
      u32x4 bx = ld16(xp + c0);
      bx[0] &= vm;
      bx[1] &= vm;
      bx[2] &= vm;
      bx[3] &= vm;
      const bf16x8 b = __builtin_bit_cast(bf16x8, bx);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const bf16x8 av = __builtin_bit_cast(bf16x8, ld16(wp + long(t) * 32 * a.cin + c0));
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b, acc[t], 0, 0, 0);
      }
    }
  }
  if (!live) return;
  const int g4 = (lane >> 5) * 4;
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int co = t * 32 + q * 8 + g4;
      float v[4];
      for (int i = 0; i < 4; ++i) {
        v[i] = acc[t][4 * q + i] + (a.bias ? a.bias[co + i] : 0.f);
        if (RELU) v[i] = fmaxf(v[i], 0.f);
      }
      if (MASK) {
        const u32x2 mk = *reinterpret_cast<const u32x2 *>(a.mask + pos * a.ldm + a.moff + co);
        if (!bf_pos(mk[0] & 0xffffu)) v[0] = 0.f;
        if (!bf_pos(mk[0] >> 16)) v[1] = 0.f;
        if (!bf_pos(mk[1] & 0xffffu)) v[2] = 0.f;
        if (!bf_pos(mk[1] >> 16)) v[3] = 0.f;
      }
      u32x2 o;
      o[0] = f2bf(v[0]) | (f2bf(v[1]) << 16);
      o[1] = f2bf(v[2]) | (f2bf(v[3]) << 16);
      *reinterpret_cast<u32x2 *>(a.y + pos * a.ldy + a.yoff + co) = o;
    }
}

// ---------------------------------------------------------------------------
// Weight gradient: partial[split][tap][co][ci] = sum_pos dy[pos][co] x[pos+tap][ci]
// ---------------------------------------------------------------------------
struct WgradArgs {
  const uint16_t *x;
  int ldx, xoff, cin;
  const uint16_t *dy;
  int ldy, yoff, cout;
  float *partial;
  int D;
  long P;
  long chunk;  // positions per split, multiple of 16
};

__global__ __launch_bounds__(64) void k_conv3_wgrad(WgradArgs a) {
  __shared__ __attribute__((aligned(16))) uint16_t sdy[16 * 32];
  __shared__ __attribute__((aligned(16))) uint16_t sx[16 * 32];
  const int lane = threadIdx.x;
  const int nci = (a.cin + 31) / 32, nco = a.cout / 32;
  int bid = blockIdx.x;
  const int cit = bid % nci;
  bid /= nci;
  const int cot = bid % nco;
  const int tap = bid / nco;
  const long p_begin = long(blockIdx.y) * a.chunk;
  const long p_end = p_begin + a.chunk < a.P ? p_begin + a.chunk : a.P;
  const int dz = tap / 9 - 1, dyy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int D = a.D;
  const long D2 = long(D) * D;
  const int row = lane >> 2, ch = (lane & 3) * 8;
  const bool ch_ok = cit * 32 + ch < a.cin;
  const int g = lane >> 5, col = lane & 31;
  f32x16 acc;
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  for (long p0 = p_begin; p0 < p_end; p0 += 16) {
    const long pos = p0 + row;
    u32x4 vdy = {0u, 0u, 0u, 0u}, vx = {0u, 0u, 0u, 0u};
    if (pos < p_end) {
      vdy = ld16(a.dy + pos * a.ldy + a.yoff + cot * 32 + ch);
      const int xq = int(pos % D), yq = int((pos / D) % D), zq = int((pos / D2) % D);
      const long bq = pos / (D2 * D);
      const int zz = zq + dz, yy = yq + dyy, xx = xq + dx;
      if (ch_ok && unsigned(zz) < unsigned(D) && unsigned(yy) < unsigned(D) &&
          unsigned(xx) < unsigned(D))
        vx = ld16(a.x + (((bq * D + zz) * D + yy) * D + xx) * a.ldx + a.xoff + cit * 32 + ch);
    }
    __syncthreads();
    *reinterpret_cast<u32x4 *>(sdy + row * 32 + ch) = vdy;
    *reinterpret_cast<u32x4 *>(sx + row * 32 + ch) = vx;
    __syncthreads();
    // K slot s of lane group g is staged position row 2s + g.
    u32x4 av, bv;
    for (int i = 0; i < 4; ++i) {
      const int r0 = 4 * i + g, r1 = 4 * i + 2 + g;
      av[i] = uint32_t(sdy[r0 * 32 + col]) | (uint32_t(sdy[r1 * 32 + col]) << 16);
      bv[i] = uint32_t(sx[r0 * 32 + col]) | (uint32_t(sx[r1 * 32 + col]) << 16);
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                  __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
  }
  const int ci = cit * 32 + col;
  if (ci >= a.cin) return;
  float *out = a.partial + ((long(blockIdx.y) * 27 + tap) * a.cout + cot * 32) * a.cin + ci;
  for (int r = 0; r < 16; ++r) {
    const int co = 8 * (r / 4) + 4 * g + (r % 4);
    out[long(co) * a.cin] = acc[r];
  }
}

// partial[split][c] = sum over the split's positions of dy[pos][off + c]
__global__ __launch_bounds__(256) void k_colsum(const uint16_t *dy, int ld, int off, int C,
                                                long P, long chunk, float *partial) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), r = threadIdx.x >> 6;
  const long b = long(blockIdx.y) * chunk;
  const long e = b + chunk < P ? b + chunk : P;
  float s = 0.f;
  if (c < C)
    for (long p = b + r; p < e; p += 4) s += bf2f(dy[p * ld + off + c]);
  red[r][threadIdx.x & 63] = s;
  __syncthreads();
  if (r == 0 && c < C)
    partial[long(blockIdx.y) * C + c] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// out[i] = sum_k partial[k n + i]
__global__ void k_reduce_splits(const float *partial, int nsplit, long n, float *out) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int k = 0; k < nsplit; ++k) s += partial[long(k) * n + i];
  out[i] = s;
}

// wf[tap][ci][co] = w[26 - tap][co][ci]
__global__ void k_flip(const uint16_t *w, int cin, int cout, uint16_t *wf) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  const long n = 27L * cin * cout;
  if (i >= n) return;
  const int co = int(i % cout);
  const int ci = int((i / cout) % cin);
  const int tap = int(i / (long(cout) * cin));
  wf[i] = w[(long(26 - tap) * cout + co) * cin + ci];
}

// ---------------------------------------------------------------------------
// Pooling / upsampling (factor 2, channels in groups of 8)
// ---------------------------------------------------------------------------
__device__ __forceinline__ long parent_of(long pos, int D) {
  const long D2 = long(D) * D;
  const int x = int(pos % D), y = int((pos / D) % D), z = int((pos / D2) % D);
  const long b = pos / (D2 * D);
  const int d = D / 2;
  return ((b * d + z / 2) * d + y / 2) * d + x / 2;
}

// y[po][c] = mean of the 8 children x[child][xoff + c]; Dout = D/2
__global__ void k_pool2(const uint16_t *x, int ldx, int xoff, uint16_t *y, int ldy, int C,
                        int Dout, long Pout) {
  const long t = long(blockIdx.x) * 256 + threadIdx.x;
  const int ng = C / 8;
  if (t >= Pout * ng) return;
  const int cg = int(t % ng);
  const long po = t / ng;
  const long d2 = long(Dout) * Dout;
  const int X = int(po % Dout), Y = int((po / Dout) % Dout), Z = int((po / d2) % Dout);
  const long b = po / (d2 * Dout);
  const int D = 2 * Dout;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 8; ++k) {
    const long src = ((b * D + 2 * Z + (k >> 2)) * D + 2 * Y + ((k >> 1) & 1)) * D + 2 * X + (k & 1);
    float f[8];
    unpack8(ld16(x + src * ldx + xoff + cg * 8), f);
    for (int i = 0; i < 8; ++i) s[i] += f[i];
  }
  for (int i = 0; i < 8; ++i) s[i] *= 0.125f;
  *reinterpret_cast<u32x4 *>(y + po * ldy + cg * 8) = pack8(s);
}

// Pooling backward into a skip tensor that already holds the gradient from
// its other consumer: dx[p][off+c] = (x[p][off+c] > 0) ? dx + dy[parent][c]/8 : 0
__global__ void k_pool2_bwd(const uint16_t *x, uint16_t *dx, int ld, int off, const uint16_t *dy,
                            int ldy, int C, int D, long P) {
  const long t = long(blockIdx.x) * 256 + threadIdx.x;
  const int ng = C / 8;
  if (t >= P * ng) return;
  const int cg = int(t % ng);
  const long p = t / ng;
  const long par = parent_of(p, D);
  float xv[8], dv[8], pv[8];
  unpack8(ld16(x + p * ld + off + cg * 8), xv);
  unpack8(ld16(dx + p * ld + off + cg * 8), dv);
  unpack8(ld16(dy + par * ldy + cg * 8), pv);
  for (int i = 0; i < 8; ++i) dv[i] = xv[i] > 0.f ? dv[i] + 0.125f * pv[i] : 0.f;
  *reinterpret_cast<u32x4 *>(dx + p * ld + off + cg * 8) = pack8(dv);
}

// y[p][yoff + c] = x[parent(p)][c]; D is the fine grid side
__global__ void k_up2(const uint16_t *x, int ldx, uint16_t *y, int ldy, int yoff, int C, int D,
                      long P) {
  const long t = long(blockIdx.x) * 256 + threadIdx.x;
  const int ng = C / 8;
  if (t >= P * ng) return;
  const int cg = int(t % ng);
  const long p = t / ng;
  *reinterpret_cast<u32x4 *>(y + p * ldy + yoff + cg * 8) = ld16(x + parent_of(p, D) * ldx + cg * 8);
}

// dx[pc][c] = (x[pc][c] > 0) ? sum over the 8 children of dy[child][dyoff + c] : 0
__global__ void k_up2_bwd(const uint16_t *x, uint16_t *dx, int ldx, const uint16_t *dy, int ldy,
                          int dyoff, int C, int Dc, long Pc) {
  const long t = long(blockIdx.x) * 256 + threadIdx.x;
  const int ng = C / 8;
  if (t >= Pc * ng) return;
  const int cg = int(t % ng);
  const long pc = t / ng;
  const long d2 = long(Dc) * Dc;
  const int X = int(pc % Dc), Y = int((pc / Dc) % Dc), Z = int((pc / d2) % Dc);
  const long b = pc / (d2 * Dc);
  const int D = 2 * Dc;
  float s[8] = {0, 0, 0, 0, 0, 0, 0, 0}, xv[8];
  for (int k = 0; k < 8; ++k) {
    const long ch = ((b * D + 2 * Z + (k >> 2)) * D + 2 * Y + ((k >> 1) & 1)) * D + 2 * X + (k & 1);
    float f[8];
    unpack8(ld16(dy + ch * ldy + dyoff + cg * 8), f);
    for (int i = 0; i < 8; ++i) s[i] += f[i];
  }
  unpack8(ld16(x + pc * ldx + cg * 8), xv);
  for (int i = 0; i < 8; ++i)
    if (!(xv[i] > 0.f)) s[i] = 0.f;
  *reinterpret_cast<u32x4 *>(dx + pc * ldx + cg * 8) = pack8(s);
}

// ---------------------------------------------------------------------------
// Input, head, loss, Adam
// ---------------------------------------------------------------------------
// in0[s][p][0..16) = (+-r_hat[0..8), tau_hat, kappa_hat, 0...): samples s < B
// carry +r_hat, s >= B carry -r_hat
__global__ void k_unet_input(const float *feat, int B, long P0, uint16_t *in0) {
  const long t = long(blockIdx.x) * 256 + threadIdx.x;
  if (t >= 2L * B * P0) return;
  const long s = t / P0, p = t % P0;
  const float *f = feat + ((s % B) * P0 + p) * kUnetFeatures;
  const float sg = s < B ? 1.0f : -1.0f;
  float v[8], w[8] = {f[kUnetFeatTau], f[kUnetFeatKappa], 0, 0, 0, 0, 0, 0};
  for (int k = 0; k < 8; ++k) v[k] = sg * f[k];
  *reinterpret_cast<u32x4 *>(in0 + t * kUnetInPad) = pack8(v);
  *reinterpret_cast<u32x4 *>(in0 + t * kUnetInPad + 8) = pack8(w);
}

// out[p][k] = b[k] + sum_c w[k][c] x[p][c]  (x: 32 channels; wb = w[8][32], b[8])
__global__ void k_head(const uint16_t *x, const float *wb, float *out, long P) {
  const long p = long(blockIdx.x) * 256 + threadIdx.x;
  if (p >= P) return;
  float f[32];
  for (int q = 0; q < 4; ++q) unpack8(ld16(x + p * 32 + 8 * q), f + 8 * q);
  for (int k = 0; k < kUnetOut; ++k) {
    float s = wb[kUnetOut * 32 + k];
    for (int c = 0; c < 32; ++c) s += wb[k * 32 + c] * f[c];
    out[p * kUnetOut + k] = s;
  }
}

// dx[p][c] = (x[p][c] > 0) ? sum_k dout[p][k] w[k][c] : 0
__global__ void k_head_bwd_dx(const uint16_t *x, const float *wb, const float *dout, uint16_t *dx,
                              long P) {
  const long p = long(blockIdx.x) * 256 + threadIdx.x;
  if (p >= P) return;
  float d[kUnetOut];
  for (int k = 0; k < kUnetOut; ++k) d[k] = dout[p * kUnetOut + k];
  for (int q = 0; q < 4; ++q) {
    float f[8], o[8];
    unpack8(ld16(x + p * 32 + 8 * q), f);
    for (int i = 0; i < 8; ++i) {
      float s = 0.f;
      for (int k = 0; k < kUnetOut; ++k) s += d[k] * wb[k * 32 + 8 * q + i];
      o[i] = f[i] > 0.f ? s : 0.f;
    }
    *reinterpret_cast<u32x4 *>(dx + p * 32 + 8 * q) = pack8(o);
  }
}

// partial[split][k 32 + c] = sum_p dout[p][k] x[p][c], partial[split][256 + k] = sum_p dout[p][k]
// (256 threads: thread = k 32 + c)
__global__ __launch_bounds__(256) void k_head_bwd_w(const uint16_t *x, const float *dout, long P,
                                                    long chunk, float *partial) {
  const long b = long(blockIdx.x) * chunk;
  const long e = b + chunk < P ? b + chunk : P;
  const int k = threadIdx.x >> 5, c = threadIdx.x & 31;
  float sw = 0.f, sb = 0.f;
  for (long p = b; p < e; ++p) {
    const float d = dout[p * kUnetOut + k];
    sw += d * bf2f(x[p * 32 + c]);
    sb += d;
  }
  partial[long(blockIdx.x) * kUnetHead + threadIdx.x] = sw;
  if (c == 0) partial[long(blockIdx.x) * kUnetHead + kUnetOut * 32 + k] = sb;
}

// The prediction of sample b at cell p:
//   gain * (skip * r_hat + (out[b][p] - out[B + b][p]) / 2)
// the odd part of the network, optionally plus a skip connection from its
// residual input (mode bit 0) and scaled by the prior g (mode bit 1, else 1).
__device__ __forceinline__ float gain_of(const float *f, int mode) {
  return (mode & 2) ? f[kUnetFeatGain] : 1.0f;
}

// Element q of a sample is output q % 8 of cell q / 8; E = cells * 8 per sample.
__device__ __forceinline__ float pred_of(const float *out, const float *feat, int B, long E,
                                         long b, long q, int mode) {
  const float *f = feat + (b * (E / kUnetOut) + q / kUnetOut) * kUnetFeatures;
  const float skip = (mode & 1) ? f[q % kUnetOut] : 0.0f;
  return gain_of(f, mode) * (skip + 0.5f * (out[b * E + q] - out[(B + b) * E + q]));
}

__global__ void k_predict(const float *out, const float *feat, int B, long E, int mode,
                          float *pred) {
  const long t = long(blockIdx.x) * 256 + threadIdx.x;
  if (t >= long(B) * E) return;
  pred[t] = pred_of(out, feat, B, E, t / E, t % E, mode);
}

// part[b][chunk] = (sum (pred - t)^2, sum t^2) over the chunk
__global__ __launch_bounds__(256) void k_loss_part(const float *out, const float *feat,
                                                   const float *target, int B, long E, long chunk,
                                                   int mode, float *part) {
  __shared__ float red[2][256];
  const int b = blockIdx.y;
  const long s0 = long(blockIdx.x) * chunk;
  const long e = s0 + chunk < E ? s0 + chunk : E;
  float num = 0.f, den = 0.f;
  for (long q = s0 + threadIdx.x; q < e; q += 256) {
    const long t = long(b) * E + q;
    const float d = pred_of(out, feat, B, E, b, q, mode) - target[t];
    num += d * d;
    den += target[t] * target[t];
  }
  red[0][threadIdx.x] = num;
  red[1][threadIdx.x] = den;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[0][threadIdx.x] += red[0][threadIdx.x + o];
      red[1][threadIdx.x] += red[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[(long(b) * gridDim.x + blockIdx.x) * 2] = red[0][0];
    part[(long(b) * gridDim.x + blockIdx.x) * 2 + 1] = red[1][0];
  }
}

// loss[0] = mean_b num_b / den_b; loss[1 + 2b] = num_b, loss[2 + 2b] = den_b
__global__ void k_loss_final(const float *part, int B, int nchunk, double *loss) {
  if (threadIdx.x != 0) return;
  double tot = 0.0;
  for (int b = 0; b < B; ++b) {
    double num = 0.0, den = 0.0;
    for (int c = 0; c < nchunk; ++c) {
      num += part[(long(b) * nchunk + c) * 2];
      den += part[(long(b) * nchunk + c) * 2 + 1];
    }
    loss[1 + 2 * b] = num;
    loss[2 + 2 * b] = den;
    tot += den > 0.0 ? num / den : 0.0;
  }
  loss[0] = tot / B;
}

// gout[b] = (pred_b - t_b) gain / (B den_b), gout[B + b] = -gout[b]
__global__ void k_loss_grad(const float *out, const float *feat, const float *target, int B,
                            long E, int mode, const double *loss, float *gout) {
  const long t = long(blockIdx.x) * 256 + threadIdx.x;
  if (t >= long(B) * E) return;
  const long b = t / E, q = t % E;
  const double den = loss[2 + 2 * b];
  const float sc = den > 0.0 ? float(1.0 / (double(B) * den)) : 0.f;
  const float d = (pred_of(out, feat, B, E, b, q, mode) - target[t]) *
                  gain_of(feat + (t / kUnetOut) * kUnetFeatures, mode) * sc;
  gout[t] = d;
  gout[(B + b) * E + q] = -d;
}

__global__ void k_adam(long n, float *p, const float *g, float *m, float *v, float lr, float b1,
                       float b2, float eps, float bc1, float bc2, uint16_t *pbf) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  p[i] = pi;
  pbf[i] = uint16_t(f2bf(pi));
}

__global__ void k_to_bf16(long n, const float *p, uint16_t *pbf) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i < n) pbf[i] = uint16_t(f2bf(p[i]));
}

inline unsigned blocks(long n, int per = 256) { return unsigned((n + per - 1) / per); }

template <int NT>
void launch_conv_nt(const ConvArgs &a, bool relu, bool mask, hipStream_t s) {
  const unsigned grid = blocks(a.P, 128);
  if (relu && !mask) k_conv3<NT, true, false><<<grid, 256, 0, s>>>(a);
  else if (!relu && mask) k_conv3<NT, false, true><<<grid, 256, 0, s>>>(a);
  else if (!relu && !mask) k_conv3<NT, false, false><<<grid, 256, 0, s>>>(a);
  else fail(OU_ERR_ARG, "conv: relu and mask together are not instantiated");
}

void launch_conv(const ConvArgs &a, bool relu, bool mask, hipStream_t s) {
  if (a.cin % 16 || a.cout % 32 || a.ldx % 8 || a.xoff % 8 || a.ldy % 4 || a.yoff % 4)
    fail(OU_ERR_ARG, "conv: channel counts or strides not aligned");
  if (mask && (a.ldm % 4 || a.moff % 4)) fail(OU_ERR_ARG, "conv: mask stride not aligned");
  switch (a.cout / 32) {
    case 1: launch_conv_nt<1>(a, relu, mask, s); break;
    case 2: launch_conv_nt<2>(a, relu, mask, s); break;
    case 3: launch_conv_nt<3>(a, relu, mask, s); break;
    case 4: launch_conv_nt<4>(a, relu, mask, s); break;
    case 6: launch_conv_nt<6>(a, relu, mask, s); break;
    default: fail(OU_ERR_ARG, "conv: unsupported output channel count");
  }
  OU_HIP(hipGetLastError());
}

// Position splits for the weight gradient: enough workgroups to fill the
// chip, at least 256 positions per split.
void wgrad_split(long P, int cin, int cout, long *chunk, int *nsplit) {
  const long tiles = 27L * (cout / 32) * ((cin + 31) / 32);
  long ns = (4096 + tiles - 1) / tiles;
  const long max_ns = (P + 255) / 256;
  if (ns > max_ns) ns = max_ns;
  if (ns < 1) ns = 1;
  long ch = (P + ns - 1) / ns;
  ch = (ch + 15) / 16 * 16;
  *chunk = ch;
  *nsplit = int((P + ch - 1) / ch);
}

void colsum_split(long P, long *chunk, int *nsplit) {
  long ns = (P + 1023) / 1024;
  if (ns > 256) ns = 256;
  if (ns < 1) ns = 1;
  *chunk = (P + ns - 1) / ns;
  *nsplit = int((P + *chunk - 1) / *chunk);
}

struct SplitMix {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    const double u1 = 1.0 - uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

}  // namespace

void *UNet::alloc(size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
    (void)hipGetLastError();
    fail(OU_ERR_NOMEM, "U-Net device allocation failed");
  }
  allocs_.push_back(p);
  return p;
}

UNet::UNet(int grid, int max_batch, uint64_t seed) : G_(grid), max_batch_(max_batch) {
  if (grid < 4 || grid % 4 || grid > 256) fail(OU_ERR_ARG, "U-Net grid must be a multiple of 4 in [4, 256]");
  if (max_batch < 1 || max_batch > 64) fail(OU_ERR_ARG, "U-Net batch must be in [1, 64]");
  const int io[kUnetConvs][2] = {{kUnetInPad, 32}, {32, 32}, {32, 64},  {64, 64}, {64, 128},
                                 {128, 128},      {192, 64}, {64, 64}, {96, 32}, {32, 32}};
  long off = 0;
  for (int i = 0; i < kUnetConvs; ++i) {
    conv_[i].cin = io[i][0];
    conv_[i].cout = io[i][1];
    conv_[i].w_off = off;
    off += 27L * io[i][0] * io[i][1];
    conv_[i].b_off = off;
    off += io[i][1];
  }
  head_w_off_ = off;
  head_b_off_ = off + kUnetOut * 32;
  n_params_ = off + kUnetHead;

  // He-normal initialisation from SplitMix64; padded input channels get 0.
  std::vector<float> h(size_t(n_params_), 0.f);
  SplitMix rng{seed};
  for (int i = 0; i < kUnetConvs; ++i) {
    const int cin_real = i == 0 ? kUnetInReal : conv_[i].cin;
    const double sd = std::sqrt(2.0 / (27.0 * cin_real));
    for (int t = 0; t < 27; ++t)
      for (int co = 0; co < conv_[i].cout; ++co)
        for (int ci = 0; ci < conv_[i].cin; ++ci)
          h[size_t(conv_[i].w_off + (long(t) * conv_[i].cout + co) * conv_[i].cin + ci)] =
              ci < cin_real ? float(sd * rng.normal()) : 0.f;
  }
  // a small head: the untrained correction stays small
  for (int c = 0; c < kUnetOut * 32; ++c)
    h[size_t(head_w_off_ + c)] = float(0.1 * std::sqrt(1.0 / 32.0) * rng.normal());

  param_ = static_cast<float *>(alloc(size_t(n_params_) * 4));
  grad_ = static_cast<float *>(alloc(size_t(n_params_) * 4));
  adam_m_ = static_cast<float *>(alloc(size_t(n_params_) * 4));
  adam_v_ = static_cast<float *>(alloc(size_t(n_params_) * 4));
  wbf_ = static_cast<uint16_t *>(alloc(size_t(n_params_) * 2));
  long max_w = 0;
  for (auto &c : conv_) max_w = std::max(max_w, 27L * c.cin * c.cout);
  wflip_ = static_cast<uint16_t *>(alloc(size_t(max_w) * 2));
  OU_HIP(hipMemset(grad_, 0, size_t(n_params_) * 4));
  OU_HIP(hipMemset(adam_m_, 0, size_t(n_params_) * 4));
  OU_HIP(hipMemset(adam_v_, 0, size_t(n_params_) * 4));

  const long S = 2L * max_batch;
  const long P0 = long(G_) * G_ * G_, P1 = P0 / 8, P2 = P0 / 64;
  auto act = [&](long P, int C) { return static_cast<uint16_t *>(alloc(size_t(S * P * C) * 2)); };
  in0_ = act(P0, kUnetInPad);
  e0a_ = act(P0, 32);
  cat0_ = act(P0, 96);
  p0_ = act(P1, 32);
  e1a_ = act(P1, 64);
  cat1_ = act(P1, 192);
  p1_ = act(P2, 64);
  ba_ = act(P2, 128);
  bb_ = act(P2, 128);
  d1a_ = act(P1, 64);
  d1b_ = act(P1, 64);
  d0a_ = act(P0, 32);
  d0b_ = act(P0, 32);
  g_e0a_ = act(P0, 32);
  g_cat0_ = act(P0, 96);
  g_p0_ = act(P1, 32);
  g_e1a_ = act(P1, 64);
  g_cat1_ = act(P1, 192);
  g_p1_ = act(P2, 64);
  g_ba_ = act(P2, 128);
  g_bb_ = act(P2, 128);
  g_d1a_ = act(P1, 64);
  g_d1b_ = act(P1, 64);
  g_d0a_ = act(P0, 32);
  g_d0b_ = act(P0, 32);
  out_ = static_cast<float *>(alloc(size_t(S * P0) * kUnetOut * 4));
  g_out_ = static_cast<float *>(alloc(size_t(S * P0) * kUnetOut * 4));

  // scratch for gradient partial sums: the largest of any layer's needs
  long need = 0;
  const int lvl[kUnetConvs] = {0, 0, 1, 1, 2, 2, 1, 1, 0, 0};
  for (int i = 0; i < kUnetConvs; ++i) {
    const long P = S * (P0 >> (3 * lvl[i]));
    long chunk;
    int ns;
    wgrad_split(P, conv_[i].cin, conv_[i].cout, &chunk, &ns);
    need = std::max(need, long(ns) * 27 * conv_[i].cin * conv_[i].cout);
    colsum_split(P, &chunk, &ns);
    need = std::max(need, long(ns) * conv_[i].cout);
  }
  need = std::max(need, long(kUnetHead) * ((S * P0 + 4095) / 4096));
  partial_size_ = need;
  partial_ = static_cast<float *>(alloc(size_t(need) * 4));
  loss_ = static_cast<double *>(alloc(size_t(1 + 2 * max_batch) * 8));
  loss_part_ = static_cast<float *>(alloc(size_t(max_batch) * 2 * 1024 * 4));
  set_params(h.data());
}

UNet::~UNet() {
  for (void *p : allocs_) (void)hipFree(p);
}

void UNet::get_params(float *h) const {
  OU_HIP(hipDeviceSynchronize());
  OU_HIP(hipMemcpy(h, param_, size_t(n_params_) * 4, hipMemcpyDeviceToHost));
}

void UNet::set_params(const float *h) {
  OU_HIP(hipMemcpy(param_, h, size_t(n_params_) * 4, hipMemcpyHostToDevice));
  refresh_bf16(nullptr);
  OU_HIP(hipDeviceSynchronize());
}

void UNet::refresh_bf16(hipStream_t s) {
  k_to_bf16<<<blocks(n_params_), 256, 0, s>>>(n_params_, param_, wbf_);
  OU_HIP(hipGetLastError());
}

void UNet::conv_forward(int layer, const uint16_t *x, int ldx, int xoff, uint16_t *y, int ldy,
                        int yoff, int D, int nb, hipStream_t s) {
  const UnetConv &c = conv_[layer];
  ConvArgs a{};
  a.x = x;
  a.ldx = ldx;
  a.xoff = xoff;
  a.cin = c.cin;
  a.w = wbf_ + c.w_off;
  a.bias = param_ + c.b_off;
  a.y = y;
  a.ldy = ldy;
  a.yoff = yoff;
  a.cout = c.cout;
  a.D = D;
  a.P = long(nb) * D * D * D;
  launch_conv(a, true, false, s);
}

void UNet::conv_backward(int layer, const uint16_t *x, int ldx, int xoff, const uint16_t *dy,
                         int ldy, int yoff, uint16_t *dx, int lddx, int dxoff,
                         const uint16_t *mask, int ldm, int moff, int D, int nb, hipStream_t s) {
  const UnetConv &c = conv_[layer];
  const long P = long(nb) * D * D * D;
  // weight gradient
  WgradArgs w{};
  w.x = x;
  w.ldx = ldx;
  w.xoff = xoff;
  w.cin = c.cin;
  w.dy = dy;
  w.ldy = ldy;
  w.yoff = yoff;
  w.cout = c.cout;
  w.D = D;
  w.P = P;
  int ns;
  wgrad_split(P, c.cin, c.cout, &w.chunk, &ns);
  const long nw = 27L * c.cin * c.cout;
  if (long(ns) * nw > partial_size_) fail(OU_ERR_SIZE, "U-Net gradient scratch too small");
  w.partial = partial_;
  dim3 grid(unsigned(27 * (c.cout / 32) * ((c.cin + 31) / 32)), unsigned(ns));
  k_conv3_wgrad<<<grid, 64, 0, s>>>(w);
  k_reduce_splits<<<blocks(nw), 256, 0, s>>>(partial_, ns, nw, grad_ + c.w_off);
  // bias gradient
  long chunk;
  colsum_split(P, &chunk, &ns);
  k_colsum<<<dim3(unsigned((c.cout + 63) / 64), unsigned(ns)), 256, 0, s>>>(dy, ldy, yoff, c.cout,
                                                                         P, chunk, partial_);
  k_reduce_splits<<<blocks(c.cout), 256, 0, s>>>(partial_, ns, c.cout, grad_ + c.b_off);
  OU_HIP(hipGetLastError());
  if (!dx) return;
  // data gradient: convolution of dy with the flipped, transposed weights
  k_flip<<<blocks(nw), 256, 0, s>>>(wbf_ + c.w_off, c.cin, c.cout, wflip_);
  ConvArgs a{};
  a.x = dy;
  a.ldx = ldy;
  a.xoff = yoff;
  a.cin = c.cout;
  a.w = wflip_;
  a.bias = nullptr;
  a.y = dx;
  a.ldy = lddx;
  a.yoff = dxoff;
  a.cout = c.cin;
  a.mask = mask;
  a.ldm = ldm;
  a.moff = moff;
  a.D = D;
  a.P = P;
  launch_conv(a, false, mask != nullptr, s);
}

void UNet::forward(const float *d_feat, int B, hipStream_t s) {
  if (B < 1 || B > max_batch_) fail(OU_ERR_SIZE, "U-Net batch exceeds the allocated maximum");
  const int G = G_, nb = 2 * B;
  const long P0 = long(G) * G * G, P1 = P0 / 8;
  k_unet_input<<<blocks(nb * P0), 256, 0, s>>>(d_feat, B, P0, in0_);
  conv_forward(0, in0_, kUnetInPad, 0, e0a_, 32, 0, G, nb, s);
  conv_forward(1, e0a_, 32, 0, cat0_, 96, 64, G, nb, s);
  k_pool2<<<blocks(nb * P1 * 4), 256, 0, s>>>(cat0_, 96, 64, p0_, 32, 32, G / 2, nb * P1);
  conv_forward(2, p0_, 32, 0, e1a_, 64, 0, G / 2, nb, s);
  conv_forward(3, e1a_, 64, 0, cat1_, 192, 128, G / 2, nb, s);
  k_pool2<<<blocks(nb * (P1 / 8) * 8), 256, 0, s>>>(cat1_, 192, 128, p1_, 64, 64, G / 4,
                                                    nb * (P1 / 8));
  conv_forward(4, p1_, 64, 0, ba_, 128, 0, G / 4, nb, s);
  conv_forward(5, ba_, 128, 0, bb_, 128, 0, G / 4, nb, s);
  k_up2<<<blocks(nb * P1 * 16), 256, 0, s>>>(bb_, 128, cat1_, 192, 0, 128, G / 2, nb * P1);
  conv_forward(6, cat1_, 192, 0, d1a_, 64, 0, G / 2, nb, s);
  conv_forward(7, d1a_, 64, 0, d1b_, 64, 0, G / 2, nb, s);
  k_up2<<<blocks(nb * P0 * 8), 256, 0, s>>>(d1b_, 64, cat0_, 96, 0, 64, G, nb * P0);
  conv_forward(8, cat0_, 96, 0, d0a_, 32, 0, G, nb, s);
  conv_forward(9, d0a_, 32, 0, d0b_, 32, 0, G, nb, s);
  k_head<<<blocks(nb * P0), 256, 0, s>>>(d0b_, param_ + head_w_off_, out_, nb * P0);
  OU_HIP(hipGetLastError());
}

void UNet::predict(const float *d_feat, int B, float *d_pred, hipStream_t s) {
  forward(d_feat, B, s);
  const long E = long(G_) * G_ * G_ * kUnetOut;
  k_predict<<<blocks(B * E), 256, 0, s>>>(out_, d_feat, B, E, mode_, d_pred);
  OU_HIP(hipGetLastError());
}

void UNet::loss_and_grad(const float *d_feat, const float *d_target, int B, hipStream_t s) {
  forward(d_feat, B, s);
  const int G = G_, nb = 2 * B;
  const long P0 = long(G) * G * G, P1 = P0 / 8, P2 = P0 / 64;
  // loss and its gradient with respect to the raw outputs
  const long E = P0 * kUnetOut;
  const int nchunk = int(std::min<long>(1024, (E + 4095) / 4096));
  const long lchunk = (E + nchunk - 1) / nchunk;
  k_loss_part<<<dim3(unsigned(nchunk), unsigned(B)), 256, 0, s>>>(out_, d_feat, d_target, B, E,
                                                                  lchunk, mode_, loss_part_);
  k_loss_final<<<1, 64, 0, s>>>(loss_part_, B, nchunk, loss_);
  k_loss_grad<<<blocks(B * E), 256, 0, s>>>(out_, d_feat, d_target, B, E, mode_, loss_, g_out_);
  // head
  k_head_bwd_dx<<<blocks(nb * P0), 256, 0, s>>>(d0b_, param_ + head_w_off_, g_out_, g_d0b_, nb * P0);
  {
    const long P = nb * P0, chunk = 4096;
    const int ns = int((P + chunk - 1) / chunk);
    k_head_bwd_w<<<unsigned(ns), 256, 0, s>>>(d0b_, g_out_, P, chunk, partial_);
    k_reduce_splits<<<blocks(kUnetHead), 256, 0, s>>>(partial_, ns, kUnetHead,
                                                      grad_ + head_w_off_);
  }
  OU_HIP(hipGetLastError());
  conv_backward(9, d0a_, 32, 0, g_d0b_, 32, 0, g_d0a_, 32, 0, d0a_, 32, 0, G, nb, s);
  conv_backward(8, cat0_, 96, 0, g_d0a_, 32, 0, g_cat0_, 96, 0, nullptr, 0, 0, G, nb, s);
  k_up2_bwd<<<blocks(nb * P1 * 8), 256, 0, s>>>(d1b_, g_d1b_, 64, g_cat0_, 96, 0, 64, G / 2, nb * P1);
  conv_backward(7, d1a_, 64, 0, g_d1b_, 64, 0, g_d1a_, 64, 0, d1a_, 64, 0, G / 2, nb, s);
  conv_backward(6, cat1_, 192, 0, g_d1a_, 64, 0, g_cat1_, 192, 0, nullptr, 0, 0, G / 2, nb, s);
  k_up2_bwd<<<blocks(nb * P2 * 16), 256, 0, s>>>(bb_, g_bb_, 128, g_cat1_, 192, 0, 128, G / 4,
                                                 nb * P2);
  conv_backward(5, ba_, 128, 0, g_bb_, 128, 0, g_ba_, 128, 0, ba_, 128, 0, G / 4, nb, s);
  conv_backward(4, p1_, 64, 0, g_ba_, 128, 0, g_p1_, 64, 0, nullptr, 0, 0, G / 4, nb, s);
  k_pool2_bwd<<<blocks(nb * P1 * 8), 256, 0, s>>>(cat1_, g_cat1_, 192, 128, g_p1_, 64, 64, G / 2,
                                                  nb * P1);
  conv_backward(3, e1a_, 64, 0, g_cat1_, 192, 128, g_e1a_, 64, 0, e1a_, 64, 0, G / 2, nb, s);
  conv_backward(2, p0_, 32, 0, g_e1a_, 64, 0, g_p0_, 32, 0, nullptr, 0, 0, G / 2, nb, s);
  k_pool2_bwd<<<blocks(nb * P0 * 4), 256, 0, s>>>(cat0_, g_cat0_, 96, 64, g_p0_, 32, 32, G, nb * P0);
  conv_backward(1, e0a_, 32, 0, g_cat0_, 96, 64, g_e0a_, 32, 0, e0a_, 32, 0, G, nb, s);
  conv_backward(0, in0_, kUnetInPad, 0, g_e0a_, 32, 0, nullptr, 0, 0, nullptr, 0, 0, G, nb, s);
}

void UNet::adam_step(float lr, hipStream_t s) {
  ++adam_t_;
  const float b1 = 0.9f, b2 = 0.999f;
  const float bc1 = float(1.0 - std::pow(double(b1), adam_t_));
  const float bc2 = float(1.0 - std::pow(double(b2), adam_t_));
  k_adam<<<blocks(n_params_), 256, 0, s>>>(n_params_, param_, grad_, adam_m_, adam_v_, lr, b1, b2,
                                           1e-8f, bc1, bc2, wbf_);
  OU_HIP(hipGetLastError());
}

}  // namespace ou
