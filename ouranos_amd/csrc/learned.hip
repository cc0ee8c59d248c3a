// Coarse-grid transfer, features and the learned FGMRES preconditioner (see
// learned.hpp), plus the FP64-residual refinement driver around the FP32
// sweep.
//
// Transfer between a fine DG Q_p field on n^3 cells and the network's G^3
// grid (f = n / G fine cells per coarse cell and axis).  Along one axis the
// fine nodes of a coarse cell sit at sub-positions s = i' p + l in [0, f p]
// (fine cell i' of the coarse cell, local node l).  A node belongs to the
// lower half of the coarse cell if 2 s < f p, to the upper half if
// 2 s > f p, and half to each if 2 s = f p; the octant weight of a node is the
// product of its three per-axis weights.
//   restriction:  octant k = kx + 2 ky + 4 kz of coarse cell C gets the
//                 weighted mean of the moment-0 nodal values in that octant;
//   prolongation: the network's eight corner values per coarse cell are
//                 averaged over the cells that share each coarse vertex, and
//                 every fine node takes the trilinear interpolant of its
//                 coarse cell's vertex values at (s_x, s_y, s_z) / (f p).
// The prolonged correction is continuous across cells, the form the
// transport solution takes in the thick diffusive limit.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>

#include "learned.hpp"
#include "vec.hpp"

// Scratch of the allocation-free prolongation (coarse_prolong_add_noalloc):
// the (G+1)^3 vertex averages and the scale, for coarse grids up to
// kNoAllocMaxGrid.  Each device holds its own copy until the process ends;
// host code reaches it only through the address hipGetSymbolAddress gives
// per device (the array's name is a host-side stand-in).
namespace {
constexpr int kNoAllocMaxGrid = 64;
constexpr long kNoAllocScratchDoubles =
    long(kNoAllocMaxGrid + 1) * (kNoAllocMaxGrid + 1) * (kNoAllocMaxGrid + 1) + 1;
}  // namespace
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace ou {

namespace {

inline unsigned blocks(long n) { return unsigned((n + 255) / 256); }

struct Transfer {
  int N, p, n0, n1, f, G;
  double inv_w3;  // 1 / (per-axis octant weight)^3
};

__device__ __forceinline__ void axis_weights(int s, int fp, double *w) {
  w[0] = 2 * s < fp ? 1.0 : (2 * s == fp ? 0.5 : 0.0);
  w[1] = 2 * s > fp ? 1.0 : (2 * s == fp ? 0.5 : 0.0);
}

// coarse[C][k] = weighted mean of the moment-0 values of v in octant k of C
__global__ void k_restrict8(const double *v, Transfer t, double *coarse) {
  const long c = long(blockIdx.x) * 256 + threadIdx.x;
  const int G = t.G;
  if (c >= long(G) * G * G) return;
  const int X = int(c % G), Y = int((c / G) % G), Z = int(c / (long(G) * G));
  const int P1 = t.p + 1, fp = t.f * t.p;
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int kz = 0; kz < t.f; ++kz)
    for (int ky = 0; ky < t.f; ++ky)
      for (int kx = 0; kx < t.f; ++kx) {
        const long cell =
            (X * t.f + kx) + long(t.n0) * ((Y * t.f + ky) + long(t.n1) * (Z * t.f + kz));
        const double *pv = v + cell * t.N;
        for (int l = 0; l < t.N; ++l) {
          const int lx = l % P1, ly = (l / P1) % P1, lz = l / (P1 * P1);
          double wx[2], wy[2], wz[2];
          axis_weights(kx * t.p + lx, fp, wx);
          axis_weights(ky * t.p + ly, fp, wy);
          axis_weights(kz * t.p + lz, fp, wz);
          const double val = pv[l];
          for (int o = 0; o < 8; ++o) acc[o] += wx[o & 1] * wy[(o >> 1) & 1] * wz[o >> 2] * val;
        }
      }
  for (int o = 0; o < 8; ++o) coarse[c * 8 + o] = acc[o] * t.inv_w3;
}

// feat[C] = (coarse[C][0..8) / s, mat[C][0..3), 0),  s = sqrt(dot / (8 P0))
__global__ void k_precond_features(const double *coarse, const double *dot, const float *mat,
                                   long P0, float *feat) {
  const long c = long(blockIdx.x) * 256 + threadIdx.x;
  if (c >= P0) return;
  const double s = std::sqrt(dot[0] / double(8 * P0));
  const double inv = s > 0.0 ? 1.0 / s : 0.0;
  float *f = feat + c * kUnetFeatures;
  for (int o = 0; o < 8; ++o) f[o] = float(coarse[c * 8 + o] * inv);
  f[kUnetFeatTau] = mat[c * 3];
  f[kUnetFeatKappa] = mat[c * 3 + 1];
  f[kUnetFeatGain] = mat[c * 3 + 2];
  f[11] = 0.0f;
}

// vertex[V] = mean over the coarse cells sharing vertex V of their corner prediction
__global__ void k_corner_to_vertex(const float *pred, int G, double *vertex) {
  const long v = long(blockIdx.x) * 256 + threadIdx.x;
  const int G1 = G + 1;
  if (v >= long(G1) * G1 * G1) return;
  const int VX = int(v % G1), VY = int((v / G1) % G1), VZ = int(v / (long(G1) * G1));
  double s = 0.0;
  int cnt = 0;
  for (int o = 0; o < 8; ++o) {
    const int X = VX - 1 + (o & 1), Y = VY - 1 + ((o >> 1) & 1), Z = VZ - 1 + (o >> 2);
    if (X < 0 || Y < 0 || Z < 0 || X >= G || Y >= G || Z >= G) continue;
    // vertex V is corner (VX - X, VY - Y, VZ - Z) of cell (X, Y, Z)
    const int k = (VX - X) + 2 * (VY - Y) + 4 * (VZ - Z);
    s += double(pred[((long(Z) * G + Y) * G + X) * 8 + k]);
    ++cnt;
  }
  vertex[v] = s / cnt;
}

// z[cell N + l] += s * trilinear interpolant of the coarse vertex values
__global__ void k_prolong_vertex(double *z, const double *vertex, const double *dot, long P0,
                                 Transfer t, long n_fine) {
  const long i = long(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n_fine) return;
  const long cell = i / t.N;
  const int l = int(i % t.N), P1 = t.p + 1;
  const int x = int(cell % t.n0), y = int((cell / t.n0) % t.n1);
  const int zc = int(cell / (long(t.n0) * t.n1));
  const int X = x / t.f, Y = y / t.f, Z = zc / t.f;
  const double fp = double(t.f * t.p);
  const double tx = ((x - X * t.f) * t.p + l % P1) / fp;
  const double ty = ((y - Y * t.f) * t.p + (l / P1) % P1) / fp;
  const double tz = ((zc - Z * t.f) * t.p + l / (P1 * P1)) / fp;
  const int G1 = t.G + 1;
  double val = 0.0;
  for (int o = 0; o < 8; ++o) {
    const int ox = o & 1, oy = (o >> 1) & 1, oz = o >> 2;
    const double w = (ox ? tx : 1.0 - tx) * (oy ? ty : 1.0 - ty) * (oz ? tz : 1.0 - tz);
    val += w * vertex[(long(Z + oz) * G1 + (Y + oy)) * G1 + (X + ox)];
  }
  const double s = std::sqrt(dot[0] / double(8 * P0));
  z[i] += s * val;
}

double *dev_doubles(long n) {
  void *p = nullptr;
  if (hipMalloc(&p, size_t(n > 0 ? n : 1) * 8) != hipSuccess) {
    (void)hipGetLastError();
    fail(OU_ERR_NOMEM, "device allocation failed");
  }
  return static_cast<double *>(p);
}

float *dev_floats(long n) {
  void *p = nullptr;
  if (hipMalloc(&p, size_t(n > 0 ? n : 1) * 4) != hipSuccess) {
    (void)hipGetLastError();
    fail(OU_ERR_NOMEM, "device allocation failed");
  }
  return static_cast<float *>(p);
}

Transfer transfer_of(const Plan &pl, int G) {
  Transfer t;
  t.N = pl.N;
  t.p = pl.p;
  t.n0 = pl.n[0];
  t.n1 = pl.n[1];
  t.f = coarse_factor(pl, G);
  t.G = G;
  const int fp = t.f * t.p;
  double w = 0.0;
  for (int i = 0; i < t.f; ++i)
    for (int l = 0; l <= t.p; ++l) {
      const int s = i * t.p + l;
      w += 2 * s < fp ? 1.0 : (2 * s == fp ? 0.5 : 0.0);
    }
  t.inv_w3 = 1.0 / (w * w * w);
  return t;
}

}  // namespace

int coarse_factor(const Plan &pl, int G) {
  if (pl.dim != 3) fail(OU_ERR_UNSUPPORTED, "the learned preconditioner needs a 3D plan");
  if (G < 4 || G % 4) fail(OU_ERR_ARG, "coarse grid must be a positive multiple of 4");
  if (pl.n[0] != pl.n[1] || pl.n[1] != pl.n[2] || pl.n[0] % G)
    fail(OU_ERR_ARG, "the mesh must be n^3 with n a multiple of the coarse grid");
  return pl.n[0] / G;
}

std::vector<float> coarse_material_features(const Plan &pl, int G) {
  const int f = coarse_factor(pl, G);
  const double H = f * std::cbrt(pl.h[0] * pl.h[1] * pl.h[2]);
  const long P0 = long(G) * G * G;
  std::vector<float> out(size_t(P0) * 3);
  for (long C = 0; C < P0; ++C) {
    const int X = int(C % G), Y = int((C / G) % G), Z = int(C / (long(G) * G));
    double st = 0.0, ss = 0.0;
    for (int kz = 0; kz < f; ++kz)
      for (int ky = 0; ky < f; ++ky)
        for (int kx = 0; kx < f; ++kx) {
          const long cell =
              (X * f + kx) + long(pl.n[0]) * ((Y * f + ky) + long(pl.n[1]) * (Z * f + kz));
          const int m = pl.mat[size_t(cell)];
          st += pl.sigma_t[size_t(m)];
          ss += pl.sigma_s[size_t(m) * (pl.L + 1)];
        }
    const double sig = st / (double(f) * f * f);
    const double c = st > 0.0 ? ss / st : 0.0;
    const double tau = sig * H;
    const double theta = tau * tau / (1.0 + tau * tau);
    const double g = c * theta / std::max(1.0 - c * theta, 1e-6);
    out[size_t(C) * 3] = float(std::log(std::max(tau, 1e-30)) / 4.0);
    out[size_t(C) * 3 + 1] = float((std::log(std::max(1.0 - c, 1e-6)) + 4.6) / 4.6);
    out[size_t(C) * 3 + 2] = float(g);
  }
  return out;
}

void coarse_restrict(const Plan &pl, int G, const double *d_v, double *d_coarse, hipStream_t s) {
  const Transfer t = transfer_of(pl, G);
  k_restrict8<<<blocks(long(G) * G * G), 256, 0, s>>>(d_v, t, d_coarse);
  OU_HIP(hipGetLastError());
}

void coarse_prolong_add(const Plan &pl, int G, const float *d_corners, const double *d_dot,
                        double *d_vertex, double *d_z, hipStream_t s) {
  const Transfer t = transfer_of(pl, G);
  const long V = long(G + 1) * (G + 1) * (G + 1);
  k_corner_to_vertex<<<blocks(V), 256, 0, s>>>(d_corners, G, d_vertex);
  k_prolong_vertex<<<blocks(pl.n_dofs), 256, 0, s>>>(d_z, d_vertex, d_dot, long(G) * G * G, t,
                                                     pl.n_dofs);
  OU_HIP(hipGetLastError());
}

namespace {

__global__ void k_set_scalar(double *dst, double v) { *dst = v; }

// The device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
double *noalloc_scratch() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double *addr[kMaxDevices] = {};
  int dev = 0;
  OU_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) fail(OU_ERR_UNSUPPORTED, "device ordinal above 63");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      fail(OU_ERR_HIP, std::string("hipGetSymbolAddress of the prolongation scratch failed: ") +
                           hipGetErrorString(e));
    }
    addr[dev] = static_cast<double *>(p);
  }
  return addr[dev];
}

}  // namespace

int coarse_noalloc_max_grid() { return kNoAllocMaxGrid; }

void coarse_prolong_add_noalloc(const Plan &pl, int G, const float *d_corners, double *d_z,
                                hipStream_t s) {
  if (G > kNoAllocMaxGrid)
    fail(OU_ERR_SIZE, "coarse grid " + std::to_string(G) +
                          " is above the allocation-free prolongation's maximum of " +
                          std::to_string(kNoAllocMaxGrid));
  const Transfer t = transfer_of(pl, G);
  double *scr = noalloc_scratch();
  const long V = long(G + 1) * (G + 1) * (G + 1);
  const long P0 = long(G) * G * G;
  double *d_vertex = scr, *d_dot = scr + V;
  // scale s = sqrt(dot / (8 G^3)) = 1, as the allocating entry point sets it;
  // every vertex value is written by k_corner_to_vertex before it is read
  k_set_scalar<<<1, 1, 0, s>>>(d_dot, 8.0 * double(P0));
  k_corner_to_vertex<<<blocks(V), 256, 0, s>>>(d_corners, G, d_vertex);
  k_prolong_vertex<<<blocks(pl.n_dofs), 256, 0, s>>>(d_z, d_vertex, d_dot, P0, t, pl.n_dofs);
  OU_HIP(hipGetLastError());
}

LearnedPreconditioner::LearnedPreconditioner(RadiativeTransfer &rt, UNet &net)
    : pl_(rt.plan()), net_(net), G_(net.grid()) {
  f_ = coarse_factor(pl_, G_);
  P0_ = long(G_) * G_ * G_;
  struct Undo {
    LearnedPreconditioner *p;
    bool done = false;
    ~Undo() {
      if (!done) p->release();
    }
  } undo{this};
  const long V = long(G_ + 1) * (G_ + 1) * (G_ + 1);
  coarse_ = dev_doubles(P0_ * 8);
  vertex_ = dev_doubles(V);
  dot_ = dev_doubles(1);
  scratch_ = dev_doubles(reduce_scratch_doubles(1));
  mat_ = dev_floats(P0_ * 3);
  feat_ = dev_floats(P0_ * kUnetFeatures);
  pred_ = dev_floats(P0_ * kUnetOut);
  const std::vector<float> m = coarse_material_features(pl_, G_);
  OU_HIP(hipMemcpy(mat_, m.data(), m.size() * 4, hipMemcpyHostToDevice));
  undo.done = true;
}

void LearnedPreconditioner::release() {
  for (void *p : {static_cast<void *>(coarse_), static_cast<void *>(vertex_),
                  static_cast<void *>(dot_), static_cast<void *>(scratch_),
                  static_cast<void *>(mat_), static_cast<void *>(feat_),
                  static_cast<void *>(pred_)})
    if (p) (void)hipFree(p);
  coarse_ = vertex_ = dot_ = scratch_ = nullptr;
  mat_ = feat_ = pred_ = nullptr;
}

LearnedPreconditioner::~LearnedPreconditioner() { release(); }

void LearnedPreconditioner::apply(const double *d_r, double *d_z, hipStream_t s) {
  const Transfer t = transfer_of(pl_, G_);
  vec_copy(pl_.vec_size, d_r, d_z, s);
  k_restrict8<<<blocks(P0_), 256, 0, s>>>(d_r, t, coarse_);
  vec_multi_dot(P0_ * 8, 1, coarse_, 0, coarse_, scratch_, dot_, s);
  k_precond_features<<<blocks(P0_), 256, 0, s>>>(coarse_, dot_, mat_, P0_, feat_);
  net_.predict(feat_, 1, pred_, s);
  coarse_prolong_add(pl_, G_, pred_, dot_, vertex_, d_z, s);
  ++applications_;
}

RefinedResult solve_refined(RadiativeTransfer &rt64, RadiativeTransfer &rt32,
                            Preconditioner *prec, double *d_phi, double tol, int max_it,
                            int restart, double inner_tol, bool verbose) {
  if (rt64.DomainSize() != rt32.DomainSize())
    fail(OU_ERR_SIZE, "FP64 and FP32 plans describe different problems");
  if (rt64.plan().precision != OU_FP64 || rt32.plan().precision != OU_FP32)
    fail(OU_ERR_ARG, "refinement needs an FP64 plan and an FP32 plan");
  // Each plan's work runs on its own stream: the FP64 residuals on the FP64
  // plan's, the inner FP32 solve on the FP32 plan's.  A plan's sweep forks
  // its octant halves onto the plan's own side streams, which were created
  // with it; another plan's stream may share a hardware queue with one of
  // them, which would serialise the halves.  The host synchronises between
  // the phases (the residual norm, and the inner solve's end).
  hipStream_t s = rt64.plan().stream, s32 = rt32.plan().stream;
  const long n = rt64.DomainSize();
  double *b = dev_doubles(n), *r = nullptr, *d = nullptr, *scr = nullptr, *dt = nullptr;
  struct Free {
    double **p[5];
    ~Free() {
      for (double **q : p)
        if (*q) (void)hipFree(*q);
    }
  } guard{{&b, &r, &d, &scr, &dt}};
  r = dev_doubles(n);
  d = dev_doubles(n);
  scr = dev_doubles(reduce_scratch_doubles(1));
  dt = dev_doubles(1);
  auto norm = [&](const double *v) {
    double h = 0.0;
    vec_multi_dot(n, 1, v, 0, v, scr, dt, s);
    OU_HIP(hipMemcpyAsync(&h, dt, 8, hipMemcpyDeviceToHost, s));
    OU_HIP(hipStreamSynchronize(s));
    return std::sqrt(h);
  };
  RefinedResult res;
  build_rhs(rt64, b, s);
  const double bnorm = norm(b);
  if (!std::isfinite(bnorm)) fail(OU_ERR_NONFINITE, "non-finite right-hand side norm");
  if (bnorm == 0.0) {
    vec_fill(n, 0.0, d_phi, s);
    res.total.converged = true;
    OU_HIP(hipStreamSynchronize(s));
    return res;
  }
  double fp64_secs = 0.0;
  // One Krylov workspace for every inner solve.  The inner solve keeps its
  // basis in FP32 (its tolerance is far above FP32 rounding; every pass over
  // the basis then reads half the bytes): 33 GB at the headline size with
  // restart 16 and a preconditioner.
  if (restart > 31) fail(OU_ERR_ARG, "restart length must be <= 31");
  GmresWorkspace ws(n, restart > 0 ? restart : 30, prec != nullptr, true);
  // A zero initial guess has residual b: the first FP64 application is skipped.
  bool zero_guess = norm(d_phi) == 0.0;
  for (int cycle = 1;; ++cycle) {
    const auto a0 = std::chrono::steady_clock::now();
    if (zero_guess) {
      vec_copy(n, b, r, s);
      zero_guess = false;
    } else {
      rt64.Apply(d_phi, r, s);
      vec_lincomb(n, 1.0, b, -1.0, r, r, s);
      ++res.outer;
    }
    const double rn = norm(r);
    if (!std::isfinite(rn))
      fail(OU_ERR_NONFINITE, "non-finite FP64 residual norm in refinement cycle " +
                                 std::to_string(cycle));
    fp64_secs += std::chrono::duration<double>(std::chrono::steady_clock::now() - a0).count();
    res.total.rel_residual = rn / bnorm;
    if (verbose)
      progress("refinement cycle %d: FP64 residual %.3e after %d FP32 applications\n", cycle,
               res.total.rel_residual, res.total.iterations);
    if (res.total.rel_residual < tol) {
      res.total.converged = true;
      break;
    }
    if (res.total.iterations >= max_it) break;
    vec_fill(n, 0.0, d, s32);
    // An inner solve that stagnates above inner_tol (the FP32 operator's own
    // accuracy can stop it) hands over to a new FP64 residual after four
    // restart cycles.  (r is complete: norm() synchronised its stream.)
    const int inner_max = std::min(max_it - res.total.iterations, 4 * std::max(restart, 1));
    const SolveResult in = solve_gmres_f32_basis(rt32, r, d, restart, inner_tol, inner_max, prec,
                                                 verbose, s32, &ws);
    OU_HIP(hipStreamSynchronize(s32));
    res.total.iterations += in.iterations;
    vec_axpby(n, 1.0, d, 1.0, d_phi, s);
    if (in.iterations == 0) break;
  }
  OU_HIP(hipStreamSynchronize(s));
  res.seconds_fp64 = fp64_secs;
  return res;
}

}  // namespace ou
