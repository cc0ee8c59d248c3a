// Solver drivers around the transport operator (src/ouranos.cc:43-209):
// source iteration, restarted GMRES, BiCGStab and flexible GMRES.  Vectors
// live in HBM; the Arnoldi process, every inner product and the residual test
// are FP64 (one GMRES form stores its Krylov basis in FP32, for inner solves).
// Small dense work (Hessenberg, Givens rotations) is on the host.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <memory>
#include <vector>

#include "radiative_transfer.hpp"
#include "vec.hpp"

namespace ou {

namespace {

struct DevBuf {
  double *p = nullptr;
  explicit DevBuf(size_t n) {
    if (hipMalloc(&p, (n ? n : 1) * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      fail(OU_ERR_NOMEM, "solver workspace allocation failed");
    }
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

// Dot products <V_i, w>, i < m, returned on the host.
struct Dotter {
  DevBuf scratch, out;
  std::vector<double> host;
  explicit Dotter(int max_m) : scratch(size_t(reduce_scratch_doubles(max_m))), out(size_t(max_m)), host(max_m) {}
  const std::vector<double> &dots(long n, int m, const double *V, long ldv, const double *w,
                                  hipStream_t s) {
    vec_multi_dot(n, m, V, ldv, w, scratch.p, out.p, s);
    return fetch(m, s);
  }
  // w -= V h (h on the device); returns <V_i, w> (i < m, when `dots`) then <w, w>.
  const std::vector<double> &axpy_dots(long n, int m, const double *V, long ldv,
                                       const double *h, double *w, bool dots, hipStream_t s) {
    vec_axpy_dots(n, m, V, ldv, h, w, dots, scratch.p, out.p, s);
    return fetch(dots ? m + 1 : 1, s);
  }
  // The FP32-basis pass: [w -= V h,] then <V_i, w> (i < m, when `dots`) and <w, w>.
  const std::vector<double> &gs_f32(long n, int m, const float *V, long ldv, const double *h,
                                    double *w, bool update, bool dots, hipStream_t s) {
    vec_gs_f32(n, m, V, ldv, h, w, update, dots, scratch.p, out.p, s);
    return fetch(dots ? m + 1 : 1, s);
  }
  const std::vector<double> &fetch(int m, hipStream_t s) {
    OU_HIP(hipMemcpyAsync(host.data(), out.p, size_t(m) * 8, hipMemcpyDeviceToHost, s));
    OU_HIP(hipStreamSynchronize(s));
    return host;
  }
  double dot(long n, const double *a, const double *b, hipStream_t s) {
    return dots(n, 1, a, 0, b, s)[0];
  }
  double norm(long n, const double *a, hipStream_t s) { return std::sqrt(dot(n, a, a, s)); }
};

// A value the device produced that is not finite ends the solve at once: no
// later iteration can recover from it, and the residual tests (x < tol) are
// false for NaN, so the loop would otherwise run to its iteration limit.
void check_finite(double v, const char *what, int it) {
  if (!std::isfinite(v)) {
    char buf[160];
    std::snprintf(buf, sizeof buf, "non-finite %s (%g) at iteration %d of the solve", what, v, it);
    fail(OU_ERR_NONFINITE, buf);
  }
}

// The small least-squares problem of restarted GMRES on the host: the
// Hessenberg columns, reduced by Givens rotations as they arrive, and the
// rotated right-hand side g (|g[j+1]| is the residual norm after column j).
struct Hessenberg {
  int m;
  std::vector<double> H, cs, sn, g, y;
  explicit Hessenberg(int m_) : m(m_), H(size_t(m_ + 1) * m_, 0.0), cs(m_), sn(m_), g(m_ + 1), y(m_) {}
  void start(double beta) {
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
  }
  // Column j: h[0..j] and the subdiagonal hn; returns |g[j+1]|.
  double add_column(int j, const double *h, double hn) {
    for (int i = 0; i <= j; ++i) H[size_t(i) * m + j] = h[i];
    H[size_t(j + 1) * m + j] = hn;
    for (int i = 0; i < j; ++i) {
      const double a = H[size_t(i) * m + j], b = H[size_t(i + 1) * m + j];
      H[size_t(i) * m + j] = cs[i] * a + sn[i] * b;
      H[size_t(i + 1) * m + j] = -sn[i] * a + cs[i] * b;
    }
    const double a = H[size_t(j) * m + j], b = H[size_t(j + 1) * m + j];
    const double r = std::hypot(a, b);
    cs[j] = a / r;
    sn[j] = b / r;
    H[size_t(j) * m + j] = r;
    H[size_t(j + 1) * m + j] = 0.0;
    g[j + 1] = -sn[j] * g[j];
    g[j] = cs[j] * g[j];
    return std::fabs(g[j + 1]);
  }
  // y solving the leading j x j triangle against g.
  const std::vector<double> &solve(int j) {
    for (int i = j - 1; i >= 0; --i) {
      double acc = g[i];
      for (int k = i + 1; k < j; ++k) acc -= H[size_t(i) * m + k] * y[k];
      y[i] = acc / H[size_t(i) * m + i];
    }
    return y;
  }
};

}  // namespace

void build_rhs(RadiativeTransfer &rt, double *d_rhs, hipStream_t s) {
  const long n = rt.DomainSize();
  rt.clear_scattering_source(s);
  vec_fill(n, 0.0, d_rhs, s);
  rt.sweep(d_rhs, true, s);
}

SolveResult solve_si(RadiativeTransfer &rt, double *d_phi, double tol, int max_it,
                     bool verbose) {
  hipStream_t s = rt.plan().stream;
  const long n = rt.DomainSize();
  DevBuf nw(n);
  Dotter dt(1);
  SolveResult res;
  for (int it = 0; it < max_it; ++it) {
    rt.compute_scattering_source(d_phi, s);
    vec_fill(n, 0.0, nw.p, s);
    rt.sweep(nw.p, true, s);
    vec_axpby(n, 1.0, nw.p, -1.0, d_phi, s);  // d_phi <- new - old
    const double num = dt.norm(n, d_phi, s);
    const double den = dt.norm(n, nw.p, s);
    check_finite(num + den, "change norm", it);
    vec_copy(n, nw.p, d_phi, s);
    res.iterations = it + 1;
    res.rel_residual = den > 0.0 ? num / den : 0.0;
    if (verbose) progress("SI iteration %d: %.3e\n", it, res.rel_residual);
    if (res.rel_residual < tol) {
      res.converged = true;
      break;
    }
  }
  OU_HIP(hipStreamSynchronize(s));
  return res;
}

SolveResult solve_si_rhs(Operator &op, const double *d_b, double *d_x, double tol, int max_it,
                         bool verbose, hipStream_t s) {
  const long n = op.DomainSize();
  DevBuf ax(n), nw(n);
  Dotter dt(1);
  SolveResult res;
  for (int it = 0; it < max_it; ++it) {
    op.Apply(d_x, ax.p, s);
    vec_lincomb(n, 1.0, d_x, -1.0, ax.p, nw.p, s);  // x - A x = D L^-1 M Sigma x
    vec_axpby(n, 1.0, d_b, 1.0, nw.p, s);           // + b
    vec_axpby(n, 1.0, nw.p, -1.0, d_x, s);          // d_x <- new - old
    const double num = dt.norm(n, d_x, s);
    const double den = dt.norm(n, nw.p, s);
    check_finite(num + den, "change norm", it);
    vec_copy(n, nw.p, d_x, s);
    res.iterations = it + 1;
    res.rel_residual = den > 0.0 ? num / den : 0.0;
    if (verbose) progress("SI iteration %d: %.3e\n", it, res.rel_residual);
    if (res.rel_residual < tol) {
      res.converged = true;
      break;
    }
  }
  OU_HIP(hipStreamSynchronize(s));
  return res;
}

GmresWorkspace::GmresWorkspace(long n, int restart, bool flexible, bool f32_basis)
    : n_(n), restart_(restart), f32_(f32_basis) {
  const size_t es = f32_basis ? sizeof(float) : sizeof(double);
  const size_t nv = size_t(n) * (restart + 1), nz = flexible ? size_t(n) * restart : 0;
  if (hipMalloc(&V_, nv * es) != hipSuccess || (nz && hipMalloc(&Z_, nz * es) != hipSuccess) ||
      (f32_basis && hipMalloc(&work_, 3 * size_t(n) * sizeof(double)) != hipSuccess)) {
    (void)hipGetLastError();
    if (V_) (void)hipFree(V_);
    if (Z_) (void)hipFree(Z_);
    V_ = Z_ = nullptr;
    fail(OU_ERR_NOMEM, "solver workspace allocation failed");
  }
}

GmresWorkspace::~GmresWorkspace() {
  if (V_) (void)hipFree(V_);
  if (Z_) (void)hipFree(Z_);
  if (work_) (void)hipFree(work_);
}

SolveResult solve_gmres(Operator &op, const double *d_b, double *d_x, int restart,
                        double tol, int max_it, Preconditioner *prec, bool verbose,
                        hipStream_t s, ArnoldiObserver *observer, GmresWorkspace *ws) {
  const long n = op.DomainSize();
  const int m = restart > 0 ? restart : 30;
  if (m > 31) fail(OU_ERR_ARG, "restart length must be <= 31");
  // the caller's workspace when it fits, else one for this solve
  std::unique_ptr<GmresWorkspace> own;
  if (!ws || !ws->fits(n, m, prec != nullptr)) {
    own.reset(new GmresWorkspace(n, m, prec != nullptr));
    ws = own.get();
  }
  DevBuf hd(size_t(m + 1));
  struct Basis {
    double *p;
  };
  const Basis V{ws->V()}, Zb{ws->Z()};
  const Basis *Z = prec ? &Zb : nullptr;
  Dotter dt(m + 1);
  int reorth = 0;  // Arnoldi steps that needed the second Gram-Schmidt pass
  Hessenberg hs(m);
  std::vector<double> h(m + 1);
  SolveResult res;
  const double bnorm = dt.norm(n, d_b, s);
  check_finite(bnorm, "right-hand side norm", 0);
  if (bnorm == 0.0) {
    vec_fill(n, 0.0, d_x, s);
    res.converged = true;
    OU_HIP(hipStreamSynchronize(s));
    return res;
  }
  int total = 0;
  // A zero initial guess needs no operator application for its residual.
  bool x_is_zero = dt.norm(n, d_x, s) == 0.0;
  while (true) {
    // r = b - A x
    if (x_is_zero) {
      vec_copy(n, d_b, V.p, s);
      x_is_zero = false;
    } else {
      op.Apply(d_x, V.p, s);
      vec_lincomb(n, 1.0, d_b, -1.0, V.p, V.p, s);
    }
    const double beta = dt.norm(n, V.p, s);
    check_finite(beta, "residual norm", total);
    res.rel_residual = beta / bnorm;
    if (verbose)
      progress("GMRES restart at %d: true residual %.3e (%d reorthogonalised steps)\n", total,
               res.rel_residual, reorth);
    if (res.rel_residual < tol) {
      res.converged = true;
      break;
    }
    if (total >= max_it) break;
    vec_scale(n, 1.0 / beta, V.p, s);
    hs.start(beta);
    int j = 0;
    for (; j < m && total < max_it; ++j) {
      double *vj = V.p + size_t(j) * n;
      // The new Arnoldi vector is built in place in the next basis slot.
      double *wj = V.p + size_t(j + 1) * n;
      if (prec) {
        double *zj = Z->p + size_t(j) * n;
        prec->apply(vj, zj, s);
        op.Apply(zj, wj, s);
        if (observer) observer->observe(total, zj, wj, s);
      } else {
        op.Apply(vj, wj, s);
        if (observer) observer->observe(total, vj, wj, s);
      }
      ++total;
      // Classical Gram-Schmidt in FP64 with the DGKS test: a second pass only
      // when the norm fell below 1/sqrt(2) of |w| (Daniel, Gragg, Kaufman and
      // Stewart, Math. Comp. 30 (1976)).  Three sweeps over the basis: the
      // first dot products (with |w|^2), the first projection fused with the
      // second pass's dot products and the new norm, and the second
      // projection fused with the final norm.
      double hn = 0.0;
      {
        const std::vector<double> &hh = dt.dots(n, j + 2, V.p, n, wj, s);
        const double wnorm = std::sqrt(std::max(hh[j + 1], 0.0));
        for (int i = 0; i <= j; ++i) h[i] = hh[i];
        OU_HIP(hipMemcpy(hd.p, hh.data(), size_t(j + 1) * 8, hipMemcpyHostToDevice));
        const std::vector<double> &h2 = dt.axpy_dots(n, j + 1, V.p, n, hd.p, wj, true, s);
        hn = std::sqrt(std::max(h2[j + 1], 0.0));
        if (hn < 0.7071067811865476 * wnorm) {
          for (int i = 0; i <= j; ++i) h[i] += h2[i];
          OU_HIP(hipMemcpy(hd.p, h2.data(), size_t(j + 1) * 8, hipMemcpyHostToDevice));
          const std::vector<double> &h3 = dt.axpy_dots(n, j + 1, V.p, n, hd.p, wj, false, s);
          hn = std::sqrt(std::max(h3[0], 0.0));
          ++reorth;
        }
      }
      {
        double hsum = hn;
        for (int i = 0; i <= j; ++i) hsum += std::fabs(h[i]);
        check_finite(hsum, "Arnoldi coefficient", total);
      }
      if (hn > 0.0) vec_scale(n, 1.0 / hn, wj, s);
      res.rel_residual = hs.add_column(j, h.data(), hn) / bnorm;
      if (verbose) progress("  it %d: %.3e\n", total, res.rel_residual);
      if (res.rel_residual < tol || hn == 0.0) {
        ++j;
        break;
      }
    }
    const std::vector<double> &y = hs.solve(j);
    // x += V y  (or Z y): multi_axpy subtracts, so pass -y.
    for (int i = 0; i < j; ++i) h[i] = -y[i];
    OU_HIP(hipStreamSynchronize(s));
    OU_HIP(hipMemcpy(hd.p, h.data(), size_t(j) * 8, hipMemcpyHostToDevice));
    vec_multi_axpy(n, j, prec ? Z->p : V.p, n, hd.p, 1.0, d_x, s);
    OU_HIP(hipStreamSynchronize(s));
  }
  res.iterations = total;
  OU_HIP(hipStreamSynchronize(s));
  return res;
}

SolveResult solve_gmres_f32_basis(Operator &op, const double *d_b, double *d_x, int restart,
                                  double tol, int max_it, Preconditioner *prec, bool verbose,
                                  hipStream_t s, GmresWorkspace *ws) {
  const long n = op.DomainSize();
  const int m = restart > 0 ? restart : 30;
  if (m > 31) fail(OU_ERR_ARG, "restart length must be <= 31");
  std::unique_ptr<GmresWorkspace> own;
  if (!ws || !ws->fits(n, m, prec != nullptr, true)) {
    own.reset(new GmresWorkspace(n, m, prec != nullptr, true));
    ws = own.get();
  }
  // FP32 basis V (and Z); FP64 work vectors: v (the current Arnoldi vector
  // as the operator or preconditioner reads it), w (the vector being
  // orthogonalised) and z (the preconditioned vector).
  float *V = ws->Vf(), *Z = ws->Zf();
  double *v = ws->work(), *w = v + n, *z = v + 2 * n;
  DevBuf hd(size_t(m + 1));
  Dotter dt(m + 1);
  int reorth = 0;
  Hessenberg hs(m);
  std::vector<double> h(m + 1);
  SolveResult res;
  const double bnorm = dt.norm(n, d_b, s);
  check_finite(bnorm, "right-hand side norm", 0);
  if (bnorm == 0.0) {
    vec_fill(n, 0.0, d_x, s);
    res.converged = true;
    OU_HIP(hipStreamSynchronize(s));
    return res;
  }
  int total = 0;
  bool x_is_zero = dt.norm(n, d_x, s) == 0.0;
  while (true) {
    // r = b - A x, in w
    if (x_is_zero) {
      vec_copy(n, d_b, w, s);
      x_is_zero = false;
    } else {
      op.Apply(d_x, w, s);
      vec_lincomb(n, 1.0, d_b, -1.0, w, w, s);
    }
    const double beta = dt.norm(n, w, s);
    check_finite(beta, "residual norm", total);
    res.rel_residual = beta / bnorm;
    if (verbose)
      progress("GMRES (FP32 basis) restart at %d: true residual %.3e (%d reorthogonalised "
               "steps)\n",
               total, res.rel_residual, reorth);
    if (res.rel_residual < tol) {
      res.converged = true;
      break;
    }
    if (total >= max_it) break;
    vec_scale_store_f32(n, 1.0 / beta, w, true, V, s);
    std::swap(v, w);
    hs.start(beta);
    int j = 0;
    for (; j < m && total < max_it; ++j) {
      if (prec) {
        prec->apply(v, z, s);
        op.Apply(z, w, s);
        vec_scale_store_f32(n, 1.0, z, false, Z + size_t(j) * n, s);
      } else {
        op.Apply(v, w, s);
      }
      ++total;
      // Classical Gram-Schmidt with the DGKS test, as in solve_gmres: the
      // first pass's dot products with |w|^2, the projection fused with the
      // second pass's dot products, and (when needed) the second projection.
      double hn = 0.0;
      {
        const std::vector<double> &hh = dt.gs_f32(n, j + 1, V, n, nullptr, w, false, true, s);
        const double wnorm = std::sqrt(std::max(hh[j + 1], 0.0));
        for (int i = 0; i <= j; ++i) h[i] = hh[i];
        OU_HIP(hipMemcpy(hd.p, hh.data(), size_t(j + 1) * 8, hipMemcpyHostToDevice));
        const std::vector<double> &h2 = dt.gs_f32(n, j + 1, V, n, hd.p, w, true, true, s);
        hn = std::sqrt(std::max(h2[j + 1], 0.0));
        if (hn < 0.7071067811865476 * wnorm) {
          for (int i = 0; i <= j; ++i) h[i] += h2[i];
          OU_HIP(hipMemcpy(hd.p, h2.data(), size_t(j + 1) * 8, hipMemcpyHostToDevice));
          const std::vector<double> &h3 = dt.gs_f32(n, j + 1, V, n, hd.p, w, true, false, s);
          hn = std::sqrt(std::max(h3[0], 0.0));
          ++reorth;
        }
      }
      {
        double hsum = hn;
        for (int i = 0; i <= j; ++i) hsum += std::fabs(h[i]);
        check_finite(hsum, "Arnoldi coefficient", total);
      }
      res.rel_residual = hs.add_column(j, h.data(), hn) / bnorm;
      if (hn > 0.0) {
        // the next Arnoldi vector: FP64 in v for the next application, FP32 in V
        vec_scale_store_f32(n, 1.0 / hn, w, true, V + size_t(j + 1) * n, s);
        std::swap(v, w);
      }
      if (verbose) progress("  it %d: %.3e\n", total, res.rel_residual);
      if (res.rel_residual < tol || hn == 0.0) {
        ++j;
        break;
      }
    }
    const std::vector<double> &y = hs.solve(j);
    OU_HIP(hipStreamSynchronize(s));
    OU_HIP(hipMemcpy(hd.p, y.data(), size_t(j) * 8, hipMemcpyHostToDevice));
    vec_multi_axpy_f32(n, j, prec ? Z : V, n, hd.p, 1.0, d_x, s);  // x += V y (or Z y)
    OU_HIP(hipStreamSynchronize(s));
  }
  res.iterations = total;
  OU_HIP(hipStreamSynchronize(s));
  return res;
}

SolveResult solve_bicgstab(Operator &op, const double *d_b, double *d_x, double tol,
                           int max_it, bool verbose, hipStream_t s) {
  const long n = op.DomainSize();
  DevBuf r(n), r0(n), p(n), v(n), sv(n), t(n);
  Dotter dt(2);
  SolveResult res;
  const double bnorm = dt.norm(n, d_b, s);
  if (bnorm == 0.0) {
    vec_fill(n, 0.0, d_x, s);
    res.converged = true;
    OU_HIP(hipStreamSynchronize(s));
    return res;
  }
  op.Apply(d_x, r.p, s);
  vec_lincomb(n, 1.0, d_b, -1.0, r.p, r.p, s);
  vec_copy(n, r.p, r0.p, s);
  vec_fill(n, 0.0, p.p, s);
  vec_fill(n, 0.0, v.p, s);
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  check_finite(bnorm, "right-hand side norm", 0);
  res.rel_residual = dt.norm(n, r.p, s) / bnorm;
  check_finite(res.rel_residual, "residual norm", 0);
  int it = 0;
  while (res.rel_residual >= tol && it < max_it) {
    ++it;
    const double rho1 = dt.dot(n, r0.p, r.p, s);
    check_finite(rho1, "rho", it);
    const double beta = (rho1 / rho) * (alpha / omega);
    rho = rho1;
    // p = r + beta (p - omega v)
    vec_axpby(n, -omega, v.p, 1.0, p.p, s);
    vec_axpby(n, 1.0, r.p, beta, p.p, s);
    op.Apply(p.p, v.p, s);
    alpha = rho / dt.dot(n, r0.p, v.p, s);
    vec_lincomb(n, 1.0, r.p, -alpha, v.p, sv.p, s);
    const double snorm = dt.norm(n, sv.p, s);
    check_finite(alpha + snorm, "alpha or |s|", it);
    if (snorm / bnorm < tol) {
      vec_axpby(n, alpha, p.p, 1.0, d_x, s);
      res.rel_residual = snorm / bnorm;
      break;
    }
    op.Apply(sv.p, t.p, s);
    const double ts = dt.dot(n, t.p, sv.p, s), tt = dt.dot(n, t.p, t.p, s);
    omega = ts / tt;
    check_finite(omega, "omega", it);
    vec_axpby(n, alpha, p.p, 1.0, d_x, s);
    vec_axpby(n, omega, sv.p, 1.0, d_x, s);
    vec_lincomb(n, 1.0, sv.p, -omega, t.p, r.p, s);
    res.rel_residual = dt.norm(n, r.p, s) / bnorm;
    check_finite(res.rel_residual, "residual norm", it);
    if (verbose) progress("BiCGStab it %d: %.3e\n", it, res.rel_residual);
  }
  res.iterations = it;
  res.converged = res.rel_residual < tol;
  OU_HIP(hipStreamSynchronize(s));
  return res;
}

}  // namespace ou
