/*
 * ou_oracle.c -- FP64 CPU restatement of the Ouranos DG-Sn transport sweep.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 * See ou_oracle.h for the routine-by-routine map onto the reference sources.
 */
#include "ou_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ========================================================================= */
/* Quadrature                                                                 */
/* ========================================================================= */

/* Gauss-Legendre nodes/weights on [-1,1], ascending.  Newton on P_n from the
 * Chebyshev-like guess cos(pi (i+3/4)/(n+1/2)); the reference ships these as
 * tables (GLC.cc:60-1286), which the tests use to pin this routine. */
void orc_gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
      }
      if (n == 1) { p1 = z; p0 = 1.0; }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-17) break;
    }
    /* recompute derivative at converged z */
    {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
      }
      if (n == 1) { p1 = z; p0 = 1.0; }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
    }
    double wi = 2.0 / ((1.0 - z * z) * dp * dp);
    x[n - 1 - i] = z;
    x[i] = -z;
    w[i] = wi;
    w[n - 1 - i] = wi;
  }
  if (n % 2 == 1) x[n / 2] = 0.0;
}

/* LS::compute_omega (LS.cc:667-682): triangular arrangement of the octant. */
static void ls_compute_omega(const double *dir, int size, double *omega) {
  int off = 0;
  for (int j = 0; j < size; ++j) {
    for (int i = 0; i < size - j; ++i) {
      omega[3 * (off + i) + 0] = dir[size - 1 - j - i];
      omega[3 * (off + i) + 1] = dir[i];
      omega[3 * (off + i) + 2] = dir[j];
    }
    off += size - j;
  }
}

/* LS::build_octant (LS.cc:14-169).  Level-symmetric octant tables S2..S12
 * (directions and weight classes as printed in the reference). Returns the
 * number of directions per octant, or -1 for an unsupported order. */
int orc_ls_octant(int sn, double *omega, double *w) {
  switch (sn) {
  case 2: {
    const double d = 0.577350269189625764509149;
    omega[0] = omega[1] = omega[2] = d;
    w[0] = 1.0;
    return 1;
  }
  case 4: {
    const double dir[2] = {0.350021174581540677777041,
                           0.868890300722201205229788};
    ls_compute_omega(dir, 2, omega);
    w[0] = w[1] = w[2] = 1.0 / 3.0;
    return 3;
  }
  case 6: {
    const double w1 = 0.176126130863383433783565, w2 = 0.157207202469949899549768;
    const double dir[3] = {0.266635401516704720331535, 0.681507726536546927403750,
                           0.926180935517489107558380};
    ls_compute_omega(dir, 3, omega);
    const double ww[6] = {w1, w2, w1, w2, w2, w1};
    memcpy(w, ww, sizeof ww);
    return 6;
  }
  case 8: {
    const double w1 = 0.120987654320987654320988, w2 = 0.0907407407407407407407407,
                 w3 = 0.0925925925925925925925926;
    const double dir[4] = {0.218217890235992381266097, 0.577350269189625764509149,
                           0.786795792469443145800830, 0.951189731211341853132399};
    ls_compute_omega(dir, 4, omega);
    const double ww[10] = {w1, w2, w2, w1, w2, w3, w2, w2, w2, w1};
    memcpy(w, ww, sizeof ww);
    return 10;
  }
  case 10: {
    const double w1 = 0.0893031479843567214704325, w2 = 0.0725291517123655242296233,
                 w3 = 0.0450437674364086390490892, w4 = 0.0539281144878369243545650;
    const double dir[5] = {0.189321326478010476671494, 0.508881755582618974382711,
                           0.694318887594384317279217, 0.839759962236684758403029,
                           0.963490981110468484701598};
    ls_compute_omega(dir, 5, omega);
    const double ww[15] = {w1, w2, w3, w2, w1, w2, w4, w4, w2, w3, w4, w3, w2, w2, w1};
    memcpy(w, ww, sizeof ww);
    return 15;
  }
  case 12: {
    const double w1 = 0.0707625899700910439766549, w2 = 0.0558811015648888075828962,
                 w3 = 0.0373376737588285824652402, w4 = 0.0502819010600571181385765,
                 w5 = 0.0258512916557503911218290;
    const double dir[6] = {0.167212652822713264084504, 0.459547634642594690016761,
                           0.628019096642130901034766, 0.760021014833664062877138,
                           0.872270543025721502340662, 0.971637719251358378302376};
    ls_compute_omega(dir, 6, omega);
    const double ww[21] = {w1, w2, w3, w3, w2, w1, w2, w4, w5, w4, w2,
                           w3, w5, w5, w3, w3, w4, w3, w2, w2, w1};
    memcpy(w, ww, sizeof ww);
    return 21;
  }
  default:
    return -1;
  }
}

/* GLC::build_octant + build_chebyshev_quadrature (GLC.cc:14-58).  Polar
 * cosines are the positive Gauss-Legendre nodes of order sn, ascending (the
 * reference table for S6 is unsorted, defect A.15; the intended ascending
 * order is used).  Level i carries sn/2-i Chebyshev azimuths. */
int orc_glc_octant(int sn, double *omega, double *w) {
  double *x = (double *)malloc(sizeof(double) * sn);
  double *wx = (double *)malloc(sizeof(double) * sn);
  orc_gauss_legendre(sn, x, wx);
  int pos = 0;
  for (int i = 0; i < sn / 2; ++i) {
    double ct = x[sn / 2 + i], pw = wx[sn / 2 + i];
    double st = sqrt(1.0 - ct * ct);
    double jmax = (double)(sn / 2 - i);
    for (int j = 0; j < sn / 2 - i; ++j) {
      double phi = M_PI / 2.0 * j / jmax + M_PI / 4.0 / jmax;
      omega[3 * pos + 0] = st * cos(phi);
      omega[3 * pos + 1] = st * sin(phi);
      omega[3 * pos + 2] = ct;
      w[pos] = pw * (M_PI / 2.0 / jmax);
      ++pos;
    }
  }
  free(x);
  free(wx);
  return pos;
}

/* Product Gauss-Legendre x Chebyshev octant (no reference counterpart: the
 * reference only ships the triangular form).  n_polar GL nodes on (-1,1), of
 * which the n_polar/2 positive ones lie in the octant; n_azim Chebyshev
 * azimuths (2j+1) pi / n_azim on the full circle, n_azim/4 per octant. */
int orc_product_octant(int n_polar, int n_azim, double *omega, double *w) {
  double *x = (double *)malloc(sizeof(double) * n_polar);
  double *wx = (double *)malloc(sizeof(double) * n_polar);
  orc_gauss_legendre(n_polar, x, wx);
  int pos = 0, nq = n_azim / 4;
  for (int i = 0; i < n_polar / 2; ++i) {
    double ct = x[n_polar / 2 + i], pw = wx[n_polar / 2 + i];
    double st = sqrt(1.0 - ct * ct);
    for (int j = 0; j < nq; ++j) {
      double phi = (2.0 * j + 1.0) * M_PI / n_azim;
      omega[3 * pos + 0] = st * cos(phi);
      omega[3 * pos + 1] = st * sin(phi);
      omega[3 * pos + 2] = ct;
      w[pos] = pw * (2.0 * M_PI / n_azim);
      ++pos;
    }
  }
  free(x);
  free(wx);
  return pos;
}

/* RTQuadrature::deploy_octant (RTQuadrature.cc:58-95).  2D: quadrants
 * (+,+) (+,-) (-,-) (-,+) with mu>0, weights scaled to sum 1.  3D (build
 * extension, fixes defect A.1): the same four quadrants, then the four
 * mirrored in mu.  Arrays must hold 4*n_oct (2D) or 8*n_oct (3D) entries. */
int orc_deploy(int dim, int n_oct, double *omega, double *w) {
  static const double sx[4] = {1, 1, -1, -1}, sy[4] = {1, -1, -1, 1};
  int nq = dim == 2 ? 4 : 8;
  for (int q = 1; q < nq; ++q)
    for (int j = 0; j < n_oct; ++j) {
      int o = q * n_oct + j;
      omega[3 * o + 0] = sx[q % 4] * omega[3 * j + 0];
      omega[3 * o + 1] = sy[q % 4] * omega[3 * j + 1];
      omega[3 * o + 2] = (q >= 4 ? -1.0 : 1.0) * omega[3 * j + 2];
      w[o] = w[j];
    }
  double s = 0.0;
  for (int j = 0; j < n_oct; ++j) s += w[j];
  for (int o = 0; o < nq * n_oct; ++o) w[o] *= (1.0 / nq) / s;
  return nq * n_oct;
}

/* Associated Legendre P_l^m(x) with the Condon-Shortley phase. */
static double assoc_legendre(int l, int m, double x) {
  double pmm = 1.0;
  if (m > 0) {
    double s = sqrt((1.0 - x) * (1.0 + x)), f = 1.0;
    for (int i = 1; i <= m; ++i) {
      pmm *= -f * s;
      f += 2.0;
    }
  }
  if (l == m) return pmm;
  double pmmp1 = x * (2.0 * m + 1.0) * pmm;
  if (l == m + 1) return pmmp1;
  double pll = 0.0;
  for (int ll = m + 2; ll <= l; ++ll) {
    pll = ((2.0 * ll - 1.0) * x * pmmp1 - (ll + m - 1.0) * pmm) / (ll - m);
    pmm = pmmp1;
    pmmp1 = pll;
  }
  return pll;
}

static double real_sh(int l, int m, int odd, double mu, double phi) {
  double f = 1.0;
  for (int k = l - m + 1; k <= l + m; ++k) f *= k; /* (l+m)!/(l-m)! */
  double N = sqrt((2.0 * l + 1.0) / (4.0 * M_PI) / f);
  double P = assoc_legendre(l, m, mu);
  if (m == 0) return odd ? 0.0 : N * P;
  return sqrt(2.0) * N * P * (odd ? sin(m * phi) : cos(m * phi));
}

/* Gauss-Jordan inverse with partial pivoting (for the Galerkin D2M). */
static int invert(int n, const double *A, double *Ainv) {
  double *a = (double *)malloc(sizeof(double) * n * 2 * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 2 * n; ++j)
      a[i * 2 * n + j] = j < n ? A[i * n + j] : (j - n == i ? 1.0 : 0.0);
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(a[i * 2 * n + k]) > fabs(a[piv * 2 * n + k])) piv = i;
    if (a[piv * 2 * n + k] == 0.0) { free(a); return -1; }
    if (piv != k)
      for (int j = 0; j < 2 * n; ++j) {
        double t = a[k * 2 * n + j];
        a[k * 2 * n + j] = a[piv * 2 * n + j];
        a[piv * 2 * n + j] = t;
      }
    double d = a[k * 2 * n + k];
    for (int j = 0; j < 2 * n; ++j) a[k * 2 * n + j] /= d;
    for (int i = 0; i < n; ++i)
      if (i != k) {
        double f = a[i * 2 * n + k];
        for (int j = 0; j < 2 * n; ++j) a[i * 2 * n + j] -= f * a[k * 2 * n + j];
      }
  }
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) Ainv[i * n + j] = a[i * 2 * n + n + j];
  free(a);
  return 0;
}

/* RTQuadrature::compute_harmonics + D2M (RTQuadrature.cc:36-51, 97-198).
 * Moment order per l: cosine harmonics m=l..0, then sine harmonics m=1..l; in
 * 2D only (l+m) even (the harmonics even in mu).  Deviations from the shipped
 * reference (intended mathematics): theta = acos(mu), not mu (defect A.7);
 * real harmonics carry the sqrt(2) of the orthonormal real basis for m>0 so
 * that D2M.M2D = I for the standard quadrature; 3D keeps every harmonic,
 * n_mom = (L+1)^2.  M2D[d,m] = sqrt(4 pi/ws) Y_m(Omega_d); D2M = M2D^T
 * diag(ws w) or M2D^{-1} for Galerkin (2D only).  Returns n_mom or -1. */
int orc_harmonics(int dim, int L, int galerkin, int sn, int n_dir,
                  const double *omega, const double *w, double ws,
                  double *M2D, double *D2M, int *mom_l) {
  if (galerkin && dim != 2) return -1;
  int n_mom = 0;
  /* count */
  for (int l = 0; l <= L; ++l) {
    for (int m = l; m >= 0; --m) {
      int keep = dim == 3 || (m + l) % 2 == 0;
      if (galerkin) keep = (l < sn) && ((m + l) % 2 == 0);
      if (keep) ++n_mom;
    }
    for (int m = 1; m <= l; ++m) {
      int keep = dim == 3 || (m + l) % 2 == 0;
      if (galerkin) keep = (l <= sn) && ((m + l) % 2 == 0);
      if (keep) ++n_mom;
    }
  }
  if (galerkin && n_mom != n_dir) return -1;
  const double wsph = sqrt(4.0 * M_PI / ws);
  for (int d = 0; d < n_dir; ++d) {
    double mu = omega[3 * d + 2];
    double phi = atan2(omega[3 * d + 1], omega[3 * d + 0]);
    int pos = 0;
    for (int l = 0; l <= L; ++l) {
      for (int m = l; m >= 0; --m) {
        int keep = dim == 3 || (m + l) % 2 == 0;
        if (galerkin) keep = (l < sn) && ((m + l) % 2 == 0);
        if (!keep) continue;
        M2D[d * n_mom + pos] = wsph * real_sh(l, m, 0, mu, phi);
        mom_l[pos] = l;
        ++pos;
      }
      for (int m = 1; m <= l; ++m) {
        int keep = dim == 3 || (m + l) % 2 == 0;
        if (galerkin) keep = (l <= sn) && ((m + l) % 2 == 0);
        if (!keep) continue;
        M2D[d * n_mom + pos] = wsph * real_sh(l, m, 1, mu, phi);
        mom_l[pos] = l;
        ++pos;
      }
    }
  }
  if (galerkin) {
    if (invert(n_dir, M2D, D2M) != 0) return -1;
  } else {
    for (int m = 0; m < n_mom; ++m)
      for (int d = 0; d < n_dir; ++d)
        D2M[m * n_dir + d] = M2D[d * n_mom + m] * ws * w[d];
  }
  return n_mom;
}

/* RTQuadrature::compute_most_normal_directions (RTQuadrature.cc:200-238).
 * flags[f*n_dir+d] = 1 when direction d has the extreme component for face
 * f: face 2a takes the largest Omega_a, face 2a+1 the smallest. */
void orc_most_normal(int dim, int n_dir, const double *omega,
                     unsigned char *flags) {
  double mn[6] = {0, 0, 0, 0, 0, 0};
  for (int d = 0; d < n_dir; ++d)
    for (int a = 0; a < dim; ++a) {
      if (omega[3 * d + a] > mn[2 * a]) mn[2 * a] = omega[3 * d + a];
      if (omega[3 * d + a] < mn[2 * a + 1]) mn[2 * a + 1] = omega[3 * d + a];
    }
  for (int f = 0; f < 2 * dim; ++f)
    for (int d = 0; d < n_dir; ++d)
      flags[f * n_dir + d] = omega[3 * d + f / 2] == mn[f];
}

/* ========================================================================= */
/* Finite elements                                                            */
/* ========================================================================= */

int orc_n_loc(int dim, int p) {
  int n = 1;
  for (int a = 0; a < dim; ++a) n *= p + 1;
  return n;
}

/* Equidistant Lagrange basis on [0,1] (FE_DGQ support points). */
static double lag(int p, int k, double t) {
  double v = 1.0;
  for (int j = 0; j <= p; ++j)
    if (j != k) v *= (t - (double)j / p) / ((double)(k - j) / p);
  return v;
}
static double dlag(int p, int k, double t) {
  double s = 0.0;
  for (int i = 0; i <= p; ++i) {
    if (i == k) continue;
    double v = 1.0 / ((double)(k - i) / p);
    for (int j = 0; j <= p; ++j)
      if (j != k && j != i) v *= (t - (double)j / p) / ((double)(k - j) / p);
    s += v;
  }
  return s;
}

/* FECell constructor (FECell.cc:10-78) on an axis-aligned cell [0,h]^dim,
 * evaluated like FEValues with the tensor Gauss rule of p+1 points:
 *   M[i][j]      = sum_q b_i b_j JxW
 *   G[d][i][j]   = sum_q b_j d_d b_i JxW                    (d < dim)
 *   E[f][i][j]   = sum_qf b_i b_j JxW                        (f < 2 dim)
 *   U[f][i][j]   = sum_qf b_i b_j^nbr JxW, nbr = the cell across face f, seen
 *                  through its opposite face f^1 (defect A.4 fixed)
 *   normals[f]   = outward unit normal (3 components per face)
 * Faces are ordered -x,+x,-y,+y,-z,+z.  Output arrays are N x N row-major. */
void orc_fe_cell(int dim, int p, const double *h, double *M, double *G,
                 double *E, double *U, double *normals) {
  const int N = orc_n_loc(dim, p), nq = p + 1;
  double *gx = (double *)malloc(sizeof(double) * nq);
  double *gw = (double *)malloc(sizeof(double) * nq);
  orc_gauss_legendre(nq, gx, gw);
  for (int q = 0; q < nq; ++q) {
    gx[q] = 0.5 * (gx[q] + 1.0);
    gw[q] *= 0.5;
  }
  memset(M, 0, sizeof(double) * N * N);
  memset(G, 0, sizeof(double) * dim * N * N);
  memset(E, 0, sizeof(double) * 2 * dim * N * N);
  memset(U, 0, sizeof(double) * 2 * dim * N * N);
  int nqv = 1;
  for (int a = 0; a < dim; ++a) nqv *= nq;
  int idx[3], qi[3];
  /* volume terms */
  for (int qv = 0; qv < nqv; ++qv) {
    double t[3] = {0, 0, 0}, jxw = 1.0;
    int r = qv;
    for (int a = 0; a < dim; ++a) {
      qi[a] = r % nq;
      r /= nq;
      t[a] = gx[qi[a]];
      jxw *= gw[qi[a]] * h[a];
    }
    for (int i = 0; i < N; ++i) {
      int ri = i;
      for (int a = 0; a < dim; ++a) { idx[a] = ri % (p + 1); ri /= p + 1; }
      double bi = 1.0, gi[3];
      for (int a = 0; a < dim; ++a) bi *= lag(p, idx[a], t[a]);
      for (int d = 0; d < dim; ++d) {
        double g = dlag(p, idx[d], t[d]) / h[d];
        for (int a = 0; a < dim; ++a)
          if (a != d) g *= lag(p, idx[a], t[a]);
        gi[d] = g;
      }
      for (int j = 0; j < N; ++j) {
        int rj = j, jd[3];
        for (int a = 0; a < dim; ++a) { jd[a] = rj % (p + 1); rj /= p + 1; }
        double bj = 1.0;
        for (int a = 0; a < dim; ++a) bj *= lag(p, jd[a], t[a]);
        M[i * N + j] += bi * bj * jxw;
        for (int d = 0; d < dim; ++d) G[(d * N + i) * N + j] += bj * gi[d] * jxw;
      }
    }
  }
  /* face terms */
  int nqf = nqv / nq;
  for (int f = 0; f < 2 * dim; ++f) {
    int fa = f / 2, side = f % 2;
    for (int c = 0; c < 3; ++c) normals[3 * f + c] = 0.0;
    normals[3 * f + fa] = side ? 1.0 : -1.0;
    for (int qf = 0; qf < nqf; ++qf) {
      double t[3] = {0, 0, 0}, tn[3], jxw = 1.0;
      int r = qf;
      for (int a = 0; a < dim; ++a) {
        if (a == fa) continue;
        int k = r % nq;
        r /= nq;
        t[a] = gx[k];
        jxw *= gw[k] * h[a];
      }
      t[fa] = side ? 1.0 : 0.0;
      for (int a = 0; a < 3; ++a) tn[a] = t[a];
      tn[fa] = side ? 0.0 : 1.0; /* neighbour's local coordinate */
      for (int i = 0; i < N; ++i) {
        int ri = i;
        for (int a = 0; a < dim; ++a) { idx[a] = ri % (p + 1); ri /= p + 1; }
        double bi = 1.0;
        for (int a = 0; a < dim; ++a) bi *= lag(p, idx[a], t[a]);
        for (int j = 0; j < N; ++j) {
          int rj = j, jd[3];
          for (int a = 0; a < dim; ++a) { jd[a] = rj % (p + 1); rj /= p + 1; }
          double bj = 1.0, bn = 1.0;
          for (int a = 0; a < dim; ++a) {
            bj *= lag(p, jd[a], t[a]);
            bn *= lag(p, jd[a], tn[a]);
          }
          E[(f * N + i) * N + j] += bi * bj * jxw;
          U[(f * N + i) * N + j] += bi * bn * jxw;
        }
      }
    }
  }
  free(gx);
  free(gw);
}

/* ========================================================================= */
/* Dense kernels                                                              */
/* ========================================================================= */

/* RadiativeTransfer::LU_decomposition (RadiativeTransfer.cc:1026-1065):
 * Crout's method with row partial pivoting; L keeps the diagonal, U is unit
 * upper triangular; whole rows are interchanged; in place, row-major. */
void orc_lu_decompose(int n, double *A, int *pivot) {
  for (int k = 0; k < n; ++k) {
    pivot[k] = k;
    double mx = fabs(A[k * n + k]);
    for (int j = k + 1; j < n; ++j)
      if (mx < fabs(A[j * n + k])) {
        mx = fabs(A[j * n + k]);
        pivot[k] = j;
      }
    if (pivot[k] != k) {
      int pv = pivot[k];
      for (int j = 0; j < n; ++j) {
        double t = A[k * n + j];
        A[k * n + j] = A[pv * n + j];
        A[pv * n + j] = t;
      }
    }
    for (int j = k + 1; j < n; ++j) A[k * n + j] /= A[k * n + k];
    for (int i = k + 1; i < n; ++i)
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= A[i * n + k] * A[k * n + j];
  }
}

/* RadiativeTransfer::LU_solve (RadiativeTransfer.cc:1067-1105): the row
 * interchanges are applied to b in sequence during forward substitution
 * with L; back substitution with the unit upper U.  b is modified. */
void orc_lu_solve(int n, const double *A, double *b, double *x,
                  const int *pivot) {
  for (int k = 0; k < n; ++k) {
    if (pivot[k] != k) {
      double t = b[k];
      b[k] = b[pivot[k]];
      b[pivot[k]] = t;
    }
    x[k] = b[k];
    for (int i = 0; i < k; ++i) x[k] -= x[i] * A[k * n + i];
    x[k] /= A[k * n + k];
  }
  for (int k = n - 1; k >= 0; --k)
    for (int i = k + 1; i < n; ++i) x[k] -= x[i] * A[k * n + i];
}

/* ========================================================================= */
/* Operator                                                                   */
/* ========================================================================= */

static long n_cells(const orc_problem *pb) {
  return (long)pb->n[0] * pb->n[1] * (pb->dim == 3 ? pb->n[2] : 1);
}

long orc_n_dofs(const orc_problem *pb) {
  return n_cells(pb) * orc_n_loc(pb->dim, pb->p);
}

typedef struct {
  int N;
  double *M, *G, *E, *U, *nrm;
} fe_t;

static void fe_build(const orc_problem *pb, fe_t *fe) {
  fe->N = orc_n_loc(pb->dim, pb->p);
  int N = fe->N, dim = pb->dim;
  fe->M = (double *)malloc(sizeof(double) * N * N);
  fe->G = (double *)malloc(sizeof(double) * dim * N * N);
  fe->E = (double *)malloc(sizeof(double) * 2 * dim * N * N);
  fe->U = (double *)malloc(sizeof(double) * 2 * dim * N * N);
  fe->nrm = (double *)malloc(sizeof(double) * 6 * 3);
  orc_fe_cell(dim, pb->p, pb->h, fe->M, fe->G, fe->E, fe->U, fe->nrm);
}
static void fe_free(fe_t *fe) {
  free(fe->M); free(fe->G); free(fe->E); free(fe->U); free(fe->nrm);
}

/* RadiativeTransfer::compute_scattering_source (RadiativeTransfer.cc:811-843)
 * S_m|K = sigma_{s,l(m)} M x_m|K.  Uses the Legendre order l(m) of moment m
 * (the shipped code passes m itself, defect A.5). */
void orc_scatter_source(const orc_problem *pb, const double *x, double *S) {
  fe_t fe;
  fe_build(pb, &fe);
  const int N = fe.N;
  const long nc = n_cells(pb), nd = nc * N;
  for (int m = 0; m < pb->n_mom; ++m)
    for (long c = 0; c < nc; ++c) {
      const double sig = pb->sigma_s[pb->mat[c] * (pb->L + 1) + pb->mom_l[m]];
      const double *xc = x + m * nd + c * N;
      double *sc = S + m * nd + c * N;
      for (int i = 0; i < N; ++i) {
        double v = 0.0;
        for (int j = 0; j < N; ++j) v += fe.M[i * N + j] * xc[j];
        sc[i] = v * sig;
      }
    }
  fe_free(&fe);
}

/* RadiativeTransfer::sweep (RadiativeTransfer.cc:878-1012) for every task of
 * the directions [d_begin,d_end), each task being one cell and one direction.
 * The task DAG scheduler (RadiativeTransfer.cc:605-778) reduces on a
 * Cartesian mesh to any order that visits upwind cells first; the nested
 * loops below, run in the direction of each sign of Omega, are such an
 * order.  phi is accumulated (not cleared), as the reference accumulates into
 * flux_moments.  with_source adds the volumetric source M q/ws and the
 * boundary inflow |n.Omega| E_f inc/ws (the group_flux != nullptr branch). */
void orc_sweep_dirs(const orc_problem *pb, const double *S, int with_source,
                    int d_begin, int d_end, double *phi) {
  fe_t fe;
  fe_build(pb, &fe);
  const int N = fe.N, dim = pb->dim, nf = 2 * dim;
  const int nx = pb->n[0], ny = pb->n[1], nz = dim == 3 ? pb->n[2] : 1;
  const long nc = n_cells(pb), nd = nc * N;
  double *psi = (double *)malloc(sizeof(double) * nd);
  double *A = (double *)malloc(sizeof(double) * N * N);
  double *b = (double *)malloc(sizeof(double) * N);
  double *xs = (double *)malloc(sizeof(double) * N);
  double *tmp = (double *)malloc(sizeof(double) * N);
  int *piv = (int *)malloc(sizeof(int) * N);
  for (int d = d_begin; d < d_end; ++d) {
    const double *om = pb->omega + 3 * d;
    int i0 = om[0] >= 0 ? 0 : nx - 1, di = om[0] >= 0 ? 1 : -1;
    int j0 = om[1] >= 0 ? 0 : ny - 1, dj = om[1] >= 0 ? 1 : -1;
    int k0 = (dim == 3 && om[2] < 0) ? nz - 1 : 0, dk = (dim == 3 && om[2] < 0) ? -1 : 1;
    for (int kk = 0, k = k0; kk < nz; ++kk, k += dk)
      for (int jj = 0, j = j0; jj < ny; ++jj, j += dj)
        for (int ii = 0, i = i0; ii < nx; ++ii, i += di) {
          const long c = i + (long)nx * (j + (long)ny * k);
          const int ijk[3] = {i, j, k};
          const double sigt = pb->sigma_t[pb->mat[c]];
          /* A = sigma_t M - sum_d Omega_d G_d */
          for (int e = 0; e < N * N; ++e) A[e] = fe.M[e] * sigt;
          for (int a = 0; a < dim; ++a)
            for (int e = 0; e < N * N; ++e) A[e] += -om[a] * fe.G[a * N * N + e];
          /* scattering source */
          for (int e = 0; e < N; ++e) b[e] = 0.0;
          for (int m = 0; m < pb->n_mom; ++m) {
            const double m2d = pb->M2D[d * pb->n_mom + m];
            for (int e = 0; e < N; ++e) b[e] += m2d * S[m * nd + c * N + e];
          }
          if (with_source) {
            double src = (pb->q ? pb->q[c] : 0.0) / pb->ws;
            for (int e = 0; e < N; ++e) {
              double v = 0.0;
              for (int f = 0; f < N; ++f) v += fe.M[e * N + f] * src;
              b[e] += v;
            }
          }
          /* surface terms */
          for (int f = 0; f < nf; ++f) {
            double ndo = 0.0;
            for (int a = 0; a < dim; ++a) ndo += om[a] * fe.nrm[3 * f + a];
            const int fa = f / 2, side = f % 2;
            if (ndo < 0.0) {
              const int nb = side ? ijk[fa] + 1 : ijk[fa] - 1;
              const int interior = nb >= 0 && nb < pb->n[fa];
              if (interior) {
                long cn = c + (side ? 1 : -1) * (fa == 0 ? 1L : fa == 1 ? (long)nx : (long)nx * ny);
                for (int e = 0; e < N; ++e) tmp[e] = -ndo * psi[cn * N + e];
                const double *Uf = fe.U + f * N * N;
                for (int e = 0; e < N; ++e) {
                  double v = 0.0;
                  for (int g = 0; g < N; ++g) v += Uf[e * N + g] * tmp[g];
                  b[e] += v;
                }
              } else if (with_source) {
                double inc = 0.0;
                if ((pb->bc[f] == ORC_BC_MOST_NORMAL && pb->most_normal &&
                     pb->most_normal[f * pb->n_dir + d]) ||
                    pb->bc[f] == ORC_BC_ISOTROPIC)
                  inc = pb->inc[f];
                inc /= pb->ws;
                const double *Ef = fe.E + f * N * N;
                for (int e = 0; e < N; ++e) {
                  double v = 0.0;
                  for (int g = 0; g < N; ++g) v += Ef[e * N + g] * inc;
                  b[e] += (-ndo) * v;
                }
              }
            } else {
              const double *Ef = fe.E + f * N * N;
              for (int e = 0; e < N * N; ++e) A[e] += ndo * Ef[e];
            }
          }
          orc_lu_decompose(N, A, piv);
          orc_lu_solve(N, A, b, xs, piv);
          for (int m = 0; m < pb->n_mom; ++m) {
            const double d2m = pb->D2M[m * pb->n_dir + d];
            for (int e = 0; e < N; ++e) phi[m * nd + c * N + e] += d2m * xs[e];
          }
          for (int e = 0; e < N; ++e) psi[c * N + e] = xs[e];
        }
  }
  free(psi); free(A); free(b); free(xs); free(tmp); free(piv);
  fe_free(&fe);
}

void orc_sweep(const orc_problem *pb, const double *S, int with_source,
               double *phi) {
  orc_sweep_dirs(pb, S, with_source, 0, pb->n_dir, phi);
}

/* RadiativeTransfer::Apply (RadiativeTransfer.cc:780-809):
 *   y = x - D2M L^-1 M2D S(x),
 * over every moment (the shipped epilogue touches moment 0 only, defect A.6). */
void orc_apply(const orc_problem *pb, const double *x, double *y) {
  const long n = orc_n_dofs(pb) * pb->n_mom;
  double *S = (double *)malloc(sizeof(double) * n);
  orc_scatter_source(pb, x, S);
  for (long i = 0; i < n; ++i) y[i] = 0.0;
  orc_sweep(pb, S, 0, y);
  for (long i = 0; i < n; ++i) y[i] = x[i] - y[i];
  free(S);
}

/* Right-hand side of the Krylov solve (ouranos.cc:149-158): one sweep with
 * a zero scattering source, the volumetric source and the boundary inflow. */
void orc_rhs(const orc_problem *pb, double *rhs) {
  const long n = orc_n_dofs(pb) * pb->n_mom;
  double *S = (double *)calloc(n, sizeof(double));
  for (long i = 0; i < n; ++i) rhs[i] = 0.0;
  orc_sweep(pb, S, 1, rhs);
  free(S);
}

/* ========================================================================= */
/* Solvers                                                                    */
/* ========================================================================= */

static double dot(long n, const double *a, const double *b) {
  double s = 0.0;
  for (long i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* Source iteration (ouranos.cc:52-108) with its intended semantics: the
 * sweep is applied to a cleared moment vector (defect A.10) and the loop is
 * reachable (defect A.9).  phi_{k+1} = D L^-1 (M Sigma phi_k + q + inflow);
 * stop when ||phi_{k+1}-phi_k|| / ||phi_{k+1}|| < tol.  phi is the initial
 * guess on entry. */
int orc_si(const orc_problem *pb, double tol, int max_it, double *phi) {
  const long n = orc_n_dofs(pb) * pb->n_mom;
  double *S = (double *)malloc(sizeof(double) * n);
  double *nw = (double *)malloc(sizeof(double) * n);
  int it, conv = 0;
  for (it = 0; it < max_it; ++it) {
    orc_scatter_source(pb, phi, S);
    for (long i = 0; i < n; ++i) nw[i] = 0.0;
    orc_sweep(pb, S, 1, nw);
    double num = 0.0;
    for (long i = 0; i < n; ++i) num += (nw[i] - phi[i]) * (nw[i] - phi[i]);
    double den = dot(n, nw, nw);
    memcpy(phi, nw, sizeof(double) * n);
    if (sqrt(num) / sqrt(den) < tol) { conv = 1; ++it; break; }
  }
  free(S);
  free(nw);
  return conv ? it : -it;
}

/* Restarted GMRES with modified Gram-Schmidt and Givens rotations, identity
 * preconditioner (ouranos.cc:160-171 calls the third-party solver).  Stops on
 * ||b - A x|| / ||b|| < tol.  x holds the initial guess on entry. */
int orc_gmres(const orc_problem *pb, const double *b, double *x, int restart,
              double tol, int max_it, double *final_rel_res) {
  const long n = orc_n_dofs(pb) * pb->n_mom;
  const int m = restart;
  double *V = (double *)malloc(sizeof(double) * n * (m + 1));
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = (double *)malloc(sizeof(double) * m);
  double *sn = (double *)malloc(sizeof(double) * m);
  double *g = (double *)malloc(sizeof(double) * (m + 1));
  double *yv = (double *)malloc(sizeof(double) * m);
  double *w = (double *)malloc(sizeof(double) * n);
  const double bnorm = sqrt(dot(n, b, b));
  int total = 0;
  double rel = 1.0;
  if (bnorm == 0.0) {
    for (long i = 0; i < n; ++i) x[i] = 0.0;
    if (final_rel_res) *final_rel_res = 0.0;
    goto done;
  }
  while (total < max_it) {
    orc_apply(pb, x, w);
    for (long i = 0; i < n; ++i) V[i] = b[i] - w[i];
    double beta = sqrt(dot(n, V, V));
    rel = beta / bnorm;
    if (rel < tol) break;
    for (long i = 0; i < n; ++i) V[i] /= beta;
    for (int i = 0; i <= m; ++i) g[i] = 0.0;
    g[0] = beta;
    int j;
    for (j = 0; j < m && total < max_it; ++j) {
      ++total;
      orc_apply(pb, V + j * n, w);
      for (int i = 0; i <= j; ++i) {
        double hij = dot(n, w, V + i * n);
        H[i * m + j] = hij;
        for (long k = 0; k < n; ++k) w[k] -= hij * V[i * n + k];
      }
      double hn = sqrt(dot(n, w, w));
      H[(j + 1) * m + j] = hn;
      if (hn != 0.0)
        for (long k = 0; k < n; ++k) V[(j + 1) * n + k] = w[k] / hn;
      for (int i = 0; i < j; ++i) {
        double t = cs[i] * H[i * m + j] + sn[i] * H[(i + 1) * m + j];
        H[(i + 1) * m + j] = -sn[i] * H[i * m + j] + cs[i] * H[(i + 1) * m + j];
        H[i * m + j] = t;
      }
      double a = H[j * m + j], bb = H[(j + 1) * m + j];
      double r = sqrt(a * a + bb * bb);
      cs[j] = a / r;
      sn[j] = bb / r;
      H[j * m + j] = r;
      H[(j + 1) * m + j] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      rel = fabs(g[j + 1]) / bnorm;
      if (rel < tol) { ++j; break; }
    }
    for (int i = j - 1; i >= 0; --i) {
      double s = g[i];
      for (int k = i + 1; k < j; ++k) s -= H[i * m + k] * yv[k];
      yv[i] = s / H[i * m + i];
    }
    for (int i = 0; i < j; ++i)
      for (long k = 0; k < n; ++k) x[k] += yv[i] * V[i * n + k];
  }
  if (final_rel_res) *final_rel_res = rel;
done:
  free(V); free(H); free(cs); free(sn); free(g); free(yv); free(w);
  return rel < tol ? total : -total;
}

/* BiCGStab (van der Vorst), identity preconditioner, as used by the golden
 * integration test (tests/RadiativeTransfer_test.cc:76-81).  Stops on
 * ||r|| / ||b|| < tol. */
int orc_bicgstab(const orc_problem *pb, const double *b, double *x,
                 double tol, int max_it, double *final_rel_res) {
  const long n = orc_n_dofs(pb) * pb->n_mom;
  double *r = (double *)malloc(sizeof(double) * n);
  double *r0 = (double *)malloc(sizeof(double) * n);
  double *pv = (double *)calloc(n, sizeof(double));
  double *v = (double *)calloc(n, sizeof(double));
  double *s = (double *)malloc(sizeof(double) * n);
  double *t = (double *)malloc(sizeof(double) * n);
  const double bnorm = sqrt(dot(n, b, b));
  orc_apply(pb, x, r);
  for (long i = 0; i < n; ++i) { r[i] = b[i] - r[i]; r0[i] = r[i]; }
  double rho = 1.0, alpha = 1.0, omega = 1.0, rel = sqrt(dot(n, r, r)) / bnorm;
  int it = 0;
  while (rel >= tol && it < max_it) {
    ++it;
    double rho1 = dot(n, r0, r);
    double beta = (rho1 / rho) * (alpha / omega);
    rho = rho1;
    for (long i = 0; i < n; ++i) pv[i] = r[i] + beta * (pv[i] - omega * v[i]);
    orc_apply(pb, pv, v);
    alpha = rho / dot(n, r0, v);
    for (long i = 0; i < n; ++i) s[i] = r[i] - alpha * v[i];
    if (sqrt(dot(n, s, s)) / bnorm < tol) {
      for (long i = 0; i < n; ++i) x[i] += alpha * pv[i];
      rel = sqrt(dot(n, s, s)) / bnorm;
      break;
    }
    orc_apply(pb, s, t);
    omega = dot(n, t, s) / dot(n, t, t);
    for (long i = 0; i < n; ++i) {
      x[i] += alpha * pv[i] + omega * s[i];
      r[i] = s[i] - omega * t[i];
    }
    rel = sqrt(dot(n, r, r)) / bnorm;
  }
  if (final_rel_res) *final_rel_res = rel;
  free(r); free(r0); free(pv); free(v); free(s); free(t);
  return rel < tol ? it : -it;
}
