// Angular quadratures and moment operators; see quadrature.hpp.
#include "quadrature.hpp"

#include <cmath>

#include "ou_common.hpp"

namespace ou {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

// Level-symmetric octant data of the reference (LS.cc:18-169): the distinct
// direction cosines, the weight classes, and the class of each direction in
// the triangular order produced by LS::compute_omega (LS.cc:667-682).
struct LSTable {
  int sn;
  int n_cos;
  double cosines[6];
  double classes[5];
  int class_of[21];
};

const LSTable kLS[] = {
    {2, 1, {0.577350269189625764509149}, {1.0}, {0}},
    {4, 2, {0.350021174581540677777041, 0.868890300722201205229788},
     {1.0 / 3.0}, {0, 0, 0}},
    {6, 3,
     {0.266635401516704720331535, 0.681507726536546927403750,
      0.926180935517489107558380},
     {0.176126130863383433783565, 0.157207202469949899549768},
     {0, 1, 0, 1, 1, 0}},
    {8, 4,
     {0.218217890235992381266097, 0.577350269189625764509149,
      0.786795792469443145800830, 0.951189731211341853132399},
     {0.120987654320987654320988, 0.0907407407407407407407407,
      0.0925925925925925925925926},
     {0, 1, 1, 0, 1, 2, 1, 1, 1, 0}},
    {10, 5,
     {0.189321326478010476671494, 0.508881755582618974382711,
      0.694318887594384317279217, 0.839759962236684758403029,
      0.963490981110468484701598},
     {0.0893031479843567214704325, 0.0725291517123655242296233,
      0.0450437674364086390490892, 0.0539281144878369243545650},
     {0, 1, 2, 1, 0, 1, 3, 3, 1, 2, 3, 2, 1, 1, 0}},
    {12, 6,
     {0.167212652822713264084504, 0.459547634642594690016761,
      0.628019096642130901034766, 0.760021014833664062877138,
      0.872270543025721502340662, 0.971637719251358378302376},
     {0.0707625899700910439766549, 0.0558811015648888075828962,
      0.0373376737588285824652402, 0.0502819010600571181385765,
      0.0258512916557503911218290},
     {0, 1, 2, 2, 1, 0, 1, 3, 4, 3, 1, 2, 4, 4, 2, 2, 3, 2, 1, 1, 0}},
};

// Orthonormal associated Legendre functions Pbar_l^m(x) for 0 <= m <= l <= L
// (Condon-Shortley phase included), by the standard stable recurrences:
//   Pbar_0^0 = 1/sqrt(4 pi),  Pbar_m^m = -sqrt((2m+1)/(2m)) s Pbar_{m-1}^{m-1},
//   Pbar_{m+1}^m = sqrt(2m+3) x Pbar_m^m,
//   Pbar_l^m = a_lm (x Pbar_{l-1}^m - Pbar_{l-2}^m / a_{l-1,m}),
//   a_lm = sqrt((4l^2-1)/(l^2-m^2)).
// Y_l^0 = Pbar_l^0, and the real harmonics for m > 0 are
// sqrt(2) Pbar_l^m cos(m phi) and sqrt(2) Pbar_l^m sin(m phi).
void normalized_legendre(int L, double x, std::vector<double> &P) {
  const int W = L + 1;
  P.assign(W * W, 0.0);
  const double s = std::sqrt(std::fmax(0.0, (1.0 - x) * (1.0 + x)));
  P[0] = 1.0 / std::sqrt(4.0 * kPi);
  for (int m = 1; m <= L; ++m)
    P[m * W + m] = -std::sqrt((2.0 * m + 1.0) / (2.0 * m)) * s * P[(m - 1) * W + m - 1];
  for (int m = 0; m < L; ++m) P[(m + 1) * W + m] = std::sqrt(2.0 * m + 3.0) * x * P[m * W + m];
  for (int m = 0; m <= L; ++m)
    for (int l = m + 2; l <= L; ++l) {
      const double a = std::sqrt((4.0 * l * l - 1.0) / (double(l) * l - double(m) * m));
      const double ap = std::sqrt((4.0 * (l - 1) * (l - 1) - 1.0) /
                                  (double(l - 1) * (l - 1) - double(m) * m));
      P[l * W + m] = a * (x * P[(l - 1) * W + m] - P[(l - 2) * W + m] / ap);
    }
}

// Square matrix inverse by LU with partial pivoting (Galerkin D2M).
bool invert(int n, const std::vector<double> &A, std::vector<double> &inv) {
  std::vector<double> a(A);
  std::vector<int> perm(n);
  for (int i = 0; i < n; ++i) perm[i] = i;
  for (int k = 0; k < n; ++k) {
    int piv = k;
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(a[i * n + k]) > std::fabs(a[piv * n + k])) piv = i;
    if (a[piv * n + k] == 0.0) return false;
    if (piv != k) {
      for (int j = 0; j < n; ++j) std::swap(a[k * n + j], a[piv * n + j]);
      std::swap(perm[k], perm[piv]);
    }
    for (int i = k + 1; i < n; ++i) {
      a[i * n + k] /= a[k * n + k];
      for (int j = k + 1; j < n; ++j) a[i * n + j] -= a[i * n + k] * a[k * n + j];
    }
  }
  inv.assign(n * n, 0.0);
  std::vector<double> col(n);
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < n; ++i) col[i] = perm[i] == c ? 1.0 : 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < i; ++j) col[i] -= a[i * n + j] * col[j];
    for (int i = n - 1; i >= 0; --i) {
      for (int j = i + 1; j < n; ++j) col[i] -= a[i * n + j] * col[j];
      col[i] /= a[i * n + i];
    }
    for (int i = 0; i < n; ++i) inv[i * n + c] = col[i];
  }
  return true;
}

}  // namespace

void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    // Tricomi's initial guess, then Newton on P_n.
    double z = std::cos(kPi * (4.0 * i + 3.0) / (4.0 * n + 2.0));
    double dpn = 1.0;
    for (int it = 0; it < 64; ++it) {
      double pm1 = 1.0, pn = z;
      for (int k = 2; k <= n; ++k) {
        const double pk = ((2 * k - 1) * z * pn - (k - 1) * pm1) / k;
        pm1 = pn;
        pn = pk;
      }
      if (n == 1) pm1 = 1.0, pn = z;
      dpn = n * (pm1 - z * pn) / (1.0 - z * z);
      const double step = pn / dpn;
      z -= step;
      if (std::fabs(step) <= 1e-16 * std::fabs(z) && it > 2) break;
    }
    double pm1 = 1.0, pn = z;
    for (int k = 2; k <= n; ++k) {
      const double pk = ((2 * k - 1) * z * pn - (k - 1) * pm1) / k;
      pm1 = pn;
      pn = pk;
    }
    if (n == 1) pm1 = 1.0, pn = z;
    dpn = n * (pm1 - z * pn) / (1.0 - z * z);
    const double wi = 2.0 / ((1.0 - z * z) * dpn * dpn);
    x[n - 1 - i] = z;
    x[i] = -z;
    w[i] = w[n - 1 - i] = wi;
  }
  if (n % 2) x[n / 2] = 0.0;
}

bool ls_octant(int sn, std::vector<double> &omega, std::vector<double> &w) {
  for (const LSTable &t : kLS) {
    if (t.sn != sn) continue;
    const int n = t.n_cos, n_oct = n * (n + 1) / 2;
    omega.assign(3 * n_oct, 0.0);
    w.assign(n_oct, 0.0);
    int pos = 0;
    // Row r holds mu = cos[r]; within a row eta runs up the cosine list.
    for (int r = 0; r < n; ++r)
      for (int c = 0; c < n - r; ++c, ++pos) {
        omega[3 * pos + 0] = t.cosines[n - 1 - r - c];
        omega[3 * pos + 1] = t.cosines[c];
        omega[3 * pos + 2] = t.cosines[r];
        w[pos] = t.classes[t.class_of[pos]];
      }
    return true;
  }
  return false;
}

void glc_octant(int sn, std::vector<double> &omega, std::vector<double> &w) {
  std::vector<double> gx, gw;
  gauss_legendre(sn, gx, gw);
  omega.clear();
  w.clear();
  const int half = sn / 2;
  for (int level = 0; level < half; ++level) {
    const double mu = gx[half + level], wp = gw[half + level];
    const double sin_t = std::sqrt(1.0 - mu * mu);
    const int n_az = half - level;
    for (int j = 0; j < n_az; ++j) {
      const double phi = (kPi / 2.0) * (j + 0.5) / n_az;
      omega.push_back(sin_t * std::cos(phi));
      omega.push_back(sin_t * std::sin(phi));
      omega.push_back(mu);
      w.push_back(wp * (kPi / 2.0) / n_az);
    }
  }
}

void product_glc_octant(int n_polar, int n_azim, std::vector<double> &omega,
                        std::vector<double> &w) {
  std::vector<double> gx, gw;
  gauss_legendre(n_polar, gx, gw);
  omega.clear();
  w.clear();
  const int per_quadrant = n_azim / 4;
  for (int i = n_polar / 2; i < n_polar; ++i) {
    const double mu = gx[i], sin_t = std::sqrt(1.0 - mu * mu);
    for (int j = 0; j < per_quadrant; ++j) {
      const double phi = kPi * (2 * j + 1) / n_azim;
      omega.push_back(sin_t * std::cos(phi));
      omega.push_back(sin_t * std::sin(phi));
      omega.push_back(mu);
      w.push_back(gw[i] * 2.0 * kPi / n_azim);
    }
  }
}

Quadrature build_quadrature(int kind, int dim, int sn, int n_polar, int n_azim,
                            int L_max, bool galerkin, double weight_sum) {
  if (dim != 2 && dim != 3) fail(OU_ERR_ARG, "dim must be 2 or 3");
  if (L_max < 0) fail(OU_ERR_ARG, "L_max must be >= 0");
  if (!(weight_sum > 0.0)) fail(OU_ERR_ARG, "weight_sum must be positive");
  if (galerkin && dim == 3)
    fail(OU_ERR_UNSUPPORTED, "Galerkin quadratures are only defined in 2D");
  std::vector<double> oct, w_oct;
  if (kind == OU_QUAD_LS) {
    if (!ls_octant(sn, oct, w_oct))
      fail(OU_ERR_UNSUPPORTED, "level-symmetric order must be one of 2,4,..,12");
  } else if (kind == OU_QUAD_GLC) {
    if (sn < 2 || sn % 2) fail(OU_ERR_ARG, "GLC order must be even and >= 2");
    glc_octant(sn, oct, w_oct);
  } else if (kind == OU_QUAD_PRODUCT_GLC) {
    if (n_polar < 2 || n_polar % 2 || n_azim < 4 || n_azim % 4)
      fail(OU_ERR_ARG, "product GLC needs even n_polar and n_azim divisible by 4");
    if (galerkin) fail(OU_ERR_UNSUPPORTED, "Galerkin product quadrature");
    product_glc_octant(n_polar, n_azim, oct, w_oct);
  } else {
    fail(OU_ERR_ARG, "unknown quadrature kind");
  }
  const int n_oct = int(w_oct.size());
  const int n_blocks = dim == 2 ? 4 : 8;
  Quadrature q;
  q.dim = dim;
  q.n_dir = n_blocks * n_oct;
  q.omega.resize(3 * q.n_dir);
  q.weight.resize(q.n_dir);
  // Quadrant order of RTQuadrature::deploy_octant (RTQuadrature.cc:58-95):
  // (+,+), (+,-), (-,-), (-,+); in 3D the same four again with mu < 0.
  const double sgn_x[4] = {1, 1, -1, -1}, sgn_y[4] = {1, -1, -1, 1};
  double w_sum = 0.0;
  for (double v : w_oct) w_sum += v;
  for (int b = 0; b < n_blocks; ++b)
    for (int j = 0; j < n_oct; ++j) {
      const int d = b * n_oct + j;
      q.omega[3 * d + 0] = sgn_x[b % 4] * oct[3 * j + 0];
      q.omega[3 * d + 1] = sgn_y[b % 4] * oct[3 * j + 1];
      q.omega[3 * d + 2] = (b < 4 ? 1.0 : -1.0) * oct[3 * j + 2];
      q.weight[d] = w_oct[j] / (w_sum * n_blocks);
    }

  // Moment list: per l, cosine harmonics m = l..0 then sine harmonics
  // m = 1..l (RTQuadrature.cc:168-196); 2D keeps the (l+m)-even ones.
  struct Mom { int l, m; bool sine; };
  std::vector<Mom> moms;
  for (int l = 0; l <= L_max; ++l) {
    for (int m = l; m >= 0; --m) {
      bool keep = dim == 3 || (l + m) % 2 == 0;
      if (galerkin) keep = l < sn && (l + m) % 2 == 0;
      if (keep) moms.push_back({l, m, false});
    }
    for (int m = 1; m <= l; ++m) {
      bool keep = dim == 3 || (l + m) % 2 == 0;
      if (galerkin) keep = l <= sn && (l + m) % 2 == 0;
      if (keep) moms.push_back({l, m, true});
    }
  }
  q.n_mom = int(moms.size());
  if (galerkin && q.n_mom != q.n_dir)
    fail(OU_ERR_ARG, "Galerkin quadrature needs as many moments as directions "
                     "(choose L_max = sn)");
  q.mom_l.resize(q.n_mom);
  for (int k = 0; k < q.n_mom; ++k) q.mom_l[k] = moms[k].l;
  q.M2D.assign(size_t(q.n_dir) * q.n_mom, 0.0);
  const double scale = std::sqrt(4.0 * kPi / weight_sum);
  std::vector<double> P;
  for (int d = 0; d < q.n_dir; ++d) {
    const double mu = q.omega[3 * d + 2];
    const double phi = std::atan2(q.omega[3 * d + 1], q.omega[3 * d + 0]);
    normalized_legendre(L_max, mu, P);
    for (int k = 0; k < q.n_mom; ++k) {
      const Mom &mo = moms[k];
      const double pl = P[mo.l * (L_max + 1) + mo.m];
      double y = pl;
      if (mo.m > 0)
        y = std::sqrt(2.0) * pl * (mo.sine ? std::sin(mo.m * phi) : std::cos(mo.m * phi));
      q.M2D[size_t(d) * q.n_mom + k] = scale * y;
    }
  }
  if (galerkin) {
    if (!invert(q.n_dir, q.M2D, q.D2M))
      fail(OU_ERR_ARG, "Galerkin M2D is singular");
  } else {
    q.D2M.assign(size_t(q.n_mom) * q.n_dir, 0.0);
    for (int k = 0; k < q.n_mom; ++k)
      for (int d = 0; d < q.n_dir; ++d)
        q.D2M[size_t(k) * q.n_dir + d] = q.M2D[size_t(d) * q.n_mom + k] * weight_sum * q.weight[d];
  }
  // Most-normal direction sets (RTQuadrature.cc:200-238): face 2a holds the
  // directions with the largest Omega_a, face 2a+1 those with the smallest.
  q.most_normal.assign(size_t(2 * dim) * q.n_dir, 0);
  for (int a = 0; a < dim; ++a) {
    double hi = 0.0, lo = 0.0;
    for (int d = 0; d < q.n_dir; ++d) {
      hi = std::fmax(hi, q.omega[3 * d + a]);
      lo = std::fmin(lo, q.omega[3 * d + a]);
    }
    for (int d = 0; d < q.n_dir; ++d) {
      q.most_normal[size_t(2 * a) * q.n_dir + d] = q.omega[3 * d + a] == hi;
      q.most_normal[size_t(2 * a + 1) * q.n_dir + d] = q.omega[3 * d + a] == lo;
    }
  }
  return q;
}

}  // namespace ou
