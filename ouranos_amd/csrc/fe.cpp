// Reference-element DG matrices; see fe.hpp.
#include "fe.hpp"

#include "ou_common.hpp"
#include "quadrature.hpp"

namespace ou {

namespace {

double lagrange(int p, int k, double t) {
  double v = 1.0;
  for (int j = 0; j <= p; ++j)
    if (j != k) v *= (t * p - j) / double(k - j);
  return v;
}

double lagrange_deriv(int p, int k, double t) {
  double s = 0.0;
  for (int i = 0; i <= p; ++i) {
    if (i == k) continue;
    double v = double(p) / double(k - i);
    for (int j = 0; j <= p; ++j)
      if (j != k && j != i) v *= (t * p - j) / double(k - j);
    s += v;
  }
  return s;
}

}  // namespace

RefElement::RefElement(int dim_, int p_, const double *h_) : dim(dim_), p(p_) {
  if (dim != 2 && dim != 3) fail(OU_ERR_ARG, "dim must be 2 or 3");
  if (p < 1 || p > 5) fail(OU_ERR_ARG, "FE order must be in 1..5");
  P1 = p + 1;
  N = dim == 2 ? P1 * P1 : P1 * P1 * P1;
  F = dim == 2 ? P1 : P1 * P1;
  for (int a = 0; a < dim; ++a) {
    if (!(h_[a] > 0.0)) fail(OU_ERR_ARG, "cell sizes must be positive");
    h[a] = h_[a];
  }
  // Exact for the degree-2p integrands with p+2 Gauss points.
  std::vector<double> x, w;
  gauss_legendre(p + 2, x, w);
  m1.assign(P1 * P1, 0.0);
  k1.assign(P1 * P1, 0.0);
  for (size_t q = 0; q < x.size(); ++q) {
    const double t = 0.5 * (x[q] + 1.0), wq = 0.5 * w[q];
    for (int i = 0; i < P1; ++i)
      for (int j = 0; j < P1; ++j) {
        m1[i * P1 + j] += wq * lagrange(p, i, t) * lagrange(p, j, t);
        k1[i * P1 + j] += wq * lagrange(p, j, t) * lagrange_deriv(p, i, t);
      }
  }
}

void RefElement::mass(std::vector<double> &M) const {
  M.assign(size_t(N) * N, 0.0);
  for (int I = 0; I < N; ++I)
    for (int J = 0; J < N; ++J) {
      double v = 1.0;
      for (int a = 0, ri = I, rj = J; a < dim; ++a, ri /= P1, rj /= P1)
        v *= h[a] * m1[(ri % P1) * P1 + rj % P1];
      M[size_t(I) * N + J] = v;
    }
}

void RefElement::grad(int axis, std::vector<double> &G) const {
  G.assign(size_t(N) * N, 0.0);
  for (int I = 0; I < N; ++I)
    for (int J = 0; J < N; ++J) {
      double v = 1.0;
      for (int a = 0, ri = I, rj = J; a < dim; ++a, ri /= P1, rj /= P1) {
        const int i = ri % P1, j = rj % P1;
        v *= a == axis ? k1[i * P1 + j] : h[a] * m1[i * P1 + j];
      }
      G[size_t(I) * N + J] = v;
    }
}

void RefElement::face(int f, bool upwind, std::vector<double> &E) const {
  const int axis = f / 2, side = f % 2;
  const int ni = side ? p : 0;                  // this cell's face nodes
  const int nj = upwind ? (side ? 0 : p) : ni;  // neighbour's (or own) nodes
  E.assign(size_t(N) * N, 0.0);
  for (int I = 0; I < N; ++I)
    for (int J = 0; J < N; ++J) {
      double v = 1.0;
      for (int a = 0, ri = I, rj = J; a < dim; ++a, ri /= P1, rj /= P1) {
        const int i = ri % P1, j = rj % P1;
        v *= a == axis ? double(i == ni && j == nj) : h[a] * m1[i * P1 + j];
      }
      E[size_t(I) * N + J] = v;
    }
}

void RefElement::face_mass(int axis, std::vector<double> &Mf) const {
  Mf.assign(size_t(F) * F, 0.0);
  int others[2], n_o = 0;
  for (int a = 0; a < dim; ++a)
    if (a != axis) others[n_o++] = a;
  for (int I = 0; I < F; ++I)
    for (int J = 0; J < F; ++J) {
      double v = 1.0;
      for (int o = 0, ri = I, rj = J; o < n_o; ++o, ri /= P1, rj /= P1)
        v *= h[others[o]] * m1[(ri % P1) * P1 + rj % P1];
      Mf[size_t(I) * F + J] = v;
    }
}

}  // namespace ou
