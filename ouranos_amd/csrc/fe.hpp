// Reference-element DG matrices on axis-aligned cells.
//
// The reference builds FECell matrices per cell with FEValues and QGauss(p+1)
// (src/FECell.cc:10-78).  On a uniform Cartesian cell every one of them is a
// Kronecker product of three 1D matrices on [0,1] with equidistant Lagrange
// nodes (the FE_DGQ support points):
//   m[i][j] = int L_i L_j,   k[i][j] = int L_j L_i'   (exact Gauss rule)
// so that, with l = lx + (p+1)(ly + (p+1) lz):
//   M        = (hz m) x (hy m) x (hx m)
//   G_x      = (hz m) x (hy m) x k            (and cyclically)
//   E_{2a+s} = e_s e_s^T along a, (h_b m) on the other axes
//   U_{2a+s} = e_s e_{1-s}^T along a, (h_b m) on the other axes
// with e_0 = (1,0,..), e_1 = (..,0,1) the nodal values at t = 0 and t = 1.
#pragma once

#include <vector>

namespace ou {

struct RefElement {
  int dim = 2, p = 1, P1 = 2, N = 4, F = 2;
  double h[3] = {1, 1, 1};
  std::vector<double> m1;  // P1 x P1
  std::vector<double> k1;  // P1 x P1

  RefElement(int dim, int p, const double *h);

  // Full N x N matrices (row-major).
  void mass(std::vector<double> &M) const;
  void grad(int axis, std::vector<double> &G) const;
  void face(int f, bool upwind, std::vector<double> &E) const;
  // F x F mass of the face normal to `axis` (area-scaled); the face dofs are
  // ordered by the remaining local indices, lower axis fastest.
  void face_mass(int axis, std::vector<double> &Mf) const;
};

}  // namespace ou
