// Angular quadratures and the moment operators M2D / D2M.
//
// Mirrors RTQuadrature + GLC + LS of the reference (src/RTQuadrature.hh:26-98,
// src/GLC.hh:21-41, src/LS.hh:19-30): an octant rule is built, mirrored into
// the quadrants (2D, mu > 0) or octants (3D), and real spherical harmonics
// give M2D[d][m] = sqrt(4 pi / ws) Y_m(Omega_d) and D2M = M2D^T diag(ws w)
// (or M2D^-1 for a Galerkin rule).
#pragma once

#include <vector>

namespace ou {

struct Quadrature {
  int dim = 2;
  int n_dir = 0;
  int n_mom = 0;
  std::vector<double> omega;        // n_dir x 3
  std::vector<double> weight;       // n_dir, sums to 1
  std::vector<double> M2D;          // n_dir x n_mom
  std::vector<double> D2M;          // n_mom x n_dir
  std::vector<int> mom_l;           // n_mom
  std::vector<unsigned char> most_normal;  // 2*dim x n_dir
};

// Gauss-Legendre rule of n points on [-1,1], nodes ascending.
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w);

// Octant rules (one octant: Omega_x, Omega_y, Omega_z > 0).
bool ls_octant(int sn, std::vector<double> &omega, std::vector<double> &w);
void glc_octant(int sn, std::vector<double> &omega, std::vector<double> &w);
void product_glc_octant(int n_polar, int n_azim, std::vector<double> &omega,
                        std::vector<double> &w);

// Full construction (kind: 0 LS, 1 triangular GLC, 2 product GLC).
Quadrature build_quadrature(int kind, int dim, int sn, int n_polar, int n_azim,
                            int L_max, bool galerkin, double weight_sum);

}  // namespace ou
