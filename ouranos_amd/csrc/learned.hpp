// The learned preconditioner around the U-Net: coarse-grid transfer between
// a 3D transport plan and the network's G^3 cell grid, the input features,
// the FGMRES right preconditioner, and the iterative-refinement driver that
// runs the FP32 sweep under FP64 residuals.
#pragma once

#include <vector>

#include "radiative_transfer.hpp"
#include "unet.hpp"

namespace ou {

// Fine cells per coarse cell along each axis (n / G, equal on all axes).
int coarse_factor(const Plan &pl, int G);

// Per coarse cell (x fastest), three FP32 features:
//   tau_hat   = ln(sigma_t H) / 4
//   kappa_hat = (ln(max(1 - c, 1e-6)) + 4.6) / 4.6
//   g         = c theta / (1 - c theta),  theta = (sigma_t H)^2 / (1 + (sigma_t H)^2)
//               (the gain of the scattered part K e of a smooth error e over its
//               residual (I - K) e when the collision probability is theta)
// with sigma_t the mean over the coarse cell's fine cells, c = sum sigma_s0 /
// sum sigma_t and H the coarse cell edge.
std::vector<float> coarse_material_features(const Plan &pl, int G);

// d_coarse[C][k] = mean of the moment-0 nodal values of d_v in octant
// k = kx + 2 ky + 4 kz of coarse cell C (nodes on a mid-plane count half to
// each side).
void coarse_restrict(const Plan &pl, int G, const double *d_v, double *d_coarse, hipStream_t s);

// d_z[moment 0] += s P(corners): corner values [G^3][8] averaged at the shared
// coarse vertices into d_vertex ((G+1)^3), then interpolated trilinearly to
// every fine node; s = sqrt(d_dot[0] / (8 G^3)).
void coarse_prolong_add(const Plan &pl, int G, const float *d_corners, const double *d_dot,
                        double *d_vertex, double *d_z, hipStream_t s);

// The same with s = 1, allocating nothing: the vertex averages and the scale
// live in a fixed device array of the library (grids up to
// coarse_noalloc_max_grid(), else OU_ERR_SIZE).  Stream-ordered; calls that
// overlap on one device share that array and race.
int coarse_noalloc_max_grid();
void coarse_prolong_add_noalloc(const Plan &pl, int G, const float *d_corners, double *d_z,
                                hipStream_t s);

// Right preconditioner z = r + P s N_odd(R r / s) on moment 0:
// R = coarse_restrict (8 octant values per coarse cell), s = RMS of R r,
// N_odd the network's odd prediction of the correction beyond the identity
// at the cell's eight corners (A^-1 r - r = K A^-1 r, the scattered part),
// P = corner values averaged at the shared coarse vertices, then trilinear
// interpolation to every fine node (a continuous correction field).  Higher
// moments pass through, and a network that predicts zero leaves plain GMRES.
class LearnedPreconditioner : public Preconditioner {
 public:
  LearnedPreconditioner(RadiativeTransfer &rt, UNet &net);
  ~LearnedPreconditioner() override;
  void apply(const double *d_r, double *d_z, hipStream_t s) override;
  int applications() const { return applications_; }

 private:
  void release();
  Plan &pl_;
  UNet &net_;
  int G_, f_;
  long P0_;
  double *coarse_ = nullptr, *vertex_ = nullptr, *dot_ = nullptr, *scratch_ = nullptr;
  float *mat_ = nullptr, *feat_ = nullptr, *pred_ = nullptr;
  int applications_ = 0;
};

struct RefinedResult {
  SolveResult total;  // iterations = FP32 (inner) operator applications
  int outer = 0;      // FP64 residual evaluations (operator applications; a zero
                      // initial guess needs none for the first cycle)
  double seconds_fp64 = 0.0;
};

// Solves (I - D L^-1 M Sigma) phi = D L^-1 (q + inflow) to `tol` in the FP64
// residual of rt64.  Each cycle computes r = b - A64 phi in FP64 and solves
// A32 d = r with FGMRES(restart) on the FP32 operator rt32 to `inner_tol`
// relative, right-preconditioned by `prec` (or unpreconditioned GMRES if
// null), then phi += d.
RefinedResult solve_refined(RadiativeTransfer &rt64, RadiativeTransfer &rt32,
                            Preconditioner *prec, double *d_phi, double tol, int max_it,
                            int restart, double inner_tol, bool verbose);

}  // namespace ou
