// C++ host interface of the engine, shaped like the reference's
// Epetra_Operator (Trilinos) and RadiativeTransfer<dim,tensor_dim>
// (src/RadiativeTransfer.hh:46-127): Apply / Label / domain and range sizes,
// plus the sweep primitives source iteration calls.  Vectors are FP64 device
// pointers of DomainSize() doubles in the moment-major layout.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>

#include "plan.hpp"
#include "vec.hpp"

namespace ou {

class Operator {
 public:
  virtual ~Operator() = default;
  // y = Op(x); returns OU_OK or an ou_status.
  virtual int Apply(const double *d_x, double *d_y, hipStream_t s) = 0;
  // Not provided by the transport operator (src/RadiativeTransfer.hh:97-114).
  virtual int ApplyInverse(const double *, double *, hipStream_t) { return OU_ERR_UNSUPPORTED; }
  virtual int SetUseTranspose(bool) { return OU_ERR_UNSUPPORTED; }
  virtual bool UseTranspose() const { return false; }
  virtual bool HasNormInf() const { return false; }
  virtual double NormInf() const { return 0.0; }
  virtual const char *Label() const = 0;
  virtual long DomainSize() const = 0;
  virtual long RangeSize() const { return DomainSize(); }
};

// Right preconditioner of FGMRES: z = P(r) (may vary between calls).
class Preconditioner {
 public:
  virtual ~Preconditioner() = default;
  virtual void apply(const double *d_r, double *d_z, hipStream_t s) = 0;
};

class RadiativeTransfer : public Operator {
 public:
  explicit RadiativeTransfer(const ou_problem_desc &desc) : plan_(create_plan(&desc)) {}
  int Apply(const double *d_x, double *d_y, hipStream_t s) override {
    apply_device(*plan_, d_x, d_y, s ? s : plan_->stream);
    return OU_OK;
  }
  const char *Label() const override { return "radiative_transfer"; }
  long DomainSize() const override { return plan_->vec_size; }

  // compute_scattering_source (src/RadiativeTransfer.cc:811-843)
  void compute_scattering_source(const double *d_x, hipStream_t s) {
    launch_scatter_source(*plan_, d_x, s ? s : plan_->stream);
  }
  // The scheduler loop of sweep() calls (src/RadiativeTransfer.cc:605-1012):
  // d_y += D2M L^-1 (M2D S [+ source + inflow]).
  void sweep(double *d_y, bool with_source, hipStream_t s) {
    sweep_device(*plan_, d_y, with_source, s ? s : plan_->stream);
  }
  void clear_scattering_source(hipStream_t s) {
    vec_zero_bytes(plan_->d_S, size_t(plan_->vec_size) * plan_->real_size(),
                   s ? s : plan_->stream);
  }
  void set_group(int g) {
    if (g != 0) fail(OU_ERR_UNSUPPORTED, "one energy group is supported (g must be 0)");
    plan_->group = g;
  }
  Plan &plan() { return *plan_; }
  const Plan &plan() const { return *plan_; }

 private:
  std::unique_ptr<Plan> plan_;
};

struct SolveResult {
  int iterations = 0;
  double rel_residual = 0.0;
  bool converged = false;
};

// Drivers of src/ouranos.cc:43-209 on device vectors (FP64 Krylov).
SolveResult solve_si(RadiativeTransfer &rt, double *d_phi, double tol, int max_it,
                     bool verbose);
// Builds the right-hand side D L^-1 (q + inflow) (src/ouranos.cc:149-158).
void build_rhs(RadiativeTransfer &rt, double *d_rhs, hipStream_t s);
// Called after every Arnoldi application with (z_j, A z_j) on the device
// (z_j = v_j without a preconditioner) and the iteration count.
struct ArnoldiObserver {
  virtual ~ArnoldiObserver() = default;
  virtual void observe(int iteration, const double *d_z, const double *d_az, hipStream_t s) = 0;
};

// The Krylov basis of solve_gmres: restart + 1 vectors V and, for the
// flexible (preconditioned) form, restart vectors Z.  At the headline size
// that is 60 GB; a caller that runs several solves of one size (the
// refinement cycles) allocates it once instead of once per solve.  With
// `f32_basis` the vectors are FP32 (for solve_gmres_f32_basis), plus three
// FP64 work vectors.
class GmresWorkspace {
 public:
  GmresWorkspace(long n, int restart, bool flexible, bool f32_basis = false);
  ~GmresWorkspace();
  GmresWorkspace(const GmresWorkspace &) = delete;
  GmresWorkspace &operator=(const GmresWorkspace &) = delete;
  bool fits(long n, int restart, bool flexible, bool f32_basis = false) const {
    return n == n_ && restart == restart_ && f32_basis == f32_ &&
           (flexible ? Z_ != nullptr : true);
  }
  double *V() const { return f32_ ? nullptr : static_cast<double *>(V_); }
  double *Z() const { return f32_ ? nullptr : static_cast<double *>(Z_); }
  float *Vf() const { return f32_ ? static_cast<float *>(V_) : nullptr; }
  float *Zf() const { return f32_ ? static_cast<float *>(Z_) : nullptr; }
  double *work() const { return work_; }  // 3 n doubles (FP32 basis only)

 private:
  long n_;
  int restart_;
  bool f32_;
  void *V_ = nullptr, *Z_ = nullptr;
  double *work_ = nullptr;
};

SolveResult solve_gmres(Operator &op, const double *d_b, double *d_x, int restart,
                        double tol, int max_it, Preconditioner *prec, bool verbose,
                        hipStream_t s, ArnoldiObserver *observer = nullptr,
                        GmresWorkspace *ws = nullptr);
// The same restarted (flexible) GMRES with its Krylov basis stored in FP32:
// every Gram-Schmidt pass reads half the bytes.  The operator's input and
// output, the vector being orthogonalised, every inner product and the
// iterate stay FP64; the stored basis carries a relative rounding of 2^-24.
// It is for inner solves whose tolerance is far above that (the FP32-operator
// solve of the refinement), not for an FP64 solve to 1e-12.
SolveResult solve_gmres_f32_basis(Operator &op, const double *d_b, double *d_x, int restart,
                                  double tol, int max_it, Preconditioner *prec, bool verbose,
                                  hipStream_t s, GmresWorkspace *ws = nullptr);
SolveResult solve_bicgstab(Operator &op, const double *d_b, double *d_x, double tol,
                           int max_it, bool verbose, hipStream_t s);
// Source iteration on A x = b: x <- x - A x + b (= D L^-1 (M Sigma x + the
// sources b was swept from)), stopping when |x_new - x| / |x_new| < tol.
SolveResult solve_si_rhs(Operator &op, const double *d_b, double *d_x, double tol, int max_it,
                         bool verbose, hipStream_t s);

}  // namespace ou
