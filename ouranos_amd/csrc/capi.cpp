// extern "C" entry points of include/ouranos_amd.h.  No exception crosses
// this boundary: every call returns OU_OK or a negative ou_status and keeps
// the message for ou_last_error().
#include <algorithm>
#include <initializer_list>
#include <chrono>
#include <cstring>
#include <string>

#include "batch.hpp"
#include "fe.hpp"
#include "inp.hpp"
#include "learned.hpp"
#include "quadrature.hpp"
#include "radiative_transfer.hpp"
#include "unet.hpp"
#include "vec.hpp"

struct ou_plan {
  std::unique_ptr<ou::RadiativeTransfer> rt;
};

// The network plus device staging for the host-array entry points.
struct ou_unet {
  std::unique_ptr<ou::UNet> net;
  float *d_feat = nullptr, *d_target = nullptr, *d_pred = nullptr;
  ~ou_unet() {
    for (float *p : {d_feat, d_target, d_pred})
      if (p) (void)hipFree(p);
  }
};

struct ou_quadrature {
  ou::Quadrature q;
};

struct ou_batch {
  std::unique_ptr<ou::Batch> b;
};

namespace ou {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace ou

namespace {

template <typename F>
int guarded(F &&f) {
  try {
    ou::set_last_error("");
    return f();
  } catch (const ou::Error &e) {
    ou::set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    ou::set_last_error("host allocation failed");
    return OU_ERR_NOMEM;
  } catch (const std::exception &e) {
    ou::set_last_error(e.what());
    return OU_ERR_ARG;
  }
}

ou::Plan &plan_of(ou_plan *p) {
  if (!p || !p->rt) ou::fail(OU_ERR_ARG, "null plan");
  return p->rt->plan();
}

ou::Batch &batch_of(const ou_batch *b) {
  if (!b || !b->b) ou::fail(OU_ERR_ARG, "null batch");
  return *b->b;
}

hipStream_t stream_of(ou_plan *p, void *s) {
  return s ? static_cast<hipStream_t>(s) : plan_of(p).stream;
}

void ensure_staging(ou::Plan &pl) {
  if (!pl.d_x) OU_HIP(hipMalloc(&pl.d_x, size_t(pl.vec_size) * 8));
  if (!pl.d_y) OU_HIP(hipMalloc(&pl.d_y, size_t(pl.vec_size) * 8));
}

// OU_ERR_TIMEOUT if the plan's persistent sweep gave up since the last check.
void check_watchdog(ou::Plan &pl) {
  if (pl.sweep_mode == OU_SWEEP_PERSISTENT) ou::check_sweep_watchdog(pl.d_error, pl.stream);
}

// Runs a solve; a watchdog trip of the persistent sweep makes the operator's
// output NaN, which the solver reports as non-finite: that is turned back into
// the watchdog's OU_ERR_TIMEOUT here, and any trip is reported after the solve.
template <typename F>
auto with_watchdog(std::initializer_list<ou::Plan *> plans, F &&f) -> decltype(f()) {
  try {
    auto r = f();
    for (ou::Plan *p : plans) check_watchdog(*p);
    return r;
  } catch (const ou::Error &e) {
    if (e.code == OU_ERR_NONFINITE)
      for (ou::Plan *p : plans) check_watchdog(*p);
    throw;
  }
}

void check_finite_host(const double *x, long n) {
  for (long i = 0; i < n; ++i)
    if (!std::isfinite(x[i])) ou::fail(OU_ERR_NONFINITE, "non-finite entry in the input vector");
}

}  // namespace

extern "C" {

const char *ou_last_error(void) { return ou::g_last_error.c_str(); }

int ou_quadrature_create(int kind, int dim, int sn, int n_polar, int n_azim, int L_max,
                         int galerkin, double weight_sum, ou_quadrature **quad) {
  return guarded([&] {
    if (!quad) ou::fail(OU_ERR_ARG, "null output pointer");
    auto *q = new ou_quadrature();
    try {
      q->q = ou::build_quadrature(kind, dim, sn, n_polar, n_azim, L_max, galerkin != 0,
                                  weight_sum);
    } catch (...) {
      delete q;
      throw;
    }
    *quad = q;
    return int(OU_OK);
  });
}

void ou_quadrature_destroy(ou_quadrature *quad) { delete quad; }

int ou_quadrature_n_dir(const ou_quadrature *quad) { return quad ? quad->q.n_dir : OU_ERR_ARG; }
int ou_quadrature_n_mom(const ou_quadrature *quad) { return quad ? quad->q.n_mom : OU_ERR_ARG; }

int ou_quadrature_copy(const ou_quadrature *quad, double *omega, double *weight, double *M2D,
                       double *D2M, int *mom_l, unsigned char *most_normal) {
  return guarded([&] {
    if (!quad) ou::fail(OU_ERR_ARG, "null quadrature");
    const ou::Quadrature &q = quad->q;
    if (omega) std::memcpy(omega, q.omega.data(), q.omega.size() * 8);
    if (weight) std::memcpy(weight, q.weight.data(), q.weight.size() * 8);
    if (M2D) std::memcpy(M2D, q.M2D.data(), q.M2D.size() * 8);
    if (D2M) std::memcpy(D2M, q.D2M.data(), q.D2M.size() * 8);
    if (mom_l) std::memcpy(mom_l, q.mom_l.data(), q.mom_l.size() * sizeof(int));
    if (most_normal) std::memcpy(most_normal, q.most_normal.data(), q.most_normal.size());
    return int(OU_OK);
  });
}

int ou_fe_matrices(int dim, int fe_order, const double *h, double *M, double *G, double *E,
                   double *U, double *normals) {
  return guarded([&] {
    if (!h) ou::fail(OU_ERR_ARG, "null cell size");
    ou::RefElement re(dim, fe_order, h);
    const size_t NN = size_t(re.N) * re.N;
    std::vector<double> t;
    if (M) {
      re.mass(t);
      std::memcpy(M, t.data(), NN * 8);
    }
    for (int a = 0; a < dim && G; ++a) {
      re.grad(a, t);
      std::memcpy(G + a * NN, t.data(), NN * 8);
    }
    for (int f = 0; f < 2 * dim; ++f) {
      if (E) {
        re.face(f, false, t);
        std::memcpy(E + f * NN, t.data(), NN * 8);
      }
      if (U) {
        re.face(f, true, t);
        std::memcpy(U + f * NN, t.data(), NN * 8);
      }
      if (normals) {
        for (int c = 0; c < 3; ++c) normals[3 * f + c] = 0.0;
        normals[3 * f + f / 2] = (f & 1) ? 1.0 : -1.0;
      }
    }
    return int(OU_OK);
  });
}

int ou_plan_create(const ou_problem_desc *desc, ou_plan **plan) {
  return guarded([&] {
    if (!plan) ou::fail(OU_ERR_ARG, "null output pointer");
    if (!desc) ou::fail(OU_ERR_ARG, "null descriptor");
    auto *p = new ou_plan();
    try {
      p->rt.reset(new ou::RadiativeTransfer(*desc));
    } catch (...) {
      delete p;
      throw;
    }
    *plan = p;
    return int(OU_OK);
  });
}

void ou_plan_destroy(ou_plan *plan) { delete plan; }

int ou_set_group(ou_plan *plan, int g) {
  return guarded([&] {
    plan_of(plan);
    plan->rt->set_group(g);
    return int(OU_OK);
  });
}

long ou_domain_size(const ou_plan *plan) {
  return plan && plan->rt ? plan->rt->DomainSize() : long(OU_ERR_ARG);
}
long ou_range_size(const ou_plan *plan) {
  return plan && plan->rt ? plan->rt->RangeSize() : long(OU_ERR_ARG);
}
long ou_n_dofs(const ou_plan *plan) {
  return plan && plan->rt ? plan->rt->plan().n_dofs : long(OU_ERR_ARG);
}
const char *ou_label(const ou_plan *plan) {
  return plan && plan->rt ? plan->rt->Label() : "";
}

int ou_apply(ou_plan *plan, const double *x, double *y) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (!x || !y) ou::fail(OU_ERR_ARG, "null vector");
    check_finite_host(x, pl.vec_size);
    ensure_staging(pl);
    const size_t bytes = size_t(pl.vec_size) * 8;
    OU_HIP(hipMemcpyAsync(pl.d_x, x, bytes, hipMemcpyHostToDevice, pl.stream));
    plan->rt->Apply(pl.d_x, pl.d_y, pl.stream);
    OU_HIP(hipMemcpyAsync(y, pl.d_y, bytes, hipMemcpyDeviceToHost, pl.stream));
    OU_HIP(hipStreamSynchronize(pl.stream));
    check_watchdog(pl);
    return int(OU_OK);
  });
}

int ou_apply_multi(ou_plan *plan, const double *x, double *y, int n_rhs) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (!x || !y) ou::fail(OU_ERR_ARG, "null vector");
    if (n_rhs < 1) ou::fail(OU_ERR_ARG, "n_rhs must be positive");
    const size_t n = size_t(pl.vec_size), bytes = n * 8;
    check_finite_host(x, long(n) * n_rhs);
    ensure_staging(pl);
    for (int r = 0; r < n_rhs; ++r) {
      OU_HIP(hipMemcpyAsync(pl.d_x, x + r * n, bytes, hipMemcpyHostToDevice, pl.stream));
      plan->rt->Apply(pl.d_x, pl.d_y, pl.stream);
      OU_HIP(hipMemcpyAsync(y + r * n, pl.d_y, bytes, hipMemcpyDeviceToHost, pl.stream));
    }
    OU_HIP(hipStreamSynchronize(pl.stream));
    check_watchdog(pl);
    return int(OU_OK);
  });
}

int ou_apply_device(ou_plan *plan, const double *d_x, double *d_y, void *stream) {
  return guarded([&] {
    plan_of(plan);
    if (!d_x || !d_y) ou::fail(OU_ERR_ARG, "null vector");
    if (d_x == d_y) ou::fail(OU_ERR_ARG, "x and y must be distinct vectors");
    return plan->rt->Apply(d_x, d_y, stream_of(plan, stream));
  });
}

int ou_scatter_source(ou_plan *plan, const double *x) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (!x) ou::fail(OU_ERR_ARG, "null vector");
    check_finite_host(x, pl.vec_size);
    ensure_staging(pl);
    OU_HIP(hipMemcpyAsync(pl.d_x, x, size_t(pl.vec_size) * 8, hipMemcpyHostToDevice,
                          pl.stream));
    plan->rt->compute_scattering_source(pl.d_x, pl.stream);
    OU_HIP(hipStreamSynchronize(pl.stream));
    return int(OU_OK);
  });
}

int ou_scatter_source_device(ou_plan *plan, const double *d_x, void *stream) {
  return guarded([&] {
    plan_of(plan);
    if (!d_x) ou::fail(OU_ERR_ARG, "null vector");
    plan->rt->compute_scattering_source(d_x, stream_of(plan, stream));
    return int(OU_OK);
  });
}

int ou_sweep(ou_plan *plan, double *y, int with_source) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (!y) ou::fail(OU_ERR_ARG, "null vector");
    ensure_staging(pl);
    const size_t bytes = size_t(pl.vec_size) * 8;
    OU_HIP(hipMemcpyAsync(pl.d_y, y, bytes, hipMemcpyHostToDevice, pl.stream));
    plan->rt->sweep(pl.d_y, with_source != 0, pl.stream);
    OU_HIP(hipMemcpyAsync(y, pl.d_y, bytes, hipMemcpyDeviceToHost, pl.stream));
    OU_HIP(hipStreamSynchronize(pl.stream));
    check_watchdog(pl);
    return int(OU_OK);
  });
}

int ou_sweep_device(ou_plan *plan, double *d_y, int with_source, void *stream) {
  return guarded([&] {
    plan_of(plan);
    if (!d_y) ou::fail(OU_ERR_ARG, "null vector");
    plan->rt->sweep(d_y, with_source != 0, stream_of(plan, stream));
    return int(OU_OK);
  });
}

int ou_set_use_transpose(ou_plan *plan, int use_transpose) {
  return guarded([&] {
    plan_of(plan);
    if (use_transpose) ou::fail(OU_ERR_UNSUPPORTED, "the transpose operator is not provided");
    return int(OU_OK);
  });
}

int ou_apply_inverse(ou_plan *plan, const double *, double *) {
  return guarded([&] {
    plan_of(plan);
    ou::fail(OU_ERR_UNSUPPORTED, "ApplyInverse is not provided by the transport operator");
    return int(OU_OK);
  });
}

int ou_use_transpose(const ou_plan *) { return 0; }
int ou_has_norm_inf(const ou_plan *) { return 0; }
double ou_norm_inf(const ou_plan *) { return 0.0; }

int ou_solve(ou_plan *plan, const ou_solver_opts *opts, double *phi, ou_solve_info *info) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (!opts || !phi) ou::fail(OU_ERR_ARG, "null argument");
    if (!(opts->tol > 0.0) || opts->max_it < 1) ou::fail(OU_ERR_ARG, "bad tolerance or max_it");
    check_finite_host(phi, pl.vec_size);
    ou::RadiativeTransfer &rt = *plan->rt;
    const long n = pl.vec_size;
    const size_t bytes = size_t(n) * 8;
    double *d_phi = nullptr, *d_rhs = nullptr;
    OU_HIP(hipMalloc(&d_phi, bytes));
    struct Free {
      double **a, **b;
      ~Free() {
        if (*a) (void)hipFree(*a);
        if (*b) (void)hipFree(*b);
      }
    } guard{&d_phi, &d_rhs};
    OU_HIP(hipMemcpy(d_phi, phi, bytes, hipMemcpyHostToDevice));
    const auto t0 = std::chrono::steady_clock::now();
    ou::SolveResult r;
    hipStream_t s = pl.stream;
    r = with_watchdog({&pl}, [&] {
      switch (opts->solver) {
        case OU_SOLVER_SI:
          return ou::solve_si(rt, d_phi, opts->tol, opts->max_it, opts->verbose != 0);
        case OU_SOLVER_GMRES:
        case OU_SOLVER_FGMRES:
        case OU_SOLVER_BICGSTAB:
          OU_HIP(hipMalloc(&d_rhs, bytes));
          ou::build_rhs(rt, d_rhs, s);
          if (opts->solver == OU_SOLVER_BICGSTAB)
            return ou::solve_bicgstab(rt, d_rhs, d_phi, opts->tol, opts->max_it,
                                      opts->verbose != 0, s);
          return ou::solve_gmres(rt, d_rhs, d_phi, opts->restart, opts->tol, opts->max_it,
                                 nullptr, opts->verbose != 0, s);
        default:
          ou::fail(OU_ERR_ARG, "unknown solver");
      }
    });
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    OU_HIP(hipMemcpy(phi, d_phi, bytes, hipMemcpyDeviceToHost));
    if (info) {
      info->iterations = r.iterations;
      info->rel_residual = r.rel_residual;
      info->seconds = secs;
      info->converged = r.converged ? 1 : 0;
      info->fp64_applications = pl.precision == OU_FP64 ? r.iterations : 0;
      info->preconditioner_applications = 0;
    }
    return r.converged ? int(OU_OK) : int(OU_ERR_NOT_CONVERGED);
  });
}

int ou_scatter_source_with(ou_plan *plan, const double *x, const double *sigma_s,
                           int accumulate) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (!x || !sigma_s) ou::fail(OU_ERR_ARG, "null argument");
    check_finite_host(x, pl.vec_size);
    check_finite_host(sigma_s, long(pl.sigma_s.size()));
    ensure_staging(pl);
    if (!pl.d_sigma_other) {
      void *p = nullptr;
      OU_HIP(hipMalloc(&p, pl.sigma_s.size() * 8));
      pl.d_sigma_other = static_cast<double *>(p);
    }
    OU_HIP(hipMemcpyAsync(pl.d_sigma_other, sigma_s, pl.sigma_s.size() * 8,
                          hipMemcpyHostToDevice, pl.stream));
    OU_HIP(hipMemcpyAsync(pl.d_x, x, size_t(pl.vec_size) * 8, hipMemcpyHostToDevice,
                          pl.stream));
    ou::launch_scatter_source_with(pl, pl.d_x, pl.d_sigma_other, accumulate != 0, pl.stream);
    OU_HIP(hipStreamSynchronize(pl.stream));
    return int(OU_OK);
  });
}

int ou_solve_rhs(ou_plan *plan, const ou_solver_opts *opts, const double *rhs, double *phi,
                 ou_solve_info *info) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (!opts || !rhs || !phi) ou::fail(OU_ERR_ARG, "null argument");
    if (!(opts->tol > 0.0) || opts->max_it < 1) ou::fail(OU_ERR_ARG, "bad tolerance or max_it");
    check_finite_host(rhs, pl.vec_size);
    check_finite_host(phi, pl.vec_size);
    ou::RadiativeTransfer &rt = *plan->rt;
    const size_t bytes = size_t(pl.vec_size) * 8;
    double *d_phi = nullptr, *d_rhs = nullptr;
    struct Free {
      double **a, **b;
      ~Free() {
        if (*a) (void)hipFree(*a);
        if (*b) (void)hipFree(*b);
      }
    } guard{&d_phi, &d_rhs};
    OU_HIP(hipMalloc(&d_phi, bytes));
    OU_HIP(hipMalloc(&d_rhs, bytes));
    OU_HIP(hipMemcpy(d_phi, phi, bytes, hipMemcpyHostToDevice));
    OU_HIP(hipMemcpy(d_rhs, rhs, bytes, hipMemcpyHostToDevice));
    const auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = pl.stream;
    ou::SolveResult r;
    r = with_watchdog({&pl}, [&] {
      switch (opts->solver) {
        case OU_SOLVER_SI:
          return ou::solve_si_rhs(rt, d_rhs, d_phi, opts->tol, opts->max_it,
                                  opts->verbose != 0, s);
        case OU_SOLVER_BICGSTAB:
          return ou::solve_bicgstab(rt, d_rhs, d_phi, opts->tol, opts->max_it,
                                    opts->verbose != 0, s);
        case OU_SOLVER_GMRES:
        case OU_SOLVER_FGMRES:
          return ou::solve_gmres(rt, d_rhs, d_phi, opts->restart, opts->tol, opts->max_it,
                                 nullptr, opts->verbose != 0, s);
        default:
          ou::fail(OU_ERR_ARG, "unknown solver");
      }
    });
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    OU_HIP(hipMemcpy(phi, d_phi, bytes, hipMemcpyDeviceToHost));
    if (info) {
      info->iterations = r.iterations;
      info->rel_residual = r.rel_residual;
      info->seconds = secs;
      info->converged = r.converged ? 1 : 0;
      info->fp64_applications = pl.precision == OU_FP64 ? r.iterations : 0;
      info->preconditioner_applications = 0;
    }
    return r.converged ? int(OU_OK) : int(OU_ERR_NOT_CONVERGED);
  });
}

int ou_device_alloc(size_t bytes, void **ptr) {
  return guarded([&] {
    if (!ptr) ou::fail(OU_ERR_ARG, "null output pointer");
    OU_HIP(hipMalloc(ptr, bytes));
    return int(OU_OK);
  });
}

int ou_device_free(void *ptr) {
  return guarded([&] {
    OU_HIP(hipFree(ptr));
    return int(OU_OK);
  });
}

int ou_memcpy_htod(void *dst, const void *src, size_t bytes) {
  return guarded([&] {
    OU_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return int(OU_OK);
  });
}

int ou_memcpy_dtoh(void *dst, const void *src, size_t bytes) {
  return guarded([&] {
    OU_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return int(OU_OK);
  });
}

int ou_memcpy_dtod(void *dst, const void *src, size_t bytes) {
  return guarded([&] {
    // A plain copy kernel (not a DMA engine), so that its memory traffic is
    // visible to the same counters as the operator's kernels.
    if (bytes % 8 == 0) {
      ou::vec_copy(long(bytes / 8), static_cast<const double *>(src), static_cast<double *>(dst),
                   nullptr);
      OU_HIP(hipDeviceSynchronize());
    } else {
      OU_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    }
    return int(OU_OK);
  });
}

int ou_device_synchronize(void) {
  return guarded([&] {
    OU_HIP(hipDeviceSynchronize());
    return int(OU_OK);
  });
}

int ou_device_name(char *buf, int len) {
  return guarded([&] {
    if (!buf || len < 1) ou::fail(OU_ERR_ARG, "bad buffer");
    hipDeviceProp_t prop;
    int dev = 0;
    OU_HIP(hipGetDevice(&dev));
    OU_HIP(hipGetDeviceProperties(&prop, dev));
    std::snprintf(buf, size_t(len), "%s (%s, %d CUs)", prop.name, prop.gcnArchName,
                  prop.multiProcessorCount);
    return int(OU_OK);
  });
}

int ou_time_apply(ou_plan *plan, const double *d_x, double *d_y, int warmup, int steps,
                  ou_timing *timing) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (!d_x || !d_y || !timing || steps < 1 || warmup < 0)
      ou::fail(OU_ERR_ARG, "bad timing arguments");
    hipStream_t s = pl.stream;
    for (int i = 0; i < warmup; ++i) ou::apply_device(pl, d_x, d_y, s);
    OU_HIP(hipStreamSynchronize(s));
    double ms_total = 0.0, ms_sweep = 0.0;
    float ms = 0.0f;
    for (int i = 0; i < steps; ++i) {
      OU_HIP(hipEventRecord(pl.ev[0], s));
      ou::launch_scatter_source(pl, d_x, s);
      ou::launch_trace_init(pl, false, s);
      ou::launch_sweep_levels(pl, nullptr, s, pl.ev[1], pl.ev[2]);
      ou::launch_epilogue(pl, d_x, d_y, true, s);
      OU_HIP(hipEventRecord(pl.ev[3], s));
      OU_HIP(hipEventSynchronize(pl.ev[3]));
      OU_HIP(hipEventElapsedTime(&ms, pl.ev[0], pl.ev[3]));
      ms_total += ms;
      OU_HIP(hipEventElapsedTime(&ms, pl.ev[1], pl.ev[2]));
      ms_sweep += ms;
    }
    check_watchdog(pl);
    int launches = 0;
    const int ns = pl.sweep_streams;
    if (pl.sweep_mode == OU_SWEEP_PERSISTENT)
      launches = 2;  // counter reset + the persistent sweep
    else if (pl.sweep_mode == OU_SWEEP_TILES)
      for (size_t l = 0; l + 1 < pl.tile_level_off.size(); ++l)
        launches += pl.tile_level_off[l + 1] > pl.tile_level_off[l];
    else
      for (int l = 0; l < pl.n_levels; ++l)
        for (int k = 0; k < ns; ++k)
          launches += pl.quarter_begin(l, 4 * (k + 1) / ns) > pl.quarter_begin(l, 4 * k / ns);
    timing->ms_total = ms_total;
    timing->ms_sweep_kernel = ms_sweep;
    timing->ms_other = ms_total - ms_sweep;
    timing->n_sweep_launches = launches;
    timing->applications = steps;
    return int(OU_OK);
  });
}

// ---------------------------------------------------------------------------
// Sweep shape and the batched launch
// ---------------------------------------------------------------------------

int ou_set_sweep_shape(ou_plan *plan, int waves, int dir_groups) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    OU_HIP(hipStreamSynchronize(pl.stream));
    if (pl.sweep_mode == OU_SWEEP_TILES)
      ou::set_tile_shape(pl, waves, dir_groups);
    else
      ou::set_sweep_shape(pl, waves, dir_groups);
    return int(OU_OK);
  });
}

int ou_sweep_shape(const ou_plan *plan, int *waves, int *dir_groups) {
  return guarded([&] {
    const ou::Plan &pl = plan_of(const_cast<ou_plan *>(plan));
    const bool tiles = pl.sweep_mode == OU_SWEEP_TILES;
    if (waves) *waves = tiles ? pl.tile_waves : pl.waves;
    if (dir_groups) *dir_groups = tiles ? pl.tile_groups : pl.dir_groups;
    return int(OU_OK);
  });
}

int ou_set_sweep_streams(ou_plan *plan, int n_streams) {
  return guarded([&] {
    if (n_streams != 1 && n_streams != 2 && n_streams != 4)
      ou::fail(OU_ERR_ARG, "sweep streams must be 1, 2 or 4");
    ou::Plan &pl = plan_of(plan);
    OU_HIP(hipStreamSynchronize(pl.stream));
    pl.sweep_streams = n_streams;
    return int(OU_OK);
  });
}

int ou_sweep_streams(const ou_plan *plan, int *n_streams) {
  return guarded([&] {
    if (!n_streams) ou::fail(OU_ERR_ARG, "null output pointer");
    *n_streams = plan_of(const_cast<ou_plan *>(plan)).sweep_streams;
    return int(OU_OK);
  });
}

int ou_set_sweep_mode(ou_plan *plan, int mode) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    OU_HIP(hipStreamSynchronize(pl.stream));
    ou::set_sweep_mode(pl, mode);
    return int(OU_OK);
  });
}

int ou_sweep_mode(const ou_plan *plan, int *mode) {
  return guarded([&] {
    if (!mode) ou::fail(OU_ERR_ARG, "null output pointer");
    *mode = plan_of(const_cast<ou_plan *>(plan)).sweep_mode;
    return int(OU_OK);
  });
}

int ou_set_sweep_wait_limit(ou_plan *plan, double seconds) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (!(seconds >= 0.0) || seconds > 3600.0)
      ou::fail(OU_ERR_ARG, "wait limit must be in [0, 3600] seconds");
    OU_HIP(hipStreamSynchronize(pl.stream));
    pl.wait_limit_s = seconds;
    return int(OU_OK);
  });
}

int ou_sweep_status(ou_plan *plan) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    ou::check_sweep_watchdog(pl.d_error, pl.stream);
    return int(OU_OK);
  });
}

int ou_batch_set_sweep_mode(ou_batch *batch, int mode) {
  return guarded([&] {
    batch_of(batch).set_sweep_mode(mode);
    return int(OU_OK);
  });
}

int ou_batch_sweep_mode(const ou_batch *batch, int *mode) {
  return guarded([&] {
    if (!mode) ou::fail(OU_ERR_ARG, "null output pointer");
    *mode = batch_of(batch).sweep_mode();
    return int(OU_OK);
  });
}

int ou_batch_create(ou_plan *const *plans, int n_plans, int waves, int dir_groups,
                    ou_batch **batch) {
  return guarded([&] {
    if (!batch) ou::fail(OU_ERR_ARG, "null output pointer");
    if (!plans || n_plans < 1) ou::fail(OU_ERR_ARG, "a batch needs at least one plan");
    std::vector<ou::RadiativeTransfer *> members(static_cast<size_t>(n_plans));
    for (int i = 0; i < n_plans; ++i) {
      plan_of(plans[i]);
      members[size_t(i)] = plans[i]->rt.get();
    }
    auto *b = new ou_batch();
    try {
      b->b.reset(new ou::Batch(std::move(members), waves, dir_groups));
    } catch (...) {
      delete b;
      throw;
    }
    *batch = b;
    return int(OU_OK);
  });
}

void ou_batch_destroy(ou_batch *batch) { delete batch; }

int ou_batch_info(const ou_batch *batch, int *waves, int *dir_groups, int *n_levels,
                  long *n_chunks) {
  return guarded([&] {
    const ou::Batch &b = batch_of(batch);
    if (waves) *waves = b.waves();
    if (dir_groups) *dir_groups = b.dir_groups();
    if (n_levels) *n_levels = b.n_levels();
    if (n_chunks) *n_chunks = b.n_chunks();
    return int(OU_OK);
  });
}

int ou_batch_set_shape(ou_batch *batch, int waves, int dir_groups) {
  return guarded([&] {
    batch_of(batch).set_shape(waves, dir_groups);
    return int(OU_OK);
  });
}

int ou_batch_set_graph(ou_batch *batch, int use_graph) {
  return guarded([&] {
    batch_of(batch).set_use_graph(use_graph != 0);
    return int(OU_OK);
  });
}

int ou_batch_set_streams(ou_batch *batch, int n_streams) {
  return guarded([&] {
    batch_of(batch).set_streams(n_streams);
    return int(OU_OK);
  });
}

int ou_batch_apply_device(ou_batch *batch, const double *const *d_x, double *const *d_y,
                          void *stream) {
  return guarded([&] {
    batch_of(batch).apply(d_x, d_y, static_cast<hipStream_t>(stream));
    return int(OU_OK);
  });
}

int ou_batch_sweep_device(ou_batch *batch, double *const *d_y, int with_source,
                          void *stream) {
  return guarded([&] {
    batch_of(batch).sweep(d_y, with_source != 0, static_cast<hipStream_t>(stream));
    return int(OU_OK);
  });
}

int ou_batch_solve_si(ou_batch *batch, double *const *d_phi, double tol, int max_it,
                      int *iterations, double *change, int *n_iterations, int verbose) {
  return guarded([&] {
    const int it =
        batch_of(batch).solve_si(d_phi, tol, max_it, iterations, change, verbose != 0);
    if (n_iterations) *n_iterations = it;
    return int(OU_OK);
  });
}

int ou_batch_tune(ou_batch *batch, int reps, ou_tune_entry *entries, int max_entries,
                  int *n_entries) {
  return guarded([&] {
    const std::vector<ou::TuneEntry> t = batch_of(batch).tune(reps);
    const int n = std::min<int>(int(t.size()), std::max(0, max_entries));
    for (int i = 0; i < n && entries; ++i)
      entries[i] = ou_tune_entry{t[size_t(i)].waves, t[size_t(i)].dir_groups, t[size_t(i)].ms,
                                 t[size_t(i)].mode};
    if (n_entries) *n_entries = int(t.size());
    return int(OU_OK);
  });
}

int ou_batch_time_apply(ou_batch *batch, const double *const *d_x, double *const *d_y,
                        int warmup, int steps, double *ms) {
  return guarded([&] {
    ou::Batch &b = batch_of(batch);
    if (!ms || steps < 1 || warmup < 0) ou::fail(OU_ERR_ARG, "bad timing arguments");
    hipStream_t s = b.stream();
    for (int i = 0; i < warmup; ++i) b.apply(d_x, d_y, s);
    OU_HIP(hipStreamSynchronize(s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    OU_HIP(hipEventCreate(&e0));
    struct Ev {
      hipEvent_t *a, *b;
      ~Ev() {
        if (*a) (void)hipEventDestroy(*a);
        if (*b) (void)hipEventDestroy(*b);
      }
    } guard{&e0, &e1};
    OU_HIP(hipEventCreate(&e1));
    OU_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < steps; ++i) b.apply(d_x, d_y, s);
    OU_HIP(hipEventRecord(e1, s));
    OU_HIP(hipEventSynchronize(e1));
    b.check_watchdog(s);
    float t = 0.0f;
    OU_HIP(hipEventElapsedTime(&t, e0, e1));
    *ms = double(t) / steps;
    return int(OU_OK);
  });
}

// ---------------------------------------------------------------------------
// Learned preconditioner
// ---------------------------------------------------------------------------

int ou_unet_create(int grid, int max_batch, unsigned long long seed, ou_unet **net) {
  return guarded([&] {
    if (!net) ou::fail(OU_ERR_ARG, "null output pointer");
    std::unique_ptr<ou_unet> u(new ou_unet());
    u->net.reset(new ou::UNet(grid, max_batch, seed));
    const size_t p0 = size_t(grid) * grid * grid * max_batch;
    OU_HIP(hipMalloc(&u->d_feat, p0 * ou::kUnetFeatures * 4));
    OU_HIP(hipMalloc(&u->d_target, p0 * ou::kUnetOut * 4));
    OU_HIP(hipMalloc(&u->d_pred, p0 * ou::kUnetOut * 4));
    *net = u.release();
    return int(OU_OK);
  });
}

void ou_unet_destroy(ou_unet *net) { delete net; }

long ou_unet_n_params(const ou_unet *net) { return net && net->net ? net->net->n_params() : 0; }

int ou_unet_grid(const ou_unet *net) { return net && net->net ? net->net->grid() : 0; }

int ou_unet_set_mode(ou_unet *net, int mode) {
  return guarded([&] {
    if (!net || !net->net) ou::fail(OU_ERR_ARG, "null network");
    net->net->set_mode(mode);
    return int(OU_OK);
  });
}

int ou_unet_mode(const ou_unet *net) { return net && net->net ? net->net->mode() : -1; }

int ou_unet_get_params(const ou_unet *net, float *params) {
  return guarded([&] {
    if (!net || !params) ou::fail(OU_ERR_ARG, "null argument");
    net->net->get_params(params);
    return int(OU_OK);
  });
}

int ou_unet_set_params(ou_unet *net, const float *params) {
  return guarded([&] {
    if (!net || !params) ou::fail(OU_ERR_ARG, "null argument");
    const long n = net->net->n_params();
    for (long i = 0; i < n; ++i)
      if (!std::isfinite(params[i])) ou::fail(OU_ERR_NONFINITE, "non-finite parameter");
    net->net->set_params(params);
    return int(OU_OK);
  });
}

namespace {

void unet_upload(ou_unet *net, const float *features, const float *target, int batch) {
  if (!net || !features) ou::fail(OU_ERR_ARG, "null argument");
  if (batch < 1 || batch > net->net->max_batch())
    ou::fail(OU_ERR_SIZE, "batch exceeds the network's maximum batch");
  const size_t p0 = size_t(net->net->grid()) * net->net->grid() * net->net->grid() * batch;
  for (size_t i = 0; i < p0 * ou::kUnetFeatures; ++i)
    if (!std::isfinite(features[i])) ou::fail(OU_ERR_NONFINITE, "non-finite feature");
  OU_HIP(hipMemcpy(net->d_feat, features, p0 * ou::kUnetFeatures * 4, hipMemcpyHostToDevice));
  if (target) {
    for (size_t i = 0; i < p0 * ou::kUnetOut; ++i)
      if (!std::isfinite(target[i])) ou::fail(OU_ERR_NONFINITE, "non-finite target");
    OU_HIP(hipMemcpy(net->d_target, target, p0 * ou::kUnetOut * 4, hipMemcpyHostToDevice));
  }
}

}  // namespace

int ou_unet_predict(ou_unet *net, const float *features, int batch, float *pred) {
  return guarded([&] {
    if (!pred) ou::fail(OU_ERR_ARG, "null output");
    unet_upload(net, features, nullptr, batch);
    net->net->predict(net->d_feat, batch, net->d_pred, nullptr);
    const size_t p0 = size_t(net->net->grid()) * net->net->grid() * net->net->grid() * batch;
    OU_HIP(hipMemcpy(pred, net->d_pred, p0 * ou::kUnetOut * 4, hipMemcpyDeviceToHost));
    return int(OU_OK);
  });
}

int ou_unet_loss_grad(ou_unet *net, const float *features, const float *target, int batch,
                      double *loss, float *grad) {
  return guarded([&] {
    if (!target || !loss) ou::fail(OU_ERR_ARG, "null argument");
    unet_upload(net, features, target, batch);
    net->net->loss_and_grad(net->d_feat, net->d_target, batch, nullptr);
    OU_HIP(hipMemcpy(loss, net->net->d_loss(), 8, hipMemcpyDeviceToHost));
    if (grad)
      OU_HIP(hipMemcpy(grad, net->net->d_grad(), size_t(net->net->n_params()) * 4,
                       hipMemcpyDeviceToHost));
    return int(OU_OK);
  });
}

int ou_unet_train_step(ou_unet *net, const float *features, const float *target, int batch,
                       float lr, double *loss) {
  return guarded([&] {
    if (!target || !loss) ou::fail(OU_ERR_ARG, "null argument");
    if (!(lr > 0.0f) || !std::isfinite(lr)) ou::fail(OU_ERR_ARG, "learning rate must be positive");
    unet_upload(net, features, target, batch);
    net->net->loss_and_grad(net->d_feat, net->d_target, batch, nullptr);
    net->net->adam_step(lr, nullptr);
    OU_HIP(hipMemcpy(loss, net->net->d_loss(), 8, hipMemcpyDeviceToHost));
    return int(OU_OK);
  });
}

int ou_coarse_features(ou_plan *plan, int grid, const double *v, double *coarse, float *material) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    ou::coarse_factor(pl, grid);
    const size_t p0 = size_t(grid) * grid * grid;
    if (v && coarse) {
      check_finite_host(v, pl.vec_size);
      ensure_staging(pl);
      OU_HIP(hipMemcpy(pl.d_x, v, size_t(pl.vec_size) * 8, hipMemcpyHostToDevice));
      ou::coarse_restrict(pl, grid, pl.d_x, pl.d_y, pl.stream);
      OU_HIP(hipStreamSynchronize(pl.stream));
      OU_HIP(hipMemcpy(coarse, pl.d_y, p0 * 8 * 8, hipMemcpyDeviceToHost));
    }
    if (material) {
      const std::vector<float> m = ou::coarse_material_features(pl, grid);
      std::memcpy(material, m.data(), m.size() * 4);
    }
    return int(OU_OK);
  });
}

int ou_coarse_prolong(ou_plan *plan, int grid, const float *corners, double *z) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    ou::coarse_factor(pl, grid);
    if (!corners || !z) ou::fail(OU_ERR_ARG, "null argument");
    check_finite_host(z, pl.vec_size);
    const size_t p0 = size_t(grid) * grid * grid;
    for (size_t i = 0; i < p0 * 8; ++i)
      if (!std::isfinite(corners[i])) ou::fail(OU_ERR_NONFINITE, "non-finite corner value");
    ensure_staging(pl);
    float *d_c = nullptr;
    double *d_aux = nullptr;
    struct Free {
      float **a;
      double **b;
      ~Free() {
        if (*a) (void)hipFree(*a);
        if (*b) (void)hipFree(*b);
      }
    } guard{&d_c, &d_aux};
    const size_t nv = size_t(grid + 1) * (grid + 1) * (grid + 1);
    OU_HIP(hipMalloc(&d_c, p0 * 8 * 4));
    OU_HIP(hipMalloc(&d_aux, (nv + 1) * 8));
    OU_HIP(hipMemcpy(d_c, corners, p0 * 8 * 4, hipMemcpyHostToDevice));
    const double dot = 8.0 * double(p0);  // scale s = 1
    OU_HIP(hipMemcpy(d_aux + nv, &dot, 8, hipMemcpyHostToDevice));
    OU_HIP(hipMemcpy(pl.d_x, z, size_t(pl.vec_size) * 8, hipMemcpyHostToDevice));
    ou::coarse_prolong_add(pl, grid, d_c, d_aux + nv, d_aux, pl.d_x, pl.stream);
    OU_HIP(hipStreamSynchronize(pl.stream));
    OU_HIP(hipMemcpy(z, pl.d_x, size_t(pl.vec_size) * 8, hipMemcpyDeviceToHost));
    return int(OU_OK);
  });
}

int ou_coarse_prolong_noalloc(ou_plan *plan, int grid, const float *d_corners, double *d_z,
                              void *stream) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (grid > ou::coarse_noalloc_max_grid())
      ou::fail(OU_ERR_SIZE, "coarse grid above the allocation-free prolongation's maximum");
    ou::coarse_factor(pl, grid);
    if (!d_corners || !d_z) ou::fail(OU_ERR_ARG, "null argument");
    ou::coarse_prolong_add_noalloc(pl, grid, d_corners, d_z, stream_of(plan, stream));
    return int(OU_OK);
  });
}

int ou_learned_pairs(ou_plan *plan, int grid, const double *e0, int n_steps,
                     unsigned long long record_mask, double *r_out, double *e_out) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    ou::RadiativeTransfer &rt = *plan->rt;
    ou::coarse_factor(pl, grid);
    if (n_steps < 1 || n_steps > 64) ou::fail(OU_ERR_ARG, "n_steps must be in [1, 64]");
    if (!r_out || !e_out) ou::fail(OU_ERR_ARG, "null output");
    const long n = pl.vec_size;
    const size_t p0 = size_t(grid) * grid * grid;
    hipStream_t s = pl.stream;
    double *d_e = nullptr, *d_ae = nullptr, *d_c = nullptr;
    struct Free {
      double **p[3];
      ~Free() {
        for (double **q : p)
          if (*q) (void)hipFree(*q);
      }
    } guard{{&d_e, &d_ae, &d_c}};
    OU_HIP(hipMalloc(&d_e, size_t(n) * 8));
    OU_HIP(hipMalloc(&d_ae, size_t(n) * 8));
    OU_HIP(hipMalloc(&d_c, p0 * 8 * 8));
    if (e0) {
      check_finite_host(e0, n);
      OU_HIP(hipMemcpy(d_e, e0, size_t(n) * 8, hipMemcpyHostToDevice));
    } else {
      ou::build_rhs(rt, d_e, s);
    }
    int rec = 0;
    for (int k = 0; k < n_steps; ++k) {
      rt.Apply(d_e, d_ae, s);
      if ((record_mask >> k) & 1ull) {
        ou::coarse_restrict(pl, grid, d_ae, d_c, s);
        OU_HIP(hipStreamSynchronize(s));
        OU_HIP(hipMemcpy(r_out + size_t(rec) * p0 * 8, d_c, p0 * 8 * 8, hipMemcpyDeviceToHost));
        ou::coarse_restrict(pl, grid, d_e, d_c, s);
        OU_HIP(hipStreamSynchronize(s));
        OU_HIP(hipMemcpy(e_out + size_t(rec) * p0 * 8, d_c, p0 * 8 * 8, hipMemcpyDeviceToHost));
        ++rec;
      }
      ou::vec_axpby(n, -1.0, d_ae, 1.0, d_e, s);
    }
    OU_HIP(hipStreamSynchronize(s));
    return rec;
  });
}

namespace {

// Records the coarse pair (R A z_j, R z_j) of every stride-th Arnoldi vector.
struct PairRecorder : ou::ArnoldiObserver {
  const ou::Plan &pl;
  int grid, stride, max_records, count = 0;
  double *d_c, *r_out, *e_out;
  PairRecorder(const ou::Plan &p, int g, int st, int mr, double *dc, double *r, double *e)
      : pl(p), grid(g), stride(st), max_records(mr), d_c(dc), r_out(r), e_out(e) {}
  void observe(int it, const double *d_z, const double *d_az, hipStream_t s) override {
    if (it % stride || count >= max_records) return;
    const size_t m = size_t(grid) * grid * grid * 8;
    ou::coarse_restrict(pl, grid, d_az, d_c, s);
    OU_HIP(hipStreamSynchronize(s));
    OU_HIP(hipMemcpy(r_out + size_t(count) * m, d_c, m * 8, hipMemcpyDeviceToHost));
    ou::coarse_restrict(pl, grid, d_z, d_c, s);
    OU_HIP(hipStreamSynchronize(s));
    OU_HIP(hipMemcpy(e_out + size_t(count) * m, d_c, m * 8, hipMemcpyDeviceToHost));
    ++count;
  }
};

}  // namespace

namespace {

// Records (R A z_j, R z_j) of n_steps Arnoldi steps of GMRES(restart) on the
// plan's source problem; z_j is the Arnoldi vector (prec null) or the
// preconditioned one (FGMRES with the network).
int learned_pairs_from_krylov(ou_plan *plan, ou::Preconditioner *prec, int grid, int n_steps,
                              int restart, int stride, int max_records, double *r_out,
                              double *e_out) {
  {
    ou::Plan &pl = plan_of(plan);
    ou::RadiativeTransfer &rt = *plan->rt;
    ou::coarse_factor(pl, grid);
    if (n_steps < 1 || stride < 1 || max_records < 1) ou::fail(OU_ERR_ARG, "bad step counts");
    if (!r_out || !e_out) ou::fail(OU_ERR_ARG, "null output");
    const long n = pl.vec_size;
    hipStream_t s = pl.stream;
    double *d_b = nullptr, *d_x = nullptr, *d_c = nullptr;
    struct Free {
      double **p[3];
      ~Free() {
        for (double **q : p)
          if (*q) (void)hipFree(*q);
      }
    } guard{{&d_b, &d_x, &d_c}};
    OU_HIP(hipMalloc(&d_b, size_t(n) * 8));
    OU_HIP(hipMalloc(&d_x, size_t(n) * 8));
    OU_HIP(hipMalloc(&d_c, size_t(grid) * grid * grid * 8 * 8));
    ou::build_rhs(rt, d_b, s);
    ou::vec_fill(n, 0.0, d_x, s);
    PairRecorder rec(pl, grid, stride, max_records, d_c, r_out, e_out);
    ou::solve_gmres(rt, d_b, d_x, restart, 1e-300, n_steps, prec, false, s, &rec);
    OU_HIP(hipStreamSynchronize(s));
    return rec.count;
  }
}

}  // namespace

int ou_learned_pairs_krylov(ou_plan *plan, int grid, int n_steps, int restart, int stride,
                            int max_records, double *r_out, double *e_out) {
  return guarded([&] {
    return learned_pairs_from_krylov(plan, nullptr, grid, n_steps, restart, stride,
                                     max_records, r_out, e_out);
  });
}

int ou_learned_pairs_fgmres(ou_plan *plan, ou_unet *net, int grid, int n_steps, int restart,
                            int stride, int max_records, double *r_out, double *e_out) {
  return guarded([&] {
    plan_of(plan);
    if (!net || !net->net) ou::fail(OU_ERR_ARG, "null network");
    if (net->net->grid() != grid) ou::fail(OU_ERR_ARG, "network grid differs from `grid`");
    ou::LearnedPreconditioner prec(*plan->rt, *net->net);
    return learned_pairs_from_krylov(plan, &prec, grid, n_steps, restart, stride, max_records,
                                     r_out, e_out);
  });
}

namespace {

void fill_info(ou_solve_info *info, const ou::SolveResult &r, double secs, int fp64, int prec) {
  if (!info) return;
  info->iterations = r.iterations;
  info->rel_residual = r.rel_residual;
  info->seconds = secs;
  info->converged = r.converged ? 1 : 0;
  info->fp64_applications = fp64;
  info->preconditioner_applications = prec;
}

}  // namespace

int ou_solve_learned(ou_plan *plan, ou_unet *net, const ou_solver_opts *opts, double *phi,
                     ou_solve_info *info) {
  return guarded([&] {
    ou::Plan &pl = plan_of(plan);
    if (!opts || !phi) ou::fail(OU_ERR_ARG, "null argument");
    if (opts->solver != OU_SOLVER_FGMRES) ou::fail(OU_ERR_ARG, "ou_solve_learned runs FGMRES");
    if (!(opts->tol > 0.0) || opts->max_it < 1) ou::fail(OU_ERR_ARG, "bad tolerance or max_it");
    if (net && !net->net) ou::fail(OU_ERR_ARG, "empty network");
    check_finite_host(phi, pl.vec_size);
    ou::RadiativeTransfer &rt = *plan->rt;
    std::unique_ptr<ou::LearnedPreconditioner> prec;
    if (net) prec.reset(new ou::LearnedPreconditioner(rt, *net->net));
    const size_t bytes = size_t(pl.vec_size) * 8;
    double *d_phi = nullptr, *d_rhs = nullptr;
    struct Free {
      double **a, **b;
      ~Free() {
        if (*a) (void)hipFree(*a);
        if (*b) (void)hipFree(*b);
      }
    } guard{&d_phi, &d_rhs};
    OU_HIP(hipMalloc(&d_phi, bytes));
    OU_HIP(hipMalloc(&d_rhs, bytes));
    OU_HIP(hipMemcpy(d_phi, phi, bytes, hipMemcpyHostToDevice));
    hipStream_t s = pl.stream;
    const auto t0 = std::chrono::steady_clock::now();
    const ou::SolveResult r = with_watchdog({&pl}, [&] {
      ou::build_rhs(rt, d_rhs, s);
      return ou::solve_gmres(rt, d_rhs, d_phi, opts->restart, opts->tol, opts->max_it,
                             prec.get(), opts->verbose != 0, s);
    });
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    OU_HIP(hipMemcpy(phi, d_phi, bytes, hipMemcpyDeviceToHost));
    fill_info(info, r, secs, pl.precision == OU_FP64 ? r.iterations : 0,
              prec ? prec->applications() : 0);
    return r.converged ? int(OU_OK) : int(OU_ERR_NOT_CONVERGED);
  });
}

int ou_solve_refined(ou_plan *plan64, ou_plan *plan32, ou_unet *net, const ou_solver_opts *opts,
                     double inner_tol, double *phi, ou_solve_info *info) {
  return guarded([&] {
    ou::Plan &p64 = plan_of(plan64);
    ou::Plan &p32 = plan_of(plan32);
    if (!opts || !phi) ou::fail(OU_ERR_ARG, "null argument");
    if (!(opts->tol > 0.0) || opts->max_it < 1 || !(inner_tol > 0.0) || !(inner_tol < 1.0))
      ou::fail(OU_ERR_ARG, "bad tolerance or max_it");
    if (net && !net->net) ou::fail(OU_ERR_ARG, "empty network");
    check_finite_host(phi, p64.vec_size);
    std::unique_ptr<ou::LearnedPreconditioner> prec;
    if (net) prec.reset(new ou::LearnedPreconditioner(*plan32->rt, *net->net));
    const size_t bytes = size_t(p64.vec_size) * 8;
    double *d_phi = nullptr;
    OU_HIP(hipMalloc(&d_phi, bytes));
    struct Free {
      double **a;
      ~Free() {
        if (*a) (void)hipFree(*a);
      }
    } guard{&d_phi};
    OU_HIP(hipMemcpy(d_phi, phi, bytes, hipMemcpyHostToDevice));
    const auto t0 = std::chrono::steady_clock::now();
    const ou::RefinedResult r = with_watchdog({&p64, &p32}, [&] {
      return ou::solve_refined(*plan64->rt, *plan32->rt, prec.get(), d_phi, opts->tol,
                               opts->max_it, opts->restart, inner_tol, opts->verbose != 0);
    });
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    OU_HIP(hipMemcpy(phi, d_phi, bytes, hipMemcpyDeviceToHost));
    fill_info(info, r.total, secs, r.outer, prec ? prec->applications() : 0);
    return r.total.converged ? int(OU_OK) : int(OU_ERR_NOT_CONVERGED);
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Input files
// ---------------------------------------------------------------------------
namespace ou {
void write_vtk(const std::string &filename, int dim, int p, const int *n_cells,
               const double *cell_size, const double *origin, int n_groups, int n_mom,
               const double *flux);
}

struct ou_parameters {
  ou::Parameters p;
};
struct ou_geometry {
  ou::Geometry g;
};
struct ou_material_properties {
  ou::MaterialProperties m;
};

namespace {
void check_index(int i, int n, const char *what) {
  if (i < 0 || i >= n) ou::fail(OU_ERR_ARG, std::string(what) + " index out of range");
}
}  // namespace

extern "C" {

int ou_parameters_read(const char *path, ou_parameters **prm) {
  return guarded([&] {
    if (!path || !prm) ou::fail(OU_ERR_ARG, "null argument");
    *prm = nullptr;
    *prm = new ou_parameters{ou::Parameters(path)};
    return int(OU_OK);
  });
}

void ou_parameters_destroy(ou_parameters *prm) { delete prm; }

int ou_parameters_get_info(const ou_parameters *prm, ou_parameters_info *info) {
  return guarded([&] {
    if (!prm || !info) ou::fail(OU_ERR_ARG, "null argument");
    const ou::Parameters &p = prm->p;
    ou_parameters_info r{};
    r.dim = p.dim;
    r.fe_order = p.fe_order;
    r.n_refinements = p.n_refinements;
    r.refinement_factor = p.refinement_factor;
    r.coarsening_factor = p.coarsening_factor;
    r.solver_type = p.solver_type == ou::INP_SI ? OU_SOLVER_SI
                    : p.solver_type == ou::INP_BICGSTAB ? OU_SOLVER_BICGSTAB
                                                          : OU_SOLVER_GMRES;
    r.max_outer_it = p.max_out_it;
    r.max_inner_it = p.max_in_it;
    r.outer_tol = p.outer_tol;
    r.inner_tol = p.inner_tol;
    r.sn_order = p.sn_order;
    r.quad_type = p.quad_type == ou::INP_LS ? OU_QUAD_LS : OU_QUAD_GLC;
    r.weight_sum = p.weight_sum;
    r.galerkin = p.galerkin ? 1 : 0;
    for (int f = 0; f < 6; ++f) {
      int bc = OU_BC_VACUUM;
      if (f < 2 * p.dim) {
        const int b = p.bc_type[size_t(f)];
        bc = b == ou::INP_ISOTROPIC     ? OU_BC_ISOTROPIC
             : b == ou::INP_MOST_NORMAL ? OU_BC_MOST_NORMAL
             : b == ou::INP_REFLECTIVE  ? OU_BC_REFLECTIVE
                                        : OU_BC_VACUUM;
      }
      r.bc_type[f] = bc;
    }
    r.n_groups = p.n_groups;
    r.n_src = p.n_src;
    *info = r;
    return int(OU_OK);
  });
}

int ou_parameters_get_filename(const ou_parameters *prm, int which, char *buf, int len,
                               int *needed) {
  return guarded([&] {
    if (!prm) ou::fail(OU_ERR_ARG, "null parameters");
    const ou::Parameters &p = prm->p;
    const std::string *s = which == 0   ? &p.geometry_filename
                           : which == 1 ? &p.xs_filename
                           : which == 2 ? &p.output_filename
                                        : nullptr;
    if (!s) ou::fail(OU_ERR_ARG, "file name selector must be 0, 1 or 2");
    if (needed) *needed = int(s->size()) + 1;
    if (buf && len > 0) {
      const size_t n = std::min(s->size(), size_t(len - 1));
      std::memcpy(buf, s->data(), n);
      buf[n] = '\0';
    }
    return int(OU_OK);
  });
}

int ou_parameters_get_inc_flux(const ou_parameters *prm, int face, int group, double *value) {
  return guarded([&] {
    if (!prm || !value) ou::fail(OU_ERR_ARG, "null argument");
    check_index(face, 2 * prm->p.dim, "face");
    check_index(group, prm->p.n_groups, "group");
    *value = prm->p.inc_flux[size_t(face)][size_t(group)];
    return int(OU_OK);
  });
}

int ou_parameters_get_src(const ou_parameters *prm, int source, int group, double *value) {
  return guarded([&] {
    if (!prm || !value) ou::fail(OU_ERR_ARG, "null argument");
    check_index(source, prm->p.n_src, "source");
    check_index(group, prm->p.n_groups, "group");
    *value = prm->p.src[size_t(source)][size_t(group)];
    return int(OU_OK);
  });
}

int ou_geometry_read(const char *path, int dim, ou_geometry **geo) {
  return guarded([&] {
    if (!path || !geo) ou::fail(OU_ERR_ARG, "null argument");
    *geo = nullptr;
    *geo = new ou_geometry{ou::Geometry(path, dim)};
    return int(OU_OK);
  });
}

void ou_geometry_destroy(ou_geometry *geo) { delete geo; }

int ou_geometry_get_info(const ou_geometry *geo, ou_geometry_info *info) {
  return guarded([&] {
    if (!geo || !info) ou::fail(OU_ERR_ARG, "null argument");
    const ou::Geometry &g = geo->g;
    ou_geometry_info r{};
    r.dim = g.dim;
    const std::vector<int> n = g.n_cells();
    const std::vector<double> h = g.cell_size();
    for (int a = 0; a < 3; ++a) {
      r.bottom_left[a] = a < g.dim ? g.bottom_left[size_t(a)] : 0.0;
      r.upper_right[a] = a < g.dim ? g.upper_right[size_t(a)] : 0.0;
      r.n_subdivisions[a] = a < g.dim ? g.n_subdivisions[size_t(a)] : 1;
      r.n_cells[a] = n[size_t(a)];
      r.cell_size[a] = h[size_t(a)];
    }
    r.n_global_refinements = g.n_global_refinements;
    r.n_materials = g.n_materials;
    *info = r;
    return int(OU_OK);
  });
}

int ou_geometry_get_ids(const ou_geometry *geo, int i, int j, int k, int *material_id,
                        int *source_id) {
  return guarded([&] {
    if (!geo) ou::fail(OU_ERR_ARG, "null geometry");
    const ou::Geometry &g = geo->g;
    check_index(i, g.n_subdivisions[0], "i");
    check_index(j, g.n_subdivisions[1], "j");
    check_index(k, g.dim == 3 ? g.n_subdivisions[2] : 1, "k");
    const size_t c = size_t(i) + size_t(g.n_subdivisions[0]) *
                                     (size_t(j) + size_t(g.n_subdivisions[1]) * size_t(k));
    if (material_id) *material_id = g.material_ids[c];
    if (source_id) *source_id = g.source_ids[c];
    return int(OU_OK);
  });
}

int ou_geometry_cell_ids(const ou_geometry *geo, int *material_id, int *source_id) {
  return guarded([&] {
    if (!geo) ou::fail(OU_ERR_ARG, "null geometry");
    std::vector<int> m, s;
    geo->g.fine_ids(m, s);
    if (material_id) std::copy(m.begin(), m.end(), material_id);
    if (source_id) std::copy(s.begin(), s.end(), source_id);
    return int(OU_OK);
  });
}

int ou_material_properties_read(const char *path, int n_materials, int n_groups,
                                ou_material_properties **mp) {
  return guarded([&] {
    if (!path || !mp) ou::fail(OU_ERR_ARG, "null argument");
    *mp = nullptr;
    *mp = new ou_material_properties{ou::MaterialProperties(path, n_materials, n_groups)};
    return int(OU_OK);
  });
}

void ou_material_properties_destroy(ou_material_properties *mp) { delete mp; }

int ou_material_properties_get_L_max(const ou_material_properties *mp, int *L_max) {
  return guarded([&] {
    if (!mp || !L_max) ou::fail(OU_ERR_ARG, "null argument");
    *L_max = mp->m.L_max;
    return int(OU_OK);
  });
}

int ou_material_properties_get_sigma_t(const ou_material_properties *mp, int mat, int group,
                                       double *value) {
  return guarded([&] {
    if (!mp || !value) ou::fail(OU_ERR_ARG, "null argument");
    check_index(mat, mp->m.n_materials, "material");
    check_index(group, mp->m.n_groups, "group");
    *value = mp->m.get_sigma_t(mat, group);
    return int(OU_OK);
  });
}

int ou_material_properties_get_sigma_s(const ou_material_properties *mp, int mat, int group,
                                       int group_to, int l, double *value) {
  return guarded([&] {
    if (!mp || !value) ou::fail(OU_ERR_ARG, "null argument");
    check_index(mat, mp->m.n_materials, "material");
    check_index(group, mp->m.n_groups, "group");
    check_index(group_to, mp->m.n_groups, "group");
    check_index(l, mp->m.L_max + 1, "Legendre order");
    *value = mp->m.get_sigma_s(mat, group, group_to, l);
    return int(OU_OK);
  });
}

int ou_write_vtk(const char *filename, int dim, int fe_order, const int *n_cells,
                 const double *cell_size, const double *origin, int n_groups, int n_mom,
                 const double *flux) {
  return guarded([&] {
    if (!filename || !n_cells || !cell_size || !origin || !flux)
      ou::fail(OU_ERR_ARG, "null argument");
    ou::write_vtk(filename, dim, fe_order, n_cells, cell_size, origin, n_groups, n_mom, flux);
    return int(OU_OK);
  });
}

}  // extern "C"
