// extern "C" entry points of include/ouranos_amd.h.  No exception crosses
// this boundary: every call returns OU_OK or a negative ou_status and keeps
// the message for ou_last_error().
#include <chrono>
#include <cstring>
#include <string>

#include "fe.hpp"
#include "quadrature.hpp"
#include "radiative_transfer.hpp"
#include "vec.hpp"

struct ou_plan {
  std::unique_ptr<ou::RadiativeTransfer> rt;
};

struct ou_quadrature {
  ou::Quadrature q;
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

hipStream_t stream_of(ou_plan *p, void *s) {
  return s ? static_cast<hipStream_t>(s) : plan_of(p).stream;
}

void ensure_staging(ou::Plan &pl) {
  if (!pl.d_x) OU_HIP(hipMalloc(&pl.d_x, size_t(pl.vec_size) * 8));
  if (!pl.d_y) OU_HIP(hipMalloc(&pl.d_y, size_t(pl.vec_size) * 8));
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
    switch (opts->solver) {
      case OU_SOLVER_SI:
        r = ou::solve_si(rt, d_phi, opts->tol, opts->max_it, opts->verbose != 0);
        break;
      case OU_SOLVER_GMRES:
      case OU_SOLVER_FGMRES:
      case OU_SOLVER_BICGSTAB:
        OU_HIP(hipMalloc(&d_rhs, bytes));
        ou::build_rhs(rt, d_rhs, s);
        if (opts->solver == OU_SOLVER_BICGSTAB)
          r = ou::solve_bicgstab(rt, d_rhs, d_phi, opts->tol, opts->max_it,
                                 opts->verbose != 0, s);
        else
          r = ou::solve_gmres(rt, d_rhs, d_phi, opts->restart, opts->tol, opts->max_it,
                              nullptr, opts->verbose != 0, s);
        break;
      default:
        ou::fail(OU_ERR_ARG, "unknown solver");
    }
    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    OU_HIP(hipMemcpy(phi, d_phi, bytes, hipMemcpyDeviceToHost));
    if (info) {
      info->iterations = r.iterations;
      info->rel_residual = r.rel_residual;
      info->seconds = secs;
      info->converged = r.converged ? 1 : 0;
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
    int launches = 0;
    for (int l = 0; l < pl.n_levels; ++l)
      launches += pl.level_chunk_off[l + 1] > pl.level_chunk_off[l];
    timing->ms_total = ms_total;
    timing->ms_sweep_kernel = ms_sweep;
    timing->ms_other = ms_total - ms_sweep;
    timing->n_sweep_launches = launches;
    timing->applications = steps;
    return int(OU_OK);
  });
}

}  // extern "C"
