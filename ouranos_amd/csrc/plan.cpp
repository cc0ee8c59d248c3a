// Plan construction: validation of the problem descriptor, the wavefront
// schedule and the device-resident data of the operator.
//
// The reference builds, per MPI rank, one FECell per cell and one Task per
// (cell, direction) with hash maps of upwind dependencies
// (RadiativeTransfer.cc:49-584, Task.cc).  On a structured Cartesian mesh the
// dependency DAG of a direction is fixed by the signs of Omega: cell
// (i,j,k) waits for its neighbours one step upwind along each axis, so the
// cells with the same i'+j'+k' (coordinates reflected into the octant's
// frame) form one level of independent tasks, shared by every direction of
// the octant.  The schedule below lists, per level, chunks of at most 64
// cells of one octant and one material.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>

#include "fe.hpp"
#include "plan.hpp"

namespace ou {

namespace {

template <typename V>
void upload(void *dst, const V &v) {
  OU_HIP(hipMemcpy(dst, v.data(), v.size() * sizeof(typename V::value_type),
                   hipMemcpyHostToDevice));
}

void *dev_alloc(size_t bytes) {
  void *p = nullptr;
  if (bytes == 0) bytes = 8;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    char buf[160];
    std::snprintf(buf, sizeof buf, "hipMalloc of %.3f GB failed", bytes / 1e9);
    fail(OU_ERR_NOMEM, buf);
  }
  return p;
}

// Converts a double table to the sweep precision and uploads it.
void *upload_real(const std::vector<double> &v, int precision) {
  if (precision == OU_FP32) {
    std::vector<float> f(v.begin(), v.end());
    void *p = dev_alloc(f.size() * 4);
    upload(p, f);
    return p;
  }
  void *p = dev_alloc(v.size() * 8);
  upload(p, v);
  return p;
}

bool finite_all(const double *p, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(p[i])) return false;
  return true;
}

}  // namespace

PersistTables chunk_dependencies(const std::vector<Chunk> &chunks, const std::vector<int> &cells,
                                 const int n[3], int dim) {
  const int nx = n[0], ny = n[1], nz = n[2];
  std::vector<std::vector<int>> by_oct(8);
  for (size_t c = 0; c < chunks.size(); ++c) by_oct[size_t(chunks[c].octant)].push_back(int(c));
  std::vector<int> owner(size_t(nx) * ny * nz, -1);
  std::vector<std::vector<int>> deps(chunks.size());
  for (int o = 0; o < 8; ++o) {
    if (by_oct[size_t(o)].empty()) continue;
    auto cells_of = [&](int c) {
      const Chunk &ch = chunks[size_t(c)];
      return std::make_pair(ch.cell_offset, ch.cell_offset + ch.count_lo + ch.count_hi);
    };
    for (int c : by_oct[size_t(o)]) {
      const auto r = cells_of(c);
      for (int q = r.first; q < r.second; ++q) owner[size_t(cells[size_t(q)])] = c;
    }
    for (int c : by_oct[size_t(o)]) {
      std::vector<int> &d = deps[size_t(c)];
      const auto r = cells_of(c);
      for (int q = r.first; q < r.second; ++q) {
        const long cell = cells[size_t(q)];
        const int ijk[3] = {int(cell % nx), int((cell / nx) % ny), int(cell / (long(nx) * ny))};
        for (int a = 0; a < dim; ++a) {
          // one step upwind: towards +a where Omega_a < 0 (octant bit a set)
          int nb[3] = {ijk[0], ijk[1], ijk[2]};
          nb[a] += (o >> a) & 1 ? 1 : -1;
          if (nb[a] < 0 || nb[a] >= n[a]) continue;
          d.push_back(owner[size_t(nb[0] + long(nx) * (nb[1] + long(ny) * nb[2]))]);
        }
      }
      std::sort(d.begin(), d.end());
      d.erase(std::unique(d.begin(), d.end()), d.end());
      // Tickets follow the chunk order: every upwind chunk must come first
      // (it does, chunks being laid out by level), or a wait could outlast
      // the items that would end it.
      if (!d.empty() && (d.front() < 0 || d.back() >= c))
        fail(OU_ERR_ARG, "sweep schedule: an upwind chunk does not precede its dependant");
    }
  }
  PersistTables t;
  t.dep_off.assign(chunks.size() + 1, 0);
  for (size_t c = 0; c < chunks.size(); ++c) {
    t.dep_off[c + 1] = t.dep_off[c] + int(deps[c].size());
    t.dep_list.insert(t.dep_list.end(), deps[c].begin(), deps[c].end());
  }
  return t;
}

void check_sweep_watchdog(int *d_error, hipStream_t s) {
  if (!d_error) return;
  OU_HIP(hipStreamSynchronize(s));
  int e = 0;
  OU_HIP(hipMemcpy(&e, d_error, sizeof(int), hipMemcpyDeviceToHost));
  if (e != 0) {
    OU_HIP(hipMemset(d_error, 0, sizeof(int)));
    fail(OU_ERR_TIMEOUT, "a wait of the persistent sweep for its upwind level gave up (the "
                         "sweep's wait limit was reached); its results are invalid");
  }
}

unsigned wait_polls(double seconds) {
  // A poll takes at least the sleep (64 clocks) and one memory round trip,
  // about half a microsecond: the poll bound ends the wait even if the wall
  // clock were read at a wrong rate.
  const double polls = std::max(0.0, seconds) * 2.0e6;
  return polls >= 4.0e9 ? 4000000000u : unsigned(polls);
}

PersistView Plan::persist_view() const {
  PersistView pv;
  pv.dep_off = d_dep_off;
  pv.dep_list = d_dep_list;
  pv.done = d_done;
  pv.ticket = d_done + size_t(n_chunks) * kMaxDirGroups;
  pv.error = d_error;
  pv.groups = dir_groups;
  pv.n_items = int(n_chunks * dir_groups);
  pv.wait_ticks = (unsigned long long)(std::max(0.0, wait_limit_s) * wall_clock_khz * 1000.0);
  pv.max_polls = wait_polls(wait_limit_s);
  return pv;
}

Plan::~Plan() {
  void *bufs[] = {d_ainv_mfma64, d_ainv_mfma, d_ainv, d_S, d_S0, d_slot_m2d, d_slot_c,
                  d_slot_d2m,  d_partial,  d_trace[0], d_trace[1], d_trace[2], d_oct_dirs,
                  d_oct_off,   d_chunks,   d_cells,    d_mat,      d_mom_l,    d_sigma_s,
                  d_source,    d_mass,     d_rowsum,   d_inflow,   d_x,        d_y,
                  d_slot_cls, d_sigma_other, d_dep_off, d_dep_list, d_done, d_error,
                  d_tiles};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : {ev_fork, ev_join[0], ev_join[1], ev_join[2], ev_join[3]})
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t s : side)
    if (s) (void)hipStreamDestroy(s);
  if (stream) (void)hipStreamDestroy(stream);
}

Plan *create_plan(const ou_problem_desc *desc) {
  if (!desc) fail(OU_ERR_ARG, "null problem descriptor");
  const ou_problem_desc &ds = *desc;
  if (ds.dim != 2 && ds.dim != 3) fail(OU_ERR_ARG, "dim must be 2 or 3");
  if (ds.fe_order < 1 || ds.fe_order > 5) fail(OU_ERR_ARG, "fe_order must be in 1..5");
  if (ds.precision != OU_FP64 && ds.precision != OU_FP32)
    fail(OU_ERR_ARG, "precision must be OU_FP64 or OU_FP32");
  for (int a = 0; a < ds.dim; ++a) {
    if (ds.n_cells[a] < 1) fail(OU_ERR_ARG, "n_cells must be >= 1 on every axis");
    if (!(ds.cell_size[a] > 0.0) || !std::isfinite(ds.cell_size[a]))
      fail(OU_ERR_ARG, "cell sizes must be positive and finite");
  }
  if (ds.n_dir < 1 || !ds.omega || !ds.M2D || !ds.D2M || !ds.mom_l)
    fail(OU_ERR_ARG, "quadrature arrays missing");
  if (ds.n_mom < 1 || ds.n_mom > 16) fail(OU_ERR_UNSUPPORTED, "n_mom must be in 1..16");
  if (ds.n_mat < 1 || !ds.sigma_t || !ds.sigma_s || !ds.material_id)
    fail(OU_ERR_ARG, "material arrays missing");
  if (ds.L < 0) fail(OU_ERR_ARG, "L must be >= 0");
  if (!(ds.weight_sum > 0.0)) fail(OU_ERR_ARG, "weight_sum must be positive");
  // The schedule and the kernels index cells and trace slots in 32 bits: the
  // cell list holds every (octant, cell) as an int, a kernel's cell id is an
  // int, and a trace value sits at unsigned (dir F + f) plane_size + slot.
  // Larger meshes are refused here, before any array is read or allocated.
  {
    const long nx = ds.n_cells[0], ny = ds.n_cells[1], nz = ds.dim == 3 ? ds.n_cells[2] : 1;
    const long n_oct = ds.dim == 2 ? 4 : 8;
    if (nx * ny * nz * n_oct > long(INT32_MAX))
      fail(OU_ERR_UNSUPPORTED, "mesh too large: octants x cells must fit a 32-bit index");
    const long P1 = ds.fe_order + 1, F = ds.dim == 2 ? P1 : P1 * P1;
    const long planes[3] = {ny * nz, nx * nz, (nx + ny - 1) * ny};
    for (int a = 0; a < ds.dim; ++a)
      if (long(ds.n_dir) * F * planes[a] > long(UINT32_MAX))
        fail(OU_ERR_UNSUPPORTED, "mesh too large: a direction set of face traces must fit "
                                 "a 32-bit index");
  }

  auto *pl = new Plan();
  try {
    Plan &p = *pl;
    p.dim = ds.dim;
    p.p = ds.fe_order;
    p.P1 = p.p + 1;
    p.N = p.dim == 2 ? p.P1 * p.P1 : p.P1 * p.P1 * p.P1;
    p.F = p.dim == 2 ? p.P1 : p.P1 * p.P1;
    p.n[0] = ds.n_cells[0];
    p.n[1] = ds.n_cells[1];
    p.n[2] = p.dim == 3 ? ds.n_cells[2] : 1;
    for (int a = 0; a < 3; ++a) p.h[a] = a < p.dim ? ds.cell_size[a] : 1.0;
    p.n_dir = ds.n_dir;
    p.n_mom = ds.n_mom;
    p.n_mat = ds.n_mat;
    p.L = ds.L;
    p.precision = ds.precision;
    p.n_cells = long(p.n[0]) * p.n[1] * p.n[2];
    p.n_dofs = p.n_cells * p.N;
    p.vec_size = p.n_dofs * p.n_mom;
    p.ws = ds.weight_sum;
    p.omega.assign(ds.omega, ds.omega + 3 * size_t(p.n_dir));
    p.M2D.assign(ds.M2D, ds.M2D + size_t(p.n_dir) * p.n_mom);
    p.D2M.assign(ds.D2M, ds.D2M + size_t(p.n_mom) * p.n_dir);
    p.mom_l.assign(ds.mom_l, ds.mom_l + p.n_mom);
    p.sigma_t.assign(ds.sigma_t, ds.sigma_t + p.n_mat);
    p.sigma_s.assign(ds.sigma_s, ds.sigma_s + size_t(p.n_mat) * (p.L + 1));
    p.mat.assign(ds.material_id, ds.material_id + p.n_cells);
    if (ds.source)
      p.source.assign(ds.source, ds.source + p.n_cells);
    else
      p.source.assign(p.n_cells, 0.0);
    if (ds.most_normal)
      p.most_normal.assign(ds.most_normal, ds.most_normal + size_t(2 * p.dim) * p.n_dir);
    for (int f = 0; f < 6; ++f) {
      p.bc[f] = f < 2 * p.dim ? ds.bc[f] : OU_BC_VACUUM;
      p.inc[f] = f < 2 * p.dim ? ds.inc_flux[f] : 0.0;
      if (p.bc[f] == OU_BC_REFLECTIVE)
        fail(OU_ERR_UNSUPPORTED, "reflective boundaries are not implemented (the reference "
                                 "parses them but never applies them)");
      if (p.bc[f] < OU_BC_VACUUM || p.bc[f] > OU_BC_MOST_NORMAL)
        fail(OU_ERR_ARG, "unknown boundary condition");
      if (p.bc[f] == OU_BC_MOST_NORMAL && !ds.most_normal)
        fail(OU_ERR_ARG, "most-normal boundary needs the most_normal flags");
    }
    if (!finite_all(p.omega.data(), p.omega.size()) ||
        !finite_all(p.M2D.data(), p.M2D.size()) ||
        !finite_all(p.D2M.data(), p.D2M.size()) ||
        !finite_all(p.sigma_t.data(), p.sigma_t.size()) ||
        !finite_all(p.sigma_s.data(), p.sigma_s.size()) ||
        !finite_all(p.source.data(), p.source.size()))
      fail(OU_ERR_NONFINITE, "non-finite value in the problem descriptor");
    for (long c = 0; c < p.n_cells; ++c)
      if (p.mat[c] < 0 || p.mat[c] >= p.n_mat) fail(OU_ERR_ARG, "material id out of range");
    for (int m = 0; m < p.n_mom; ++m)
      if (p.mom_l[m] < 0) fail(OU_ERR_ARG, "negative moment order");
    // Moment 0 must be the isotropic harmonic (constant over directions):
    // the volumetric source is folded into it.
    for (int d = 1; d < p.n_dir; ++d)
      if (std::fabs(p.M2D[size_t(d) * p.n_mom] - p.M2D[0]) > 1e-12 * std::fabs(p.M2D[0]))
        fail(OU_ERR_ARG, "moment 0 of M2D must be the isotropic harmonic");
    if (p.M2D[0] == 0.0) fail(OU_ERR_ARG, "moment 0 of M2D is zero");

    p.mm = p.n_mom <= 1 ? 1 : p.n_mom <= 4 ? 4 : p.n_mom <= 10 ? 10 : 16;
    {
      int dev = 0;
      OU_HIP(hipGetDevice(&dev));
      OU_HIP(hipDeviceGetAttribute(&p.max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    }
    p.use_mfma = p.precision == OU_FP32 && p.N == 27;
    p.use_mfma64 = p.precision == OU_FP64 && p.N == 27;
    // Waves per workgroup: 12 for the FP32 matrix-core Q2 solve, 4 elsewhere
    // (the FP64 matrix-core Q2 solve spills registers beyond 4).
    p.waves = p.use_mfma ? 12 : 4;
    if (const char *env = std::getenv("OURANOS_SWEEP_WAVES")) p.waves = std::atoi(env);
    if (!sweep_supported(p.dim, p.p, p.mm, p.precision, p.waves, p.max_lds)) p.waves = 4;
    if (!sweep_supported(p.dim, p.p, p.mm, p.precision, p.waves, p.max_lds))
      fail(OU_ERR_UNSUPPORTED,
           "no GPU sweep kernel for this (dim, fe_order, n_mom): supported are "
           "2D Q1-Q3 and 3D Q1-Q2 with up to 16 moments, 2D Q4-Q5 with up to 10 "
           "(L <= 3) and 3D Q3 with up to 4 (L <= 1)");

    // ---- directions per octant -------------------------------------------
    p.n_oct = p.dim == 2 ? 4 : 8;
    p.dir_oct.resize(p.n_dir);
    std::vector<double> omega_abs(3 * size_t(p.n_dir));
    for (int d = 0; d < p.n_dir; ++d) {
      int o = 0;
      for (int a = 0; a < p.dim; ++a) {
        if (p.omega[3 * d + a] < 0.0) o |= 1 << a;
        omega_abs[3 * d + a] = std::fabs(p.omega[3 * d + a]);
      }
      p.dir_oct[d] = o;
    }
    p.oct_off.assign(p.n_oct + 1, 0);
    for (int d = 0; d < p.n_dir; ++d) p.oct_off[p.dir_oct[d] + 1]++;
    for (int o = 0; o < p.n_oct; ++o) p.oct_off[o + 1] += p.oct_off[o];
    p.oct_dirs.resize(p.n_dir);
    {
      std::vector<int> fill(p.oct_off.begin(), p.oct_off.end() - 1);
      for (int d = 0; d < p.n_dir; ++d) p.oct_dirs[fill[p.dir_oct[d]]++] = d;
    }
    // Inverse classes: in the canonical frame of its octant the cell matrix of
    // a direction depends on |Omega_a| (a < dim) only, so directions that
    // mirror into each other share one inverse per material.
    {
      std::vector<std::array<double, 3>> keys;
      p.dir_cls.assign(p.n_dir, -1);
      for (int d = 0; d < p.n_dir; ++d) {
        std::array<double, 3> k = {0.0, 0.0, 0.0};
        for (int a = 0; a < p.dim; ++a) k[a] = omega_abs[3 * d + a];
        int found = -1;
        for (size_t c = 0; c < keys.size(); ++c)
          if (keys[c] == k) {
            found = int(c);
            break;
          }
        if (found < 0) {
          found = int(keys.size());
          keys.push_back(k);
        }
        p.dir_cls[d] = found;
      }
      p.n_cls = int(keys.size());
      p.cls_omega.resize(3 * size_t(p.n_cls));
      for (int c = 0; c < p.n_cls; ++c)
        for (int a = 0; a < 3; ++a) p.cls_omega[3 * c + a] = keys[c][a];
    }

    // ---- wavefront schedule ------------------------------------------------
    const int nx = p.n[0], ny = p.n[1], nz = p.n[2];
    p.n_levels = nx + ny + nz - 2;
    std::vector<std::vector<Chunk>> per_level(p.n_levels);
    std::vector<int> cells;
    cells.reserve(size_t(p.n_cells) * p.n_oct);
    std::vector<std::vector<long>> bucket(p.n_levels);
    for (int o = 0; o < p.n_oct; ++o) {
      if (p.oct_off[o + 1] == p.oct_off[o]) continue;
      for (auto &b : bucket) b.clear();
      // Canonical coordinates (kk, jj, ii) in ascending order: within a
      // level, consecutive cells then share kk and have consecutive jj.
      for (int kk = 0; kk < nz; ++kk)
        for (int jj = 0; jj < ny; ++jj)
          for (int ii = 0; ii < nx; ++ii) {
            const int i = (o & 1) ? nx - 1 - ii : ii;
            const int j = (o & 2) ? ny - 1 - jj : jj;
            const int k = (o & 4) ? nz - 1 - kk : kk;
            bucket[ii + jj + kk].push_back(i + long(nx) * (j + long(ny) * k));
          }
      for (int l = 0; l < p.n_levels; ++l) {
        std::vector<long> &b = bucket[l];
        std::stable_sort(b.begin(), b.end(),
                         [&](long a, long c) { return p.mat[a] < p.mat[c]; });
        // Half chunks: runs of up to 32 cells of one material.  Consecutive
        // halves pair into chunks; with MFMA cell solves the two halves of a
        // chunk may carry different materials, otherwise a chunk ends where
        // its material does.
        struct Half {
          int material;
          size_t begin, count;
        };
        std::vector<Half> halves;
        for (size_t s = 0; s < b.size();) {
          const int m = p.mat[b[s]];
          size_t e = s;
          while (e < b.size() && p.mat[b[e]] == m) ++e;
          for (size_t q = s; q < e; q += 32) halves.push_back({m, q, std::min<size_t>(32, e - q)});
          s = e;
        }
        for (size_t h = 0; h < halves.size();) {
          const Half &lo = halves[h];
          Chunk c{o, lo.material, lo.material, int(lo.count), 0, int(cells.size())};
          for (size_t q = 0; q < lo.count; ++q) cells.push_back(int(b[lo.begin + q]));
          ++h;
          // a partial first half only pairs with another material on the
          // matrix-core path; a full first half pairs with the next half of
          // its own material on every path
          if (h < halves.size() && (p.use_mfma || (lo.count == 32 && halves[h].material == lo.material))) {
            const Half &hi = halves[h];
            c.material_hi = hi.material;
            c.count_hi = int(hi.count);
            for (size_t q = 0; q < hi.count; ++q) cells.push_back(int(b[hi.begin + q]));
            ++h;
          }
          per_level[l].push_back(c);
        }
      }
    }
    // Each level lists its chunks by octant quarter (octants q*n_oct/4 ..):
    // octants are independent, so the parts of a level may be launched on
    // separate streams, and one part's level fills the CUs another's leaves
    // idle while its last workgroups drain.  Within a quarter, chunks are
    // ordered so that all chunks of one
    // material (every octant uses the same inverse classes) land on the same
    // XCD: workgroup i of a launch runs on XCD i mod 8, each XCD with its own
    // L2.  Chunks are bucketed by material mod 8 and dealt out so that launch
    // position p draws from bucket p mod 8 (from the fullest bucket once that
    // one is empty).  Positions count from the start of the launch in the
    // default two-part split: a level's launch for octants 0-3 (4-7) holds
    // quarters 0 and 1 (2 and 3), so quarter 1 (3) continues after quarter 0 (2).
    constexpr int kXcds = 8;
    std::vector<Chunk> chunks;
    p.level_chunk_off.assign(p.n_levels + 1, 0);
    p.level_quarter.assign(4 * size_t(p.n_levels), 0);
    for (int l = 0; l < p.n_levels; ++l) {
      p.level_chunk_off[l] = int(chunks.size());
      for (int q = 0; q < 4; ++q) {
        p.level_quarter[4 * l + q] = int(chunks.size());
        std::vector<std::vector<Chunk>> xb(kXcds);
        size_t n_half = 0;
        for (const Chunk &c : per_level[l])
          if (c.octant * 4 / p.n_oct == q) {
            xb[c.material % kXcds].push_back(c);
            ++n_half;
          }
        std::vector<size_t> next(kXcds, 0);
        const size_t base = (q & 1) ? chunks.size() - size_t(p.level_quarter[4 * l + q - 1]) : 0;
        for (size_t pos = 0; pos < n_half; ++pos) {
          int b = int((base + pos) % kXcds);
          if (next[b] == xb[b].size()) {
            size_t best = 0;
            for (int q = 0; q < kXcds; ++q)
              if (xb[q].size() - next[q] > best) {
                best = xb[q].size() - next[q];
                b = q;
              }
          }
          chunks.push_back(xb[b][next[b]++]);
        }
      }
    }
    p.level_chunk_off[p.n_levels] = int(chunks.size());
    p.n_chunks = long(chunks.size());
    // upwind chunks of every chunk (the persistent sweep's dependencies)
    p.deps = chunk_dependencies(chunks, cells, p.n, p.dim);

    // ---- direction groups ----------------------------------------------------
    // Enough workgroups per level launch to give each CU about two (a 2D level
    // or a small mesh has few chunks), but at least one full round of
    // directions per workgroup: each group re-reads its chunk's source and
    // writes its own partial moments.
    {
      int dev = 0, cus = 256, khz = 0;
      if (hipGetDevice(&dev) == hipSuccess) {
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess &&
            khz > 0)
          p.wall_clock_khz = double(khz);
      }
      p.cus = std::max(cus, 1);
      int max_dirs = 1;
      for (int o = 0; o < p.n_oct; ++o) max_dirs = std::max(max_dirs, p.oct_off[o + 1] - p.oct_off[o]);
      const double per_level = double(std::max<long>(p.n_chunks, 1)) / std::max(p.n_levels, 1);
      const int g_fill = int(std::ceil(2.0 * std::max(cus, 1) / per_level));
      const int g_dirs = std::max(1, max_dirs / std::max(p.waves, 1));
      // A sweep workgroup stages the tables of its direction group's range in
      // LDS (at most kMaxSlotsPerGroup directions): octants with more
      // directions need several groups.
      p.min_dir_groups = (max_dirs + kMaxSlotsPerGroup - 1) / kMaxSlotsPerGroup;
      if (p.min_dir_groups > kMaxDirGroups)
        fail(OU_ERR_UNSUPPORTED, "more than " + std::to_string(kMaxSlotsPerGroup * kMaxDirGroups) +
                                     " directions in one octant");
      p.dir_groups = std::max(p.min_dir_groups, std::min({g_fill, g_dirs, kMaxDirGroups}));
      p.partial_groups = p.dir_groups;
      for (int o = 0; o < p.n_oct; ++o)
        if (p.oct_off[o + 1] > p.oct_off[o]) p.oct_mask |= 1u << o;

      // ---- tile schedule (OU_SWEEP_TILES) -------------------------------------
      // Tiles of 8 x 8 (2D) or 4 x 4 x 4 (3D) canonical cells per octant, listed
      // by tile level.  Each tile's cells must share one material (its
      // inverses sit in registers for the whole tile); otherwise, or without a
      // tile kernel for this element, the plan has no tile mode.
      const int twm = tile_waves_max(p.dim, p.p, p.mm, p.precision, p.max_lds);
      if (twm > 0) {
        const int E = p.dim == 2 ? kTileEdge2 : kTileEdge3;
        const int ntx = (nx + E - 1) / E, nty = (ny + E - 1) / E;
        const int ntz = p.dim == 3 ? (nz + E - 1) / E : 1;
        std::vector<std::vector<TileItem>> by_level(size_t(ntx + nty + ntz - 2));
        bool uniform = true;
        for (int o = 0; o < p.n_oct && uniform; ++o) {
          if (p.oct_off[o + 1] == p.oct_off[o]) continue;
          for (int tk = 0; tk < ntz && uniform; ++tk)
            for (int tj = 0; tj < nty && uniform; ++tj)
              for (int ti = 0; ti < ntx && uniform; ++ti) {
                int m = -1;
                for (int c = 0; c < (p.dim == 3 ? E : 1) && uniform; ++c)
                  for (int b = 0; b < E && uniform; ++b)
                    for (int a = 0; a < E; ++a) {
                      const int ii = ti * E + a, jj = tj * E + b, kk = tk * E + c;
                      if (ii >= nx || jj >= ny || kk >= nz) continue;
                      const int i = (o & 1) ? nx - 1 - ii : ii;
                      const int j = (o & 2) ? ny - 1 - jj : jj;
                      const int k = (p.dim == 3 && (o & 4)) ? nz - 1 - kk : kk;
                      const int mc = p.mat[size_t(i) + size_t(nx) * (size_t(j) + size_t(ny) * k)];
                      if (m < 0) m = mc;
                      if (mc != m) {
                        uniform = false;
                        break;
                      }
                    }
                by_level[size_t(ti + tj + tk)].push_back(TileItem{o, ti, tj, tk, m});
              }
        }
        if (uniform && p.oct_mask != 0) {
          std::vector<TileItem> tiles;
          p.tile_level_off.assign(1, 0);
          size_t widest = 1;
          for (const auto &lv : by_level) {
            tiles.insert(tiles.end(), lv.begin(), lv.end());
            p.tile_level_off.push_back(int(tiles.size()));
            widest = std::max(widest, lv.size());
          }
          p.d_tiles = static_cast<TileItem *>(dev_alloc(tiles.size() * sizeof(TileItem)));
          upload(p.d_tiles, tiles);
          p.tiles_ok = true;
          // About one workgroup per CU at the widest tile level, each with at
          // most kMaxSlotsPerGroup directions, preferably in one round of
          // dpw directions per wave.
          const int dpw = p.dim == 2 ? 8 : 4;
          const int mid = int(widest);
          const bool q2 = p.dim == 3 && p.p == 2;  // the matrix-core tile kernel
          int g = std::min(kMaxDirGroups, std::max(1, (p.cus + mid - 1) / mid));
          g = std::min(g, (max_dirs + dpw - 1) / dpw);
          // The Q2 kernel sweeps its group in rounds and keeps one set of FP64
          // accumulators for all of them: more groups only add partial moments.
          if (!q2) g = std::max(g, (max_dirs + dpw * twm - 1) / (dpw * twm));
          g = std::max(p.min_dir_groups, std::min(g, kMaxDirGroups));
          const int glen = (max_dirs + g - 1) / g;
          p.tile_groups = g;
          const int need = (glen + dpw - 1) / dpw;  // waves for one round
          if (q2) {
            // 3 or 4 waves: the fewer idle direction slots over the rounds
            auto idle = [&](int w) { return (glen + dpw * w - 1) / (dpw * w) * dpw * w - glen; };
            p.tile_waves = (twm >= 4 && idle(4) < idle(3)) ? 4 : 3;
          } else {
            p.tile_waves = std::min(twm, need <= 1 ? 1 : need <= 2 ? 2 : 4);
          }
          p.partial_groups = std::max(p.partial_groups, p.tile_groups);
          // The Q2 matrix-core tile sweep is opt-in: on the 128^3 headline it
          // measured 116 ms per sweep against 54.6 ms in level launches.
          if (!q2) p.sweep_mode = OU_SWEEP_TILES;
        }
      }
    }

    // ---- boundary inflow per (direction, axis) ----------------------------
    std::vector<double> inflow(3 * size_t(p.n_dir), 0.0);
    for (int d = 0; d < p.n_dir; ++d)
      for (int a = 0; a < p.dim; ++a) {
        const int f = 2 * a + (p.omega[3 * d + a] >= 0.0 ? 0 : 1);
        double v = 0.0;
        if (p.bc[f] == OU_BC_ISOTROPIC ||
            (p.bc[f] == OU_BC_MOST_NORMAL && p.most_normal[size_t(f) * p.n_dir + d]))
          v = p.inc[f];
        inflow[3 * d + a] = v / p.ws;
      }

    // ---- element data -------------------------------------------------------
    RefElement re(p.dim, p.p, p.h);
    std::vector<double> mass;
    re.mass(mass);
    std::vector<double> rowsum(p.N, 0.0);
    for (int i = 0; i < p.N; ++i)
      for (int j = 0; j < p.N; ++j) rowsum[i] += mass[size_t(i) * p.N + j];
    for (int k = 0; k < p.P1 * p.P1; ++k) p.m1[k] = re.m1[k];

    // ---- per-slot direction tables (slot g = direction oct_dirs[g]) ------------
    // The face mass of axis a is (area_a) m1 (x) m1 in 3D and h_b m1 in 2D
    // (area_a = product of the cell sizes along the other axes); the kernel
    // applies the m1 factors and takes |Omega_a| area_a from this table.
    std::vector<double> slot_m2d(size_t(p.n_dir) * p.mm, 0.0), slot_d2m(size_t(p.n_dir) * p.mm, 0.0),
        slot_c(size_t(p.n_dir) * 4, 0.0);
    for (int g = 0; g < p.n_dir; ++g) {
      const int d = p.oct_dirs[g];
      for (int m = 0; m < p.n_mom; ++m) {
        slot_m2d[size_t(g) * p.mm + m] = p.M2D[size_t(d) * p.n_mom + m];
        slot_d2m[size_t(g) * p.mm + m] = p.D2M[size_t(m) * p.n_dir + d];
      }
      for (int a = 0; a < p.dim; ++a) {
        double area = 1.0;
        for (int b2 = 0; b2 < p.dim; ++b2)
          if (b2 != a) area *= p.h[b2];
        slot_c[size_t(g) * 4 + a] = omega_abs[3 * d + a] * area;
      }
    }

    // ---- device buffers -----------------------------------------------------
    OU_HIP(hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking));
    for (auto &s : p.side) OU_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (auto &e : p.ev) OU_HIP(hipEventCreate(&e));
    OU_HIP(hipEventCreateWithFlags(&p.ev_fork, hipEventDisableTiming));
    for (auto &e : p.ev_join) OU_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    const size_t rs = p.real_size();
    p.d_ainv = dev_alloc(size_t(p.n_mat) * p.n_cls * p.N * p.N * rs);
    if (p.use_mfma)
      p.d_ainv_mfma = static_cast<float *>(dev_alloc(size_t(p.n_mat) * p.n_cls * 64 * 16 * 4));
    if (p.use_mfma64)
      p.d_ainv_mfma64 = static_cast<double *>(dev_alloc(size_t(p.n_mat) * p.n_cls * 64 * 14 * 8));
    p.d_S = dev_alloc(size_t(p.vec_size) * rs);
    OU_HIP(hipMemset(p.d_S, 0, size_t(p.vec_size) * rs));
    p.d_S0 = dev_alloc(size_t(p.n_dofs) * rs);
    p.d_slot_m2d = upload_real(slot_m2d, p.precision);
    p.d_slot_c = upload_real(slot_c, p.precision);
    p.d_slot_d2m = static_cast<double *>(dev_alloc(slot_d2m.size() * 8));
    upload(p.d_slot_d2m, slot_d2m);
    p.d_partial = static_cast<double *>(
        dev_alloc(size_t(p.n_oct) * p.partial_groups * p.vec_size * 8));
    OU_HIP(hipMemset(p.d_partial, 0, size_t(p.n_oct) * p.partial_groups * p.vec_size * 8));
    p.plane_size[0] = long(ny) * nz;
    p.plane_size[1] = long(nx) * nz;
    p.plane_size[2] = long(nx + ny - 1) * ny;
    for (int a = 0; a < p.dim; ++a)
      p.d_trace[a] = dev_alloc(size_t(p.n_dir) * p.F * p.plane_size[a] * rs);
    p.d_oct_dirs = static_cast<int *>(dev_alloc(p.oct_dirs.size() * 4));
    upload(p.d_oct_dirs, p.oct_dirs);
    {
      std::vector<int> slot_cls(p.n_dir);
      for (int g = 0; g < p.n_dir; ++g) slot_cls[g] = p.dir_cls[p.oct_dirs[g]];
      p.d_slot_cls = static_cast<int *>(dev_alloc(slot_cls.size() * 4));
      upload(p.d_slot_cls, slot_cls);
    }
    p.d_oct_off = static_cast<int *>(dev_alloc(p.oct_off.size() * 4));
    upload(p.d_oct_off, p.oct_off);
    p.d_chunks = static_cast<Chunk *>(dev_alloc(chunks.size() * sizeof(Chunk)));
    upload(p.d_chunks, chunks);
    p.d_cells = static_cast<int *>(dev_alloc(cells.size() * 4));
    upload(p.d_cells, cells);
    p.d_mat = static_cast<int *>(dev_alloc(p.mat.size() * 4));
    upload(p.d_mat, p.mat);
    p.d_mom_l = static_cast<int *>(dev_alloc(p.mom_l.size() * 4));
    upload(p.d_mom_l, p.mom_l);
    p.d_sigma_s = static_cast<double *>(dev_alloc(p.sigma_s.size() * 8));
    upload(p.d_sigma_s, p.sigma_s);
    p.d_source = static_cast<double *>(dev_alloc(p.source.size() * 8));
    upload(p.d_source, p.source);
    p.d_mass = static_cast<double *>(dev_alloc(mass.size() * 8));
    upload(p.d_mass, mass);
    p.d_rowsum = static_cast<double *>(dev_alloc(rowsum.size() * 8));
    upload(p.d_rowsum, rowsum);
    p.d_inflow = static_cast<double *>(dev_alloc(inflow.size() * 8));
    upload(p.d_inflow, inflow);
    p.d_dep_off = static_cast<int *>(dev_alloc(p.deps.dep_off.size() * 4));
    upload(p.d_dep_off, p.deps.dep_off);
    p.d_dep_list = static_cast<int *>(dev_alloc(p.deps.dep_list.size() * 4));
    upload(p.d_dep_list, p.deps.dep_list);
    {
      const size_t nc = size_t(p.n_chunks) * kMaxDirGroups + 1;
      p.d_done = static_cast<unsigned *>(dev_alloc(nc * 4));
      OU_HIP(hipMemset(p.d_done, 0, nc * 4));
      p.d_error = static_cast<int *>(dev_alloc(4));
      OU_HIP(hipMemset(p.d_error, 0, 4));
    }

    launch_build_inverses(p);
    return pl;
  } catch (...) {
    delete pl;
    throw;
  }
}

void set_sweep_shape(Plan &pl, int waves, int dir_groups) {
  if (dir_groups < 1 || dir_groups > kMaxDirGroups)
    fail(OU_ERR_ARG, "direction groups per octant must be in [1, 8]");
  if (dir_groups < pl.min_dir_groups)
    fail(OU_ERR_ARG, "this plan's octants hold more than " +
                         std::to_string(kMaxSlotsPerGroup) + " directions: at least " +
                         std::to_string(pl.min_dir_groups) + " direction groups are needed");
  if (!sweep_supported(pl.dim, pl.p, pl.mm, pl.precision, waves, pl.max_lds))
    fail(OU_ERR_UNSUPPORTED, "no sweep kernel with this many waves for this element");
  if (dir_groups > pl.partial_groups) {
    const size_t bytes = size_t(pl.n_oct) * dir_groups * pl.vec_size * 8;
    OU_HIP(hipDeviceSynchronize());
    void *fresh = dev_alloc(bytes);
    (void)hipFree(pl.d_partial);
    pl.d_partial = static_cast<double *>(fresh);
    pl.partial_groups = dir_groups;
  }
  if (dir_groups != pl.dir_groups) {
    // The epilogue sums the slots of all octants x groups; an octant without
    // directions has no chunks, so nothing writes its slots.  Its slots must
    // hold zeros in the new (octant, group) layout, whose offsets differ from
    // the old one's.
    OU_HIP(hipDeviceSynchronize());
    OU_HIP(hipMemset(pl.d_partial, 0, size_t(pl.n_oct) * pl.partial_groups * pl.vec_size * 8));
  }
  pl.waves = waves;
  pl.dir_groups = dir_groups;
}

// The partial moments hold n_oct x groups slots per mode; the epilogue sums
// only the slots of octants with directions, so a layout change needs no
// clearing, only room.
static void reserve_partial_groups(Plan &pl, int groups) {
  if (groups <= pl.partial_groups) return;
  const size_t bytes = size_t(pl.n_oct) * groups * pl.vec_size * 8;
  OU_HIP(hipDeviceSynchronize());
  void *fresh = dev_alloc(bytes);
  (void)hipFree(pl.d_partial);
  pl.d_partial = static_cast<double *>(fresh);
  pl.partial_groups = groups;
}

void set_tile_shape(Plan &pl, int waves, int groups) {
  if (!pl.tiles_ok)
    fail(OU_ERR_UNSUPPORTED, "this plan has no tile sweep (2D Q1-Q2 with one material per 8x8 "
                             "tile, 3D Q1, or 3D Q2 in FP32 with at most 4 moments, with one "
                             "material per 4x4x4 tile)");
  if (groups < std::max(1, pl.min_dir_groups) || groups > kMaxDirGroups)
    fail(OU_ERR_ARG, "direction groups per quadrant must be in [" +
                         std::to_string(std::max(1, pl.min_dir_groups)) + ", 8]");
  if (!tile_sweep_supported(pl.dim, pl.p, pl.mm, pl.precision, waves, pl.max_lds))
    fail(OU_ERR_UNSUPPORTED, "no tile sweep kernel with this many waves for this element");
  reserve_partial_groups(pl, groups);
  pl.tile_waves = waves;
  pl.tile_groups = groups;
}

void set_sweep_mode(Plan &pl, int mode) {
  if (mode != OU_SWEEP_LEVELS && mode != OU_SWEEP_PERSISTENT && mode != OU_SWEEP_TILES)
    fail(OU_ERR_ARG, "sweep mode must be OU_SWEEP_LEVELS, OU_SWEEP_PERSISTENT or OU_SWEEP_TILES");
  if (mode == OU_SWEEP_TILES && !pl.tiles_ok)
    fail(OU_ERR_UNSUPPORTED, "this plan has no tile sweep (2D Q1-Q2 with one material per 8x8 "
                             "tile, 3D Q1, or 3D Q2 in FP32 with at most 4 moments, with one "
                             "material per 4x4x4 tile)");
  pl.sweep_mode = mode;
}

}  // namespace ou
