// The persistent multi-instance launch (merged level schedule, captured HIP
// graphs, device pointer table) and its autotuner.
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <set>
#include <sstream>

#include "vec.hpp"

namespace ou {

namespace {

// ---------------------------------------------------------------------------
// Tuning cache: key -> (sweep mode, waves, direction groups).  Kept for the
// process; when OURANOS_TUNE_CACHE names a file, entries are read from it on
// first use and every new result is appended ("<key> | <mode> <waves>
// <groups>" per line).
// ---------------------------------------------------------------------------
struct TuneChoice {
  int mode = OU_SWEEP_LEVELS, waves = 0, groups = 0;
};
std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune_cache;
bool g_tune_loaded = false;

void load_tune_cache_locked() {
  if (g_tune_loaded) return;
  g_tune_loaded = true;
  const char *path = std::getenv("OURANOS_TUNE_CACHE");
  if (!path) return;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    const size_t bar = line.rfind(" | ");
    if (bar == std::string::npos || bar == 0) continue;
    std::istringstream rest(line.substr(bar + 3));
    TuneChoice c;
    if (!(rest >> c.mode >> c.waves >> c.groups)) continue;
    if (c.waves > 0 && c.groups > 0) g_tune_cache[line.substr(0, bar)] = c;
  }
}

bool tune_lookup(const std::string &key, TuneChoice &c) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  load_tune_cache_locked();
  auto it = g_tune_cache.find(key);
  if (it == g_tune_cache.end()) return false;
  c = it->second;
  return true;
}

void tune_store(const std::string &key, const TuneChoice &c) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  load_tune_cache_locked();
  g_tune_cache[key] = c;
  if (const char *path = std::getenv("OURANOS_TUNE_CACHE")) {
    std::ofstream out(path, std::ios::app);
    out << key << " | " << c.mode << ' ' << c.waves << ' ' << c.groups << '\n';
  }
}

template <typename T>
T *dev_upload(const std::vector<T> &v) {
  T *p = nullptr;
  OU_HIP(hipMalloc(&p, std::max<size_t>(1, v.size()) * sizeof(T)));
  if (!v.empty()) OU_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return p;
}

int max_dirs_per_octant(const Plan &pl) {
  int m = 0;
  for (int o = 0; o < pl.n_oct; ++o) m = std::max(m, pl.oct_off[o + 1] - pl.oct_off[o]);
  return m;
}

}  // namespace

Batch::Batch(std::vector<RadiativeTransfer *> members, int waves, int dir_groups)
    : members_(std::move(members)) {
  try {
    if (members_.empty()) fail(OU_ERR_ARG, "empty batch");
    for (RadiativeTransfer *m : members_)
      if (!m) fail(OU_ERR_ARG, "null batch member");
    {
      std::set<const RadiativeTransfer *> uniq(members_.begin(), members_.end());
      if (uniq.size() != members_.size())
        fail(OU_ERR_ARG, "a plan appears twice in the batch (members share no state)");
    }
    OU_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    OU_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
    OU_HIP(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    OU_HIP(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
    bl_.side = side_;
    bl_.ev_fork = ev_fork_;
    bl_.ev_join = ev_join_;
    OU_HIP(hipEventCreateWithFlags(&io_done_, hipEventDisableTiming));
    OU_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_io_), 2 * members_.size() * sizeof(double *)));
    OU_HIP(hipMalloc(reinterpret_cast<void **>(&d_io_), 2 * members_.size() * sizeof(double *)));
    OU_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_active_), members_.size() * sizeof(int)));
    OU_HIP(hipMalloc(reinterpret_cast<void **>(&d_active_), members_.size() * sizeof(int)));
    std::fill(h_active_, h_active_ + members_.size(), 1);
    upload_active(stream_);
    OU_HIP(hipStreamSynchronize(stream_));
    bl_.active = d_active_;
    build_schedule();
    if (waves > 0 && dir_groups > 0) {
      set_shape(waves, dir_groups);
    } else {
      TuneChoice c;
      if (tune_lookup(tune_key(), c) && (c.mode != OU_SWEEP_TILES || bl_.tiles_ok)) {
        set_sweep_mode(c.mode);
        set_shape(c.waves, c.groups);
      } else {
        tune(3);
      }
    }
  } catch (...) {
    release();
    throw;
  }
}

Batch::~Batch() { release(); }

void Batch::release() {
  drop_graphs();
  for (void *p : {static_cast<void *>(d_dep_off_), static_cast<void *>(d_dep_list_),
                  static_cast<void *>(d_done_), static_cast<void *>(d_error_)})
    if (p) (void)hipFree(p);
  d_dep_off_ = d_dep_list_ = d_error_ = nullptr;
  d_done_ = nullptr;
  for (void *p : {static_cast<void *>(d_chunks_), static_cast<void *>(d_chunk_view_),
                  static_cast<void *>(d_tiles_), static_cast<void *>(d_tile_view_),
                  static_cast<void *>(d_views_), static_cast<void *>(d_members_),
                  static_cast<void *>(d_io_), static_cast<void *>(d_si_work_),
                  static_cast<void *>(d_si_partial_), static_cast<void *>(d_si_norms_),
                  static_cast<void *>(d_active_)})
    if (p) (void)hipFree(p);
  d_active_ = nullptr;
  if (h_active_) (void)hipHostFree(h_active_);
  h_active_ = nullptr;
  d_si_work_ = nullptr;
  d_si_partial_ = nullptr;
  d_si_norms_ = nullptr;
  if (h_si_norms_) (void)hipHostFree(h_si_norms_);
  h_si_norms_ = nullptr;
  d_chunks_ = nullptr;
  d_chunk_view_ = nullptr;
  d_tiles_ = nullptr;
  d_tile_view_ = nullptr;
  d_views_ = nullptr;
  d_members_ = nullptr;
  d_io_ = nullptr;
  if (h_io_) (void)hipHostFree(h_io_);
  h_io_ = nullptr;
  if (io_done_) (void)hipEventDestroy(io_done_);
  io_done_ = nullptr;
  for (hipEvent_t e : {ev_fork_, ev_join_})
    if (e) (void)hipEventDestroy(e);
  ev_fork_ = ev_join_ = nullptr;
  if (side_) (void)hipStreamDestroy(side_);
  side_ = nullptr;
  if (stream_) (void)hipStreamDestroy(stream_);
  stream_ = nullptr;
}

void Batch::build_schedule() {
  const Plan &p0 = members_[0]->plan();
  bl_.n_members = size();
  bl_.dim = p0.dim;
  bl_.p = p0.p;
  bl_.N = p0.N;
  bl_.mm = p0.mm;
  bl_.precision = p0.precision;
  int n_levels = 0;
  long max_elems = 0;
  std::vector<std::vector<Chunk>> host(members_.size());
  for (size_t u = 0; u < members_.size(); ++u) {
    const Plan &pl = members_[u]->plan();
    if (pl.dim != p0.dim || pl.p != p0.p || pl.mm != p0.mm || pl.precision != p0.precision)
      fail(OU_ERR_ARG, "batch members must share dimension, element order, moment count "
                       "and precision");
    if (pl.group != 0) fail(OU_ERR_UNSUPPORTED, "batch members must use energy group 0");
    n_levels = std::max(n_levels, pl.n_levels);
    max_elems = std::max({max_elems, pl.vec_size, pl.n_cells * pl.N});
    for (int a = 0; a < pl.dim; ++a)
      max_elems = std::max(max_elems, long(pl.F) * pl.plane_size[a] * pl.n_dir);
    host[u].resize(size_t(pl.n_chunks));
    if (pl.n_chunks > 0)
      OU_HIP(hipMemcpy(host[u].data(), pl.d_chunks, size_t(pl.n_chunks) * sizeof(Chunk),
                       hipMemcpyDeviceToHost));
  }
  std::vector<Chunk> chunks;
  std::vector<int> view, local_index;
  bl_.level_off.assign(1, 0);
  bl_.level_split.assign(size_t(n_levels), 0);
  for (int l = 0; l < n_levels; ++l) {
    // every member's first octant half (quarters 0-1), then every second half
    for (int half = 0; half < 2; ++half) {
      if (half == 1) bl_.level_split[size_t(l)] = int(chunks.size());
      for (size_t u = 0; u < members_.size(); ++u) {
        const Plan &pl = members_[u]->plan();
        if (l >= pl.n_levels) continue;
        for (int c = pl.quarter_begin(l, 2 * half); c < pl.quarter_begin(l, 2 * half + 2); ++c) {
          chunks.push_back(host[u][size_t(c)]);
          view.push_back(int(u));
          local_index.push_back(c);
        }
      }
    }
    bl_.level_off.push_back(int(chunks.size()));
  }
  bl_.max_elems = max_elems;
  d_chunks_ = dev_upload(chunks);
  d_chunk_view_ = dev_upload(view);
  bl_.chunks = d_chunks_;
  bl_.chunk_view = d_chunk_view_;
  // The persistent launch's dependency lists: each member's own, with its
  // chunks renumbered into the merged list.
  {
    if (long(chunks.size()) * kMaxDirGroups > long(INT32_MAX))
      fail(OU_ERR_UNSUPPORTED, "too many chunks for the batch's persistent launch");
    std::vector<std::vector<int>> merged_of(members_.size());
    for (size_t u = 0; u < members_.size(); ++u)
      merged_of[u].assign(size_t(members_[u]->plan().n_chunks), -1);
    for (size_t i = 0; i < chunks.size(); ++i)
      merged_of[size_t(view[i])][size_t(local_index[i])] = int(i);
    std::vector<int> dep_off(chunks.size() + 1, 0), dep_list;
    for (size_t i = 0; i < chunks.size(); ++i) {
      const PersistTables &t = members_[size_t(view[i])]->plan().deps;
      const int c = local_index[i];
      for (int k = t.dep_off[size_t(c)]; k < t.dep_off[size_t(c) + 1]; ++k)
        dep_list.push_back(merged_of[size_t(view[i])][size_t(t.dep_list[size_t(k)])]);
      dep_off[i + 1] = int(dep_list.size());
    }
    d_dep_off_ = dev_upload(dep_off);
    d_dep_list_ = dev_upload(dep_list);
    const size_t nd = chunks.size() * kMaxDirGroups + 1;
    OU_HIP(hipMalloc(reinterpret_cast<void **>(&d_done_), nd * sizeof(unsigned)));
    OU_HIP(hipMemset(d_done_, 0, nd * sizeof(unsigned)));
    OU_HIP(hipMalloc(reinterpret_cast<void **>(&d_error_), sizeof(int)));
    OU_HIP(hipMemset(d_error_, 0, sizeof(int)));
    bl_.cus = p0.cus;
    PersistView &pv = bl_.persist;
    pv.dep_off = d_dep_off_;
    pv.dep_list = d_dep_list_;
    pv.done = d_done_;
    pv.ticket = d_done_ + chunks.size() * kMaxDirGroups;
    pv.error = d_error_;
    pv.wait_ticks =
        (unsigned long long)(std::max(0.0, p0.wait_limit_s) * p0.wall_clock_khz * 1000.0);
    pv.max_polls = wait_polls(p0.wait_limit_s);
  }
  bl_.waves = p0.waves;
  bl_.dir_groups = p0.dir_groups;

  // The merged tile schedule (OU_SWEEP_TILES): tile level l of the batch
  // holds tile level l of every member.
  bl_.tiles_ok = true;
  for (const RadiativeTransfer *m : members_) bl_.tiles_ok = bl_.tiles_ok && m->plan().tiles_ok;
  if (bl_.tiles_ok) {
    size_t n_tl = 0;
    std::vector<std::vector<TileItem>> th(members_.size());
    for (size_t u = 0; u < members_.size(); ++u) {
      const Plan &pl = members_[u]->plan();
      n_tl = std::max(n_tl, pl.tile_level_off.size() - 1);
      th[u].resize(size_t(pl.tile_level_off.back()));
      if (!th[u].empty())
        OU_HIP(hipMemcpy(th[u].data(), pl.d_tiles, th[u].size() * sizeof(TileItem),
                         hipMemcpyDeviceToHost));
    }
    std::vector<TileItem> tiles;
    std::vector<int> tview;
    bl_.tile_level_off.assign(1, 0);
    for (size_t l = 0; l < n_tl; ++l) {
      for (size_t u = 0; u < members_.size(); ++u) {
        const Plan &pl = members_[u]->plan();
        if (l + 1 >= pl.tile_level_off.size()) continue;
        for (int i = pl.tile_level_off[l]; i < pl.tile_level_off[l + 1]; ++i) {
          tiles.push_back(th[u][size_t(i)]);
          tview.push_back(int(u));
        }
      }
      bl_.tile_level_off.push_back(int(tiles.size()));
    }
    d_tiles_ = dev_upload(tiles);
    d_tile_view_ = dev_upload(tview);
    bl_.tiles = d_tiles_;
    bl_.tile_view = d_tile_view_;
    bl_.tile_waves = p0.tile_waves;
    bl_.tile_groups = p0.tile_groups;
  }
}

void Batch::build_views() {
  const size_t n = members_.size();
  std::vector<SweepView> views(2 * n);
  std::vector<MemberView> mv(n);
  seen_partial_.assign(n, nullptr);
  const bool tiles = bl_.sweep_mode == OU_SWEEP_TILES;
  const int groups = tiles ? bl_.tile_groups : bl_.dir_groups;
  for (size_t u = 0; u < n; ++u) {
    const Plan &pl = members_[u]->plan();
    views[u] = pl.view();
    views[u].dir_groups = groups;
    views[n + u] = views[u];
    views[n + u].S0 = pl.d_S0;
    MemberView &m = mv[u];
    m.n_cells = pl.n_cells;
    m.n_dofs = pl.n_dofs;
    m.vec_size = pl.vec_size;
    m.n_mom = pl.n_mom;
    m.L = pl.L;
    m.n_dir = pl.n_dir;
    m.n_parts = pl.n_oct * groups;
    m.groups = groups;
    m.oct_mask = pl.oct_mask;
    m.mat = pl.d_mat;
    m.sigma_s = pl.d_sigma_s;
    m.mom_l = pl.d_mom_l;
    m.mass = pl.d_mass;
    m.source = pl.d_source;
    m.rowsum = pl.d_rowsum;
    m.s0_factor = 1.0 / (pl.ws * pl.M2D[0]);
    m.S = pl.d_S;
    m.S0 = pl.d_S0;
    for (int a = 0; a < 3; ++a) {
      m.trace[a] = pl.d_trace[a];
      m.trace_per_dir[a] = a < pl.dim ? long(pl.F) * pl.plane_size[a] : 0;
    }
    m.inflow = pl.d_inflow;
    m.partial = pl.d_partial;
    seen_partial_[u] = pl.d_partial;
  }
  if (!d_views_) OU_HIP(hipMalloc(reinterpret_cast<void **>(&d_views_), 2 * n * sizeof(SweepView)));
  if (!d_members_) OU_HIP(hipMalloc(reinterpret_cast<void **>(&d_members_), n * sizeof(MemberView)));
  // The tables may be read by work still in flight on any stream.
  OU_HIP(hipDeviceSynchronize());
  OU_HIP(hipMemcpy(d_views_, views.data(), 2 * n * sizeof(SweepView), hipMemcpyHostToDevice));
  OU_HIP(hipMemcpy(d_members_, mv.data(), n * sizeof(MemberView), hipMemcpyHostToDevice));
  bl_.views = d_views_;
  bl_.views_src = d_views_ + n;
  bl_.members = d_members_;
}

void Batch::drop_graphs() {
  for (hipGraphExec_t &e : exec_)
    if (e) {
      (void)hipGraphExecDestroy(e);
      e = nullptr;
    }
}

void Batch::set_shape(int waves, int dir_groups) {
  OU_HIP(hipStreamSynchronize(stream_));
  if (bl_.sweep_mode == OU_SWEEP_TILES) {
    for (RadiativeTransfer *m : members_) set_tile_shape(m->plan(), waves, dir_groups);
    bl_.tile_waves = waves;
    bl_.tile_groups = dir_groups;
  } else {
    for (RadiativeTransfer *m : members_) set_sweep_shape(m->plan(), waves, dir_groups);
    bl_.waves = waves;
    bl_.dir_groups = dir_groups;
  }
  batch_prepare(bl_);
  drop_graphs();
  build_views();
}

void Batch::set_sweep_mode(int mode) {
  if (mode != OU_SWEEP_LEVELS && mode != OU_SWEEP_PERSISTENT && mode != OU_SWEEP_TILES)
    fail(OU_ERR_ARG,
         "sweep mode must be OU_SWEEP_LEVELS, OU_SWEEP_PERSISTENT or OU_SWEEP_TILES");
  if (mode == OU_SWEEP_TILES && !bl_.tiles_ok)
    fail(OU_ERR_UNSUPPORTED, "a member of this batch has no tile sweep");
  OU_HIP(hipStreamSynchronize(stream_));
  const bool was_tiles = bl_.sweep_mode == OU_SWEEP_TILES;
  bl_.sweep_mode = mode;
  drop_graphs();
  // the views carry the mode's direction groups
  if (was_tiles != (mode == OU_SWEEP_TILES)) set_shape(waves(), dir_groups());
}

void Batch::check_watchdog(hipStream_t s) {
  if (bl_.sweep_mode == OU_SWEEP_PERSISTENT) check_sweep_watchdog(d_error_, s ? s : stream_);
}

void Batch::set_streams(int n) {
  if (n != 1 && n != 2) fail(OU_ERR_ARG, "batch streams must be 1 or 2");
  OU_HIP(hipStreamSynchronize(stream_));
  bl_.side = n == 2 ? side_ : nullptr;
  drop_graphs();
}

// A member plan may have been reshaped (or its partial moments reallocated)
// through another path since the views were written: restore the batch's
// shape and rewrite the tables.
void Batch::ensure_current() {
  bool stale = d_views_ == nullptr;
  const bool tiles = bl_.sweep_mode == OU_SWEEP_TILES;
  for (size_t u = 0; u < members_.size(); ++u) {
    const Plan &pl = members_[u]->plan();
    const bool shape_differs = tiles ? (pl.tile_waves != bl_.tile_waves ||
                                        pl.tile_groups != bl_.tile_groups)
                                     : (pl.waves != bl_.waves || pl.dir_groups != bl_.dir_groups);
    if (shape_differs || (u < seen_partial_.size() && pl.d_partial != seen_partial_[u]))
      stale = true;
  }
  if (stale) set_shape(waves(), dir_groups());
}

void Batch::upload_io(const double *const *xs, double *const *ys, hipStream_t s) {
  const size_t n = members_.size();
  if (!ys) fail(OU_ERR_ARG, "null output pointer table");
  std::set<const double *> outs;
  for (size_t u = 0; u < n; ++u) {
    if (!ys[u]) fail(OU_ERR_ARG, "null output vector in the batch");
    if (xs && !xs[u]) fail(OU_ERR_ARG, "null input vector in the batch");
    if (xs && xs[u] == ys[u]) fail(OU_ERR_ARG, "x and y must be distinct vectors");
    if (!outs.insert(ys[u]).second) fail(OU_ERR_ARG, "two members share an output vector");
  }
  // The pinned table is rewritten only after the previous copy has run.
  if (io_pending_) OU_HIP(hipEventSynchronize(io_done_));
  for (size_t u = 0; u < n; ++u) {
    h_io_[u] = xs ? xs[u] : nullptr;
    h_io_[n + u] = ys[u];
  }
  OU_HIP(hipMemcpyAsync(d_io_, h_io_, 2 * n * sizeof(double *), hipMemcpyHostToDevice, s));
  OU_HIP(hipEventRecord(io_done_, s));
  io_pending_ = true;
}

void Batch::upload_active(hipStream_t s) {
  OU_HIP(hipMemcpyAsync(d_active_, h_active_, members_.size() * sizeof(int),
                        hipMemcpyHostToDevice, s));
}

void Batch::enqueue(int op, hipStream_t s) {
  const double *const *xs = d_io_;
  double *const *ys = const_cast<double *const *>(d_io_ + members_.size());
  if (op == 0)
    batch_apply(bl_, xs, ys, s);
  else if (op == 3)
    batch_si_step(bl_, xs, ys, d_si_partial_, kSiBlocks, d_si_norms_, s);
  else
    batch_sweep(bl_, ys, op == 2, s);
}

void Batch::run(int op, hipStream_t s) {
  if (!use_graph_) {
    enqueue(op, s);
    return;
  }
  if (!exec_[op]) {
    hipGraph_t g = nullptr;
    OU_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    try {
      enqueue(op, stream_);
    } catch (...) {
      (void)hipStreamEndCapture(stream_, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    OU_HIP(hipStreamEndCapture(stream_, &g));
    const hipError_t e = hipGraphInstantiate(&exec_[op], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    OU_HIP(e);
  }
  OU_HIP(hipGraphLaunch(exec_[op], s));
}

void Batch::apply(const double *const *d_xs, double *const *d_ys, hipStream_t s) {
  if (!d_xs) fail(OU_ERR_ARG, "null input pointer table");
  s = s ? s : stream_;
  ensure_current();
  upload_io(d_xs, d_ys, s);
  run(0, s);
}

void Batch::sweep(double *const *d_ys, bool with_source, hipStream_t s) {
  s = s ? s : stream_;
  ensure_current();
  upload_io(nullptr, d_ys, s);
  run(with_source ? 2 : 1, s);
}

int Batch::solve_si(double *const *d_phi, double tol, int max_it, int *iterations,
                    double *change, bool verbose) {
  const size_t n = members_.size();
  if (!d_phi || !iterations || !change) fail(OU_ERR_ARG, "null argument");
  if (max_it < 1 || !(tol >= 0.0)) fail(OU_ERR_ARG, "bad tolerance or iteration limit");
  ensure_current();
  if (!d_si_work_) {
    long total = 0;
    for (const RadiativeTransfer *m : members_) total += m->plan().vec_size;
    OU_HIP(hipMalloc(&d_si_work_, size_t(total) * 8));
    OU_HIP(hipMalloc(&d_si_partial_, n * kSiBlocks * 2 * sizeof(double)));
    OU_HIP(hipMalloc(&d_si_norms_, n * 2 * sizeof(double)));
    OU_HIP(hipHostMalloc(reinterpret_cast<void **>(&h_si_norms_), n * 2 * sizeof(double)));
  }
  std::vector<double *> cur(n), nxt(n);
  for (size_t u = 0, off = 0; u < n; off += members_[u]->plan().vec_size, ++u) {
    if (!d_phi[u]) fail(OU_ERR_ARG, "null initial guess in the batch");
    cur[u] = d_phi[u];
    nxt[u] = d_si_work_ + off;
    iterations[u] = -1;
    change[u] = 0.0;
  }
  // A converged member keeps its last iterate: its flag is cleared, so the
  // following steps skip its chunks and element-wise work, and its iterate
  // is what single-plan source iteration returns at the same tolerance.
  int it = 0;
  try {
    while (it < max_it) {
      upload_io(cur.data(), nxt.data(), stream_);
      run(3, stream_);
      OU_HIP(hipMemcpyAsync(h_si_norms_, d_si_norms_, n * 2 * sizeof(double),
                            hipMemcpyDeviceToHost, stream_));
      OU_HIP(hipStreamSynchronize(stream_));
      check_watchdog(stream_);
      ++it;
      size_t done = 0;
      double worst = 0.0;
      bool cleared = false;
      for (size_t u = 0; u < n; ++u) {
        if (iterations[u] >= 0) {
          ++done;
          continue;
        }
        const double d2 = h_si_norms_[2 * u], y2 = h_si_norms_[2 * u + 1];
        change[u] = y2 > 0.0 ? std::sqrt(d2 / y2) : 0.0;
        worst = std::max(worst, change[u]);
        std::swap(cur[u], nxt[u]);
        if (change[u] < tol) {
          iterations[u] = it;
          h_active_[u] = 0;
          cleared = true;
          ++done;
        }
      }
      if (verbose && (it % 25 == 0 || done == n || it == max_it))
        progress("batched SI iteration %d: largest open change %.3e, %zu of %zu members converged\n",
                 it, worst, done, n);
      if (done == n) break;
      if (cleared) upload_active(stream_);  // ordered before the next step on this stream
    }
  } catch (...) {
    (void)hipStreamSynchronize(stream_);
    std::fill(h_active_, h_active_ + n, 1);
    (void)hipMemcpy(d_active_, h_active_, n * sizeof(int), hipMemcpyHostToDevice);
    throw;
  }
  std::fill(h_active_, h_active_ + n, 1);
  upload_active(stream_);
  // the last iterate of a member may sit in the work buffer
  for (size_t u = 0; u < n; ++u)
    if (cur[u] != d_phi[u])
      OU_HIP(hipMemcpyAsync(d_phi[u], cur[u], size_t(members_[u]->plan().vec_size) * 8,
                            hipMemcpyDeviceToDevice, stream_));
  OU_HIP(hipStreamSynchronize(stream_));
  return it;
}

std::string Batch::tune_key() const {
  const Plan &p0 = members_[0]->plan();
  long cells = 0;
  for (const RadiativeTransfer *m : members_) cells += m->plan().n_cells;
  std::ostringstream k;
  k << "dim=" << p0.dim << ",p=" << p0.p << ",mm=" << p0.mm << ",prec=" << p0.precision
    << ",dirs/oct=" << max_dirs_per_octant(p0) << ",members=" << members_.size()
    << ",cells=" << cells << ",levels=" << n_levels();
  return k.str();
}

std::vector<TuneEntry> Batch::tune(int reps) {
  if (reps < 1) reps = 1;
  const Plan &p0 = members_[0]->plan();
  int min_g = 1;
  for (const RadiativeTransfer *m : members_) min_g = std::max(min_g, m->plan().min_dir_groups);
  const int max_g = std::max(min_g, std::min(4, std::max(1, max_dirs_per_octant(p0))));
  // Scratch vectors: one shared input, one output per member.
  long max_vec = 0, sum_vec = 0;
  for (const RadiativeTransfer *m : members_) {
    max_vec = std::max(max_vec, m->plan().vec_size);
    sum_vec += m->plan().vec_size;
  }
  double *x = nullptr, *y = nullptr;
  OU_HIP(hipMalloc(&x, size_t(max_vec) * 8));
  if (hipMalloc(&y, size_t(sum_vec) * 8) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(x);
    fail(OU_ERR_NOMEM, "autotuner scratch allocation failed");
  }
  std::vector<const double *> xs(members_.size(), x);
  std::vector<double *> ys(members_.size());
  for (size_t u = 0, off = 0; u < members_.size(); off += members_[u]->plan().vec_size, ++u)
    ys[u] = y + off;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::vector<TuneEntry> out;
  try {
    vec_fill(max_vec, 1.0, x, stream_);
    OU_HIP(hipEventCreate(&e0));
    OU_HIP(hipEventCreate(&e1));
    auto time_shape = [&](int mode, int w, int g) {
      set_shape(w, g);
      apply(xs.data(), ys.data(), stream_);  // warm-up, captures the graph
      OU_HIP(hipStreamSynchronize(stream_));
      OU_HIP(hipEventRecord(e0, stream_));
      for (int r = 0; r < reps; ++r) apply(xs.data(), ys.data(), stream_);
      OU_HIP(hipEventRecord(e1, stream_));
      OU_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      OU_HIP(hipEventElapsedTime(&ms, e0, e1));
      out.push_back({w, g, double(ms) / reps, mode});
    };
    // the chunk sweep's (waves, groups), then the tile sweep's
    set_sweep_mode(OU_SWEEP_LEVELS);
    for (int w : {2, 4, 6, 8, 12}) {
      if (!sweep_supported(p0.dim, p0.p, p0.mm, p0.precision, w, p0.max_lds)) continue;
      for (int g = min_g; g <= max_g; ++g) time_shape(OU_SWEEP_LEVELS, w, g);
    }
    if (bl_.tiles_ok) {
      set_sweep_mode(OU_SWEEP_TILES);
      for (int w : {1, 2, 3, 4}) {  // 3: the 3D Q2 tile sweep's 12-direction rounds
        if (!tile_sweep_supported(p0.dim, p0.p, p0.mm, p0.precision, w, p0.max_lds)) continue;
        for (int g = min_g; g <= max_g; ++g) time_shape(OU_SWEEP_TILES, w, g);
      }
    }
  } catch (...) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(stream_);
    (void)hipFree(x);
    (void)hipFree(y);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  OU_HIP(hipStreamSynchronize(stream_));
  (void)hipFree(x);
  (void)hipFree(y);
  if (out.empty()) fail(OU_ERR_UNSUPPORTED, "no sweep shape to tune for this batch");
  const TuneEntry best =
      *std::min_element(out.begin(), out.end(),
                        [](const TuneEntry &a, const TuneEntry &b) { return a.ms < b.ms; });
  set_sweep_mode(best.mode);
  set_shape(best.waves, best.dir_groups);
  tune_store(tune_key(), TuneChoice{best.mode, best.waves, best.dir_groups});
  return out;
}

}  // namespace ou
