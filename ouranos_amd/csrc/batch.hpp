// Many independent transport operators swept as one: the persistent
// multi-instance launch and its autotuner.
//
// The members' wavefront levels are merged (level l of the batch holds
// level l of every member), so a batch of any number of uneven meshes costs
// max(levels) sweep launches instead of sum(levels), and every launch has the
// chunks of all members to spread over the chip.  The element-wise stages
// (scattering source, trace set-up, epilogue) are one launch each for the
// whole batch.  The whole sequence is captured once per operation as a HIP
// graph and replayed; the per-member vector pointers live in a device table
// written before each replay, so one graph serves any vectors.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "radiative_transfer.hpp"

namespace ou {

struct TuneEntry {
  int waves = 0, dir_groups = 0;
  double ms = 0.0;  // mean time of one batched Apply
  int mode = OU_SWEEP_LEVELS;  // the sweep mode the shape belongs to
};

class Batch {
 public:
  // All members must share dimension, element order, moment count and
  // precision.  waves / dir_groups <= 0 ask the autotuner (or its cache).
  Batch(std::vector<RadiativeTransfer *> members, int waves, int dir_groups);
  ~Batch();
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  int size() const { return int(members_.size()); }
  // the shape of the sweep mode in use
  int waves() const { return bl_.sweep_mode == OU_SWEEP_TILES ? bl_.tile_waves : bl_.waves; }
  int dir_groups() const {
    return bl_.sweep_mode == OU_SWEEP_TILES ? bl_.tile_groups : bl_.dir_groups;
  }
  int n_levels() const { return int(bl_.level_off.size()) - 1; }
  long n_chunks() const { return bl_.level_off.back(); }
  hipStream_t stream() const { return stream_; }

  // y_i = x_i - D L^-1 M S x_i for every member i (d_xs, d_ys: host arrays of
  // device pointers, one per member).
  void apply(const double *const *d_xs, double *const *d_ys, hipStream_t s);
  // y_i += D L^-1 (M2D S_i [+ M q_i/ws + inflow]) for every member.
  void sweep(double *const *d_ys, bool with_source, hipStream_t s);

  // Source iteration of every member from the initial guesses d_phi (device
  // vectors, updated in place): phi_i <- D L^-1 (M2D S(phi_i) + M q_i/ws +
  // inflow), one captured graph per iteration for the whole batch, until every
  // member's change |phi_new - phi| / |phi_new| is below tol or max_it
  // iterations.  iterations[i]: the iteration at which member i first met tol
  // (-1 if it never did); change[i]: its last change.  Returns the number of
  // iterations run.  Each member's iterates equal its single-plan source
  // iteration's bit-for-bit.
  int solve_si(double *const *d_phi, double tol, int max_it, int *iterations, double *change,
               bool verbose);

  // The shape of the sweep mode in use (OU_SWEEP_TILES: the tile sweep's).
  void set_shape(int waves, int dir_groups);
  // OU_SWEEP_LEVELS, OU_SWEEP_PERSISTENT (one persistent launch) or
  // OU_SWEEP_TILES (merged tile levels; every member needs a tile sweep).
  void set_sweep_mode(int mode);
  int sweep_mode() const { return bl_.sweep_mode; }
  // OU_ERR_TIMEOUT if a persistent-sweep wait gave up since the last check
  // (synchronises s, or the batch's stream).
  void check_watchdog(hipStream_t s);
  void set_use_graph(bool on) { use_graph_ = on; }
  // 2: each level's second octant half on a side stream (the default); 1: one
  // launch per level.
  void set_streams(int n);
  int streams() const { return bl_.side ? 2 : 1; }
  bool use_graph() const { return use_graph_; }
  // Times every supported (waves, direction groups) shape of the chunk sweep
  // and, where every member has one, of the tile sweep on this batch, keeps
  // the fastest (mode and shape) and records it in the tuning cache.
  std::vector<TuneEntry> tune(int reps);
  // The cache key of this batch's shape class.
  std::string tune_key() const;

 private:
  void build_schedule();
  void build_views();
  void ensure_current();
  void upload_io(const double *const *xs, double *const *ys, hipStream_t s);
  void run(int op, hipStream_t s);
  void enqueue(int op, hipStream_t s);
  void drop_graphs();
  void release();
  void upload_active(hipStream_t s);

  std::vector<RadiativeTransfer *> members_;
  BatchLaunch bl_;
  hipStream_t stream_ = nullptr;
  hipStream_t side_ = nullptr;  // the second octant half's level launches
  hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
  Chunk *d_chunks_ = nullptr;
  int *d_dep_off_ = nullptr, *d_dep_list_ = nullptr, *d_error_ = nullptr;
  unsigned *d_done_ = nullptr;  // persistent launch: completion flags, then the ticket
  int *d_chunk_view_ = nullptr;
  TileItem *d_tiles_ = nullptr;  // merged tile levels
  int *d_tile_view_ = nullptr;
  SweepView *d_views_ = nullptr;     // [2][members]: plain, with source
  MemberView *d_members_ = nullptr;  // [members]
  const double **d_io_ = nullptr;    // [2][members]: x pointers, then y pointers
  const double **h_io_ = nullptr;    // pinned staging of the table
  hipEvent_t io_done_ = nullptr;
  bool io_pending_ = false;
  std::vector<double *> seen_partial_;
  bool use_graph_ = true;
  // apply, sweep, sweep + source, source iteration step
  hipGraphExec_t exec_[4] = {nullptr, nullptr, nullptr, nullptr};
  // source iteration: second iterate per member, norm partials and sums
  double *d_si_work_ = nullptr, *d_si_partial_ = nullptr, *d_si_norms_ = nullptr;
  double *h_si_norms_ = nullptr;
  // [members] flags read by every batched kernel; all ones outside solve_si,
  // which clears a member's flag once it has converged
  int *d_active_ = nullptr;
  int *h_active_ = nullptr;  // pinned staging of the flags
  static constexpr int kSiBlocks = 64;  // partial-sum blocks per member
};

}  // namespace ou
