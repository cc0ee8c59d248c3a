// The sweep plan: everything ou_plan_create builds once per problem, and
// the launch sequence of one operator application.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "ou_common.hpp"

namespace ou {

constexpr int kMaxDirGroups = 8;
// (p+1)^2 entries of the 1D reference mass matrix, up to order 5.
constexpr int kMaxP1Sq = 36;
// Directions of one direction group whose tables a sweep workgroup stages in
// LDS: an octant with more directions needs several groups.
constexpr int kMaxSlotsPerGroup = 64;

// One chunk: up to 64 cells of one sweep level of one octant.  Lanes 0-31
// (count_lo cells, material `material`) and lanes 32-63 (count_hi cells,
// material `material_hi`) each use one inverse matrix; the two halves differ
// only in plans whose cell solves run on the matrix cores (one 32-column MFMA
// block per half).  The chunk's cells are cell_offset .. + count_lo + count_hi
// of the cell list, lanes 0.. then lanes 32...
struct Chunk {
  int octant;
  int material;
  int material_hi;
  int count_lo;
  int count_hi;
  int cell_offset;  // into the cell list
};

// The tile sweep (OU_SWEEP_TILES): a tile is a block of cells of one octant in
// its canonical frame, 8 x 8 in 2D, 4 x 4 x 4 in 3D (edge E: canonical cells
// (E ti + a, E tj + b[, E tk + c])), all of one material.  Tiles of tile level
// ti + tj (+ tk) depend only on tiles of the previous tile level.
constexpr int kTileEdge2 = 8, kTileEdge3 = 4;
struct TileItem {
  int octant, ti, tj, tk, material;
};

// Device-side views handed to the kernels (plain pointers, by value).
struct SweepView {
  int n[3];
  int n_dir, n_mom;
  long n_dofs;
  // Inverses per (material, class): a class is a set of directions with equal
  // |Omega_a| on every axis, which share one matrix in their octants'
  // canonical frames.
  int n_cls;
  const int *slot_cls;     // [slot] class of direction oct_dirs[slot]
  const void *ainv;        // [mat][cls][N][N] in the sweep precision (canonical frame)
  const float *ainv_mfma;  // [mat][cls][64 lanes][16]: MFMA A-operand layout (FP32, N=27)
  const double *ainv_mfma64;  // [mat][cls][64 lanes][14]: FP64 MFMA A-operand layout (N=27)
  const void *S;           // [m][dof] scattering source (sweep precision)
  const void *S0;          // moment-0 override for the source sweep, or null
  const int *oct_dirs;     // directions grouped by octant ("slots")
  const int *oct_off;      // [n_oct+1]
  // Per slot g (direction oct_dirs[g]), padded to MM moments:
  const void *slot_m2d;    // [g][MM] M2D row (sweep precision)
  const double *slot_d2m;  // [g][MM] D2M column
  const void *slot_c;      // [g][4] |Omega_a| x (area of the face normal to a) (sweep precision)
  float m1f[kMaxP1Sq];     // 1D reference mass matrix on [0,1], (p+1) x (p+1)
  double m1d[kMaxP1Sq];
  // Per axis: [dir][F][slot].  Slot of a cell in the octant's canonical
  // coordinates (ii,jj,kk): x -> jj + ny kk, y -> ii + nx kk, z -> (ii+jj) ny + jj
  // (skewed so that the cells of one level with equal kk and consecutive jj
  // use consecutive slots on every axis).
  void *trace[3];
  long plane_size[3];
  double *partial;         // [oct][group][m][dof]
  const Chunk *chunks;
  const int *cells;
  int dir_groups;          // workgroups per (chunk, octant), each a range of directions
};

// The persistent sweep's launch state (one launch sweeps every level).
// Item k = (chunk k / groups, direction group k % groups), in the chunks'
// level order.  The item of chunk c and group g waits until every upwind
// chunk of c (dep_list[dep_off[c] .. dep_off[c+1]): the chunks of the
// previous level of its octant that hold a neighbour one step upwind of one
// of its cells) has set its flag done[chunk * groups + g], then sets its own
// once its traces are stored.
struct PersistView {
  const int *dep_off;   // [chunk + 1]
  const int *dep_list;
  unsigned *done;       // [chunks * groups], cleared before each launch
  unsigned *ticket;     // next item, cleared before each launch
  int *error;           // sticky: 1 once a wait gave up (cleared by the host)
  int n_items;
  int groups;
  unsigned long long wait_ticks;  // a wait gives up after this many wall-clock ticks
  unsigned max_polls;             // ... or after this many polls
};

// Host-side dependency lists of a persistent schedule (CSR over chunks).
struct PersistTables {
  std::vector<int> dep_off, dep_list;
};

struct Plan {
  // problem
  int dim = 2, p = 1, P1 = 2, N = 4, F = 2;
  int n[3] = {1, 1, 1};
  double h[3] = {1, 1, 1};
  int n_dir = 0, n_mom = 0, n_mat = 0, L = 0, precision = OU_FP64;
  long n_cells = 0, n_dofs = 0, vec_size = 0;
  double ws = 1.0;
  std::vector<double> omega, M2D, D2M, sigma_t, sigma_s, source;
  std::vector<int> mom_l, mat;
  std::vector<unsigned char> most_normal;
  int bc[6] = {0, 0, 0, 0, 0, 0};
  double inc[6] = {0, 0, 0, 0, 0, 0};
  int group = 0;

  // schedule
  int n_oct = 4, n_levels = 0, waves = 4, mm = 1;
  int dir_groups = 1;        // direction groups per octant (one workgroup each)
  int min_dir_groups = 1;    // fewest groups that keep a group within kMaxSlotsPerGroup
  int partial_groups = 1;    // groups d_partial is allocated for
  std::vector<int> oct_dirs, oct_off, dir_oct;
  int n_cls = 0;
  std::vector<int> dir_cls;          // [dir] inverse class
  std::vector<double> cls_omega;     // [cls][3] |Omega| of the class
  int *d_slot_cls = nullptr;
  std::vector<int> level_chunk_off;  // [n_levels+1]
  // Each level lists its chunks by octant quarter (quarter q holds octants
  // q*n_oct/4 ..); level_quarter[4*l + q] is the first chunk of quarter q
  // (level_chunk_off[l] for q = 0, level_chunk_off[l+1] for q = 4 is implied).
  std::vector<int> level_quarter;
  int sweep_streams = 2;  // launches per level: 1, 2 (octant halves) or 4 (quarters)
  int quarter_begin(int l, int q) const {
    return q == 0 ? level_chunk_off[l] : q == 4 ? level_chunk_off[l + 1] : level_quarter[4 * l + q];
  }
  long n_chunks = 0;

  // How a sweep is launched: OU_SWEEP_LEVELS (one launch per level),
  // OU_SWEEP_PERSISTENT (one persistent launch) or OU_SWEEP_TILES (2D: one
  // launch per tile level).
  int sweep_mode = OU_SWEEP_LEVELS;
  PersistTables deps;                  // upwind chunks of every chunk (host)
  int *d_dep_off = nullptr, *d_dep_list = nullptr;
  unsigned *d_done = nullptr;          // [n_chunks * kMaxDirGroups + 1]: flags, ticket
  int *d_error = nullptr;              // the persistent sweep's sticky error word
  double wait_limit_s = 10.0;          // a persistent-sweep wait gives up after this long
  double wall_clock_khz = 100000.0;    // rate of the device's wall clock
  int cus = 256;

  // Tile sweep (OU_SWEEP_TILES): tiles by tile level, and its own workgroup
  // shape (waves of 8 directions each in 2D, 4 in 3D; direction groups).
  bool tiles_ok = false;             // every tile holds one material and a kernel exists
  std::vector<int> tile_level_off;   // [tile levels + 1] into d_tiles
  TileItem *d_tiles = nullptr;
  int tile_waves = 1, tile_groups = 1;
  unsigned oct_mask = 0;             // octants that hold directions
  // direction groups of the partial-moment layout the active mode writes
  int groups_in_use() const { return sweep_mode == OU_SWEEP_TILES ? tile_groups : dir_groups; }

  // device buffers
  hipStream_t stream = nullptr;
  // the other octant parts' launches, forked from and joined back into the
  // sweep's stream by events
  hipStream_t side[3] = {nullptr, nullptr, nullptr};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[4] = {nullptr, nullptr, nullptr, nullptr};  // side[0..2], then stream
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  void *d_ainv = nullptr, *d_S = nullptr, *d_S0 = nullptr;
  float *d_ainv_mfma = nullptr;
  bool use_mfma = false;
  double *d_ainv_mfma64 = nullptr;
  bool use_mfma64 = false;
  int max_lds = 65536;
  void *d_slot_m2d = nullptr, *d_slot_c = nullptr;
  double *d_slot_d2m = nullptr, *d_partial = nullptr;
  double m1[kMaxP1Sq] = {};
  void *d_trace[3] = {nullptr, nullptr, nullptr};
  long plane_size[3] = {0, 0, 0};
  int *d_oct_dirs = nullptr, *d_oct_off = nullptr;
  Chunk *d_chunks = nullptr;
  int *d_cells = nullptr, *d_mat = nullptr, *d_mom_l = nullptr;
  double *d_sigma_s = nullptr, *d_source = nullptr, *d_mass = nullptr;
  double *d_sigma_other = nullptr;  // a caller's sigma_s table (ou_scatter_source_with)
  double *d_rowsum = nullptr;
  double *d_inflow = nullptr;  // [dir][3] inflow value on the upwind face of each axis
  double *d_x = nullptr, *d_y = nullptr;  // staging for the host-pointer API
  bool S_is_zero = true;

  ~Plan();
  size_t real_size() const { return precision == OU_FP32 ? 4 : 8; }
  SweepView view() const;
  PersistView persist_view() const;
};

// One member of a batch, as the batched element-wise kernels see it.
struct MemberView {
  long n_cells, n_dofs, vec_size;
  int n_mom, L, n_dir, n_parts;  // n_parts = octants x direction groups
  int groups;                     // direction groups per octant
  unsigned oct_mask;              // octants that hold directions (others' parts are skipped)
  const int *mat;
  const double *sigma_s;
  const int *mom_l;
  const double *mass;
  const double *source;
  const double *rowsum;
  double s0_factor;
  void *S, *S0;
  void *trace[3];
  long trace_per_dir[3];
  const double *inflow;
  const double *partial;
};

// A batch's merged level schedule on the device: level l of the batch is
// chunks [level_off[l], level_off[l+1]) of `chunks`, chunk i belonging to the
// member whose view is views[chunk_view[i]].
struct BatchLaunch {
  int n_members = 0, dim = 3, p = 1, N = 8, mm = 1, precision = OU_FP32;
  int waves = 4, dir_groups = 1;
  std::vector<int> level_off;  // host
  // host, [n_levels]: first chunk of level l in the second octant half of
  // every member (each merged level lists all members' first halves first)
  std::vector<int> level_split;
  // With side set, a level's second half is launched on it (forked from and
  // joined back into the launch stream by the two events).
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  const Chunk *chunks = nullptr;
  const int *chunk_view = nullptr;
  const int *active = nullptr;           // [member]: 0 = skip the member's work
  const SweepView *views = nullptr;      // S0 = null (Apply, plain sweep)
  const SweepView *views_src = nullptr;  // S0 = member source (sweep with source)
  const MemberView *members = nullptr;
  long max_elems = 0;                    // largest element count of any stage / member
  // one persistent launch for every member's levels instead of level launches
  int sweep_mode = OU_SWEEP_LEVELS;
  PersistView persist = {};              // groups and n_items set by the launch
  int cus = 256;
  // OU_SWEEP_TILES: tile level l of the batch is tiles [tile_level_off[l],
  // tile_level_off[l+1]) of `tiles`, tile i of the member at views[tile_view[i]]
  bool tiles_ok = false;                 // every member has a tile sweep
  std::vector<int> tile_level_off;       // host
  const TileItem *tiles = nullptr;
  const int *tile_view = nullptr;
  int tile_waves = 1, tile_groups = 1;
};

Plan *create_plan(const ou_problem_desc *desc);
// Upwind chunks of every chunk of one plan (chunks laid out by level, their
// cells in `cells`): for each cell, the chunk of the same octant that holds
// its neighbour one step upwind along each axis.
PersistTables chunk_dependencies(const std::vector<Chunk> &chunks, const std::vector<int> &cells,
                                 const int n[3], int dim);
// Sticky error word of the persistent sweep: if a wait gave up since the last
// check, clears it and fails with OU_ERR_TIMEOUT (synchronises `s` first).
void check_sweep_watchdog(int *d_error, hipStream_t s);
// Poll bound of a persistent-sweep wait of `seconds`.
unsigned wait_polls(double seconds);
bool sweep_supported(int dim, int p, int mm, int precision, int waves, int max_lds);
// Sets the sweep workgroup shape (waves per workgroup, direction groups per
// octant).  Growing the group count reallocates the partial moments after a
// device synchronisation.
void set_sweep_shape(Plan &pl, int waves, int dir_groups);
// The tile sweep's shape: waves per workgroup (8 directions each) and
// direction groups per quadrant.
void set_tile_shape(Plan &pl, int waves, int groups);
bool tile_sweep_supported(int dim, int p, int mm, int precision, int waves, int max_lds);
// Largest tile-sweep wave count (1, 2 or 4) with a kernel that fits max_lds.
int tile_waves_max(int dim, int p, int mm, int precision, int max_lds);
// Switches the launch mode (OU_SWEEP_LEVELS, OU_SWEEP_PERSISTENT, OU_SWEEP_TILES).
void set_sweep_mode(Plan &pl, int mode);

// Launch sequences (stream-ordered, no host synchronisation).
void launch_build_inverses(Plan &pl);
void launch_scatter_source(Plan &pl, const double *d_x, hipStream_t s);
// S_m (+)= sigma[mat][l(m)] M x_m with a caller's table (n_mat x (L+1), on the
// device): a scattering source from another group's cross sections.
void launch_scatter_source_with(Plan &pl, const double *d_x, const double *d_sigma,
                                bool accumulate, hipStream_t s);
void launch_source_moment0(Plan &pl, hipStream_t s);
void launch_trace_init(Plan &pl, bool with_source, hipStream_t s);
void launch_sweep_levels(Plan &pl, const void *S0, hipStream_t s,
                         hipEvent_t before = nullptr, hipEvent_t after = nullptr);
void launch_epilogue(Plan &pl, const double *d_x, double *d_y, bool apply,
                     hipStream_t s);
// One launch per tile level (sweep_tile.hip).
void launch_sweep_tiles(Plan &pl, const SweepView &v, hipStream_t s, hipEvent_t before,
                        hipEvent_t after);
// The batch's merged tile levels (launch false only sets kernel attributes).
void launch_batch_tiles(const BatchLaunch &b, const SweepView *views, hipStream_t s, bool launch);

// Whole operations.
void apply_device(Plan &pl, const double *d_x, double *d_y, hipStream_t s);
void sweep_device(Plan &pl, double *d_y, bool with_source, hipStream_t s);

// Batched operations over every member (stream-ordered; capturable: kernel
// launches only).  d_xs / d_ys are device arrays of per-member vector
// pointers.  batch_prepare sets kernel attributes and must run outside
// stream capture before the first batched launch of a shape.
void batch_prepare(const BatchLaunch &b);
void batch_apply(const BatchLaunch &b, const double *const *d_xs, double *const *d_ys,
                 hipStream_t s);
void batch_sweep(const BatchLaunch &b, double *const *d_ys, bool with_source, hipStream_t s);
// One source iteration of every member: S_i = S(x_i), y_i = D L^-1 (M2D S_i +
// M q_i/ws + inflow), and d_norms[2i], d_norms[2i+1] = |y_i - x_i|^2, |y_i|^2
// (d_partial: partial_blocks x 2 doubles per member).
void batch_si_step(const BatchLaunch &b, const double *const *d_xs, double *const *d_ys,
                   double *d_partial, int partial_blocks, double *d_norms, hipStream_t s);

}  // namespace ou
