// ob_internal.hpp -- shared types of liborbit_mi355x (host + device).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/orbit_mi355x.h"

namespace ob {

// ---------------------------------------------------------------------------
// Errors: every C entry point returns a status; the message is kept per thread.
void set_error(const std::string& msg);
struct Error {
  std::string msg;
};
#define OB_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::ob::Error{std::string(#call) + ": " + hipGetErrorString(e_)};          \
  } while (0)
#define OB_REQUIRE(cond, msg)              \
  do {                                     \
    if (!(cond)) throw ::ob::Error{(msg)}; \
  } while (0)

hipStream_t stream();
// A second stream (default priority) for work that overlaps the main stream; joined back
// with events before its results are used.
hipStream_t aux_stream();
int device_cu_count();
void ensure_init();
// Look up, once per device and with that device current, the device address of the
// fixed scratch array of the no-allocation entry points (ob_diag.hip).
void diag_scratch_setup(int device);

// Kernel timing with HIP events on the library stream (ob_profile_enable/_report).  A
// ProfScope brackets one launch when profiling is on and costs nothing otherwise.
bool profiling();
void prof_mark(const char* name, bool begin, hipStream_t s);
struct ProfScope {
  const char* name;
  hipStream_t s;
  explicit ProfScope(const char* n, hipStream_t st = nullptr) : name(n), s(st ? st : stream()) {
    if (profiling()) prof_mark(name, true, s);
  }
  ~ProfScope() {
    if (profiling()) prof_mark(name, false, s);
  }
};

// Status words: a pinned host int that a kernel sets to 1 (through its device address)
// when it finds a condition the reference treats as fatal.  The host reads and clears the
// words at sync points (check_statuses(true) waits for the stream first); nothing is
// copied or allocated per step.
struct StatusBox {
  int* host;
  int* dev;
};
StatusBox status_box_create(const char* what);
void status_box_destroy(StatusBox b);
void check_statuses(bool wait);

// ---------------------------------------------------------------------------
// Fused map program.  Each op carries host-precomputed constants so the device does
// exactly the reference's per-particle arithmetic (teapotbase.cc).
enum OpKind : int32_t {
  OP_DRIFT = 1,
  OP_QUAD1,
  OP_QUAD2,
  OP_QUADFRINGE_IN,
  OP_QUADFRINGE_OUT,
  OP_MULTP,
  OP_MULTPFRINGE_IN,
  OP_MULTPFRINGE_OUT,
  OP_WRAPBUNCH,
  OP_ROTATEXY,
  OP_KICK,
  OP_SC2P5DRB_KICK,
  OP_SC3D_KICK,
  OP_BEND1,
  OP_BEND3,
  OP_BEND4,
  OP_BENDFRINGE_IN,
  OP_BENDFRINGE_OUT,
  OP_WEDGEROTATE,
  OP_WEDGEDRIFT,
  OP_WEDGEBEND,  // wedgebend and wedgebendCF
  OP_SOLN,
  OP_RINGRF,
};

struct Op {
  int32_t kind;
  int32_t i0, i1, i2;  // integer parameters (pole, skew, ...)
  double p[12];        // double parameters / device pointers
};
static_assert(sizeof(Op) == 112, "Op layout");

constexpr int kMaxOpsPerLaunch = 24;
struct OpProgram {
  int32_t nops;
  int32_t want_extrema;  // write per-block min/max of x, y, z after the last op
  int32_t pad[2];
  Op ops[kMaxOpsPerLaunch];
};

// Space-charge state shared between the 2.5D kernels and the kick op.
struct SC2DState {
  double ext[6];                         // xMin xMax yMin yMax zMin zMax of the bunch
  double xMin, xMax, yMin, yMax, dx, dy; // rho/phi grid (Grid2D::setGridX/Y)
  double zMin, zMax, dz;                 // z line-density grid (Grid1D::setGridZ)
  double total, x_c, y_c, a_bunch;       // bunchAnalysis moments (:229-254)
  double scale_coeff;                    // log(dx_solver/dx_rho)
  double factor;                         // transverse kick factor (:129-133)
  double lsf_in, a2;                     // longitudinal RB coefficients (:138-141)
  int32_t status;                        // 1: zero-size bunch (reference aborts, :176-189)
  int32_t skip;                          // 1: do not kick
  // SpaceChargeCalc2p5D (SpaceChargeCalc2p5D.cc:152-172, 206-216): the calculator's x/y
  // ratio and its solver's box change from step to step.  Each step reads entry [par] and
  // its grid pass writes [par ^ 1], so every workgroup of that pass reads the same values.
  double ratio[2];
  double sol[2][4];                      // solver box xMin xMax yMin yMax
  double sdx, sdy;                       // solver grid steps (PoissonSolverFFT2D dx_, dy_)
  int32_t long_kick;                     // 1: 2p5Drb longitudinal kick, 0: transverse only
  int32_t rebuilds;                      // Green function rebuilds so far
  // SpaceChargeCalc2p5D with a boundary: only particles inside it are kicked
  // (SpaceChargeCalc2p5D.cc:109).  Type 0 none, 1 point-set box, 2 circle, 3 ellipse,
  // 4 rectangle; bnd = box xMin xMax yMin yMax, then the shape's two sizes.
  int32_t bnd_type, pad2;
  double bnd[6];
};

// Boundary2D inside test (BaseBoundary2D.cc:139-147, ShapedBoundary2D.cc:284-304).
template <typename P>
__host__ __device__ inline bool boundary_inside(int type, P b, double x, double y) {
  switch (type) {
    case 1: return x > b[0] && x < b[1] && y > b[2] && y < b[3];
    case 2: return x * x + y * y < b[4] * b[4];
    case 3: {
      const double xx = x / b[4], yy = y / b[5];
      return (xx * xx + yy * yy) < 1.0;
    }
    case 4: return fabs(x / b[4]) < 0.5 && fabs(y / b[5]) < 0.5;
    default: return true;
  }
}

struct SC3DState {
  double ext[6];
  double xMin, xMax, yMin, yMax, zMin, zMax, dx, dy, dz;  // rest-frame z limits
  double zc;                                              // (zMax+zMin)/2 rest frame
  int32_t status, skip;
};

// Device pointers travel inside Op::p as raw 64-bit patterns.
inline double ptr_bits(const void* p) {
  double d;
  std::memcpy(&d, &p, sizeof(d));
  return d;
}

struct DeviceBunchView {
  double* c[6];
  const double* msize;  // nullptr if no per-particle macrosize
  int64_t n;
};

// Launch the fused map program over a bunch; partial extrema (6 per block) go to
// ext_partials when prog.want_extrema.  Returns the number of blocks used.
int launch_push(const DeviceBunchView& v, const OpProgram& prog, double* ext_partials,
                hipStream_t s);

// Tile sort of a 2.5D deposit: T x T tiles of TSC centre nodes, work items of at most
// `item` particles of one tile, scratch copies of the in-grid particles sorted by tile.
struct Tile2D {
  int32_t T, shift, tx, ty, ntiles, item, max_items, zrows;  // T = 1 << shift
  uint32_t* counts;   // [ntiles], zero between steps
  uint32_t* cursor;   // [ntiles] scatter positions taken, cleared before each scatter
  int4* items;        // [max_items] (tile, begin, end, 0)
  uint32_t* nitems;   // [1]
  double2* sxy;       // [cap] (x, y) of the in-grid particles, sorted by tile
  double* sm;         // [cap] their macrosizes (per-particle macrosize only)
  double* zpart;      // [zrows][nz] per-block z line densities (single-bunch path)
  double* mpart;      // [max_items][5] rho moments per deposit item
  uint32_t* icounts;  // [hist_rows][ntiles] tile counts per batch work item (batch path)
  int32_t hist_rows, pad;
};

}  // namespace ob
struct ob_denoiser;
namespace ob {
// ---- batches of bunches, each with its own SpaceChargeCalc2p5Drb ----------------------
// Device-resident record of one bunch of a batch.
struct BatchBunch {
  DeviceBunchView v;
  int32_t item_begin, item_end;  // its work items (particle ranges); partials are per item
  int32_t rows_begin, cols_begin;  // its first block in the batched FFT row / column passes
  SC2DState* st;
  double *rho, *phi, *zg, *zd;
  double2 *spec, *ghat_t, *tw;
  int32_t nx, ny, nz, M0, M1, H, nmax, n_long_avg;
  double xy_ratio, dx_solver, macrosize, r0, mass;
  double gamma_D;  // (beta*beta)*pow(gamma,3) of the synchronous particle
  Tile2D tg;                      // deposit tile sort (buffers owned by the calculator)
  int32_t slot_begin, slot_end;   // its blocks of the batched tile-deposit pass
  int32_t moments_from_rho, pad;  // 1: rho changes after the deposit (denoiser)
};
struct BatchItem {  // particles [begin, end) of one bunch
  int32_t bunch, pad;
  int64_t begin, end;
};
struct BatchKick {  // deferred 2.5D kicks of the whole batch (common factors, op p[4], p[5])
  int32_t active, pad;
  double p4, p5;
};
// Push of all bunches: the deferred kick (if active) then prog; per-item extrema partials.
void launch_push_batch(const BatchBunch* bb, const BatchItem* items, int nitems,
                       const OpProgram& prog, const BatchKick& kick, double* partials,
                       int threads, hipStream_t s);
// SpaceChargeCalc2p5Drb steps of all bunches after launch_push_batch with extrema.
struct BatchShape {
  int nb, nitems, row_blocks, col_blocks, max_M, dep_slots;
  size_t hist_lds, dep_lds;  // dynamic LDS of the hist/scatter and tile-deposit passes
};
// Denoiser applied to one bunch's rho between the deposit and the moments.
struct DenoiseJob {
  ob_denoiser* d;
  double* rho;
  int nx, ny;
  const int32_t* skip;
};
void launch_sc2d_batch(const BatchBunch* bb, const BatchItem* items, const BatchShape& sh,
                       const double* partials, double length, double pipe_radius, hipStream_t s,
                       const std::vector<DenoiseJob>& jobs = {});
// Consecutive jobs with the same denoiser run as one batched inference (up to its
// max_batch, at most 64 images per launch).
void dn_denoise_rhos(const std::vector<DenoiseJob>& jobs);
}  // namespace ob
struct ob_boundary2d;
namespace ob {
// Boundary2D::addBoundaryPotential (BaseBoundary2D.cc:349-458) on a device phi grid whose
// geometry is in st (skipped when st->skip); the boundary must be initialized.
void boundary_add_potential(ob_boundary2d* b, double* phi, int nx, int ny, const SC2DState* st,
                            hipStream_t s);
// The same for each z slice of a z-fastest [nx][ny][nz] stack (its own coefficients).
void boundary_add_potential_slices(ob_boundary2d* b, double* phi, int nx, int ny, int nz,
                                   const SC2DState* st, hipStream_t s);
// Type (1 point set, 2 circle, 3 ellipse, 4 rectangle) and the six numbers of
// SC2DState::bnd for an initialized boundary.
int boundary_describe(ob_boundary2d* b, double* bnd6);
}  // namespace ob
struct ob_sc2p5drb;
namespace ob {
// Calculator fields of a batch record; mark a calculator as tracked; its status word
// (1: the last step found a zero-extent bunch).
void sc2d_batch_record(ob_sc2p5drb* c, BatchBunch& r);
// Size the calculator's tile-sort buffers for a bunch of n particles (per-particle
// macrosize or not) with work items of `item` particles; returns the tile sort.
// hist_rows: batch work items of the bunch, each of which keeps its row of tile counts.
Tile2D sc2d_tile_reserve(ob_sc2p5drb* c, int64_t n, bool per_particle_m, int item,
                         int zrows = 1, int hist_rows = 0);
// Whether `t` still describes the calculator's tile buffers and fits such a bunch.
bool sc2d_tile_same(ob_sc2p5drb* c, const Tile2D& t, int64_t n, bool per_particle_m);
void sc2d_mark_tracked(ob_sc2p5drb* c);
int sc2d_read_status(ob_sc2p5drb* c);
int push_grid_blocks(int64_t n);
// The calculator's denoiser job (false: none set).
bool sc2d_denoise_job(ob_sc2p5drb* c, DenoiseJob* job);
// Replace a [nx][ny] FP64 rho grid by the denoiser's estimate (same total charge); nothing
// happens where *skip is set.  dn_check_shape: the denoiser's image size must be nx x ny.
void dn_denoise_rho(ob_denoiser* d, double* rho, int nx, int ny, const int32_t* skip);
void dn_check_shape(ob_denoiser* d, int nx, int ny);

}  // namespace ob

struct ob_bunch;
namespace ob {
// Bunch internals used by the space-charge calculators.
DeviceBunchView bunch_device_view(ob_bunch* b);
// Queue an op; flushes first when the queue is full.
void bunch_queue(ob_bunch* b, const Op& op);
// Apply queued ops; if ext_partials != nullptr compute partial extrema too.
int bunch_flush(ob_bunch* b, double* ext_partials);
int64_t bunch_size(ob_bunch* b);
// Whether the bunch belongs to a batch (its flush then applies the batch's work).
bool bunch_in_batch(ob_bunch* b);
// Tile sorting of the device arrays: begin returns the current and spare buffers (the
// permutation is initialised to the identity when the bunch is in user order); the caller
// writes the reordered particles into the spare buffers and commits, which swaps them in.
struct SortBuffers {
  double *cur, *alt;   // 6 * cap each
  int64_t cap;
  int32_t *perm, *perm_alt;
  double *msize, *msize_alt;  // null when the bunch has no macrosize attribute
  int64_t n;
};
SortBuffers bunch_sort_begin(ob_bunch* b);
void bunch_sort_commit(ob_bunch* b);
struct SyncInfo {
  double mass, charge, r0, macrosize, T, p, beta, gamma, time;
};
SyncInfo bunch_sync(ob_bunch* b);
// Take the queued (not yet applied) map ops of a bunch.
std::vector<Op> bunch_take_queue(ob_bunch* b);
// Host-side synchronous time (advanced by the drift maps when queued).
void bunch_set_sync_time(ob_bunch* b, double t);
// Synchronous particle kinematics (T, p, beta, gamma, time) of src onto dst.
void bunch_copy_kinematics(ob_bunch* dst, const ob_bunch* src);
}  // namespace ob
struct ob_batch;
namespace ob {
// A batched bunch's flush applies the whole batch's pending work (ob_batch.hip).
void bunch_set_owner(ob_bunch* b, ob_batch* owner);
void batch_flush_owner(ob_batch* B);
// Live bunches; apply the queued SC kicks whose op p[0] is `state` (before it is freed).
extern std::vector<ob_bunch*> g_live_bunches;
void flush_kicks_reading(const void* state);
// A 2.5D calculator is going away: batches using it apply their pending kicks and drop it.
void batch_forget_calc(ob_sc2p5drb* c);
// Queue one map of the public op list (ob_map_op) exactly as its entry point does.
int queue_map_op(ob_bunch* b, const ob_map_op& o);
// A kernel outside the map queue has written the device coordinates.
void bunch_device_written(ob_bunch* b);
}  // namespace ob
