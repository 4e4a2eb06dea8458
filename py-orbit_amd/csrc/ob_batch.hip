// ob_batch.hip -- many independent bunches tracked together (no reference counterpart).
//
// A batch holds B bunches of arbitrary, uneven sizes, each with its own
// SpaceChargeCalc2p5Drb (grids may differ), that share one lattice and one synchronous
// particle.  Instead of B launches per phase, every phase covers all bunches in one launch:
//   push     one workgroup per work item (a chunk of one bunch's particles): deferred SC
//            kick, the map program, per-item extrema;
//   setup    one workgroup per bunch (extrema of its items, grid sizing, clearing);
//   deposit  tile sort per bunch: hist and scatter one workgroup per work item, then one
//            workgroup per tile work item accumulating rho in LDS;
//   post     one workgroup per bunch; FFT row / column / row passes over all bunches'
//            rows and columns (the FFT length is dispatched per workgroup).
// The per-bunch records and the work items live in device memory and change only when a
// bunch changes size, so a stretch of tracking (a lattice cell) can be captured once into
// a hipGraph and replayed.  The work-item size and the tile-deposit item size are chosen
// by timing the candidates on the batch's own data (autotune).
//
// Results are those of tracking each bunch alone: the same device bodies, the same host
// constants (computed once, for the shared synchronous particle).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "ob_internal.hpp"

struct ob_batch {
  std::vector<ob_bunch*> bunches;
  std::vector<ob_sc2p5drb*> calcs;
  ob_bunch* leader = nullptr;          // queues the maps (host constants) for everyone
  std::vector<ob::Op> queue;           // maps waiting to be applied to all bunches
  ob::BatchKick kick{};                // deferred SC kicks of all bunches
  int chunk = 16384;                   // particles per work item
  int item = 2048;                     // particles per tile-deposit work item
  bool tuned = false, fixed = false;   // fixed: tuning set by the caller
  // device records
  std::vector<ob::BatchBunch> h_bb;
  std::vector<ob::BatchItem> h_items;
  ob::BatchBunch* d_bb = nullptr;
  ob::BatchItem* d_items = nullptr;
  double* d_partials = nullptr;
  int cap_bb = 0, cap_items = 0;
  ob::BatchShape shape{};
  std::vector<int64_t> seen_n;         // sizes / pointers the records were built for
  std::vector<const double*> seen_ptr;
  int built_chunk = 0, built_item = -1;
  int version = 0;                     // incremented by every rebuild of the records
  // graph
  int graph_version = -1;              // records version the graph was captured with
  bool capturing = false;
  hipGraphExec_t exec = nullptr;
  double graph_dt = 0.0;               // synchronous time advanced by one replay
};

namespace ob {
namespace {

constexpr int kBatchThreads = 256;

void check_same_species(const std::vector<ob_bunch*>& bs) {
  const SyncInfo a = bunch_sync(bs[0]);
  for (size_t k = 1; k < bs.size(); k++) {
    const SyncInfo s = bunch_sync(bs[k]);
    OB_REQUIRE(s.mass == a.mass && s.charge == a.charge && s.r0 == a.r0 && s.p == a.p,
               "batch: all bunches must share mass, charge, classical radius and the "
               "synchronous momentum");
  }
}

// (Re)build the device records when the bunches or the tuning changed.
void build_records(ob_batch* B, bool force) {
  const int nb = (int)B->bunches.size();
  bool same = !force && B->built_chunk == B->chunk && B->built_item == B->item &&
              (int)B->seen_n.size() == nb;
  std::vector<DeviceBunchView> views(nb);
  for (int b = 0; b < nb; b++) {
    views[b] = bunch_device_view(B->bunches[b]);
    if (same && (B->seen_n[b] != views[b].n || B->seen_ptr[b] != views[b].c[0])) same = false;
    // the calculator's tile buffers move if it tracked a larger bunch on its own
    if (same && !B->calcs.empty() &&
        !sc2d_tile_same(B->calcs[b], B->h_bb[b].tg, views[b].n, views[b].msize != nullptr))
      same = false;
    if (same) {  // RingRF changed the synchronous momentum: new kick factors
      const SyncInfo sy = bunch_sync(B->bunches[b]);
      if (B->h_bb[b].gamma_D != (sy.beta * sy.beta) * pow(sy.gamma, 3)) same = false;
    }
    DenoiseJob dj;  // a denoiser set or cleared since: the moments come from elsewhere
    if (same && !B->calcs.empty() &&
        (sc2d_denoise_job(B->calcs[b], &dj) ? 1 : 0) != B->h_bb[b].moments_from_rho)
      same = false;
  }
  if (same) return;
  OB_REQUIRE(!B->capturing, "batch: bunch sizes changed during a graph capture");
  B->h_bb.assign(nb, BatchBunch{});
  B->h_items.clear();
  int rows = 0, cols = 0, max_M = 16, slots = 0;
  size_t hist_lds = 0, dep_lds = 0;
  for (int b = 0; b < nb; b++) {
    BatchBunch& r = B->h_bb[b];
    std::memset(&r, 0, sizeof(r));
    r.v = views[b];
    const SyncInfo sy = bunch_sync(B->bunches[b]);
    r.macrosize = sy.macrosize;
    r.r0 = sy.r0;
    r.mass = sy.mass;
    r.gamma_D = (sy.beta * sy.beta) * pow(sy.gamma, 3);
    r.item_begin = (int)B->h_items.size();
    for (int64_t i = 0; i < r.v.n; i += B->chunk)
      B->h_items.push_back(BatchItem{b, 0, i, std::min<int64_t>(r.v.n, i + B->chunk)});
    r.item_end = (int)B->h_items.size();
    if (!B->calcs.empty()) {
      sc2d_batch_record(B->calcs[b], r);
      r.rows_begin = rows;
      r.cols_begin = cols;
      rows += (r.nx + 1) / 2;
      cols += r.H;
      max_M = std::max(max_M, std::max(r.M0, r.M1));
      r.tg = sc2d_tile_reserve(B->calcs[b], r.v.n, r.v.msize != nullptr, B->item, 1,
                               r.item_end - r.item_begin);
      r.slot_begin = slots;
      slots += (int)(r.v.n / r.tg.item + r.tg.ntiles + 1);
      r.slot_end = slots;
      hist_lds = std::max(hist_lds, std::max(sizeof(double) * r.nz + sizeof(uint32_t) * r.tg.ntiles,
                                             2 * sizeof(uint32_t) * r.tg.ntiles));
      DenoiseJob dj;
      r.moments_from_rho = sc2d_denoise_job(B->calcs[b], &dj) ? 1 : 0;
      dep_lds = std::max(dep_lds, sizeof(double) * (r.tg.T + 2) * (r.tg.T + 2));
    }
  }
  if (nb > B->cap_bb) {
    if (B->d_bb) OB_HIP(hipFree(B->d_bb));
    OB_HIP(hipMalloc((void**)&B->d_bb, sizeof(BatchBunch) * nb));
    B->cap_bb = nb;
  }
  const int ni = (int)B->h_items.size();
  if (ni > B->cap_items) {
    if (B->d_items) OB_HIP(hipFree(B->d_items));
    if (B->d_partials) OB_HIP(hipFree(B->d_partials));
    OB_HIP(hipMalloc((void**)&B->d_items, sizeof(BatchItem) * std::max(ni, 1)));
    OB_HIP(hipMalloc((void**)&B->d_partials, sizeof(double) * 6 * std::max(ni, 1)));
    B->cap_items = ni;
  }
  hipStream_t s = stream();
  OB_HIP(hipMemcpyAsync(B->d_bb, B->h_bb.data(), sizeof(BatchBunch) * nb, hipMemcpyHostToDevice, s));
  if (ni)
    OB_HIP(hipMemcpyAsync(B->d_items, B->h_items.data(), sizeof(BatchItem) * ni,
                          hipMemcpyHostToDevice, s));
  OB_HIP(hipStreamSynchronize(s));  // host vectors may change before the copies run
  B->shape = BatchShape{nb, ni, rows, cols, max_M, slots, hist_lds, dep_lds};
  B->seen_n.resize(nb);
  B->seen_ptr.resize(nb);
  for (int b = 0; b < nb; b++) {
    B->seen_n[b] = views[b].n;
    B->seen_ptr[b] = views[b].c[0];
  }
  B->built_chunk = B->chunk;
  B->built_item = B->item;
  B->version++;
}

// Move the leader's queued maps into the batch queue; keep every bunch's clock equal.
void collect(ob_batch* B) {
  std::vector<Op> q = bunch_take_queue(B->leader);
  B->queue.insert(B->queue.end(), q.begin(), q.end());
  // the maps advance the synchronous time (drift, bend1, soln) and RingRF may change the
  // synchronous momentum: every bunch of the batch follows the leader
  for (ob_bunch* b : B->bunches)
    if (b != B->leader) bunch_copy_kinematics(b, B->leader);
}

// Apply the pending kick and maps to all bunches (extrema per work item if asked).
void flush(ob_batch* B, bool extrema) {
  collect(B);
  build_records(B, false);
  hipStream_t s = stream();
  size_t pos = 0;
  bool first = true;
  do {
    OpProgram prog;
    std::memset(&prog, 0, sizeof(prog));
    const size_t take = std::min(B->queue.size() - pos, (size_t)kMaxOpsPerLaunch);
    prog.nops = (int32_t)take;
    for (size_t k = 0; k < take; k++) prog.ops[k] = B->queue[pos + k];
    pos += take;
    const bool last = pos >= B->queue.size();
    prog.want_extrema = (last && extrema) ? 1 : 0;
    BatchKick k = first ? B->kick : BatchKick{};
    launch_push_batch(B->d_bb, B->d_items, B->shape.nitems, prog, k, B->d_partials,
                      kBatchThreads, s);
    first = false;
  } while (pos < B->queue.size());
  B->queue.clear();
  B->kick = BatchKick{};
  for (ob_bunch* b : B->bunches) bunch_device_written(b);
}

float time_ms(hipEvent_t a, hipEvent_t b) {
  OB_HIP(hipEventSynchronize(b));
  float ms = 0.f;
  OB_HIP(hipEventElapsedTime(&ms, a, b));
  return ms;
}

// Tuning results of this process by device and batch shape (bunch count, then every
// bunch's particle count, grid and macro-size attribute): a batch of the same shape takes
// them instead of timing the candidates again.
std::map<std::vector<int64_t>, std::pair<int, int>>& tuning_cache() {
  static std::map<std::vector<int64_t>, std::pair<int, int>> m;
  return m;
}
std::vector<int64_t> tuning_key(const ob_batch* B) {
  int dev = 0;
  OB_HIP(hipGetDevice(&dev));
  std::vector<int64_t> k{dev, (int64_t)B->h_bb.size(), B->calcs.empty() ? 0 : 1};
  for (const BatchBunch& r : B->h_bb) {
    k.push_back(r.v.n);
    k.push_back(r.nx);
    k.push_back(r.ny);
    k.push_back(r.nz);
    k.push_back(r.v.msize != nullptr ? 1 : 0);
  }
  return k;
}

// Time the candidates on the batch's data: work-item size with an identity push (reads
// and writes every particle unchanged), then the tile-deposit item size with repeated
// setup + deposit (setup clears the grids, so the repeats are idempotent).  The small
// candidates serve batches of few, small bunches (a single 100k bunch: 512 particles per
// item gives it 196 push workgroups instead of 49).
void autotune(ob_batch* B) {
  build_records(B, false);
  const std::vector<int64_t> key = tuning_key(B);
  const auto hit = tuning_cache().find(key);
  if (hit != tuning_cache().end()) {
    B->chunk = hit->second.first;
    B->item = hit->second.second;
    build_records(B, false);
    B->tuned = true;
    return;
  }
  hipStream_t s = stream();
  hipEvent_t e0, e1;
  OB_HIP(hipEventCreate(&e0));
  OB_HIP(hipEventCreate(&e1));
  OpProgram idp;
  std::memset(&idp, 0, sizeof(idp));
  idp.want_extrema = 1;
  float best = 1e30f;
  int best_chunk = B->chunk;
  for (int c : {512, 1024, 2048, 4096, 16384, 65536}) {
    B->chunk = c;
    build_records(B, false);
    launch_push_batch(B->d_bb, B->d_items, B->shape.nitems, idp, BatchKick{}, B->d_partials,
                      kBatchThreads, s);
    OB_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < 3; r++)
      launch_push_batch(B->d_bb, B->d_items, B->shape.nitems, idp, BatchKick{}, B->d_partials,
                        kBatchThreads, s);
    OB_HIP(hipEventRecord(e1, s));
    const float t = time_ms(e0, e1);
    if (t < best) {
      best = t;
      best_chunk = c;
    }
  }
  B->chunk = best_chunk;
  if (!B->calcs.empty()) {
    best = 1e30f;
    int best_item = B->item;
    for (int item : {256, 1024, 4096}) {
      B->item = item;
      build_records(B, false);
      launch_push_batch(B->d_bb, B->d_items, B->shape.nitems, idp, BatchKick{}, B->d_partials,
                        kBatchThreads, s);
      BatchShape sh = B->shape;
      launch_sc2d_batch(B->d_bb, B->d_items, sh, B->d_partials, 0.0, 1.0, s);
      OB_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < 3; r++) launch_sc2d_batch(B->d_bb, B->d_items, sh, B->d_partials, 0.0, 1.0, s);
      OB_HIP(hipEventRecord(e1, s));
      const float t = time_ms(e0, e1);
      if (t < best) {
        best = t;
        best_item = item;
      }
    }
    B->item = best_item;
  }
  build_records(B, false);
  OB_HIP(hipEventDestroy(e0));
  OB_HIP(hipEventDestroy(e1));
  B->tuned = true;
  tuning_cache()[key] = {B->chunk, B->item};
}

}  // namespace

void batch_flush_owner(ob_batch* B) {
  OB_REQUIRE(!B->capturing, "a batched bunch was read during a graph capture");
  flush(B, false);
}

std::vector<ob_batch*> g_live_batches;

void batch_forget_calc(ob_sc2p5drb* c) {
  for (ob_batch* B : g_live_batches) {
    if (std::find(B->calcs.begin(), B->calcs.end(), c) == B->calcs.end()) continue;
    // The pending kick (and a captured graph) read this calculator's buffers: apply the
    // kick now, then drop every calculator of the batch; SC tracking needs a new set.
    std::string err;
    try {
      if (!B->capturing) flush(B, false);
    } catch (const Error& e) {
      err = e.msg;
    }
    B->calcs.clear();
    B->kick = BatchKick{};
    B->graph_version = -1;
    B->built_chunk = 0;
    OB_REQUIRE(err.empty(), err);
  }
}

}  // namespace ob

#define OB_API extern "C" __attribute__((visibility("default")))
#define OB_TRY try {
#define OB_CATCH               \
  }                            \
  catch (const ob::Error& e) { \
    ob::set_error(e.msg);      \
    return 1;                  \
  }                            \
  return 0;

OB_API int ob_batch_create(ob_bunch** bunches, int nb, ob_batch** out) {
  OB_TRY OB_REQUIRE(out != nullptr && bunches != nullptr && nb >= 1, "invalid batch arguments");
  ob::ensure_init();
  ob_batch* B = new ob_batch();
  try {
    for (int k = 0; k < nb; k++) {
      OB_REQUIRE(bunches[k] != nullptr, "null bunch handle");
      for (int j = 0; j < k; j++) OB_REQUIRE(bunches[j] != bunches[k], "a bunch appears twice");
      B->bunches.push_back(bunches[k]);
    }
    ob::check_same_species(B->bunches);
    for (ob_bunch* b : B->bunches) {
      ob::bunch_flush(b, nullptr);
      if (!B->leader && ob::bunch_size(b) > 0) B->leader = b;
    }
    OB_REQUIRE(B->leader != nullptr, "batch: all bunches are empty");
    for (ob_bunch* b : B->bunches) ob::bunch_set_owner(b, B);
    ob::g_live_batches.push_back(B);
  } catch (...) {
    delete B;
    throw;
  }
  *out = B;
  OB_CATCH
}

OB_API int ob_batch_destroy(ob_batch* B) {
  OB_TRY if (!B) return 0;
  if (!B->capturing) ob::flush(B, false);  // nothing pending is lost
  auto& lb = ob::g_live_batches;
  lb.erase(std::remove(lb.begin(), lb.end(), B), lb.end());
  for (ob_bunch* b : B->bunches) ob::bunch_set_owner(b, nullptr);
  (void)hipStreamSynchronize(ob::stream());
  if (B->exec) (void)hipGraphExecDestroy(B->exec);
  (void)hipFree(B->d_bb);
  (void)hipFree(B->d_items);
  (void)hipFree(B->d_partials);
  delete B;
  OB_CATCH
}

OB_API int ob_batch_set_sc2p5drb(ob_batch* B, ob_sc2p5drb** calcs) {
  OB_TRY OB_REQUIRE(B != nullptr && calcs != nullptr, "null handle");
  OB_REQUIRE(!B->capturing, "batch: cannot change calculators during a capture");
  std::vector<ob_sc2p5drb*> cs(calcs, calcs + B->bunches.size());
  for (size_t k = 0; k < cs.size(); k++) {
    ob_sc2p5drb* c = cs[k];
    OB_REQUIRE(c != nullptr, "null calculator handle");
    for (size_t j = 0; j < k; j++) OB_REQUIRE(cs[j] != c, "a calculator appears twice");
    ob::BatchBunch r{};
    ob::sc2d_batch_record(c, r);
    OB_REQUIRE(r.M0 <= 1024 && r.M1 <= 1024, "batch: grids up to 512 points per axis");
  }
  B->calcs = cs;
  B->tuned = B->fixed;
  B->built_chunk = 0;  // force a rebuild
  OB_CATCH
}

OB_API int ob_batch_track_sequence(ob_batch* B, const ob_map_op* ops, int nops) {
  OB_TRY OB_REQUIRE(B != nullptr, "null handle");
  OB_REQUIRE(nops >= 0 && (nops == 0 || ops != nullptr), "invalid op list");
  for (int k = 0; k < nops; k++) {
    // a graph replays the recorded constants: a momentum change would not be repeated
    OB_REQUIRE(!(B->capturing && ops[k].kind == OB_MAP_RINGRF && ops[k].c != 0.),
               "batch: RingRF with a non-zero synchronous phase cannot be captured");
    const int rc = ob::queue_map_op(B->leader, ops[k]);
    if (rc != 0) return rc;
    ob::collect(B);
  }
  OB_CATCH
}

OB_API int ob_batch_flush(ob_batch* B) {
  OB_TRY OB_REQUIRE(B != nullptr, "null handle");
  ob::flush(B, false);
  OB_CATCH
}

OB_API int ob_batch_sc2p5drb_track(ob_batch* B, double length, double pipe_radius) {
  OB_TRY OB_REQUIRE(B != nullptr, "null handle");
  OB_REQUIRE(!B->calcs.empty(), "batch: no SpaceChargeCalc2p5Drb set");
  for (ob_bunch* b : B->bunches)
    OB_REQUIRE(ob::bunch_size(b) >= 2, "batch: every bunch needs at least 2 particles");
  if (!B->tuned) {
    OB_REQUIRE(!B->capturing, "batch: tune before capturing");
    ob::collect(B);
    ob::autotune(B);
  }
  ob::flush(B, true);
  std::vector<ob::DenoiseJob> jobs;
  for (ob_sc2p5drb* c : B->calcs) {
    ob::DenoiseJob j;
    if (ob::sc2d_denoise_job(c, &j)) jobs.push_back(j);
  }
  ob::launch_sc2d_batch(B->d_bb, B->d_items, B->shape, B->d_partials, length, pipe_radius,
                        ob::stream(), jobs);
  const ob::SyncInfo sy = ob::bunch_sync(B->leader);
  B->kick.active = 1;
  B->kick.p4 = 2 * log(pipe_radius);                               // :139
  B->kick.p5 = -length * sy.r0 * sy.mass / (sy.gamma * sy.gamma);  // :141
  for (ob_sc2p5drb* c : B->calcs) ob::sc2d_mark_tracked(c);
  OB_CATCH
}

// Waits for the batch's work; fails if a bunch had zero extent in its last SC step.
OB_API int ob_batch_synchronize(ob_batch* B) {
  OB_TRY OB_REQUIRE(B != nullptr, "null handle");
  OB_REQUIRE(!B->capturing, "batch: synchronize during a capture");
  OB_HIP(hipStreamSynchronize(ob::stream()));
  for (size_t k = 0; k < B->calcs.size(); k++)
    OB_REQUIRE(ob::sc2d_read_status(B->calcs[k]) == 0,
               "SpaceChargeCalc2p5Drb::bunchAnalysis: bunch " + std::to_string(k) +
                   " has zero extent; cannot calculate space charge");
  OB_CATCH
}

OB_API int ob_batch_set_tuning(ob_batch* B, int chunk, int deposit_item) {
  OB_TRY OB_REQUIRE(B != nullptr, "null handle");
  OB_REQUIRE(chunk >= 256 && chunk <= (1 << 24), "work-item size out of range");
  OB_REQUIRE(deposit_item >= 64 && deposit_item <= (1 << 20), "deposit work-item size out of range");
  B->chunk = chunk;
  B->item = deposit_item;
  B->tuned = B->fixed = true;
  OB_CATCH
}

OB_API int ob_batch_get_tuning(ob_batch* B, int* chunk, int* deposit_item, int* tuned) {
  OB_TRY OB_REQUIRE(B != nullptr, "null handle");
  if (chunk) *chunk = B->chunk;
  if (deposit_item) *deposit_item = B->item;
  if (tuned) *tuned = B->tuned ? 1 : 0;
  OB_CATCH
}

OB_API int ob_batch_capture_begin(ob_batch* B) {
  OB_TRY OB_REQUIRE(B != nullptr, "null handle");
  OB_REQUIRE(!B->capturing, "batch: already capturing");
  if (!B->tuned && !B->calcs.empty()) {
    ob::collect(B);
    ob::autotune(B);
  }
  ob::flush(B, false);  // nothing pending inside the graph's boundaries
  OB_HIP(hipStreamSynchronize(ob::stream()));
  if (B->exec) {
    OB_HIP(hipGraphExecDestroy(B->exec));
    B->exec = nullptr;
  }
  B->graph_dt = ob::bunch_sync(B->leader).time;
  B->graph_version = B->version;
  OB_HIP(hipStreamBeginCapture(ob::stream(), hipStreamCaptureModeRelaxed));
  B->capturing = true;
  OB_CATCH
}

OB_API int ob_batch_capture_end(ob_batch* B) {
  OB_TRY OB_REQUIRE(B != nullptr && B->capturing, "batch: not capturing");
  hipGraph_t g = nullptr;
  try {
    ob::flush(B, false);
  } catch (...) {
    (void)hipStreamEndCapture(ob::stream(), &g);
    if (g) (void)hipGraphDestroy(g);
    B->capturing = false;
    throw;
  }
  OB_HIP(hipStreamEndCapture(ob::stream(), &g));
  B->capturing = false;
  OB_HIP(hipGraphInstantiate(&B->exec, g, nullptr, nullptr, 0));
  OB_HIP(hipGraphDestroy(g));
  B->graph_dt = ob::bunch_sync(B->leader).time - B->graph_dt;
  // The capture recorded the work without running it: the clocks advanced by one pass.
  const double t = ob::bunch_sync(B->leader).time - B->graph_dt;
  for (ob_bunch* b : B->bunches) ob::bunch_set_sync_time(b, t);
  OB_CATCH
}

// Run the captured work `count` times (asynchronously on the library stream).
OB_API int ob_batch_replay(ob_batch* B, int count) {
  OB_TRY OB_REQUIRE(B != nullptr && B->exec != nullptr, "batch: nothing captured");
  OB_REQUIRE(count >= 0, "negative replay count");
  ob::flush(B, false);
  OB_REQUIRE(B->version == B->graph_version,
             "batch: bunch sizes or buffers changed since the capture; capture again");
  for (int k = 0; k < count; k++) OB_HIP(hipGraphLaunch(B->exec, ob::stream()));
  const double t = ob::bunch_sync(B->leader).time + count * B->graph_dt;
  for (ob_bunch* b : B->bunches) {
    ob::bunch_set_sync_time(b, t);
    ob::bunch_device_written(b);
  }
  for (ob_sc2p5drb* c : B->calcs) ob::sc2d_mark_tracked(c);
  OB_CATCH
}
