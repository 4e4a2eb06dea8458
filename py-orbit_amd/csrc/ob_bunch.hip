// ob_bunch.hip -- library runtime, device-resident SoA bunch and the teapot_base entry
// points.
//
// The bunch keeps six FP64 coordinate arrays (x, xp, y, yp, z, dE) in HBM and a host
// mirror used only by per-particle accessors and bulk copies.  teapot_base calls do not
// launch anything: they append an Op (with host-precomputed constants) to the bunch's
// queue, and the queue is applied by one fused kernel when the bunch is next needed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "ob_internal.hpp"

namespace ob {

// ---------------------------------------------------------------------------
// errors and runtime
static thread_local std::string t_last_error;
void set_error(const std::string& msg) { t_last_error = msg; }

namespace {
int g_device = -1;
hipStream_t g_stream = nullptr;
hipStream_t g_aux_stream = nullptr;
int g_cus = 0;

struct Box {
  int* host;  // pinned
  std::string what;
};
std::vector<Box> g_boxes;
}  // namespace

hipStream_t stream() { return g_stream; }
hipStream_t aux_stream() { return g_aux_stream; }
int device_cu_count() { return g_cus; }

static void init_device(int device) {
  if (g_stream != nullptr) {
    OB_REQUIRE(device == g_device, "ob_init: library already bound to another device");
    return;
  }
  int n = 0;
  OB_HIP(hipGetDeviceCount(&n));
  OB_REQUIRE(n > 0, "ob_init: no HIP device visible");
  OB_REQUIRE(device >= 0 && device < n, "ob_init: device index out of range");
  OB_HIP(hipSetDevice(device));
  OB_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  OB_HIP(hipStreamCreateWithFlags(&g_aux_stream, hipStreamNonBlocking));
  hipDeviceProp_t prop;
  OB_HIP(hipGetDeviceProperties(&prop, device));
  g_cus = prop.multiProcessorCount;
  diag_scratch_setup(device);
  g_device = device;
}

void ensure_init() {
  if (g_stream == nullptr) init_device(0);
}

// ---------------------------------------------------------------------------
// kernel profiler
namespace {
struct ProfRec {
  const char* name;
  hipEvent_t a, b;
};
bool g_prof = false;
std::vector<ProfRec> g_prof_recs;
std::vector<hipEvent_t> g_prof_pool;
hipEvent_t prof_event() {
  if (!g_prof_pool.empty()) {
    hipEvent_t e = g_prof_pool.back();
    g_prof_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  OB_HIP(hipEventCreate(&e));
  return e;
}
}  // namespace

bool profiling() { return g_prof; }

void prof_mark(const char* name, bool begin, hipStream_t s) {
  if (begin) {
    ProfRec r{name, prof_event(), nullptr};
    OB_HIP(hipEventRecord(r.a, s));
    g_prof_recs.push_back(r);
  } else {
    for (size_t i = g_prof_recs.size(); i-- > 0;)
      if (g_prof_recs[i].b == nullptr && g_prof_recs[i].name == name) {
        g_prof_recs[i].b = prof_event();
        OB_HIP(hipEventRecord(g_prof_recs[i].b, s));
        return;
      }
  }
}

StatusBox status_box_create(const char* what) {
  ensure_init();
  StatusBox b;
  OB_HIP(hipHostMalloc((void**)&b.host, sizeof(int), hipHostMallocDefault));
  *b.host = 0;
  OB_HIP(hipHostGetDevicePointer((void**)&b.dev, b.host, 0));
  g_boxes.push_back(Box{b.host, what});
  return b;
}

void status_box_destroy(StatusBox b) {
  for (size_t i = 0; i < g_boxes.size(); i++)
    if (g_boxes[i].host == b.host) {
      g_boxes.erase(g_boxes.begin() + i);
      break;
    }
  (void)hipHostFree(b.host);
}

void check_statuses(bool wait) {
  if (g_boxes.empty()) return;
  if (wait) OB_HIP(hipStreamSynchronize(g_stream));
  std::string err;
  for (Box& b : g_boxes) {
    volatile int* v = b.host;
    if (*v != 0) {
      *v = 0;
      if (err.empty()) err = b.what;
    }
  }
  OB_REQUIRE(err.empty(), err);
}

}  // namespace ob

// ---------------------------------------------------------------------------
// Bunch
struct ob_bunch {
  // attributes (Bunch.cc:404-425) and synchronous particle (SyncPart.cc)
  double mass = 0.938272013, charge = 1.0, r0 = 1.534698e-18, macrosize = 0.0;
  double T = 0.0, p = 0.0, beta = 0.0, gamma = 1.0, time = 0.0;

  int64_t n = 0, cap = 0;
  double* d = nullptr;        // 6 * cap doubles (SoA)
  double* d_msize = nullptr;  // per-particle macrosize (cap) or null
  // Space-charge calculators keep the device arrays sorted by grid tile; perm[k] is the
  // user index of device slot k (valid when `permuted`).  Host-visible order never
  // changes: reads and raw device access restore it first.
  double* d_alt = nullptr;
  double* d_msize_alt = nullptr;
  int32_t* d_perm = nullptr;
  int32_t* d_perm_alt = nullptr;
  bool permuted = false;
  std::vector<double> h[6];   // host mirror
  std::vector<double> h_msize;
  std::vector<uint8_t> alive; // host-side flags for deleteParticle/compress
  bool has_msize = false;
  bool host_valid = true, dev_valid = true;

  std::vector<ob::Op> queue;
  ob_batch* owner = nullptr;  // batch applying this bunch's maps, if any
};

namespace ob {
namespace {

void sync_set_momentum(ob_bunch* b, double pz) {  // SyncPart.cc:133-138, 236-317
  const double m = b->mass;
  const double p2 = 0.0 * 0.0 + 0.0 * 0.0 + pz * pz;
  const double p_abs = sqrt(p2);
  const double m2 = m * m;
  const double w2 = m2 + p2;
  const double w = sqrt(w2);
  b->T = p2 / (sqrt(m2 + p2) + m);
  b->p = p_abs;
  b->beta = p_abs / w;
  b->gamma = w / m;
}

void dev_reserve(ob_bunch* b, int64_t want) {
  if (want <= b->cap) return;
  int64_t cap = b->cap ? b->cap : 1024;
  while (cap < want) cap *= 2;
  double* nd = nullptr;
  OB_HIP(hipMalloc((void**)&nd, sizeof(double) * 6 * cap));
  if (b->d) OB_HIP(hipFree(b->d));
  b->d = nd;
  if (b->d_msize) {
    OB_HIP(hipFree(b->d_msize));
    b->d_msize = nullptr;
  }
  for (void** p : {(void**)&b->d_alt, (void**)&b->d_msize_alt, (void**)&b->d_perm,
                   (void**)&b->d_perm_alt})
    if (*p) {
      OB_HIP(hipFree(*p));
      *p = nullptr;
    }
  b->permuted = false;
  b->cap = cap;
  b->dev_valid = false;
}

__global__ void unpermute_kernel(const double* __restrict__ src, double* __restrict__ dst,
                                 int64_t cap, const int32_t* __restrict__ perm,
                                 const double* __restrict__ msrc, double* __restrict__ mdst,
                                 int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t p = perm[i];
#pragma unroll
    for (int k = 0; k < 6; k++) dst[k * cap + p] = src[k * cap + i];
    if (msrc) mdst[p] = msrc[i];
  }
}

__global__ void iota_kernel(int32_t* __restrict__ perm, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    perm[i] = (int32_t)i;
}

void ensure_alt(ob_bunch* b) {
  if (!b->d_alt) OB_HIP(hipMalloc((void**)&b->d_alt, sizeof(double) * 6 * b->cap));
  if (!b->d_perm) OB_HIP(hipMalloc((void**)&b->d_perm, sizeof(int32_t) * b->cap));
  if (!b->d_perm_alt) OB_HIP(hipMalloc((void**)&b->d_perm_alt, sizeof(int32_t) * b->cap));
  if (b->has_msize && !b->d_msize_alt)
    OB_HIP(hipMalloc((void**)&b->d_msize_alt, sizeof(double) * b->cap));
}

// Put the device arrays back in user order (pending maps are unaffected: they are
// elementwise).
void restore_order(ob_bunch* b) {
  if (!b->permuted) return;
  ensure_alt(b);
  const int blocks = (int)std::min<int64_t>((b->n + 255) / 256, 4096);
  if (b->n) {
    ProfScope ps("unpermute");
    hipLaunchKernelGGL(unpermute_kernel, dim3(blocks), dim3(256), 0, stream(), b->d, b->d_alt,
                       b->cap, b->d_perm, b->has_msize ? b->d_msize : nullptr,
                       b->has_msize ? b->d_msize_alt : nullptr, b->n);
    OB_HIP(hipGetLastError());
  }
  std::swap(b->d, b->d_alt);
  if (b->has_msize) std::swap(b->d_msize, b->d_msize_alt);
  b->permuted = false;
}

void upload(ob_bunch* b) {
  if (b->dev_valid) return;
  dev_reserve(b, b->n);
  hipStream_t s = stream();
  for (int k = 0; k < 6; k++)
    if (b->n)
      OB_HIP(hipMemcpyAsync(b->d + (size_t)k * b->cap, b->h[k].data(), sizeof(double) * b->n,
                            hipMemcpyHostToDevice, s));
  if (b->has_msize) {
    if (!b->d_msize) OB_HIP(hipMalloc((void**)&b->d_msize, sizeof(double) * b->cap));
    if (b->n)
      OB_HIP(hipMemcpyAsync(b->d_msize, b->h_msize.data(), sizeof(double) * b->n,
                            hipMemcpyHostToDevice, s));
  }
  OB_HIP(hipStreamSynchronize(s));  // host vectors may change right after
  b->dev_valid = true;
  b->permuted = false;
}

void download(ob_bunch* b) {
  if (b->host_valid) return;
  bunch_flush(b, nullptr);
  restore_order(b);
  hipStream_t s = stream();
  for (int k = 0; k < 6; k++) {
    b->h[k].resize(b->n);
    if (b->n)
      OB_HIP(hipMemcpyAsync(b->h[k].data(), b->d + (size_t)k * b->cap, sizeof(double) * b->n,
                            hipMemcpyDeviceToHost, s));
  }
  OB_HIP(hipStreamSynchronize(s));
  check_statuses(true);
  b->host_valid = true;
}

// Host is about to modify the mirror: make it current and mark the device stale.
void host_write(ob_bunch* b) {
  download(b);
  b->dev_valid = false;
}

}  // namespace

DeviceBunchView bunch_device_view(ob_bunch* b) {
  upload(b);
  DeviceBunchView v;
  for (int k = 0; k < 6; k++) v.c[k] = b->d + (size_t)k * b->cap;
  v.msize = b->has_msize ? b->d_msize : nullptr;
  v.n = b->n;
  return v;
}

int64_t bunch_size(ob_bunch* b) { return b->n; }
bool bunch_in_batch(ob_bunch* b) { return b->owner != nullptr; }

SortBuffers bunch_sort_begin(ob_bunch* b) {
  upload(b);
  ensure_alt(b);
  if (!b->permuted && b->n) {
    const int blocks = (int)std::min<int64_t>((b->n + 255) / 256, 4096);
    hipLaunchKernelGGL(iota_kernel, dim3(blocks), dim3(256), 0, stream(), b->d_perm, b->n);
    OB_HIP(hipGetLastError());
  }
  SortBuffers s;
  s.cur = b->d;
  s.alt = b->d_alt;
  s.cap = b->cap;
  s.perm = b->d_perm;
  s.perm_alt = b->d_perm_alt;
  s.msize = b->has_msize ? b->d_msize : nullptr;
  s.msize_alt = b->has_msize ? b->d_msize_alt : nullptr;
  s.n = b->n;
  return s;
}

void bunch_sort_commit(ob_bunch* b) {
  std::swap(b->d, b->d_alt);
  std::swap(b->d_perm, b->d_perm_alt);
  if (b->has_msize) std::swap(b->d_msize, b->d_msize_alt);
  b->permuted = true;
  b->host_valid = false;
}

void bunch_device_written(ob_bunch* b) { b->host_valid = false; }

std::vector<Op> bunch_take_queue(ob_bunch* b) {
  std::vector<Op> q;
  q.swap(b->queue);
  return q;
}

void bunch_set_sync_time(ob_bunch* b, double t) { b->time = t; }
void bunch_copy_kinematics(ob_bunch* dst, const ob_bunch* src) {
  dst->T = src->T;
  dst->p = src->p;
  dst->beta = src->beta;
  dst->gamma = src->gamma;
  dst->time = src->time;
}

void bunch_set_owner(ob_bunch* b, ob_batch* owner) { b->owner = owner; }

std::vector<ob_bunch*> g_live_bunches;

// A calculator is about to free its buffers: apply every queued space-charge kick that
// reads them (the kick belongs to a trackBunch call that has already happened).
void flush_kicks_reading(const void* state) {
  for (ob_bunch* b : g_live_bunches) {
    for (const Op& op : b->queue) {
      const double want = ptr_bits(state);
      if ((op.kind == OP_SC2P5DRB_KICK || op.kind == OP_SC3D_KICK) &&
          std::memcmp(&op.p[0], &want, sizeof(double)) == 0) {
        try {
          bunch_flush(b, nullptr);
        } catch (const Error&) {
          // The kick cannot be applied; it must not outlive the buffers it reads.
          b->queue.clear();
          throw;
        }
        break;
      }
    }
  }
}

SyncInfo bunch_sync(ob_bunch* b) {
  return SyncInfo{b->mass, b->charge, b->r0, b->macrosize, b->T, b->p, b->beta, b->gamma, b->time};
}

int bunch_flush(ob_bunch* b, double* ext_partials) {
  if (b->owner) {  // maps and kicks of a batched bunch are the batch's
    OB_REQUIRE(ext_partials == nullptr,
               "this bunch belongs to a batch: track it through the batch");
    batch_flush_owner(b->owner);
    return 0;
  }
  if (b->queue.empty() && ext_partials == nullptr) return 0;
  DeviceBunchView v = bunch_device_view(b);
  size_t pos = 0;
  int blocks = 0;
  do {
    OpProgram prog;
    std::memset(&prog, 0, sizeof(prog));
    size_t take = std::min(b->queue.size() - pos, (size_t)kMaxOpsPerLaunch);
    prog.nops = (int32_t)take;
    for (size_t k = 0; k < take; k++) prog.ops[k] = b->queue[pos + k];
    pos += take;
    const bool last = pos >= b->queue.size();
    prog.want_extrema = (last && ext_partials) ? 1 : 0;
    blocks = launch_push(v, prog, prog.want_extrema ? ext_partials : nullptr, stream());
  } while (pos < b->queue.size());
  b->queue.clear();
  b->host_valid = false;
  return blocks;
}

void bunch_queue(ob_bunch* b, const Op& op) {
  if (b->n == 0) return;  // maps over an empty bunch are no-ops
  b->queue.push_back(op);
  b->host_valid = false;
  if (b->queue.size() >= 256 && !b->owner) bunch_flush(b, nullptr);
}

}  // namespace ob

// ---------------------------------------------------------------------------
// C ABI
#define OB_API extern "C" __attribute__((visibility("default")))
#define OB_TRY try {
#define OB_CATCH                                   \
  }                                                \
  catch (const ob::Error& e) {                     \
    ob::set_error(e.msg);                          \
    return 1;                                      \
  }                                                \
  catch (const std::exception& e) {                \
    ob::set_error(std::string("std: ") + e.what()); \
    return 1;                                      \
  }                                                \
  return 0;
#define OB_CHECK_B(b) OB_REQUIRE((b) != nullptr, "null bunch handle")

OB_API const char* ob_last_error(void) { return ob::t_last_error.c_str(); }
OB_API const char* ob_version(void) { return "orbit_mi355x 0.1 (gfx950, FP64)"; }

OB_API int ob_init(int device) {
  OB_TRY ob::init_device(device);
  OB_CATCH
}

OB_API int ob_synchronize(void) {
  OB_TRY ob::ensure_init();
  OB_HIP(hipStreamSynchronize(ob::stream()));
  OB_HIP(hipStreamSynchronize(ob::aux_stream()));
  ob::check_statuses(true);
  OB_CATCH
}

OB_API int ob_profile_enable(int on) {
  OB_TRY ob::ensure_init();
  OB_HIP(hipStreamSynchronize(ob::stream()));
  OB_HIP(hipStreamSynchronize(ob::aux_stream()));
  for (auto& r : ob::g_prof_recs) {
    ob::g_prof_pool.push_back(r.a);
    if (r.b) ob::g_prof_pool.push_back(r.b);
  }
  ob::g_prof_recs.clear();
  ob::g_prof = on != 0;
  OB_CATCH
}

OB_API int ob_profile_report(char* buf, int buflen) {
  OB_TRY ob::ensure_init();
  OB_HIP(hipStreamSynchronize(ob::stream()));
  OB_HIP(hipStreamSynchronize(ob::aux_stream()));
  std::vector<std::string> names;
  std::vector<double> tot, mn, mx;
  std::vector<long> cnt;
  for (auto& r : ob::g_prof_recs) {
    if (!r.b) continue;
    float ms = 0.f;
    OB_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    size_t k = 0;
    while (k < names.size() && names[k] != r.name) k++;
    if (k == names.size()) {
      names.push_back(r.name);
      tot.push_back(0.0);
      cnt.push_back(0);
      mn.push_back(1e300);
      mx.push_back(0.0);
    }
    tot[k] += ms;
    cnt[k]++;
    mn[k] = std::min(mn[k], (double)ms);
    mx[k] = std::max(mx[k], (double)ms);
  }
  std::string s;
  char line[256];
  for (size_t k = 0; k < names.size(); k++) {
    snprintf(line, sizeof(line), "%s %ld %.6f %.6f %.6f\n", names[k].c_str(), cnt[k], tot[k],
             mn[k], mx[k]);
    s += line;
  }
  OB_REQUIRE((int)s.size() < buflen, "profile report buffer too small");
  std::memcpy(buf, s.c_str(), s.size() + 1);
  OB_CATCH
}

OB_API int ob_bunch_create(ob_bunch** out) {
  OB_TRY OB_REQUIRE(out != nullptr, "null output pointer");
  ob::ensure_init();
  *out = new ob_bunch();
  ob::g_live_bunches.push_back(*out);
  OB_CATCH
}

OB_API int ob_bunch_destroy(ob_bunch* b) {
  OB_TRY if (!b) return 0;
  OB_REQUIRE(b->owner == nullptr, "bunch belongs to a batch: destroy the batch first");
  auto& lb = ob::g_live_bunches;
  lb.erase(std::remove(lb.begin(), lb.end(), b), lb.end());
  (void)hipStreamSynchronize(ob::stream());
  for (void* p : {(void*)b->d, (void*)b->d_msize, (void*)b->d_alt, (void*)b->d_msize_alt,
                  (void*)b->d_perm, (void*)b->d_perm_alt})
    if (p) (void)hipFree(p);
  delete b;
  OB_CATCH
}

OB_API int ob_bunch_set_mass(ob_bunch* b, double v) {
  OB_TRY OB_CHECK_B(b);
  b->mass = v;
  OB_CATCH
}
OB_API int ob_bunch_set_charge(ob_bunch* b, double v) {
  OB_TRY OB_CHECK_B(b);
  b->charge = v;
  OB_CATCH
}
OB_API int ob_bunch_set_classical_radius(ob_bunch* b, double v) {
  OB_TRY OB_CHECK_B(b);
  b->r0 = v;
  OB_CATCH
}
OB_API int ob_bunch_set_macrosize(ob_bunch* b, double v) {
  OB_TRY OB_CHECK_B(b);
  b->macrosize = v;
  OB_CATCH
}
OB_API int ob_bunch_get_attrs(ob_bunch* b, double* out) {
  OB_TRY OB_CHECK_B(b);
  out[0] = b->mass;
  out[1] = b->charge;
  out[2] = b->r0;
  out[3] = b->macrosize;
  OB_CATCH
}
OB_API int ob_bunch_set_kin_energy(ob_bunch* b, double T) {  // wrap_syncpart.cc:70-92
  OB_TRY OB_CHECK_B(b);
  const double m = b->mass;
  ob::sync_set_momentum(b, sqrt(T * (T + 2.0 * m)));
  OB_CATCH
}
OB_API int ob_bunch_set_sync_momentum(ob_bunch* b, double p) {  // SyncPart::setMomentum
  OB_TRY OB_CHECK_B(b);
  OB_REQUIRE(p >= 0.0, "synchronous momentum must be >= 0");
  ob::sync_set_momentum(b, p);
  OB_CATCH
}
OB_API int ob_bunch_get_sync(ob_bunch* b, double* out) {
  OB_TRY OB_CHECK_B(b);
  out[0] = b->T;
  out[1] = b->p;
  out[2] = b->beta;
  out[3] = b->gamma;
  out[4] = b->time;
  out[5] = b->mass;
  OB_CATCH
}
OB_API int ob_bunch_set_sync_time(ob_bunch* b, double t) {
  OB_TRY OB_CHECK_B(b);
  b->time = t;
  OB_CATCH
}

OB_API int ob_bunch_add_particle(ob_bunch* b, double x, double xp, double y, double yp, double z,
                                 double dE) {
  OB_TRY OB_CHECK_B(b);
  ob::host_write(b);
  const double v[6] = {x, xp, y, yp, z, dE};
  for (int k = 0; k < 6; k++) b->h[k].push_back(v[k]);
  b->alive.push_back(1);
  if (b->has_msize) b->h_msize.push_back(0.0);
  b->n++;
  OB_CATCH
}

OB_API int ob_bunch_set_coords(ob_bunch* b, int64_t n, const double* x, const double* xp,
                               const double* y, const double* yp, const double* z,
                               const double* dE) {
  OB_TRY OB_CHECK_B(b);
  OB_REQUIRE(n >= 0, "negative particle count");
  b->queue.clear();
  const double* src[6] = {x, xp, y, yp, z, dE};
  for (int k = 0; k < 6; k++) OB_REQUIRE(n == 0 || src[k] != nullptr, "null coordinate array");
  ob::dev_reserve(b, n);
  b->n = n;
  b->permuted = false;
  hipStream_t s = ob::stream();
  for (int k = 0; k < 6; k++)
    if (n)
      OB_HIP(hipMemcpyAsync(b->d + (size_t)k * b->cap, src[k], sizeof(double) * n,
                            hipMemcpyHostToDevice, s));
  OB_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < 6; k++) b->h[k].clear();
  b->alive.assign(n, 1);
  b->has_msize = false;
  b->h_msize.clear();
  b->dev_valid = true;
  b->host_valid = false;
  OB_CATCH
}

OB_API int ob_bunch_get_coords(ob_bunch* b, double* x, double* xp, double* y, double* yp,
                               double* z, double* dE) {
  OB_TRY OB_CHECK_B(b);
  double* dst[6] = {x, xp, y, yp, z, dE};
  if (b->host_valid || !b->dev_valid) {
    ob::download(b);
    for (int k = 0; k < 6; k++)
      if (dst[k] && b->n) std::memcpy(dst[k], b->h[k].data(), sizeof(double) * b->n);
    return 0;
  }
  ob::bunch_flush(b, nullptr);
  ob::restore_order(b);
  hipStream_t s = ob::stream();
  for (int k = 0; k < 6; k++)
    if (dst[k] && b->n)
      OB_HIP(hipMemcpyAsync(dst[k], b->d + (size_t)k * b->cap, sizeof(double) * b->n,
                            hipMemcpyDeviceToHost, s));
  OB_HIP(hipStreamSynchronize(s));
  ob::check_statuses(true);
  OB_CATCH
}

OB_API int ob_bunch_get_coord(ob_bunch* b, int64_t i, int k, double* v) {
  OB_TRY OB_CHECK_B(b);
  OB_REQUIRE(i >= 0 && i < b->n, "particle index out of range");
  OB_REQUIRE(k >= 0 && k < 6, "coordinate index out of range");
  ob::download(b);
  *v = b->h[k][i];
  OB_CATCH
}

OB_API int ob_bunch_set_coord(ob_bunch* b, int64_t i, int k, double v) {
  OB_TRY OB_CHECK_B(b);
  OB_REQUIRE(i >= 0 && i < b->n, "particle index out of range");
  OB_REQUIRE(k >= 0 && k < 6, "coordinate index out of range");
  ob::host_write(b);
  b->h[k][i] = v;
  OB_CATCH
}

OB_API int ob_bunch_get_size(ob_bunch* b, int64_t* n) {
  OB_TRY OB_CHECK_B(b);
  *n = b->n;
  OB_CATCH
}

OB_API int ob_bunch_delete_particle(ob_bunch* b, int64_t i) {
  OB_TRY OB_CHECK_B(b);
  OB_REQUIRE(i >= 0 && i < b->n, "particle index out of range");
  if ((int64_t)b->alive.size() != b->n) b->alive.assign(b->n, 1);
  b->alive[i] = 0;
  OB_CATCH
}

OB_API int ob_bunch_compress(ob_bunch* b) {  // Bunch.cc:704-763 (order kept)
  OB_TRY OB_CHECK_B(b);
  if ((int64_t)b->alive.size() != b->n) {
    b->alive.assign(b->n, 1);
    return 0;
  }
  bool any_dead = false;
  for (int64_t i = 0; i < b->n; i++) any_dead |= (b->alive[i] == 0);
  if (!any_dead) return 0;
  ob::host_write(b);
  int64_t w = 0;
  for (int64_t i = 0; i < b->n; i++) {
    if (!b->alive[i]) continue;
    for (int k = 0; k < 6; k++) b->h[k][w] = b->h[k][i];
    if (b->has_msize) b->h_msize[w] = b->h_msize[i];
    w++;
  }
  for (int k = 0; k < 6; k++) b->h[k].resize(w);
  if (b->has_msize) b->h_msize.resize(w);
  b->n = w;
  b->alive.assign(w, 1);
  OB_CATCH
}

OB_API int ob_bunch_set_macrosizes(ob_bunch* b, int64_t n, const double* m) {
  OB_TRY OB_CHECK_B(b);
  if (m == nullptr) {
    b->has_msize = false;
    b->h_msize.clear();
    return 0;
  }
  OB_REQUIRE(n == b->n, "macrosize array length differs from the bunch size");
  ob::download(b);
  ob::bunch_flush(b, nullptr);
  ob::restore_order(b);  // m is in user order
  b->h_msize.assign(m, m + n);
  b->has_msize = true;
  if (b->dev_valid) {
    if (!b->d_msize) OB_HIP(hipMalloc((void**)&b->d_msize, sizeof(double) * b->cap));
    if (n)
      OB_HIP(hipMemcpyAsync(b->d_msize, m, sizeof(double) * n, hipMemcpyHostToDevice,
                            ob::stream()));
    OB_HIP(hipStreamSynchronize(ob::stream()));
  }
  OB_CATCH
}

OB_API int ob_bunch_get_macrosizes(ob_bunch* b, int64_t n, double* m) {
  OB_TRY OB_CHECK_B(b);
  OB_REQUIRE(b->has_msize, "bunch has no macrosize attribute");
  OB_REQUIRE(n == b->n, "macrosize array length differs from the bunch size");
  if (n) std::memcpy(m, b->h_msize.data(), sizeof(double) * n);
  OB_CATCH
}

OB_API int ob_bunch_device_coords(ob_bunch* b, double** ptrs6) {
  OB_TRY OB_CHECK_B(b);
  ob::bunch_flush(b, nullptr);
  ob::restore_order(b);
  ob::DeviceBunchView v = ob::bunch_device_view(b);
  for (int k = 0; k < 6; k++) ptrs6[k] = v.c[k];
  b->host_valid = false;
  OB_CATCH
}

OB_API int ob_bunch_flush(ob_bunch* b) {
  OB_TRY OB_CHECK_B(b);
  ob::bunch_flush(b, nullptr);
  OB_CATCH
}

// ---------------------------------------------------------------------------
// teapot_base (teapotbase.cc).  Host work per call = the reference's pre-loop code.
namespace {
ob::Op mkop(int kind) {
  ob::Op op;
  std::memset(&op, 0, sizeof(op));
  op.kind = kind;
  return op;
}
double factorial(int n) {  // teapotbase.cc:45-57
  double f = 1.0;
  for (int i = 1; i <= n; i++) f = i * f;
  return f;
}
void drift_op(ob_bunch* b, double length) {  // teapotbase.cc:158-183
  const double v = 2.99792458e+8 * b->beta;
  b->time = b->time + length / v;
  ob::Op op = mkop(ob::OP_DRIFT);
  op.p[0] = length;
  op.p[1] = 1.0 / (b->gamma * b->gamma);
  op.p[2] = 1.0 / (b->p * b->beta);
  ob::bunch_queue(b, op);
}
}  // namespace

OB_API int ob_tpb_rotatexy(ob_bunch* b, double angle) {
  OB_TRY OB_CHECK_B(b);
  ob::Op op = mkop(ob::OP_ROTATEXY);
  op.p[0] = cos(angle);
  op.p[1] = sin(angle);
  ob::bunch_queue(b, op);
  OB_CATCH
}

OB_API int ob_tpb_drift(ob_bunch* b, double length) {
  OB_TRY OB_CHECK_B(b);
  drift_op(b, length);
  OB_CATCH
}

OB_API int ob_tpb_wrapbunch(ob_bunch* b, double length) {
  OB_TRY OB_CHECK_B(b);
  ob::Op op = mkop(ob::OP_WRAPBUNCH);
  op.p[0] = length;
  ob::bunch_queue(b, op);
  OB_CATCH
}

OB_API int ob_tpb_kick(ob_bunch* b, double kx, double ky, double kE, int) {
  OB_TRY OB_CHECK_B(b);
  ob::Op op = mkop(ob::OP_KICK);
  op.p[0] = kx;
  op.p[1] = ky;
  op.p[2] = kE;
  ob::bunch_queue(b, op);
  OB_CATCH
}

OB_API int ob_tpb_multp(ob_bunch* b, int pole, double kl, int skew, int) {
  OB_TRY OB_CHECK_B(b);
  OB_REQUIRE(pole >= 0 && pole < 47, "multp: pole out of range");
  if (b->charge == 0.) return 0;  // teapotbase.cc:338-340
  ob::Op op = mkop(ob::OP_MULTP);
  op.i0 = pole;
  op.i1 = skew;
  op.p[0] = kl / factorial(pole);
  ob::bunch_queue(b, op);
  OB_CATCH
}

static int multpfringe(ob_bunch* b, int pole, double kl, int skew, bool out) {
  OB_REQUIRE(pole >= 0 && pole < 47, "multpfringe: pole out of range");
  if (b->charge == 0.) return 0;  // teapotbase.cc:397-399
  ob::Op op = mkop(out ? ob::OP_MULTPFRINGE_OUT : ob::OP_MULTPFRINGE_IN);
  op.i0 = pole;
  op.i1 = skew;
  const double klfactlp1 = kl / (4.0 * factorial(pole + 2));
  double kre = skew ? 0.0 : klfactlp1, kim = skew ? klfactlp1 : 0.0;
  if (!out) {  // -kterm (teapotbase.cc:459-464)
    kre = -kre;
    kim = -kim;
  }
  op.p[0] = 1.0 / (b->p * b->beta);
  op.p[1] = kre;
  op.p[2] = kim;
  ob::bunch_queue(b, op);
  return 0;
}

OB_API int ob_tpb_multpfringeIN(ob_bunch* b, int pole, double kl, int skew, int) {
  OB_TRY OB_CHECK_B(b);
  multpfringe(b, pole, kl, skew, false);
  OB_CATCH
}
OB_API int ob_tpb_multpfringeOUT(ob_bunch* b, int pole, double kl, int skew, int) {
  OB_TRY OB_CHECK_B(b);
  multpfringe(b, pole, kl, skew, true);
  OB_CATCH
}

OB_API int ob_tpb_quad1(ob_bunch* b, double length, double kq, int) {  // teapotbase.cc:620-700
  OB_TRY OB_CHECK_B(b);
  if (kq == 0. || b->charge == 0.) {
    drift_op(b, length);
    return 0;
  }
  const double v = 2.99792458e+8 * b->beta;
  if (length > 0.) b->time = b->time + length / v;
  const double gamma2i = 1.0 / (b->gamma * b->gamma);
  const double dp_p_coeff = 1.0 / (b->p * b->beta);
  double sqrt_kq, kqlength, cx, sx, cy, sy;
  double m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (kq > 0.) {
    sqrt_kq = pow(kq, 0.5);
    kqlength = sqrt_kq * length;
    cx = cos(kqlength); sx = sin(kqlength); cy = cosh(kqlength); sy = sinh(kqlength);
    m[0] = cx; m[1] = sx / sqrt_kq; m[2] = -sx * sqrt_kq; m[3] = cx;
    m[4] = cy; m[5] = sy / sqrt_kq; m[6] = sy * sqrt_kq; m[7] = cy;
  } else {
    sqrt_kq = pow(-kq, 0.5);
    kqlength = sqrt_kq * length;
    cx = cosh(kqlength); sx = sinh(kqlength); cy = cos(kqlength); sy = sin(kqlength);
    m[0] = cx; m[1] = sx / sqrt_kq; m[2] = sx * sqrt_kq; m[3] = cx;
    m[4] = cy; m[5] = sy / sqrt_kq; m[6] = -sy * sqrt_kq; m[7] = cy;
  }
  ob::Op op = mkop(ob::OP_QUAD1);
  for (int k = 0; k < 8; k++) op.p[k] = m[k];
  op.p[8] = gamma2i;
  op.p[9] = length;
  op.p[10] = dp_p_coeff;
  ob::bunch_queue(b, op);
  OB_CATCH
}

OB_API int ob_tpb_quad2(ob_bunch* b, double length) {
  OB_TRY OB_CHECK_B(b);
  ob::Op op = mkop(ob::OP_QUAD2);
  op.p[0] = length;
  op.p[1] = 1.0 / (b->gamma * b->gamma);
  op.p[2] = 1.0 / (b->p * b->beta);
  ob::bunch_queue(b, op);
  OB_CATCH
}

static void quadfringe(ob_bunch* b, double kq, bool out) {
  if (b->charge == 0.) return;  // teapotbase.cc:786-788
  ob::Op op = mkop(out ? ob::OP_QUADFRINGE_OUT : ob::OP_QUADFRINGE_IN);
  op.p[0] = 1.0 / (b->p * b->beta);
  op.p[1] = kq;
  ob::bunch_queue(b, op);
}
OB_API int ob_tpb_quadfringeIN(ob_bunch* b, double kq, int) {
  OB_TRY OB_CHECK_B(b);
  quadfringe(b, kq, false);
  OB_CATCH
}
OB_API int ob_tpb_quadfringeOUT(ob_bunch* b, double kq, int) {
  OB_TRY OB_CHECK_B(b);
  quadfringe(b, kq, true);
  OB_CATCH
}

// quad3 (teapotbase.cc:761-764) does nothing.
OB_API int ob_tpb_quad3(ob_bunch* b, double, double, int) {
  OB_TRY OB_CHECK_B(b);
  OB_CATCH
}

// drifti (teapotbase.cc:120-141): one particle, through the host mirror.
OB_API int ob_tpb_drifti(ob_bunch* b, int64_t i, double length) {
  OB_TRY OB_CHECK_B(b);
  OB_REQUIRE(i >= 0 && i < b->n, "particle index out of range");
  ob::download(b);
  const double gamma2i = 1.0 / (b->gamma * b->gamma);
  const double dp_p_coeff = 1.0 / (b->p * b->beta);
  double* a[6];
  for (int k = 0; k < 6; k++) a[k] = b->h[k].data();
  const double dp_p = a[5][i] * dp_p_coeff;
  const double KNL = 1.0 / (1.0 + dp_p);
  a[0][i] += KNL * length * a[1][i];
  a[2][i] += KNL * length * a[3][i];
  double phifac = (a[1][i] * a[1][i] + a[3][i] * a[3][i] + dp_p * dp_p * gamma2i) / 2.0;
  phifac = (phifac * KNL - dp_p * gamma2i) * KNL;
  a[4][i] -= length * phifac;
  ob::host_write(b);
  OB_CATCH
}

OB_API int ob_tpb_wedgerotate(ob_bunch* b, double e, int frinout) {  // teapotbase.cc:923-966
  OB_TRY OB_CHECK_B(b);
  ob::Op op = mkop(ob::OP_WEDGEROTATE);
  op.i0 = frinout;
  op.p[0] = cos(e);
  op.p[1] = sin(e);
  op.p[2] = 1.0 / (b->p * b->beta);
  ob::bunch_queue(b, op);
  OB_CATCH
}

OB_API int ob_tpb_wedgedrift(ob_bunch* b, double e, int inout) {  // teapotbase.cc:986-1012
  OB_TRY OB_CHECK_B(b);
  ob::Op op = mkop(ob::OP_WEDGEDRIFT);
  op.i0 = inout;
  op.p[0] = cos(e) / sin(e);
  op.p[1] = 1.0 / (b->p * b->beta);
  op.p[2] = 1.0 / (b->gamma * b->gamma);
  ob::bunch_queue(b, op);
  OB_CATCH
}

// wedgebendCF (teapotbase.cc:1447-1509); wedgebend (:1037-1080) is the case n = 0.
OB_API int ob_tpb_wedgebendCF(ob_bunch* b, double e, int inout, double rho, int n,
                              const int* poles, const double* kls, const int* skews,
                              int nsteps, int) {
  OB_TRY OB_CHECK_B(b);
  OB_REQUIRE(n >= 0 && n <= 8, "wedgebendCF: at most 8 multipole terms");
  OB_REQUIRE(n == 0 || (poles && kls && skews), "wedgebendCF: null multipole arrays");
  int nst = nsteps / 2;
  if (nst < 1) nst = 1;
  ob::Op op = mkop(ob::OP_WEDGEBEND);
  const int nm = b->charge == 0. ? 0 : n;  // multpi does nothing for neutral particles
  int skewbits = 0, polebits = 0;
  for (int l = 0; l < nm; l++) {
    OB_REQUIRE(poles[l] >= 0 && poles[l] < 16, "wedgebendCF: pole out of range (0..15)");
    polebits |= poles[l] << (4 * l);
    if (skews[l]) skewbits |= 1 << l;
    op.p[4 + l] = kls[l];
  }
  op.i0 = (inout ? 1 : 0) | (nm << 1) | (skewbits << 8);
  op.i1 = nst;
  op.i2 = polebits;
  op.p[0] = cos(e) / sin(e);
  op.p[1] = rho;
  op.p[2] = 1.0 / (b->p * b->beta);
  op.p[3] = 1.0 / (b->gamma * b->gamma);
  ob::bunch_queue(b, op);
  OB_CATCH
}

OB_API int ob_tpb_wedgebend(ob_bunch* b, double e, int inout, double rho, int nsteps) {
  return ob_tpb_wedgebendCF(b, e, inout, rho, 0, nullptr, nullptr, nullptr, nsteps, 1);
}

OB_API int ob_tpb_bend1(ob_bunch* b, double length, double th) {  // teapotbase.cc:1099-1145
  OB_TRY OB_CHECK_B(b);
  const double v = 2.99792458e+8 * b->beta;
  if (length > 0.) b->time = b->time + length / v;
  const double betasq = b->beta * b->beta;
  const double rho = length / th, cx = cos(th), sx = sin(th);
  ob::Op op = mkop(ob::OP_BEND1);
  op.p[0] = cx;                             // m11
  op.p[1] = rho * sx;                       // m12
  op.p[2] = rho * (1.0 - cx);               // m16
  op.p[3] = -sx / rho;                      // m21
  op.p[4] = cx;                             // m22
  op.p[5] = sx;                             // m26
  op.p[6] = -sx;                            // m51
  op.p[7] = -rho * (1.0 - cx);              // m52
  op.p[8] = -betasq * length + rho * sx;    // m56
  op.p[9] = length;
  op.p[10] = 1.0 / (b->p * b->beta);
  ob::bunch_queue(b, op);
  OB_CATCH
}

OB_API int ob_tpb_bend2(ob_bunch* b, double length) {  // teapotbase.cc:1164-1190 (= quad2)
  return ob_tpb_quad2(b, length);
}

OB_API int ob_tpb_bend3(ob_bunch* b, double th) {  // teapotbase.cc:1208-1233
  OB_TRY OB_CHECK_B(b);
  ob::Op op = mkop(ob::OP_BEND3);
  op.p[0] = th;
  op.p[1] = 1.0 / (b->gamma * b->gamma);
  op.p[2] = 1.0 / (b->p * b->beta);
  ob::bunch_queue(b, op);
  OB_CATCH
}

OB_API int ob_tpb_bend4(ob_bunch* b, double th) {  // teapotbase.cc:1251-1274
  OB_TRY OB_CHECK_B(b);
  ob::Op op = mkop(ob::OP_BEND4);
  op.p[0] = th;
  op.p[1] = 1.0 / (b->p * b->beta);
  ob::bunch_queue(b, op);
  OB_CATCH
}

static void bendfringe(ob_bunch* b, double rho, bool out) {  // teapotbase.cc:1292-1348
  ob::Op op = mkop(out ? ob::OP_BENDFRINGE_OUT : ob::OP_BENDFRINGE_IN);
  op.p[0] = rho;
  op.p[1] = 1.0 / (b->p * b->beta);
  ob::bunch_queue(b, op);
}
OB_API int ob_tpb_bendfringeIN(ob_bunch* b, double rho) {
  OB_TRY OB_CHECK_B(b);
  bendfringe(b, rho, false);
  OB_CATCH
}
OB_API int ob_tpb_bendfringeOUT(ob_bunch* b, double rho) {
  OB_TRY OB_CHECK_B(b);
  bendfringe(b, rho, true);
  OB_CATCH
}

// soln (teapotbase.cc:1367-1425).  The zero-field test is on |B| as a double (abs
// resolves to the double overload there).
OB_API int ob_tpb_soln(ob_bunch* b, double length, double B, int) {
  OB_TRY OB_CHECK_B(b);
  if (fabs(B) < 1.0e-100 || b->charge == 0.) {
    drift_op(b, length);
    return 0;
  }
  const double v = 2.99792458e+8 * b->beta;
  if (length > 0.) b->time = b->time + length / v;
  ob::Op op = mkop(ob::OP_SOLN);
  op.p[0] = B * b->charge;
  op.p[1] = length;
  op.p[2] = 1.0 / (b->gamma * b->gamma);
  op.p[3] = 1.0 / (b->p * b->beta);
  ob::bunch_queue(b, op);
  OB_CATCH
}

// RingRF (teapotbase.cc:1527-1566): a non-zero synchronous phase changes the synchronous
// momentum by the energy gain (SyncPart::setMomentum, SyncPart.cc:133-138, 314-317).
OB_API int ob_tpb_RingRF(ob_bunch* b, double ring_length, int harmonic_numb, double voltage,
                         double phase_s, int useCharge) {
  OB_TRY OB_CHECK_B(b);
  double charge = +1.0;
  if (useCharge == 1) charge = b->charge;
  const double coeff = charge;
  double Factor = 0.;
  if (ring_length > 0.) Factor = 2.0 * 3.14159265358979323846 / ring_length;
  const double p_synch_in = b->p;
  if (phase_s != 0.) {
    double kin_e = b->T;
    kin_e += coeff * voltage * sin(phase_s);
    ob::sync_set_momentum(b, sqrt(kin_e * (kin_e + 2.0 * b->mass)));
  }
  ob::Op op = mkop(ob::OP_RINGRF);
  op.p[0] = harmonic_numb * Factor;
  op.p[1] = phase_s;
  op.p[2] = voltage;
  op.p[3] = coeff;
  op.p[4] = p_synch_in / b->p;
  ob::bunch_queue(b, op);
  OB_CATCH
}

namespace ob {
// Queue one map of the public op list through its single entry point (host constants).
int queue_map_op(ob_bunch* b, const ob_map_op& o) {
  switch (o.kind) {
    case OB_MAP_ROTATEXY: return ob_tpb_rotatexy(b, o.a);
    case OB_MAP_DRIFT: return ob_tpb_drift(b, o.a);
    case OB_MAP_WRAPBUNCH: return ob_tpb_wrapbunch(b, o.a);
    case OB_MAP_KICK: return ob_tpb_kick(b, o.a, o.b, o.c, o.ic);
    case OB_MAP_MULTP: return ob_tpb_multp(b, o.ia, o.a, o.ib, o.ic);
    case OB_MAP_MULTPFRINGE_IN: return ob_tpb_multpfringeIN(b, o.ia, o.a, o.ib, o.ic);
    case OB_MAP_MULTPFRINGE_OUT: return ob_tpb_multpfringeOUT(b, o.ia, o.a, o.ib, o.ic);
    case OB_MAP_QUAD1: return ob_tpb_quad1(b, o.a, o.b, o.ic);
    case OB_MAP_QUAD2: return ob_tpb_quad2(b, o.a);
    case OB_MAP_QUADFRINGE_IN: return ob_tpb_quadfringeIN(b, o.b, o.ic);
    case OB_MAP_QUADFRINGE_OUT: return ob_tpb_quadfringeOUT(b, o.b, o.ic);
    case OB_MAP_BEND1: return ob_tpb_bend1(b, o.a, o.b);
    case OB_MAP_BEND2: return ob_tpb_bend2(b, o.a);
    case OB_MAP_BEND3: return ob_tpb_bend3(b, o.a);
    case OB_MAP_BEND4: return ob_tpb_bend4(b, o.a);
    case OB_MAP_BENDFRINGE_IN: return ob_tpb_bendfringeIN(b, o.a);
    case OB_MAP_BENDFRINGE_OUT: return ob_tpb_bendfringeOUT(b, o.a);
    case OB_MAP_WEDGEROTATE: return ob_tpb_wedgerotate(b, o.a, o.ia);
    case OB_MAP_WEDGEDRIFT: return ob_tpb_wedgedrift(b, o.a, o.ia);
    case OB_MAP_WEDGEBEND: return ob_tpb_wedgebend(b, o.a, o.ia, o.b, o.ib);
    case OB_MAP_WEDGEBENDCF:
      return ob_tpb_wedgebendCF(b, o.a, o.ia, o.b, o.n, o.poles, o.kls, o.skews, o.ib, o.ic);
    case OB_MAP_SOLN: return ob_tpb_soln(b, o.a, o.b, o.ic);
    case OB_MAP_RINGRF: return ob_tpb_RingRF(b, o.a, o.ia, o.b, o.c, o.ic);
    default: throw Error{"unknown map kind " + std::to_string(o.kind)};
  }
}
}  // namespace ob

// A whole map sequence in one call: each op goes through the same host-side constants as
// its single entry point above, then the sequence runs as one fused launch.
OB_API int ob_track_sequence(ob_bunch* b, const ob_map_op* ops, int nops) {
  OB_TRY OB_CHECK_B(b);
  OB_REQUIRE(nops >= 0 && (nops == 0 || ops != nullptr), "invalid op list");
  for (int k = 0; k < nops; k++) {
    const int rc = ob::queue_map_op(b, ops[k]);
    if (rc != 0) return rc;  // message already set by the map entry point
  }
  ob::bunch_flush(b, nullptr);
  OB_CATCH
}

// ---------------------------------------------------------------------------
// Calibration: a plain streaming copy of n doubles (8n bytes read, 8n written), used to
// scale hardware byte counters.  No reference counterpart.
namespace {
__global__ __launch_bounds__(256) void copy_kernel(const double* __restrict__ src,
                                                   double* __restrict__ dst, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}
// The wider yardstick: 16-byte loads, U of them in flight per thread before the stores;
// NT: non-temporal stores (they bypass the L2 allocate on write).
template <int U, bool NT>
__global__ __launch_bounds__(256) void copy_wide_kernel(const double2* __restrict__ src,
                                                        double2* __restrict__ dst, int64_t n2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + (U - 1) * stride < n2; i += U * stride) {
    double2 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = src[i + u * stride];
#pragma unroll
    for (int u = 0; u < U; u++) {
      if constexpr (NT) {
        __builtin_nontemporal_store(v[u].x, &dst[i + u * stride].x);
        __builtin_nontemporal_store(v[u].y, &dst[i + u * stride].y);
      } else {
        dst[i + u * stride] = v[u];
      }
    }
  }
  for (; i < n2; i += stride) dst[i] = src[i];
}
}  // namespace

// Best of the wide copy variants over a sweep of grid sizes: ms per copy of n doubles
// (n even) and the variant that reached it (*variant = 4 * grid-size index + 2 * NT + (U==8)).
OB_API int ob_calibrate_copy_wide(int64_t n, int reps, double* ms_per_copy, int* variant) {
  OB_TRY OB_REQUIRE(n > 0 && n % 2 == 0 && reps > 0, "n must be positive and even, reps > 0");
  ob::ensure_init();
  double *a = nullptr, *b2 = nullptr;
  OB_HIP(hipMalloc((void**)&a, sizeof(double) * n));
  OB_HIP(hipMalloc((void**)&b2, sizeof(double) * n));
  hipStream_t s = ob::stream();
  const int64_t n2 = n / 2;
  hipLaunchKernelGGL(copy_kernel, dim3(1024), dim3(256), 0, s, b2, a, n);  // touch
  hipEvent_t e0, e1;
  OB_HIP(hipEventCreate(&e0));
  OB_HIP(hipEventCreate(&e1));
  const int grids[4] = {1024, 2048, 4096, 8192};
  double best = 1e30;
  int best_v = -1;
  for (int gi = 0; gi < 4; gi++)
    for (int v = 0; v < 4; v++) {
      const int blocks = (int)std::min<int64_t>((n2 + 255) / 256, grids[gi]);
      auto launch = [&]() {
        const double2* src = (const double2*)a;
        double2* dst = (double2*)b2;
        switch (v) {
          case 0: hipLaunchKernelGGL((copy_wide_kernel<4, false>), dim3(blocks), dim3(256), 0, s, src, dst, n2); break;
          case 1: hipLaunchKernelGGL((copy_wide_kernel<8, false>), dim3(blocks), dim3(256), 0, s, src, dst, n2); break;
          case 2: hipLaunchKernelGGL((copy_wide_kernel<4, true>), dim3(blocks), dim3(256), 0, s, src, dst, n2); break;
          default: hipLaunchKernelGGL((copy_wide_kernel<8, true>), dim3(blocks), dim3(256), 0, s, src, dst, n2); break;
        }
      };
      launch();  // warm
      OB_HIP(hipEventRecord(e0, s));
      for (int r = 0; r < reps; r++) {
        ob::ProfScope ps("copy_wide_kernel");
        launch();
      }
      OB_HIP(hipEventRecord(e1, s));
      OB_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      OB_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms / reps < best) {
        best = ms / reps;
        best_v = 4 * gi + v;
      }
    }
  OB_HIP(hipGetLastError());
  *ms_per_copy = best;
  *variant = best_v;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b2);
  OB_CATCH
}

OB_API int ob_calibrate_copy(int64_t n, int reps, double* ms_per_copy) {
  OB_TRY OB_REQUIRE(n > 0 && reps > 0, "n and reps must be positive");
  ob::ensure_init();
  double *a = nullptr, *b2 = nullptr;
  OB_HIP(hipMalloc((void**)&a, sizeof(double) * n));
  OB_HIP(hipMalloc((void**)&b2, sizeof(double) * n));
  hipStream_t s = ob::stream();
  hipLaunchKernelGGL(copy_kernel, dim3(1024), dim3(256), 0, s, b2, a, n);  // touch
  hipEvent_t e0, e1;
  OB_HIP(hipEventCreate(&e0));
  OB_HIP(hipEventCreate(&e1));
  const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8 * 1024);
  OB_HIP(hipEventRecord(e0, s));
  for (int r = 0; r < reps; r++) {
    ob::ProfScope ps("copy_kernel");
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, s, a, b2, n);
  }
  OB_HIP(hipEventRecord(e1, s));
  OB_HIP(hipEventSynchronize(e1));
  float ms = 0.f;
  OB_HIP(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_copy = ms / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b2);
  OB_CATCH
}
