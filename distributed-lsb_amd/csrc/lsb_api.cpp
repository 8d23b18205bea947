// C ABI of the sort (include/lsb.h): argument checks, pass planning,
// workspace management and the launch sequence that replaces mySort
// (mpi/mpi_lsbsort.cpp:591-597).
//
// One sort is, all on the caller's stream and without host synchronisation:
//   k_zero      counters
//   k_histogram every digit's counts in one read of the keys
//   k_scan_plan bucket starts per digit, skip mask, pass plan
//   k_onesweep  once per digit (a skipped digit's launch exits at once)
//   k_copy_back result B -> A if an odd number of passes ran
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <stdio.h>
#include <string.h>
#include <utility>
#include <vector>
#include <math.h>

#include "../../include/lsb.h"
#include "lsb_device.hpp"
#include "lsb_internal.hpp"

namespace {

thread_local int g_last_hip = 0;

#define LSB_HIP(x)                      \
  do {                                  \
    hipError_t e_ = (x);                \
    if (e_ != hipSuccess) {             \
      g_last_hip = (int)e_;             \
      (void)hipGetLastError();          \
      return LSB_EHIP;                  \
    }                                   \
  } while (0)

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct Plan {
  uint32_t key_bits, d, tile, passes, nb, ntiles, region_tiles, regions;
  lsb::KeyOrder ko;
  const lsb::SortVariant* v;
  // workspace offsets
  size_t counts, regnext, starts, plan, tickets, status, total;
};

lsb_config defaults() {
  lsb_config c;
  memset(&c, 0, sizeof c);
  c.key_bits = 64;
  return c;
}

// Key order from LSB_KEYS_SIGNED / LSB_KEYS_FLOAT (at most one; floats are
// binary32 or binary64, so key_bits 32 or 64).
int make_key_order(uint32_t flags, uint32_t kb, lsb::KeyOrder* ko) {
  const bool sgn = (flags & LSB_KEYS_SIGNED) != 0, flt = (flags & LSB_KEYS_FLOAT) != 0;
  ko->sign = ko->fsign = ko->full = 0;
  if (sgn && flt) return LSB_EINVAL;
  if (flt && kb != 32 && kb != 64) return LSB_EINVAL;
  if (!sgn && !flt) return LSB_OK;
  ko->sign = 1ull << (kb - 1);
  ko->full = kb >= 64 ? ~0ull : ((1ull << kb) - 1);
  ko->fsign = flt ? ko->sign : 0;
  return LSB_OK;
}

// Tuned table (lsb_autotune / lsb_tune_load), guarded by g_tune_mu.
std::mutex g_tune_mu;
std::vector<lsb_tune_entry> g_tune;

// Built-in table: lsb_autotune's own output on one MI355X (uniform keys,
// 2^12 .. 2^30, both key-width classes).  Variants are named by (digit_bits,
// tile, form) and resolved to ids when used, like a loaded table.  A table
// made by lsb_autotune or lsb_tune_load takes precedence; lsb_tune_clear
// returns to this one.
const lsb_tune_entry kBuiltinTune[] = {
    {12, 64, 8, 2048, 0, 0, 0.0671f, 0.0678f},
    {13, 64, 8, 2048, 3, 0, 0.0689f, 0.0726f},
    {14, 64, 8, 2048, 3, 0, 0.0727f, 0.0768f},
    {15, 64, 8, 2048, 3, 0, 0.0814f, 0.0829f},
    {16, 64, 8, 2048, 3, 0, 0.0857f, 0.0948f},
    {17, 64, 8, 2048, 3, 0, 0.0961f, 0.1086f},
    {18, 64, 8, 2048, 3, 0, 0.1114f, 0.1208f},
    {19, 64, 8, 4096, 3, 0, 0.1455f, 0.1495f},
    {20, 64, 8, 8192, 3, 0, 0.1777f, 0.1799f},
    {21, 64, 8, 8192, 3, 0, 0.1960f, 0.2246f},
    {22, 64, 8, 8192, 3, 0, 0.3044f, 0.3286f},
    {23, 64, 8, 8192, 3, 0, 0.5250f, 0.6053f},
    {24, 64, 8, 8192, 3, 0, 1.0219f, 1.1940f},
    {25, 64, 8, 8192, 3, 0, 1.9267f, 2.2351f},
    {26, 64, 8, 8192, 3, 0, 3.7347f, 4.3239f},
    {27, 64, 8, 8192, 3, 0, 7.3250f, 8.5104f},
    {28, 64, 8, 8192, 3, 0, 14.4841f, 17.4149f},
    {29, 64, 8, 8192, 3, 0, 29.1784f, 35.0985f},
    {30, 64, 8, 8192, 3, 0, 58.8061f, 70.4506f},
    {12, 32, 8, 2048, 3, 0, 0.0401f, 0.0412f},
    {13, 32, 8, 2048, 0, 0, 0.0417f, 0.0417f},
    {14, 32, 8, 2048, 3, 0, 0.0426f, 0.0434f},
    {15, 32, 8, 2048, 3, 0, 0.0473f, 0.0478f},
    {16, 32, 8, 2048, 3, 0, 0.0498f, 0.0542f},
    {17, 32, 8, 2048, 3, 0, 0.0571f, 0.0637f},
    {18, 32, 8, 2048, 3, 0, 0.0652f, 0.0703f},
    {19, 32, 8, 2048, 3, 0, 0.0815f, 0.0821f},
    {20, 32, 8, 8192, 3, 0, 0.0956f, 0.0981f},
    {21, 32, 8, 8192, 3, 0, 0.1049f, 0.1191f},
    {22, 32, 8, 8192, 3, 0, 0.1601f, 0.1724f},
    {23, 32, 8, 8192, 3, 0, 0.2634f, 0.2971f},
    {24, 32, 8, 8192, 3, 0, 0.5076f, 0.6040f},
    {25, 32, 8, 8192, 3, 0, 0.9753f, 1.1542f},
    {26, 32, 8, 8192, 3, 0, 1.8945f, 2.2264f},
    {27, 32, 8, 8192, 3, 0, 3.6853f, 4.3765f},
    {28, 32, 8, 8192, 3, 0, 7.3920f, 8.8008f},
    {29, 32, 8, 8192, 3, 0, 14.8625f, 17.6192f},
    {30, 32, 8, 8192, 3, 0, 29.6654f, 35.0282f},
};

uint32_t variant_id(uint32_t d, uint32_t tile, uint32_t persistent) {
  int nv = 0;
  const lsb::SortVariant* vs = lsb::sort_variants(&nv);
  for (int i = 0; i < nv; ++i)
    if (vs[i].digit_bits == d && vs[i].tile == tile && vs[i].persistent == persistent)
      return (uint32_t)i + 1;
  return 0;
}

// Automatic variant: the tuned entry of the nearest size class with the
// same key-width class (<= 32 bits or wider), from the current table or
// else the built-in one; 0 if neither names a compiled variant.
uint32_t auto_variant(size_t n, uint32_t key_bits) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  const bool narrow = key_bits <= 32;
  uint32_t lg = 0;
  while (lg < 63 && ((size_t)1 << (lg + 1)) <= n) ++lg;
  const lsb_tune_entry* table = g_tune.empty() ? kBuiltinTune : g_tune.data();
  const size_t count =
      g_tune.empty() ? sizeof(kBuiltinTune) / sizeof(kBuiltinTune[0]) : g_tune.size();
  const lsb_tune_entry* best = nullptr;
  for (size_t i = 0; i < count; ++i) {
    const lsb_tune_entry& e = table[i];
    if ((e.key_bits <= 32) != narrow) continue;
    const uint32_t dist = e.log2n > lg ? e.log2n - lg : lg - e.log2n;
    const uint32_t bd = best ? (best->log2n > lg ? best->log2n - lg : lg - best->log2n) : ~0u;
    if (dist < bd) best = &e;
  }
  if (!best) return 0u;
  return g_tune.empty() ? variant_id(best->digit_bits, best->tile, best->persistent)
                        : best->variant;
}

void auto_choice(size_t n, uint32_t key_bits, uint32_t* d, uint32_t* tile, uint32_t* variant) {
  *variant = auto_variant(n, key_bits);
  *d = 8;
  *tile = 256 * 16;
}

int make_plan(size_t n, const lsb_config& c, Plan* p) {
  p->key_bits = c.key_bits ? c.key_bits : 64;
  if (p->key_bits > 64) return LSB_EINVAL;
  if (make_key_order(c.flags, p->key_bits, &p->ko)) return LSB_EINVAL;
  uint32_t d = c.digit_bits, tile = c.tile, variant = c.variant;
  if (d == 0 && tile == 0 && variant == 0) {
    auto_choice(n, p->key_bits, &d, &tile, &variant);
  } else if (d == 0) {
    d = 8;
  }
  int nv = 0;
  const lsb::SortVariant* vs = lsb::sort_variants(&nv);
  const lsb::SortVariant* pick = nullptr;
  if (variant) {
    if (variant > (uint32_t)nv) return LSB_EINVAL;
    pick = &vs[variant - 1];
    if ((c.digit_bits && c.digit_bits != pick->digit_bits) || (c.tile && c.tile != pick->tile))
      return LSB_EINVAL;
    d = pick->digit_bits;
  } else {
    for (int i = 0; i < nv; ++i) {
      if (vs[i].digit_bits != d) continue;
      if (tile == 0 || vs[i].tile == tile) {
        pick = &vs[i];
        break;
      }
    }
  }
  if (!pick) return LSB_EINVAL;
  p->v = pick;
  p->d = d;
  p->tile = pick->tile;
  p->passes = (p->key_bits + d - 1) / d;
  p->nb = 1u << d;
  p->ntiles = (uint32_t)((n + p->tile - 1) / p->tile);
  p->regions = p->v->persistent == 1 ? 1u : (uint32_t)lsb::kRegions;
  p->region_tiles = (p->ntiles + p->regions - 1) / p->regions;
  size_t off = 0;
  // counts: [P][R][2^d] words, then the lean histogram's two OR words
  p->counts = off;  off += align_up((size_t)p->passes * lsb::kRegions * p->nb * 4 + 8);
  p->regnext = off; off += align_up((size_t)p->passes * lsb::kRegions * p->nb * 4);
  p->starts = off;  off += align_up((size_t)p->passes * p->nb * 4);
  p->plan = off;    off += align_up((size_t)(4 * p->passes + 4) * 4);
  p->tickets = off; off += align_up((size_t)p->passes * lsb::kRegions * 4);
  p->status = off;  off += align_up(2 * (size_t)p->ntiles * p->nb * 4);
  p->total = off;
  return LSB_OK;
}

bool overlaps(const void* a, const void* b, size_t bytes) {
  const char* x = static_cast<const char*>(a);
  const char* y = static_cast<const char*>(b);
  return x < y + bytes && y < x + bytes;
}

// Host checks that need no HIP call.
int check_host_args(const lsb_elem* A, const lsb_elem* B, size_t n) {
  if (n > LSB_MAX_N) return LSB_ERANGE;
  if (n <= 1) return LSB_OK;
  if (!A || !B) return LSB_EINVAL;
  if (((uintptr_t)A | (uintptr_t)B) & 15u) return LSB_EINVAL;
  if (overlaps(A, B, n * sizeof(lsb_elem))) return LSB_EINVAL;
  return LSB_OK;
}

// The pointer must be device memory holding `bytes` from ptr on.
int check_device_range(const void* ptr, size_t bytes) {
  hipPointerAttribute_t attr;
  if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
    (void)hipGetLastError();
    return LSB_EINVAL;
  }
  if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged) return LSB_EINVAL;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)ptr) == hipSuccess && base) {
    const char* b = static_cast<const char*>(base);
    const char* p = static_cast<const char*>(ptr);
    if (p < b || (size_t)(p - b) + bytes > size) return LSB_EINVAL;
  } else {
    (void)hipGetLastError();
  }
  return LSB_OK;
}

// Library-managed workspaces, one per (device, stream), grown on demand.
}  // namespace

namespace lsb {

// One memory pool per device for the library's own stream-ordered
// allocations (workspaces, the server's scratch).  It keeps freed memory
// for reuse instead of returning it at every synchronisation, and being the
// library's own pool it leaves the device's default pool as the caller set
// it.
static std::mutex g_pool_mu;
static std::map<int, hipMemPool_t> g_pools;

hipError_t pool_alloc(void** p, size_t bytes, hipStream_t s) {
  std::mutex& mu = g_pool_mu;
  std::map<int, hipMemPool_t>& pools = g_pools;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  hipMemPool_t pool = nullptr;
  {
    std::lock_guard<std::mutex> lock(mu);
    auto it = pools.find(dev);
    if (it == pools.end()) {
      hipMemPoolProps props;
      memset(&props, 0, sizeof props);
      props.allocType = hipMemAllocationTypePinned;
      props.handleTypes = hipMemHandleTypeNone;
      props.location.type = hipMemLocationTypeDevice;
      props.location.id = dev;
      if ((e = hipMemPoolCreate(&pool, &props)) != hipSuccess) return e;
      uint64_t keep = UINT64_MAX;
      if ((e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep)) !=
          hipSuccess) {
        (void)hipMemPoolDestroy(pool);
        return e;
      }
      pools[dev] = pool;
    } else {
      pool = it->second;
    }
  }
  return hipMallocFromPoolAsync(p, bytes, pool, s);
}

// Returns the pool's unused memory on device dev to the device.
hipError_t pool_trim(int dev) {
  std::lock_guard<std::mutex> lock(g_pool_mu);
  auto it = g_pools.find(dev);
  return it == g_pools.end() ? hipSuccess : hipMemPoolTrimTo(it->second, 0);
}

int check_sort_args(const void* A, const void* B, size_t n) {
  return check_host_args(static_cast<const lsb_elem*>(A), static_cast<const lsb_elem*>(B), n);
}

void note_hip_error(hipError_t e) {
  g_last_hip = (int)e;
  (void)hipGetLastError();
}

}  // namespace lsb

namespace {

// Each entry has its own mutex, held by a sort from the moment it takes the
// workspace until its launches are enqueued: a second thread on the same
// stream that grows the workspace then frees the old buffer (stream-ordered)
// only after those launches.  g_ws_mu guards the map alone; entries are
// never erased (lsb_release empties them), so a reference stays valid.
struct Workspace {
  std::mutex mu;
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::mutex g_ws_mu;
std::map<std::pair<int, void*>, Workspace> g_ws;

template <class W>
W& cache_entry(std::map<std::pair<int, void*>, W>& m, int dev, void* stream) {
  std::lock_guard<std::mutex> lock(g_ws_mu);
  return m[std::make_pair(dev, stream)];
}

// Takes the (device, stream) workspace, grown to `need` bytes, and returns
// it with its entry locked in *hold.
int get_workspace(void* stream, size_t need, void** out, std::unique_lock<std::mutex>* hold) {
  int dev = 0;
  LSB_HIP(hipGetDevice(&dev));
  Workspace& w = cache_entry(g_ws, dev, stream);
  *hold = std::unique_lock<std::mutex>(w.mu);
  if (w.bytes < need) {
    // Stream-ordered: the old buffer is released after the sorts already
    // queued on this stream, so growing waits for nothing.
    const size_t bytes = need + need / 4;
    void* p = nullptr;
    const hipError_t he = lsb::pool_alloc(&p, bytes, (hipStream_t)stream);
    if (he != hipSuccess) {
      g_last_hip = (int)he;
      (void)hipGetLastError();
      return LSB_ENOMEM;
    }
    if (w.ptr) LSB_HIP(hipFreeAsync(w.ptr, (hipStream_t)stream));
    w.ptr = p;
    w.bytes = bytes;
  }
  *out = w.ptr;
  return LSB_OK;
}

// Events of the last LSB_PROFILE call on this thread: [0] before the
// histogram, [1] after it, [2] after the scan, [3 + q] after pass q, then
// one after the copy-back.
struct ProfileState {
  std::vector<hipEvent_t> ev;
  uint32_t passes = 0, d = 0, tile = 0;
  const uint32_t* plan = nullptr;
  hipStream_t stream = nullptr;
  bool valid = false;
};
thread_local ProfileState g_prof;

// Events of the last LSB_PROFILE call on this thread that ran the MSD-first
// form, in launch order: [0] before the set-up, [1] after it, then one after
// each histogram, plan and scatter launch of wide step 0, level 0, wide
// step 1 and levels 1 .. levels-1, and one after the finish.
struct MsdProfileState {
  std::vector<hipEvent_t> ev;
  uint32_t levels = 0;
  const lsb::MsdLevel* lvl = nullptr;   // the call's level counters (workspace)
  const lsb::MsdLevel* wlvl = nullptr;
  const lsb::MsdRun* runs = nullptr;
  const uint32_t* run_ctr = nullptr;
  uint32_t cap_runs = 0;
  bool valid = false;
};
thread_local MsdProfileState g_mprof;

// Wide steps launched with `levels` levels: step 0, and step s after level
// s-1 while a level s exists to take its large buckets.
uint32_t msd_wide_steps(uint32_t levels) {
  return std::min<uint32_t>((uint32_t)lsb::kMsdWideSteps, levels);
}
// Events of a profiled MSD sort: start, set-up, three per wide step and
// level, finish.
size_t msd_prof_event_count(uint32_t levels) {
  return 2 + 3 * (size_t)(levels + msd_wide_steps(levels)) + 1;
}

// g_mprof.ev with at least `count` events.
int msd_prof_events(size_t count, hipEvent_t** out) {
  while (g_mprof.ev.size() < count) {
    hipEvent_t e;
    LSB_HIP(hipEventCreate(&e));
    g_mprof.ev.push_back(e);
  }
  *out = g_mprof.ev.data();
  return LSB_OK;
}

int prof_mark(bool on, size_t i, hipStream_t s) {
  if (!on) return LSB_OK;
  while (g_prof.ev.size() <= i) {
    hipEvent_t e;
    LSB_HIP(hipEventCreate(&e));
    g_prof.ev.push_back(e);
  }
  LSB_HIP(hipEventRecord(g_prof.ev[i], s));
  return LSB_OK;
}

int enqueue_sort(lsb_elem* A, lsb_elem* B, uint32_t n, const Plan& p, char* ws,
                 uint32_t flags, hipStream_t s) {
  const bool prof = (flags & LSB_PROFILE) != 0;
  int rc;
  uint32_t* counts = reinterpret_cast<uint32_t*>(ws + p.counts);
  uint32_t* regnext = reinterpret_cast<uint32_t*>(ws + p.regnext);
  uint32_t* starts = reinterpret_cast<uint32_t*>(ws + p.starts);
  uint32_t* plan = reinterpret_cast<uint32_t*>(ws + p.plan);
  uint32_t* tickets = reinterpret_cast<uint32_t*>(ws + p.tickets);
  uint32_t* status = reinterpret_cast<uint32_t*>(ws + p.status);
  const size_t slice = (size_t)p.ntiles * p.nb;

  if (prof) {
    g_prof.valid = g_mprof.valid = false;
    if (p.passes > 16) return LSB_EINVAL;
  }
  // Region passes count each later digit for their successor, so their
  // histogram is the lean one: the first kLeanDigits digits plus the OR mask
  // that finds single-bucket digits (d = 8 only).
  const size_t count_words = (size_t)p.passes * lsb::kRegions * p.nb;
  const bool lean = p.v->persistent == 3 && p.d == 8 && p.passes > (uint32_t)lsb::kLeanDigits;
  const uint32_t counted = lean ? (uint32_t)lsb::kLeanDigits : p.passes;
  uint32_t* orr = lean ? counts + count_words : nullptr;
  if ((rc = prof_mark(prof, 0, s))) return rc;
  lsb::launch_zero(counts, count_words + 2, s);
  LSB_HIP(lsb::launch_histogram(p.d, A, n, p.key_bits, p.ko, p.passes, counted, counts,
                                p.region_tiles * p.tile, p.regions, status, slice, orr, s));
  if ((rc = prof_mark(prof, 1, s))) return rc;
  if (lean)
    LSB_HIP(lsb::launch_plan_first(p.d, A, n, p.key_bits, p.ko, counts, p.region_tiles * p.tile,
                                   orr, plan, tickets, regnext, p.passes, counted, flags, s));
  else
    LSB_HIP(lsb::launch_scan_plan(p.d, counts, starts, plan, tickets, regnext, n, p.passes,
                                  flags, s));
  if ((rc = prof_mark(prof, 2, s))) return rc;
  for (uint32_t q = 0; q < p.passes; ++q) {
    lsb::PassArgs a;
    a.A = A;
    a.B = B;
    a.n = n;
    a.pass = q;
    a.shift = q * p.d;
    const uint32_t w = p.key_bits - a.shift < p.d ? p.key_bits - a.shift : p.d;
    a.mask = (w >= 32) ? 0xFFFFFFFFu : ((1u << w) - 1u);
    a.plan = plan;
    a.starts = starts + (size_t)q * p.nb;
    a.tickets = tickets;
    a.status = status;
    a.ntiles = p.ntiles;
    a.passes = p.passes;
    a.key_bits = p.key_bits;
    a.ko = p.ko;
    a.region_tiles = p.region_tiles;
    a.regcount = counts;
    a.regnext = regnext;
    p.v->scatter(a, s);
    if ((rc = prof_mark(prof, 3 + q, s))) return rc;
  }
  lsb::launch_copy_back(A, B, n, plan + 4 * p.passes, s);
  if ((rc = prof_mark(prof, 3 + p.passes, s))) return rc;
  LSB_HIP(hipGetLastError());
  if (prof) {
    g_prof.passes = p.passes;
    g_prof.d = p.d;
    g_prof.tile = p.tile;
    g_prof.plan = plan;
    g_prof.stream = s;
    g_prof.valid = true;
  }
  return LSB_OK;
}


// ------------------------------------------------------------ hipGraphs
// Captured launch sequences, keyed by everything the kernels' arguments
// depend on.  Capture runs on a library stream (capture only records), the
// instantiated graph is launched on the caller's stream.
std::mutex g_graph_mu;
std::map<std::vector<uint64_t>, hipGraphExec_t> g_graphs;
std::map<int, hipStream_t> g_capture_streams;
constexpr size_t kMaxGraphs = 256;

template <class Enqueue>
int launch_graph(const std::vector<uint64_t>& key, hipStream_t s, Enqueue enqueue) {
  std::lock_guard<std::mutex> lock(g_graph_mu);
  auto it = g_graphs.find(key);
  hipGraphExec_t exec = nullptr;
  if (it != g_graphs.end()) {
    exec = it->second;
  } else {
    int dev = 0;
    LSB_HIP(hipGetDevice(&dev));
    hipStream_t cs = g_capture_streams[dev];
    if (!cs) {
      LSB_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      g_capture_streams[dev] = cs;
    }
    if (g_graphs.size() >= kMaxGraphs) {
      LSB_HIP(hipDeviceSynchronize());  // no evicted graph may still run
      for (auto& kv : g_graphs) (void)hipGraphExecDestroy(kv.second);
      g_graphs.clear();
    }
    LSB_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(cs);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cs, &g);
    if (rc) {
      if (e == hipSuccess && g) (void)hipGraphDestroy(g);
      return rc;
    }
    LSB_HIP(e);
    const hipError_t ie = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    LSB_HIP(ie);
    g_graphs[key] = exec;
  }
  LSB_HIP(hipGraphLaunch(exec, s));
  return LSB_OK;
}

// ------------------------------------------------------- segmented sort
struct SegLayout {
  const lsb::SegVariant* v;
  bool msd;  // large segments sorted MSD-first (else by LSD passes)
  uint32_t key_bits, d, tile, passes, maxp, nb;
  lsb::KeyOrder ko;
  size_t cap_n;
  uint32_t cap_segs, max_tiles;
  size_t hdr, meta, counts, plan, tickets, status, tileseg, perm, total;
  // MSD-first form
  uint32_t cap_entries, cap_runs, cap_wide, cap_wide_tiles;
  size_t ment[2], mcounts[2], morr[2], mtile[2], mstatus, mlvl, mruns, mrunctr, mrg;
  size_t went, wcounts, worr, wtile, wstatus, wlvl;
};

// The MSD-first form sorts the large segments unless the caller asks for the
// LSD passes (LSB_FORM_LSD), names a digit width other than 8 or a tile, or
// sets LSB_NO_SKIP (run every digit pass).  Unless asked for (LSB_FORM_MSD)
// it is used for keys wider than 32 bits only: at 32 bits and below the LSD
// form's 4 passes move as many bytes as MSD's partitions and finish, run as
// fast on uniform keys and skip the constant digits of skewed ones (Zipf
// keys: 34 against 25 Gelem/s over a serving stream's jobs).
bool wants_msd(const lsb_config& c) {
  if ((c.flags & (LSB_FORM_LSD | LSB_NO_SKIP)) || !(c.digit_bits == 0 || c.digit_bits == 8) ||
      c.tile != 0)
    return false;
  return (c.flags & LSB_FORM_MSD) || c.key_bits == 0 || c.key_bits > 32;
}

// lsb_sort takes the MSD-first form (one segment of the batched sort) where
// wants_msd would and, unless asked for, only with the digit width left
// automatic and from LSB_MSD_MIN_N elements on.  There two 8-bit or 10-bit
// partitions and the in-LDS finish move about 112 B per element against the
// LSD form's 264 B (eight passes of 32 B and the 8 B histogram); below
// 2^24 elements the LSD passes run fast enough that the partitions' extra
// launches and finish do not pay.
bool plain_msd(size_t n, const lsb_config& c) {
  if (c.variant != 0 || !wants_msd(c)) return false;
  return (c.flags & LSB_FORM_MSD) || (c.digit_bits == 0 && n >= LSB_MSD_MIN_N);
}

// Capacities of the MSD-first form for batches of at most cap_n elements.
// Every range of every level holds more than kMsdFinishCap elements (the
// large segments, then buckets too big to finish) and the ranges of a level
// are disjoint, so a level has at most cap_n / (kMsdFinishCap + 1) of them,
// and at most cap_n / kMsdTile + that many tiles.  Runs: along one range's
// buckets, two consecutive runs not separated by a large bucket hold more
// than kMsdFinishCap elements together, so runs <= 2 cap_n / kMsdFinishCap
// + (number of partitioned ranges + large buckets) over all levels and wide
// steps (a wide range has at most 2^kMsdWideBits large buckets).
int make_msd_layout(size_t cap_n, SegLayout* L) {
  // Level 0 also holds one dead entry (and its tiles) per wide segment
  // besides the wide segments' large buckets: twice the bound.
  const uint64_t ents = 2 * (cap_n / (lsb::kMsdFinishCap + 1)) + 2;
  const uint64_t tiles = 2 * (cap_n / lsb::kMsdTile) + ents + 1;
  const uint64_t wide = cap_n / (lsb::kMsdWideMin + 1) + 1;
  const uint64_t wtiles = cap_n / lsb::kMsdTile + wide;
  const uint64_t runs = 2 * (cap_n / lsb::kMsdFinishCap) + 2 * lsb::kMsdMaxLevels * ents +
                        lsb::kMsdWideSteps * ((1u << lsb::kMsdWideBits) + 2) * wide + 64;
  if (tiles > 0x7FFFFFFFull || runs > 0xFFFFFFFFull) return LSB_ERANGE;
  L->cap_entries = (uint32_t)ents;
  L->max_tiles = (uint32_t)tiles;
  L->cap_runs = (uint32_t)runs;
  L->cap_wide = (uint32_t)wide;
  L->cap_wide_tiles = (uint32_t)wtiles;
  size_t off = L->perm;  // after the header and the segment table
  for (int i = 0; i < 2; ++i) {
    L->ment[i] = off;    off += align_up(ents * sizeof(lsb::MsdEntry));
    L->mcounts[i] = off; off += align_up(ents * 256 * 4);
    L->morr[i] = off;    off += align_up(ents * 16);
    L->mtile[i] = off;   off += align_up(tiles * 4);
  }
  L->mstatus = off; off += align_up(tiles * 256 * 4);
  L->mlvl = off;    off += align_up(lsb::kMsdMaxLevels * sizeof(lsb::MsdLevel));
  L->mruns = off;   off += align_up(runs * sizeof(lsb::MsdRun));
  L->mrunctr = off; off += align_up(16);
  L->mrg = off;     off += align_up((lsb::kRegions * 256 + lsb::kRegions) * 4);
  const size_t wnb = (size_t)1 << lsb::kMsdWideBits;
  L->went = off;    off += align_up(wide * sizeof(lsb::MsdEntry));
  L->wcounts = off; off += align_up(wide * wnb * 4);
  L->worr = off;    off += align_up(wide * 16);
  L->wtile = off;   off += align_up(wtiles * 4);
  L->wstatus = off; off += align_up(wtiles * wnb * 4);
  L->wlvl = off;    off += align_up(lsb::kMsdWideSteps * sizeof(lsb::MsdLevel));
  L->total = off;
  return LSB_OK;
}

int make_seg_layout(size_t cap_n, uint32_t cap_segs, const lsb_config& c, SegLayout* L,
                    size_t typical_n = 0) {
  L->key_bits = c.key_bits ? c.key_bits : 64;
  if (L->key_bits > 64) return LSB_EINVAL;
  if (make_key_order(c.flags, L->key_bits, &L->ko)) return LSB_EINVAL;
  L->msd = wants_msd(c);
  if (L->msd) {
    int nv = 0;
    const lsb::SegVariant* vs = lsb::seg_variants(&nv);
    L->v = &vs[0];
    L->d = 8;
    L->tile = lsb::kMsdTile;
    L->passes = (L->key_bits + 7) / 8;
    L->maxp = 8;
    L->nb = 256;
    L->cap_n = cap_n;
    L->cap_segs = cap_segs ? cap_segs : 1;
    L->hdr = 0;
    L->meta = align_up(sizeof(lsb::SegHeader));
    L->perm = L->meta + align_up((size_t)L->cap_segs * sizeof(lsb::SegLarge));
    L->counts = L->plan = L->tickets = L->status = L->tileseg = 0;
    return make_msd_layout(cap_n, L);
  }
  uint32_t d = c.digit_bits ? c.digit_bits : 8;
  uint32_t want_tile = c.tile;
  if (!c.digit_bits && !c.tile && typical_n) {
    // follow the tuned plain-sort choice for an array of typical_n elements
    const uint32_t v = auto_variant(typical_n, L->key_bits);
    if (v) {
      int nsv = 0;
      const lsb::SortVariant* sv = lsb::sort_variants(&nsv);
      if (v <= (uint32_t)nsv) {
        d = sv[v - 1].digit_bits;
        want_tile = sv[v - 1].tile;
      }
    }
  }
  int nv = 0;
  const lsb::SegVariant* vs = lsb::seg_variants(&nv);
  L->v = nullptr;
  for (int i = 0; i < nv && want_tile; ++i)
    if (vs[i].digit_bits == d && vs[i].tile == want_tile) {
      L->v = &vs[i];
      break;
    }
  for (int i = 0; i < nv && !L->v; ++i)
    if (vs[i].digit_bits == d && (c.tile == 0 || vs[i].tile == c.tile)) {
      L->v = &vs[i];
      break;
    }
  if (!L->v) return LSB_EINVAL;
  L->d = d;
  L->tile = L->v->tile;
  L->passes = (L->key_bits + d - 1) / d;
  L->maxp = (64 + d - 1) / d;
  L->nb = 1u << d;
  L->cap_n = cap_n;
  L->cap_segs = cap_segs ? cap_segs : 1;
  const uint64_t tiles = cap_n / L->tile + L->cap_segs;
  if (tiles > 0x7FFFFFFFull) return LSB_ERANGE;
  L->max_tiles = (uint32_t)tiles;
  size_t off = 0;
  L->hdr = off;     off += align_up(sizeof(lsb::SegHeader));
  L->meta = off;    off += align_up((size_t)L->cap_segs * sizeof(lsb::SegLarge));
  L->counts = off;  off += align_up((size_t)L->cap_segs * L->maxp * L->nb * 4);
  L->plan = off;    off += align_up((size_t)L->cap_segs * 8);
  L->tickets = off; off += align_up(8 * lsb::kRegions * 4);
  L->status = off;  off += align_up(2 * (size_t)L->max_tiles * L->nb * 4);
  L->tileseg = off; off += align_up((size_t)L->max_tiles * 4);
  L->perm = off;    off += align_up((size_t)L->max_tiles * 4);
  L->total = off;
  return LSB_OK;
}

// Library-managed segmented workspaces, one per (device, stream), with the
// capacity their layout was made for, and a pinned staging buffer for the
// per-call batch shape.
struct SegWorkspace {
  std::mutex mu;  // held by a batched sort from staging until its launches are enqueued
  void* ptr = nullptr;
  size_t bytes = 0, cap_n = 0;
  uint32_t cap_segs = 0;
  void* stage = nullptr;
  size_t stage_bytes = 0;
  hipEvent_t staged = nullptr;
};
std::map<std::pair<int, void*>, SegWorkspace> g_seg_ws;

uint32_t device_cus() {
  static const int cus = [] {
    int dev = 0, c = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        c <= 0) {
      (void)hipGetLastError();
      c = 256;
    }
    return c;
  }();
  return (uint32_t)cus;
}

// The kernels reach every segment through the table in the workspace, so
// the sequence depends on the workspace and its layout only.
int enqueue_segmented(const SegLayout& L, char* ws, uint32_t flags, hipStream_t s) {
  const lsb::SegHeader* hdr = reinterpret_cast<const lsb::SegHeader*>(ws + L.hdr);
  const lsb::SegLarge* large = reinterpret_cast<const lsb::SegLarge*>(ws + L.meta);
  const lsb::SegSmall* small = reinterpret_cast<const lsb::SegSmall*>(ws + L.meta);
  uint32_t* counts = reinterpret_cast<uint32_t*>(ws + L.counts);
  uint32_t* plan = reinterpret_cast<uint32_t*>(ws + L.plan);
  uint32_t* tickets = reinterpret_cast<uint32_t*>(ws + L.tickets);
  uint32_t* status = reinterpret_cast<uint32_t*>(ws + L.status);
  uint32_t* tileseg = reinterpret_cast<uint32_t*>(ws + L.tileseg);
  uint32_t* perm = reinterpret_cast<uint32_t*>(ws + L.perm);
  const uint32_t cus = device_cus();
  if (L.msd) {
    lsb::MsdArgs m;
    m.hdr = hdr;
    m.large = large;
    for (int i = 0; i < 2; ++i) {
      m.ent[i] = reinterpret_cast<lsb::MsdEntry*>(ws + L.ment[i]);
      m.counts[i] = reinterpret_cast<uint32_t*>(ws + L.mcounts[i]);
      m.orr[i] = reinterpret_cast<uint32_t*>(ws + L.morr[i]);
      m.tilemap[i] = reinterpret_cast<uint32_t*>(ws + L.mtile[i]);
    }
    m.status = reinterpret_cast<uint32_t*>(ws + L.mstatus);
    m.lvl = reinterpret_cast<lsb::MsdLevel*>(ws + L.mlvl);
    m.runs = reinterpret_cast<lsb::MsdRun*>(ws + L.mruns);
    m.run_ctr = reinterpret_cast<uint32_t*>(ws + L.mrunctr);
    m.rg = reinterpret_cast<uint32_t*>(ws + L.mrg);
    m.cap_entries = L.cap_entries;
    m.cap_tiles = L.max_tiles;
    m.cap_runs = L.cap_runs;
    m.went = reinterpret_cast<lsb::MsdEntry*>(ws + L.went);
    m.wcounts = reinterpret_cast<uint32_t*>(ws + L.wcounts);
    m.worr = reinterpret_cast<uint32_t*>(ws + L.worr);
    m.wtile = reinterpret_cast<uint32_t*>(ws + L.wtile);
    m.wstatus = reinterpret_cast<uint32_t*>(ws + L.wstatus);
    m.wlvl = reinterpret_cast<lsb::MsdLevel*>(ws + L.wlvl);
    m.cap_wide = L.cap_wide;
    m.cap_wide_tiles = L.cap_wide_tiles;
    m.key_bits = L.key_bits;
    m.ko = L.ko;
    m.level = 0;
    // levels 0 .. P: the last one only copies ranges left in B
    const uint32_t levels = std::min<uint32_t>(L.passes + 1, (uint32_t)lsb::kMsdMaxLevels);
    const uint32_t steps = msd_wide_steps(levels);
    hipEvent_t* ev = nullptr;  // LSB_PROFILE: one event after each launch
    if (flags & LSB_PROFILE) {
      g_prof.valid = g_mprof.valid = false;
      const int rc = msd_prof_events(msd_prof_event_count(levels), &ev);
      if (rc) return rc;
      LSB_HIP(hipEventRecord(ev[0], s));
    }
    LSB_HIP(lsb::launch_seg_small(hdr, small, L.key_bits, L.ko, 4 * cus, s));
    LSB_HIP(lsb::launch_msd_init(m, 4 * cus, s));
    if (ev) LSB_HIP(hipEventRecord(ev[1], s));
    size_t ei = 2;  // next event: three per wide step or level, in launch order
    const auto next3 = [&]() {
      hipEvent_t* e = ev ? ev + ei : nullptr;
      ei += 3;
      return e;
    };
    m.wstep = 0;
    LSB_HIP(lsb::launch_msd_wide(m, cus, s, next3()));
    // after level k the wide step k+1 for its children of the wide size range
    for (uint32_t k = 0; k < levels; ++k) {
      m.level = k;
      LSB_HIP(lsb::launch_msd_level(m, cus, s, next3()));
      if (k + 1 < steps) {
        m.wstep = k + 1;
        LSB_HIP(lsb::launch_msd_wide(m, cus, s, next3()));
      }
    }
    LSB_HIP(lsb::launch_msd_finish(m, cus, s));
    if (ev) {
      LSB_HIP(hipEventRecord(ev[ei], s));
      g_mprof.levels = levels;
      g_mprof.lvl = m.lvl;
      g_mprof.wlvl = m.wlvl;
      g_mprof.runs = m.runs;
      g_mprof.run_ctr = m.run_ctr;
      g_mprof.cap_runs = m.cap_runs;
      g_mprof.valid = true;
    }
    return LSB_OK;
  }
  LSB_HIP(lsb::launch_seg_prepare(L.d, hdr, large, L.tile, L.key_bits, L.ko, L.passes, counts,
                                  plan, tickets, status, tileseg, perm, flags, 4 * cus, s));
  LSB_HIP(lsb::launch_seg_small(hdr, small, L.key_bits, L.ko, 4 * cus, s));
  // whole groups of blocks (kRegions of them, each taking its own tickets)
  uint32_t grid = L.v->blocks_per_cu() * cus;
  grid = grid < (uint32_t)lsb::kRegions ? (uint32_t)lsb::kRegions : grid - grid % lsb::kRegions;
  for (uint32_t q = 0; q < L.passes; ++q) {
    lsb::SegPassArgs a;
    a.hdr = hdr;
    a.large = large;
    a.seg_plan = plan;
    a.tile_seg = tileseg;
    a.perm = perm;
    a.counts = counts;
    a.tickets = tickets;
    a.status = status;
    a.max_tiles = L.max_tiles;
    a.pass = q;
    a.shift = q * L.d;
    const uint32_t w = L.key_bits - a.shift < L.d ? L.key_bits - a.shift : L.d;
    a.mask = (1u << w) - 1u;
    a.ko = L.ko;
    LSB_HIP(L.v->scatter(a, grid, s));
  }
  LSB_HIP(lsb::launch_seg_copy(hdr, large, plan, tileseg, L.tile, 4 * cus, s));
  return LSB_OK;
}

// One segment of a batch: sorted in place in a, with b as its scratch.
struct SegEntry {
  lsb_elem* a;
  lsb_elem* b;
  uint64_t n;
};

int sort_entries(const std::vector<SegEntry>& ent, uint64_t total_n, const lsb_config& c);

}  // namespace

extern "C" {

void lsb_config_init(lsb_config* cfg) {
  if (cfg) *cfg = defaults();
}

const char* lsb_version(void) { return "lsb-mi355x 0.1.0 (gfx950)"; }

const char* lsb_strerror(int code) {
  switch (code) {
    case LSB_OK: return "success";
    case LSB_EINVAL: return "invalid argument";
    case LSB_ENOMEM: return "out of device memory or workspace too small";
    case LSB_EHIP: return "HIP runtime error";
    case LSB_ERANGE: return "n exceeds LSB_MAX_N";
    case LSB_EBUSY: return "queue or memory budget full";
    case LSB_ENOTFOUND: return "unknown job";
    default: return "unknown error";
  }
}

int lsb_last_hip_error(void) { return g_last_hip; }

int lsb_variants(uint32_t* digit_bits, uint32_t* tiles, uint32_t* persistent, int cap) {
  int nv = 0;
  const lsb::SortVariant* vs = lsb::sort_variants(&nv);
  for (int i = 0; i < nv && i < cap; ++i) {
    if (digit_bits) digit_bits[i] = vs[i].digit_bits;
    if (tiles) tiles[i] = vs[i].tile;
    if (persistent) persistent[i] = vs[i].persistent;
  }
  return nv;
}

int lsb_plan(size_t n, const lsb_config* cfg, uint32_t* digit_bits, uint32_t* tile,
             uint32_t* passes) {
  const lsb_config c = cfg ? *cfg : defaults();
  if (n > LSB_MAX_N) return LSB_ERANGE;
  Plan p;
  const int rc = make_plan(n, c, &p);
  if (rc) return rc;
  if (digit_bits) *digit_bits = p.d;
  if (tile) *tile = p.tile;
  if (passes) *passes = p.passes;
  return LSB_OK;
}

size_t lsb_workspace_size(size_t n, const lsb_config* cfg) {
  const lsb_config c = cfg ? *cfg : defaults();
  if (n <= 1 || n > LSB_MAX_N) return 0;
  Plan p;
  if (make_plan(n, c, &p)) return 0;
  if (plain_msd(n, c)) {
    lsb_config m = c;
    m.flags |= LSB_FORM_MSD;
    SegLayout L;
    if (make_seg_layout(n, 1, m, &L)) return 0;
    return L.total;
  }
  return p.total;
}

int lsb_sort_form(size_t n, const lsb_config* cfg) {
  const lsb_config c = cfg ? *cfg : defaults();
  if (n > LSB_MAX_N) return LSB_ERANGE;
  Plan p;
  const int rc = make_plan(n > 1 ? n : 2, c, &p);
  if (rc) return rc;
  return plain_msd(n, c) ? (int)LSB_FORM_MSD : (int)LSB_FORM_LSD;
}

int lsb_sort(lsb_elem* A, lsb_elem* B, size_t n, const lsb_config* cfg) {
  const lsb_config c = cfg ? *cfg : defaults();
  int rc = check_host_args(A, B, n);
  if (rc) return rc;
  Plan p;
  rc = make_plan(n > 1 ? n : 2, c, &p);
  if (rc) return rc;
  if (n <= 1) return LSB_OK;
  const size_t bytes = n * sizeof(lsb_elem);
  if ((rc = check_device_range(A, bytes)) || (rc = check_device_range(B, bytes))) return rc;
  if ((c.flags & LSB_GRAPH) && (c.flags & LSB_PROFILE)) return LSB_EINVAL;
  if (plain_msd(n, c)) {
    // one segment of the batched sort's MSD-first form
    lsb_config m = c;
    m.flags |= LSB_FORM_MSD;
    return sort_entries({SegEntry{A, B, n}}, n, m);
  }
  hipStream_t s = (hipStream_t)c.stream;
  // Arrays that fit one workgroup are sorted in LDS by a single launch
  // (auto configuration only; an explicit digit width or tile, or
  // profiling, runs the pass pipeline).
  if (c.digit_bits == 0 && c.tile == 0 && c.variant == 0 && !(c.flags & LSB_PROFILE) &&
      n <= lsb::small_sort_capacity()) {
    LSB_HIP(lsb::launch_small_sort(A, (uint32_t)n, p.key_bits, p.ko, s));
    if (c.flags & LSB_SYNC) LSB_HIP(hipStreamSynchronize(s));
    return LSB_OK;
  }
  void* ws = c.workspace;
  std::unique_lock<std::mutex> hold;  // the library workspace, until enqueued
  if (ws) {
    if (c.workspace_bytes < p.total || ((uintptr_t)ws & 255u)) return LSB_ENOMEM;
  } else if ((rc = get_workspace(c.stream, p.total, &ws, &hold))) {
    return rc;
  }
  if (c.flags & LSB_GRAPH) {
    if (c.flags & LSB_PROFILE) return LSB_EINVAL;
    int dev = 0;
    LSB_HIP(hipGetDevice(&dev));
    const std::vector<uint64_t> key = {0u, (uint64_t)(uintptr_t)A, (uint64_t)(uintptr_t)B, n,
                                       (uint64_t)(uintptr_t)ws, p.d, p.tile, p.v->persistent,
                                       p.key_bits, (uint64_t)(c.flags & ~LSB_SYNC),
                                       (uint64_t)dev};
    char* w = static_cast<char*>(ws);
    rc = launch_graph(key, s, [&](hipStream_t cs) {
      return enqueue_sort(A, B, (uint32_t)n, p, w, c.flags, cs);
    });
  } else {
    rc = enqueue_sort(A, B, (uint32_t)n, p, static_cast<char*>(ws), c.flags, s);
  }
  if (hold.owns_lock()) hold.unlock();
  if (rc) return rc;
  if (c.flags & LSB_SYNC) LSB_HIP(hipStreamSynchronize(s));
  return LSB_OK;
}


// The largest layout over the compiled segmented variants of the digit
// width (and tile, when the config names one): the tuned choice for a batch
// depends on its segment sizes, and any of them must fit.
size_t lsb_segmented_workspace_size(size_t total_n, uint32_t nseg, const lsb_config* cfg) {
  const lsb_config c = cfg ? *cfg : defaults();
  SegLayout L;
  if (make_seg_layout(total_n, nseg, c, &L)) return 0;
  size_t most = L.total;
  int nv = 0;
  const lsb::SegVariant* vs = lsb::seg_variants(&nv);
  for (int i = 0; i < nv; ++i) {
    if (vs[i].digit_bits != L.d || (c.tile && vs[i].tile != c.tile)) continue;
    lsb_config v = c;
    v.digit_bits = vs[i].digit_bits;
    v.tile = vs[i].tile;
    SegLayout M;
    if (make_seg_layout(total_n, nseg, v, &M) == LSB_OK) most = std::max(most, M.total);
  }
  return most;
}

}  // extern "C"

namespace {

// Sorts every entry on its own (shared body of lsb_sort_segmented and
// lsb_sort_batch).  total_n = sum of the entries' sizes.
int sort_entries(const std::vector<SegEntry>& ent, uint64_t total_n, const lsb_config& c) {
  const uint32_t nseg = (uint32_t)ent.size();
  uint32_t n_large = 0, n_tiny = 0, n_small = 0;
  // The tile pipeline runs the tiles of every large segment in one
  // persistent launch per pass, so its variant follows the tuned choice for
  // one array of all their elements together (typical_n).
  size_t typical_n = 0;
  for (const SegEntry& e : ent) {
    if (e.n <= 1) continue;
    if (e.n <= lsb::seg_tiny_capacity()) ++n_tiny;
    else if (e.n <= lsb::seg_small_capacity()) ++n_small;
    else { ++n_large; typical_n += e.n; }
  }
  if (n_large + n_tiny + n_small == 0) return LSB_OK;
  int rc;
  SegLayout L;
  if ((rc = make_seg_layout(total_n, nseg, c, &L, typical_n))) return rc;
  hipStream_t s = (hipStream_t)c.stream;
  int dev = 0;
  LSB_HIP(hipGetDevice(&dev));

  SegWorkspace& W = cache_entry(g_seg_ws, dev, c.stream);
  std::unique_lock<std::mutex> hold(W.mu);
  char* ws = static_cast<char*>(c.workspace);
  if (ws) {
    // A caller workspace sized by lsb_segmented_workspace_size holds the
    // largest layout of the digit width; if the tuned tile needs more than
    // the caller gave, the default tile's layout is used instead.
    if (L.total > c.workspace_bytes && (rc = make_seg_layout(total_n, nseg, c, &L))) return rc;
    if (c.workspace_bytes < L.total || ((uintptr_t)ws & 255u)) return LSB_ENOMEM;
  } else {
    // The layout of a library workspace follows its capacity, not the batch,
    // so a captured graph stays valid for every batch that fits.  It is
    // regrown when the batch exceeds the capacity or the digit width / tile
    // asks for a larger layout.
    const size_t cap_n = std::max<size_t>(total_n, W.cap_n);
    const uint32_t cap_s = std::max<uint32_t>(nseg, W.cap_segs);
    if ((rc = make_seg_layout(cap_n, cap_s, c, &L, typical_n))) return rc;
    if (!W.ptr || cap_n > W.cap_n || cap_s > W.cap_segs || L.total > W.bytes) {
      const size_t grow_n = cap_n > W.cap_n ? std::max<size_t>(cap_n, W.cap_n + W.cap_n / 4) : cap_n;
      const uint32_t grow_s =
          cap_s > W.cap_segs ? std::max<uint32_t>(cap_s, W.cap_segs + W.cap_segs / 4) : cap_s;
      if ((rc = make_seg_layout(grow_n, grow_s, c, &L, typical_n))) return rc;
      // Stream-ordered, like get_workspace: no wait for earlier batches.
      void* p = nullptr;
      const hipError_t he = lsb::pool_alloc(&p, L.total, s);
      if (he != hipSuccess) {
        g_last_hip = (int)he;
        (void)hipGetLastError();
        return LSB_ENOMEM;
      }
      if (W.ptr) LSB_HIP(hipFreeAsync(W.ptr, s));
      W.ptr = p;
      W.bytes = L.total;
      W.cap_n = grow_n;
      W.cap_segs = grow_s;
    }
    ws = static_cast<char*>(W.ptr);
  }

  // Stage the batch's segment table in pinned memory and upload it on the
  // stream (outside any captured graph: replays read the new table).
  const size_t meta_bytes = L.meta - L.hdr + (size_t)n_large * sizeof(lsb::SegLarge) +
                            (size_t)(n_tiny + n_small) * sizeof(lsb::SegSmall);
  if (W.stage_bytes < meta_bytes) {
    if (W.stage) {
      if (W.staged) LSB_HIP(hipEventSynchronize(W.staged));
      LSB_HIP(hipHostFree(W.stage));
      W.stage = nullptr;
    }
    LSB_HIP(hipHostMalloc(&W.stage, meta_bytes * 2, hipHostMallocDefault));
    W.stage_bytes = meta_bytes * 2;
  }
  if (!W.staged) LSB_HIP(hipEventCreateWithFlags(&W.staged, hipEventDisableTiming));
  LSB_HIP(hipEventSynchronize(W.staged));  // the previous upload has left the buffer
  char* h = static_cast<char*>(W.stage);
  lsb::SegHeader* hdr = reinterpret_cast<lsb::SegHeader*>(h);
  lsb::SegLarge* lg = reinterpret_cast<lsb::SegLarge*>(h + (L.meta - L.hdr));
  lsb::SegSmall* tiny = reinterpret_cast<lsb::SegSmall*>(lg + n_large);
  lsb::SegSmall* sml = tiny + n_tiny;
  uint32_t il = 0, it = 0, is = 0, tiles = 0;
  for (const SegEntry& e : ent) {
    if (e.n <= 1) continue;
    const uint64_t a = (uint64_t)(uintptr_t)e.a, b = (uint64_t)(uintptr_t)e.b;
    if (e.n <= lsb::seg_tiny_capacity()) {
      tiny[it++] = lsb::SegSmall{a, (uint32_t)e.n, 0u};
    } else if (e.n <= lsb::seg_small_capacity()) {
      sml[is++] = lsb::SegSmall{a, (uint32_t)e.n, 0u};
    } else {
      lg[il++] = lsb::SegLarge{a, b, (uint32_t)e.n, tiles, 0u, lsb::kNotWide};
      tiles += (uint32_t)((e.n + L.tile - 1) / L.tile);
    }
  }
  hdr->n_large = n_large;
  hdr->n_tiny = n_tiny;
  hdr->n_small = n_small;
  hdr->total_tiles = tiles;
  hdr->n_wide = hdr->wide_tiles = 0;
  if (tiles > L.max_tiles) return LSB_ENOMEM;
  if (L.msd) {
    // Segments whose 8-bit buckets would need a second partition are first
    // partitioned on a kMsdWideBits-bit digit (their buckets then finish in LDS);
    // they get compact tile ranges of their own for that partition.
    for (uint32_t i = 0; i < n_large; ++i) {
      if (L.key_bits >= lsb::kMsdWideKeyBits && lg[i].n > lsb::kMsdWideMin &&
          lg[i].n <= lsb::kMsdWideMax) {
        lg[i].wide = hdr->n_wide++;
        lg[i].perm_pos = hdr->wide_tiles;
        hdr->wide_tiles += (lg[i].n + lsb::kMsdTile - 1) / lsb::kMsdTile;
      }
    }
    if (hdr->n_wide > L.cap_wide || hdr->wide_tiles > L.cap_wide_tiles) return LSB_ENOMEM;
  } else {
  // The scatter passes give each group of blocks whole segments: largest
  // first onto the least-loaded group; inside a group the segments keep
  // their order.  When that leaves one group with much more than its share
  // (few segments, or one dominant one), all blocks share one group instead,
  // as a group's blocks cannot help with another group's segments.
  {
    const auto ntl = [&](uint32_t i) { return (uint32_t)((lg[i].n + L.tile - 1) / L.tile); };
    std::vector<uint32_t> order(n_large), group(n_large);
    for (uint32_t i = 0; i < n_large; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t x, uint32_t y) { return ntl(x) > ntl(y); });
    uint32_t load[lsb::kRegions] = {};
    for (uint32_t i : order) {
      int g = 0;
      for (int k = 1; k < lsb::kRegions; ++k)
        if (load[k] < load[g]) g = k;
      group[i] = (uint32_t)g;
      load[g] += ntl(i);
    }
    uint32_t most = 0;
    for (int g = 0; g < lsb::kRegions; ++g) most = std::max(most, load[g]);
    // grouped only while the busiest group stays within 1/16 of an even split
    const bool grouped = (uint64_t)most * lsb::kRegions * 16 <= (uint64_t)tiles * 17;
    for (int g = 0; g < lsb::kRegions; ++g) hdr->group_off[g] = hdr->group_tiles[g] = 0;
    if (grouped) {
      hdr->groups = lsb::kRegions;
      uint32_t next[lsb::kRegions];
      uint32_t off = 0;
      for (int g = 0; g < lsb::kRegions; ++g) {
        hdr->group_off[g] = next[g] = off;
        hdr->group_tiles[g] = load[g];
        off += load[g];
      }
      for (uint32_t i = 0; i < n_large; ++i) {
        lg[i].perm_pos = next[group[i]];
        next[group[i]] += ntl(i);
      }
    } else {
      hdr->groups = 1;
      hdr->group_tiles[0] = tiles;
      for (uint32_t i = 0; i < n_large; ++i) lg[i].perm_pos = lg[i].tile_begin;
    }
  }
  }
  LSB_HIP(hipMemcpyAsync(ws + L.hdr, h, meta_bytes, hipMemcpyHostToDevice, s));
  LSB_HIP(hipEventRecord(W.staged, s));

  if (c.flags & LSB_GRAPH) {
    const std::vector<uint64_t> key = {1u, (uint64_t)(uintptr_t)ws, L.cap_n, L.cap_segs, L.d,
                                       L.tile, L.key_bits, (uint64_t)(c.flags & ~LSB_SYNC),
                                       (uint64_t)dev, (uint64_t)L.msd};
    rc = launch_graph(key, s, [&](hipStream_t cs) {
      return enqueue_segmented(L, ws, c.flags, cs);
    });
  } else {
    rc = enqueue_segmented(L, ws, c.flags, s);
  }
  if (rc) return rc;
  if (c.flags & LSB_SYNC) {
    LSB_HIP(hipStreamSynchronize(s));
    if (L.msd) {
      // the MSD lists stayed within their capacities (a proven bound; a
      // violation would mean wrong results, so it is reported)
      // (on the caller's stream: a copy on the legacy stream would be an
      // implicit dependency for any graph capture another thread has open)
      uint32_t overflow = 0;
      LSB_HIP(hipMemcpyAsync(&overflow, ws + L.mrunctr + 8, 4, hipMemcpyDeviceToHost, s));
      LSB_HIP(hipStreamSynchronize(s));
      if (overflow) return LSB_ENOMEM;
    }
  }
  return LSB_OK;
}

}  // namespace

extern "C" {

int lsb_sort_segmented(lsb_elem* A, lsb_elem* B, const uint64_t* seg_offsets, uint32_t nseg,
                       const lsb_config* cfg) {
  const lsb_config c = cfg ? *cfg : defaults();
  if (nseg == 0) return LSB_OK;
  if (!seg_offsets) return LSB_EINVAL;
  if (c.flags & LSB_PROFILE) return LSB_EINVAL;
  if ((c.key_bits ? c.key_bits : 64) > 64) return LSB_EINVAL;
  const uint64_t total_n = seg_offsets[nseg];
  bool any = false;
  for (uint32_t i = 0; i < nseg; ++i) {
    if (seg_offsets[i + 1] < seg_offsets[i]) return LSB_EINVAL;
    const uint64_t n = seg_offsets[i + 1] - seg_offsets[i];
    if (n > LSB_MAX_N) return LSB_ERANGE;
    any |= n > 1;
  }
  if (!any) return LSB_OK;
  if (!A || !B || (((uintptr_t)A | (uintptr_t)B) & 15u)) return LSB_EINVAL;
  if (overlaps(A, B, total_n * sizeof(lsb_elem))) return LSB_EINVAL;
  int rc;
  const size_t bytes = total_n * sizeof(lsb_elem);
  if ((rc = check_device_range(A, bytes)) || (rc = check_device_range(B, bytes))) return rc;
  std::vector<SegEntry> ent(nseg);
  for (uint32_t i = 0; i < nseg; ++i)
    ent[i] = SegEntry{A + seg_offsets[i], B + seg_offsets[i], seg_offsets[i + 1] - seg_offsets[i]};
  return sort_entries(ent, total_n, c);
}

int lsb_sort_batch(lsb_elem* const* As, lsb_elem* const* Bs, const uint64_t* ns, uint32_t nseg,
                   const lsb_config* cfg) {
  const lsb_config c = cfg ? *cfg : defaults();
  if (nseg == 0) return LSB_OK;
  if (!As || !ns) return LSB_EINVAL;
  if (c.flags & LSB_PROFILE) return LSB_EINVAL;
  if ((c.key_bits ? c.key_bits : 64) > 64) return LSB_EINVAL;
  uint64_t total_n = 0;
  std::vector<SegEntry> ent(nseg);
  for (uint32_t i = 0; i < nseg; ++i) {
    const uint64_t n = ns[i];
    if (n > LSB_MAX_N) return LSB_ERANGE;
    lsb_elem* b = Bs ? Bs[i] : nullptr;
    if (n > 1 && (!As[i] || ((uintptr_t)As[i] & 15u))) return LSB_EINVAL;
    if (n > lsb::seg_small_capacity()) {
      if (!b || ((uintptr_t)b & 15u)) return LSB_EINVAL;
      if (As[i] < b + n && b < As[i] + n) return LSB_EINVAL;  // A and its scratch overlap
    }
    ent[i] = SegEntry{As[i], b, n};
    total_n += n;
  }
  return sort_entries(ent, total_n, c);
}

int lsb_release(void* stream) {
  int dev = 0;
  LSB_HIP(hipGetDevice(&dev));
  const hipStream_t s = (hipStream_t)stream;
  LSB_HIP(hipStreamSynchronize(s));  // nothing queued on it may still use them
  std::vector<uint64_t> freed;
  int rc = LSB_OK;
  auto note = [&](hipError_t e) {
    if (e != hipSuccess && rc == LSB_OK) {
      g_last_hip = (int)e;
      rc = LSB_EHIP;
    }
  };
  // Entries stay in the maps, emptied, so no other thread can hold a
  // reference to a destroyed one.
  {
    Workspace& w = cache_entry(g_ws, dev, stream);
    std::lock_guard<std::mutex> lock(w.mu);
    if (w.ptr) {
      freed.push_back((uint64_t)(uintptr_t)w.ptr);
      note(hipFreeAsync(w.ptr, s));
    }
    w.ptr = nullptr;
    w.bytes = 0;
  }
  {
    SegWorkspace& W = cache_entry(g_seg_ws, dev, stream);
    std::lock_guard<std::mutex> lock(W.mu);
    if (W.ptr) {
      freed.push_back((uint64_t)(uintptr_t)W.ptr);
      note(hipFreeAsync(W.ptr, s));
    }
    if (W.staged) note(hipEventDestroy(W.staged));
    if (W.stage) note(hipHostFree(W.stage));
    W.ptr = nullptr;
    W.bytes = W.cap_n = 0;
    W.cap_segs = 0;
    W.stage = nullptr;
    W.stage_bytes = 0;
    W.staged = nullptr;
  }
  if (!freed.empty()) {
    // Graphs captured against a freed workspace (its address is part of
    // the key: entry 4 of a plain sort's key, entry 1 of a segmented one's)
    // were only ever launched on this stream.
    std::lock_guard<std::mutex> glock(g_graph_mu);
    for (auto it = g_graphs.begin(); it != g_graphs.end();) {
      const std::vector<uint64_t>& k = it->first;
      const uint64_t ws = k[0] == 0u ? k[4] : k[1];
      if (std::find(freed.begin(), freed.end(), ws) != freed.end()) {
        note(hipGraphExecDestroy(it->second));
        it = g_graphs.erase(it);
      } else {
        ++it;
      }
    }
  }
  note(hipStreamSynchronize(s));
  note(lsb::pool_trim(dev));
  if (rc != LSB_OK) (void)hipGetLastError();
  return rc;
}

// ------------------------------------------------------------- autotuner

int lsb_autotune(uint32_t lo_log2, uint32_t hi_log2, uint32_t key_bits, void* stream) {
  if (lo_log2 < 10 || hi_log2 < lo_log2 || hi_log2 > 30 || key_bits > 64) return LSB_EINVAL;
  const uint32_t kb = key_bits ? key_bits : 64;
  const size_t nmax = (size_t)1 << hi_log2;
  hipStream_t s = (hipStream_t)stream;
  lsb_elem *pristine = nullptr, *A = nullptr, *B = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = LSB_OK;
  auto cleanup = [&]() {
    if (pristine) (void)hipFree(pristine);
    if (A) (void)hipFree(A);
    if (B) (void)hipFree(B);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipGetLastError();
  };
  hipError_t he = hipMalloc(&pristine, nmax * sizeof(lsb_elem));
  if (he == hipSuccess) he = hipMalloc(&A, nmax * sizeof(lsb_elem));
  if (he == hipSuccess) he = hipMalloc(&B, nmax * sizeof(lsb_elem));
  if (he != hipSuccess) {
    g_last_hip = (int)he;
    cleanup();
    return LSB_ENOMEM;
  }
  he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    g_last_hip = (int)he;
    cleanup();
    return LSB_EHIP;
  }
  lsb_gen_spec g;
  memset(&g, 0, sizeof g);
  g.kind = LSB_GEN_UNIFORM;
  g.ranks = 64;
  g.key_bits = kb;
  g.seed_base = 0x5eed;
  if ((rc = lsb_generate(pristine, nmax, &g, stream))) {
    cleanup();
    return rc;
  }
  int nv = 0;
  const lsb::SortVariant* vs = lsb::sort_variants(&nv);
  std::vector<lsb_tune_entry> table;
  for (uint32_t lg = lo_log2; lg <= hi_log2 && rc == LSB_OK; ++lg) {
    const size_t n = (size_t)1 << lg;
    lsb_tune_entry best;
    memset(&best, 0, sizeof best);
    best.ms = 1e30f;
    best.runner_up_ms = 1e30f;
    for (int v = 0; v < nv && rc == LSB_OK; ++v) {
      lsb_config c = defaults();
      c.key_bits = kb;
      c.variant = (uint32_t)v + 1;
      c.stream = stream;
      std::vector<float> ts;
      for (int rep = 0; rep < 4 && rc == LSB_OK; ++rep) {  // first run warms up
        if (hipMemcpyAsync(A, pristine, n * sizeof(lsb_elem), hipMemcpyDeviceToDevice, s) !=
                hipSuccess ||
            hipEventRecord(e0, s) != hipSuccess) {
          rc = LSB_EHIP;
          break;
        }
        if ((rc = lsb_sort(A, B, n, &c))) break;
        float ms = 0;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = LSB_EHIP;
          break;
        }
        if (rep) ts.push_back(ms);
      }
      if (rc) break;
      std::sort(ts.begin(), ts.end());
      const float med = ts[ts.size() / 2];
      if (med < best.ms) {
        best.runner_up_ms = best.ms;
        best.ms = med;
        best.variant = (uint32_t)v + 1;
        best.digit_bits = vs[v].digit_bits;
        best.tile = vs[v].tile;
        best.persistent = vs[v].persistent;
      } else if (med < best.runner_up_ms) {
        best.runner_up_ms = med;
      }
    }
    best.log2n = lg;
    best.key_bits = kb;
    table.push_back(best);
  }
  cleanup();
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(g_tune_mu);
  // replace entries of the same key-width class and sizes
  const bool narrow = kb <= 32;
  std::vector<lsb_tune_entry> keep;
  for (const lsb_tune_entry& e : g_tune)
    if ((e.key_bits <= 32) != narrow || e.log2n < lo_log2 || e.log2n > hi_log2) keep.push_back(e);
  for (const lsb_tune_entry& e : table) keep.push_back(e);
  g_tune.swap(keep);
  return LSB_OK;
}

int lsb_tune_table(lsb_tune_entry* out, int cap) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  for (int i = 0; i < (int)g_tune.size() && i < cap; ++i) out[i] = g_tune[i];
  return (int)g_tune.size();
}

int lsb_tune_clear(void) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune.clear();
  return LSB_OK;
}

int lsb_tune_save(const char* path) {
  if (!path) return LSB_EINVAL;
  std::lock_guard<std::mutex> lock(g_tune_mu);
  FILE* f = fopen(path, "w");
  if (!f) return LSB_EINVAL;
  fprintf(f, "# lsb tune table: log2n key_bits digit_bits tile persistent ms runner_up_ms\n");
  for (const lsb_tune_entry& e : g_tune)
    fprintf(f, "%u %u %u %u %u %.6f %.6f\n", e.log2n, e.key_bits, e.digit_bits, e.tile,
            e.persistent, e.ms, e.runner_up_ms);
  fclose(f);
  return LSB_OK;
}

// Entries name variants by (digit_bits, tile, persistent), so a table stays
// valid if the compiled variant list is reordered; unknown ones are dropped.
int lsb_tune_load(const char* path) {
  if (!path) return LSB_EINVAL;
  FILE* f = fopen(path, "r");
  if (!f) return LSB_EINVAL;
  int nv = 0;
  const lsb::SortVariant* vs = lsb::sort_variants(&nv);
  std::vector<lsb_tune_entry> t;
  char line[256];
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    lsb_tune_entry e;
    memset(&e, 0, sizeof e);
    if (sscanf(line, "%u %u %u %u %u %f %f", &e.log2n, &e.key_bits, &e.digit_bits, &e.tile,
               &e.persistent, &e.ms, &e.runner_up_ms) != 7)
      continue;
    for (int v = 0; v < nv; ++v)
      if (vs[v].digit_bits == e.digit_bits && vs[v].tile == e.tile &&
          vs[v].persistent == e.persistent)
        e.variant = (uint32_t)v + 1;
    if (e.variant) t.push_back(e);
  }
  fclose(f);
  std::lock_guard<std::mutex> lock(g_tune_mu);
  g_tune.swap(t);
  return LSB_OK;
}

int lsb_profile_read(lsb_profile* out) {
  if (!out) return LSB_EINVAL;
  memset(out, 0, sizeof *out);
  if (!g_prof.valid) return LSB_EINVAL;
  const uint32_t P = g_prof.passes;
  std::vector<hipEvent_t>& ev = g_prof.ev;
  LSB_HIP(hipEventSynchronize(ev[3 + P]));
  uint32_t plan[4 * 16 + 4];
  LSB_HIP(hipMemcpy(plan, g_prof.plan, (4 * P + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  out->passes = P;
  out->digit_bits = g_prof.d;
  out->tile = g_prof.tile;
  for (uint32_t q = 0; q < P; ++q)
    if (plan[4 * q]) out->executed_mask |= 1u << q;
  LSB_HIP(hipEventElapsedTime(&out->histogram_ms, ev[0], ev[1]));
  LSB_HIP(hipEventElapsedTime(&out->scan_ms, ev[1], ev[2]));
  for (uint32_t q = 0; q < P; ++q)
    LSB_HIP(hipEventElapsedTime(&out->pass_ms[q], ev[2 + q], ev[3 + q]));
  LSB_HIP(hipEventElapsedTime(&out->copy_ms, ev[2 + P], ev[3 + P]));
  LSB_HIP(hipEventElapsedTime(&out->total_ms, ev[0], ev[3 + P]));
  return LSB_OK;
}

static_assert(LSB_MSD_MAX_LEVELS == lsb::kMsdMaxLevels, "lsb_msd_profile level rows");
static_assert(LSB_MSD_WIDE_STEPS == lsb::kMsdWideSteps, "lsb_msd_profile wide-step rows");

int lsb_msd_profile_read(lsb_msd_profile* out) {
  if (!out) return LSB_EINVAL;
  memset(out, 0, sizeof *out);
  if (!g_mprof.valid) return LSB_EINVAL;
  const uint32_t K = g_mprof.levels, S = msd_wide_steps(K);
  hipEvent_t* ev = g_mprof.ev.data();
  const size_t last = msd_prof_event_count(K) - 1;
  LSB_HIP(hipEventSynchronize(ev[last]));
  lsb::MsdLevel lvl[lsb::kMsdMaxLevels], wlvl[lsb::kMsdWideSteps];
  uint32_t ctr[4];
  LSB_HIP(hipMemcpy(lvl, g_mprof.lvl, sizeof lvl, hipMemcpyDeviceToHost));
  LSB_HIP(hipMemcpy(wlvl, g_mprof.wlvl, sizeof wlvl, hipMemcpyDeviceToHost));
  LSB_HIP(hipMemcpy(ctr, g_mprof.run_ctr, sizeof ctr, hipMemcpyDeviceToHost));
  const uint32_t nr = std::min(ctr[0], g_mprof.cap_runs);
  std::vector<lsb::MsdRun> runs(nr);
  if (nr)
    LSB_HIP(hipMemcpy(runs.data(), g_mprof.runs, nr * sizeof(lsb::MsdRun),
                      hipMemcpyDeviceToHost));
  out->levels = K;
  out->runs = nr;
  for (const lsb::MsdRun& r : runs) out->finish_elems += r.n;
  out->wide_steps = S;
  for (uint32_t w = 0; w < S; ++w) out->wide_elems[w] = wlvl[w].moved;
  for (uint32_t k = 0; k < K; ++k) out->level_elems[k] = lvl[k].moved;
  const auto span = [&](size_t j, float* ms) {  // launch j: from event j-1 to event j
    return hipEventElapsedTime(ms, ev[j - 1], ev[j]);
  };
  LSB_HIP(span(1, &out->setup_ms));
  // the launch order of enqueue_segmented: wide step 0, then each level
  // followed by its wide step
  size_t j = 2;
  for (int i = 0; i < 3; ++i) LSB_HIP(span(j++, &out->wide_ms[0][i]));
  for (uint32_t k = 0; k < K; ++k) {
    for (int i = 0; i < 3; ++i) LSB_HIP(span(j++, &out->level_ms[k][i]));
    if (k + 1 < S)
      for (int i = 0; i < 3; ++i) LSB_HIP(span(j++, &out->wide_ms[k + 1][i]));
  }
  LSB_HIP(span(last, &out->finish_ms));
  LSB_HIP(hipEventElapsedTime(&out->total_ms, ev[0], ev[last]));
  return LSB_OK;
}

int lsb_debug_histogram(const lsb_elem* A, size_t n, const lsb_config* cfg, uint32_t* counts) {
  const lsb_config c = cfg ? *cfg : defaults();
  if (n > LSB_MAX_N) return LSB_ERANGE;
  if (!counts || (n && !A)) return LSB_EINVAL;
  Plan p;
  int rc = make_plan(n > 1 ? n : 2, c, &p);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)c.stream;
  // One region spanning the whole array: regcount[p][0][b] is then the
  // global count, the other regions stay empty.  The caller's buffer holds
  // P rows, so the kRegions-strided result goes through a scratch copy.
  uint32_t* tmp = nullptr;
  const size_t words = (size_t)p.passes * lsb::kRegions * p.nb;
  LSB_HIP(hipMallocAsync(&tmp, words * 4, s));
  lsb::launch_zero(tmp, words, s);
  if (n) LSB_HIP(lsb::launch_histogram(p.d, A, (uint32_t)n, p.key_bits, p.ko, p.passes, p.passes,
                                       tmp, (uint32_t)n, 1u, nullptr, 0, nullptr, s));
  for (uint32_t q = 0; q < p.passes; ++q)
    LSB_HIP(hipMemcpyAsync(counts + (size_t)q * p.nb, tmp + (size_t)q * lsb::kRegions * p.nb,
                           p.nb * 4, hipMemcpyDeviceToDevice, s));
  LSB_HIP(hipFreeAsync(tmp, s));
  LSB_HIP(hipGetLastError());
  if (c.flags & LSB_SYNC) LSB_HIP(hipStreamSynchronize(s));
  return LSB_OK;
}

// ------------------------------------------------------------- workloads

int lsb_generate(lsb_elem* A, size_t n, const lsb_gen_spec* spec, void* stream) {
  if (!spec) return LSB_EINVAL;
  if (n == 0) return LSB_OK;
  if (!A || ((uintptr_t)A & 15u)) return LSB_EINVAL;
  if (spec->kind > LSB_GEN_ZIPF || spec->key_bits > 64) return LSB_EINVAL;
  if (spec->kind == LSB_GEN_ZIPF && (spec->zipf_m == 0 || !(spec->zipf_s > 0.0)))
    return LSB_EINVAL;
  const uint32_t ranks = spec->ranks ? spec->ranks : 1;
  const uint64_t per = (n + ranks - 1) / ranks;
  const uint32_t kb = spec->key_bits ? spec->key_bits : 64;
  const uint64_t mask = kb >= 64 ? ~0ull : ((1ull << kb) - 1);
  hipStream_t s = (hipStream_t)stream;
  double* d_cdf = nullptr;
  uint64_t* d_table = nullptr;
  if (spec->kind == LSB_GEN_ZIPF) {
    const uint32_t m = spec->zipf_m;
    std::vector<double> cdf(m);
    std::vector<uint64_t> table(m);
    double acc = 0.0;
    for (uint32_t k = 0; k < m; ++k) {
      acc += pow((double)(k + 1), -spec->zipf_s);
      cdf[k] = acc;
    }
    for (uint32_t k = 0; k < m; ++k) cdf[k] /= acc;
    cdf[m - 1] = 1.0;
    lsb::u128 st = lsb::pcg_seed_state(spec->zipf_table_seed);
    for (uint32_t k = 0; k < m; ++k) {
      st = st * lsb::pcg_mul() + lsb::pcg_inc();
      table[k] = lsb::pcg_output(st);
    }
    LSB_HIP(hipMalloc(&d_cdf, m * sizeof(double)));
    LSB_HIP(hipMalloc(&d_table, m * sizeof(uint64_t)));
    LSB_HIP(hipMemcpyAsync(d_cdf, cdf.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
    LSB_HIP(hipMemcpyAsync(d_table, table.data(), m * sizeof(uint64_t), hipMemcpyHostToDevice, s));
  }
  hipError_t e = lsb::launch_generate(A, n, spec->kind, per, spec->seed_base, mask,
                                      spec->val_base, d_cdf, d_table, spec->zipf_m, s);
  if (d_cdf) {
    LSB_HIP(hipStreamSynchronize(s));
    LSB_HIP(hipFree(d_cdf));
    LSB_HIP(hipFree(d_table));
  }
  LSB_HIP(e);
  return LSB_OK;
}

int lsb_verify(const lsb_elem* out, const uint64_t* in_keys, size_t n, uint64_t val_base,
               uint32_t key_bits, uint32_t key_flags, uint64_t* bad_count, int64_t* first_bad,
               void* stream) {
  if (!bad_count || !first_bad) return LSB_EINVAL;
  *bad_count = 0;
  *first_bad = -1;
  if (key_bits > 64 || (key_flags & ~(LSB_KEYS_SIGNED | LSB_KEYS_FLOAT))) return LSB_EINVAL;
  const uint32_t kb = key_bits ? key_bits : 64;
  lsb::KeyOrder ko;
  if (make_key_order(key_flags, kb, &ko)) return LSB_EINVAL;
  if (n == 0) return LSB_OK;
  if (!out || !in_keys) return LSB_EINVAL;
  const uint64_t mask = kb >= 64 ? ~0ull : ((1ull << kb) - 1);
  hipStream_t s = (hipStream_t)stream;
  const size_t words = (n + 31) / 32;
  char* buf = nullptr;
  LSB_HIP(hipMalloc(&buf, align_up(words * 4) + 16));
  uint32_t* bitmap = reinterpret_cast<uint32_t*>(buf);
  unsigned long long* res = reinterpret_cast<unsigned long long*>(buf + align_up(words * 4));
  hipError_t e = lsb::launch_verify(out, in_keys, n, val_base, mask, ko, bitmap, res, s);
  unsigned long long host[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(host, res, sizeof host, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(buf);
  LSB_HIP(e);
  *bad_count = host[0];
  *first_bad = host[0] ? (int64_t)host[1] : -1;
  return LSB_OK;
}

int lsb_debug_scatter_floor(const lsb_elem* src, lsb_elem* dst, size_t n, uint32_t runs,
                            void* stream) {
  const uint32_t T = lsb::scatter_floor_tile();
  if (n == 0) return LSB_OK;
  if (!src || !dst || (((uintptr_t)src | (uintptr_t)dst) & 15u)) return LSB_EINVAL;
  if (n % T || n / T > 0xFFFFFFFFull) return LSB_EINVAL;
  if (runs == 0 || (runs & (runs - 1)) || runs > T) return LSB_EINVAL;
  if (overlaps(src, dst, n * sizeof(lsb_elem))) return LSB_EINVAL;
  LSB_HIP(lsb::launch_scatter_floor(src, dst, n, runs, device_cus(), (hipStream_t)stream));
  return LSB_OK;
}

int lsb_extract_keys(const lsb_elem* A, size_t n, uint64_t* keys, void* stream) {
  if (n == 0) return LSB_OK;
  if (!A || !keys) return LSB_EINVAL;
  LSB_HIP(lsb::launch_extract_keys(A, n, keys, (hipStream_t)stream));
  return LSB_OK;
}

}  // extern "C"
