// Serving loop: continuous batching of sort jobs resident in HBM.
//
// Clients submit jobs (device pointer A, n, key width); the server sorts A in
// place, like one rank's mySort (mpi/mpi_lsbsort.cpp:591-597) per job.  A
// scheduler thread
//   * admits jobs in arrival order against a memory budget that charges each
//     job its doubled buffers (A + B, 32 bytes per element);
//   * coalesces admitted small jobs of one key width and key order into a
//     batch sorted in place by one batched launch sequence (lsb_sort_batch,
//     replayed from a captured hipGraph) with a per-stream scratch slab --
//     no staging copies;
//   * launches larger jobs as plain sorts, with a free slab as scratch when
//     the job fits one (else scratch from the library's memory pool);
//   * spreads launches over several streams so jobs that arrive while others
//     run overlap on the device;
//   * records each job's latency (completion observed - submission) for
//     p50/p99.
// Nothing here has a counterpart in the reference, whose programs sort one
// array per run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string.h>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/lsb.h"
#include "lsb_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double now_s() {
  return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

struct Job {
  uint64_t id;
  lsb_elem* A;
  size_t n;
  uint32_t key_bits;
  uint32_t key_flags;  // LSB_KEYS_SIGNED / LSB_KEYS_FLOAT or 0
  double submitted, done_at = 0;
  int state = 0;  // 0 pending, 1 admitted/running, 2 done, <0 error code
  int rc = 0;
  bool waited = false;  // held at the head of the queue for budget memory
};

struct Launch {
  std::vector<uint64_t> jobs;
  hipEvent_t ev = nullptr;
  int stream = 0;
  void* scratch = nullptr;  // plain job's B (freed stream-ordered)
  uint64_t charge = 0;      // budget bytes released at completion
  bool batch = false;
};

struct StreamSlot {
  hipStream_t s = nullptr;
  lsb_elem* slab_b = nullptr;  // scratch of coalesced batches
  hipEvent_t slab_free = nullptr;  // recorded after the last job that used the slab
  int inflight = 0;
};

}  // namespace

// Finished jobs (done or failed) kept for lsb_server_poll / wait / latency,
// and latencies kept for the percentiles: the most recent kRetain of each.
constexpr size_t kRetain = size_t(1) << 18;

struct lsb_server {
  lsb_server_config cfg;
  std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  std::deque<uint64_t> pending;
  std::unordered_map<uint64_t, Job> jobs;
  std::vector<Launch> inflight;
  std::vector<StreamSlot> slots;
  // key class (key_bits | key_flags << 8) -> admitted small jobs
  std::map<uint32_t, std::vector<uint64_t>> open_batch;
  std::map<uint32_t, uint64_t> open_elems;
  std::map<uint32_t, double> batch_opened;
  uint64_t next_id = 1;
  uint64_t used = 0;
  int next_slot = 0;
  std::atomic<bool> stop{false};
  std::thread worker;
  int device = 0;
  int error = 0;
  std::deque<uint64_t> finished;  // finished job ids, oldest first
  // stats
  std::vector<double> lat_us;  // ring of the latest kRetain latencies
  size_t lat_next = 0;
  uint64_t done_jobs = 0, done_elems = 0, batches = 0, batched_jobs = 0, plain_jobs = 0;
  uint64_t pool_plain_jobs = 0, admission_waits = 0;
  double first_submit = 0, last_done = 0;

  void run();
  bool poll_completions();
  bool admit();
  bool flush_batches();
  int launch_plain(Job& j);
  int launch_batch(uint32_t key_class, std::vector<uint64_t>& ids);
  int pick_slot(bool need_slab);
  void retire(uint64_t id);
  void record_latency(double us);
};

namespace {

int hip_rc(hipError_t e) {
  if (e == hipSuccess) return LSB_OK;
  (void)hipGetLastError();
  return LSB_EHIP;
}

}  // namespace

void lsb_server::retire(uint64_t id) {
  finished.push_back(id);
  if (finished.size() > kRetain) {
    jobs.erase(finished.front());
    finished.pop_front();
  }
}

void lsb_server::record_latency(double us) {
  if (lat_us.size() < kRetain) {
    lat_us.push_back(us);
  } else {
    lat_us[lat_next] = us;
    lat_next = (lat_next + 1) % kRetain;
  }
}

int lsb_server::pick_slot(bool need_slab) {
  // Least-loaded stream; a batch also needs the slot's scratch slab free.
  int best = -1;
  for (int k = 0; k < (int)slots.size(); ++k) {
    const int i = (next_slot + k) % (int)slots.size();
    if (need_slab && hipEventQuery(slots[i].slab_free) != hipSuccess) continue;
    if (best < 0 || slots[i].inflight < slots[best].inflight) best = i;
  }
  if (best >= 0) next_slot = (best + 1) % (int)slots.size();
  (void)hipGetLastError();
  return best;
}

int lsb_server::launch_plain(Job& j) {
  // Scratch: a free slab when the job fits one, else the memory pool.  A job
  // that fits a slab waits for one (LSB_EBUSY, nothing changed) rather than
  // allocate: fresh device allocations of gigabytes have taken from 10 ms to
  // over half a second and stall the scheduler meanwhile.
  const bool use_slab = j.n > 1 && j.n <= cfg.batch_max_total;
  int si = pick_slot(use_slab);
  if (si < 0) return LSB_EBUSY;
  StreamSlot& sl = slots[si];
  Launch L;
  L.stream = si;
  L.charge = 32ull * j.n;
  L.jobs.push_back(j.id);
  int rc = LSB_OK;
  if (j.n > 1) {
    if (use_slab)
      L.scratch = sl.slab_b;
    else
      rc = hip_rc(lsb::pool_alloc(&L.scratch, j.n * sizeof(lsb_elem), sl.s));
    if (rc == LSB_OK) {
      lsb_config c;
      lsb_config_init(&c);
      c.key_bits = j.key_bits;
      c.flags = j.key_flags;
      c.stream = sl.s;
      rc = lsb_sort(j.A, static_cast<lsb_elem*>(L.scratch), j.n, &c);
      if (use_slab) {
        // the slab is busy until this sort is done
        const int er = hip_rc(hipEventRecord(sl.slab_free, sl.s));
        if (rc == LSB_OK) rc = er;
      } else {
        (void)hipFreeAsync(L.scratch, sl.s);
      }
      L.scratch = nullptr;
    }
  }
  if (rc == LSB_OK) rc = hip_rc(hipEventCreateWithFlags(&L.ev, hipEventDisableTiming));
  if (rc == LSB_OK) rc = hip_rc(hipEventRecord(L.ev, sl.s));
  if (rc != LSB_OK) {
    j.state = rc;
    j.done_at = now_s();
    used -= L.charge;
    if (L.ev) (void)hipEventDestroy(L.ev);
    retire(j.id);
    return rc;
  }
  ++sl.inflight;
  ++plain_jobs;
  if (j.n > 1 && !use_slab) ++pool_plain_jobs;
  inflight.push_back(L);
  return LSB_OK;
}

int lsb_server::launch_batch(uint32_t key_class, std::vector<uint64_t>& ids) {
  const int si = pick_slot(true);
  if (si < 0) return LSB_EBUSY;  // every scratch slab still busy: retry later
  StreamSlot& sl = slots[si];
  // Each job is sorted in place in its own buffer; its scratch is a slice of
  // the slot's slab.
  std::vector<lsb_elem*> as(ids.size()), bs(ids.size());
  std::vector<uint64_t> ns(ids.size());
  uint64_t off = 0;
  for (size_t k = 0; k < ids.size(); ++k) {
    const Job& j = jobs[ids[k]];
    as[k] = j.A;
    bs[k] = sl.slab_b + off;
    ns[k] = j.n;
    off += j.n;
  }
  lsb_config c;
  lsb_config_init(&c);
  c.key_bits = key_class & 0xFFu;
  c.flags = key_class >> 8;
  c.stream = sl.s;
  if (cfg.use_graphs) c.flags |= LSB_GRAPH;
  int rc = lsb_sort_batch(as.data(), bs.data(), ns.data(), (uint32_t)ids.size(), &c);
  Launch L;
  L.stream = si;
  L.batch = true;
  L.jobs = ids;
  for (uint64_t id : ids) L.charge += 32ull * jobs[id].n;
  if (rc == LSB_OK) rc = hip_rc(hipEventRecord(sl.slab_free, sl.s));
  if (rc == LSB_OK) rc = hip_rc(hipEventCreateWithFlags(&L.ev, hipEventDisableTiming));
  if (rc == LSB_OK) rc = hip_rc(hipEventRecord(L.ev, sl.s));
  if (rc != LSB_OK) {
    const double t = now_s();
    for (uint64_t id : ids) {
      jobs[id].state = rc;
      jobs[id].done_at = t;
    }
    used -= L.charge;
    if (L.ev) (void)hipEventDestroy(L.ev);
    for (uint64_t id : ids) retire(id);
    return rc;
  }
  ++sl.inflight;
  ++batches;
  batched_jobs += ids.size();
  inflight.push_back(L);
  return LSB_OK;
}

bool lsb_server::poll_completions() {
  bool any = false;
  for (size_t i = 0; i < inflight.size();) {
    Launch& L = inflight[i];
    const hipError_t q = hipEventQuery(L.ev);
    if (q == hipErrorNotReady) {
      ++i;
      continue;
    }
    const int rc = q == hipSuccess ? LSB_OK : LSB_EHIP;
    (void)hipGetLastError();
    const double t = now_s();
    for (uint64_t id : L.jobs) {
      Job& j = jobs[id];
      j.state = rc == LSB_OK ? 2 : rc;
      j.done_at = t;
      if (rc == LSB_OK) {
        record_latency((t - j.submitted) * 1e6);
        ++done_jobs;
        done_elems += j.n;
        last_done = std::max(last_done, t);
      }
    }
    for (uint64_t id : L.jobs) retire(id);
    used -= L.charge;
    --slots[L.stream].inflight;
    (void)hipEventDestroy(L.ev);
    inflight[i] = inflight.back();
    inflight.pop_back();
    any = true;
  }
  if (any) cv_done.notify_all();
  return any;
}

bool lsb_server::admit() {
  bool any = false;
  while (!pending.empty()) {
    Job& j = jobs[pending.front()];
    const uint64_t charge = 32ull * j.n;
    if (used + charge > cfg.memory_budget && used > 0) {  // FIFO: wait for memory
      if (!j.waited) {
        j.waited = true;
        ++admission_waits;
      }
      break;
    }
    if (j.n <= cfg.batch_max_elems) {
      // jobs are batched with others of the same key width and key order
      const uint32_t cls = j.key_bits | (j.key_flags << 8);
      auto& b = open_batch[cls];
      uint64_t& tot = open_elems[cls];
      if (!b.empty() && tot + j.n > cfg.batch_max_total) {
        if (launch_batch(cls, b) == LSB_EBUSY) break;  // every slab busy
        b.clear();
        tot = 0;
      }
      pending.pop_front();
      used += charge;
      j.state = 1;
      if (b.empty()) batch_opened[cls] = now_s();
      b.push_back(j.id);
      tot += j.n;
    } else {
      used += charge;
      j.state = 1;
      if (launch_plain(j) == LSB_EBUSY) {  // every slab busy: retry later
        used -= charge;
        j.state = 0;
        break;
      }
      pending.pop_front();
    }
    any = true;
  }
  return any;
}

// An open batch is launched when a stream is idle (nothing to wait for),
// when the batch window has passed, or when nothing else is pending.  A
// batch of one job runs as a plain sort.
bool lsb_server::flush_batches() {
  bool any = false;
  bool idle_stream = false;
  for (auto& sl : slots) idle_stream |= sl.inflight == 0;
  const double t = now_s();
  for (auto& kv : open_batch) {
    std::vector<uint64_t>& b = kv.second;
    if (b.empty()) continue;
    const bool due = idle_stream || pending.empty() ||
                     t - batch_opened[kv.first] >= cfg.batch_window_us * 1e-6;
    if (!due) continue;
    if (b.size() == 1) {
      if (launch_plain(jobs[b[0]]) == LSB_EBUSY) continue;
    } else if (launch_batch(kv.first, b) == LSB_EBUSY) {
      continue;
    }
    b.clear();
    open_elems[kv.first] = 0;
    any = true;
  }
  return any;
}

void lsb_server::run() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lock(mu);
  while (true) {
    bool progress = poll_completions();
    progress |= admit();
    // Everything admitted so far is batched now: small jobs arriving while
    // the device works on earlier launches join the next batch.
    progress |= flush_batches();
    const bool idle = pending.empty() && inflight.empty() &&
                      std::all_of(open_batch.begin(), open_batch.end(),
                                  [](const std::pair<const uint32_t, std::vector<uint64_t>>& kv) {
                                    return kv.second.empty();
                                  });
    if (stop && idle) break;
    if (!progress) {
      if (idle) {
        cv_submit.wait_for(lock, std::chrono::milliseconds(5));
      } else {
        lock.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lock.lock();
      }
    }
  }
}

namespace {

// Checks of one job that need no server state.
int check_job(const lsb_elem* A, size_t n, uint32_t key_bits, uint32_t key_flags) {
  if (n > LSB_MAX_N) return LSB_ERANGE;
  if (key_bits > 64 || (n > 1 && (!A || ((uintptr_t)A & 15u)))) return LSB_EINVAL;
  if (key_flags != 0 && key_flags != LSB_KEYS_SIGNED && key_flags != LSB_KEYS_FLOAT)
    return LSB_EINVAL;
  if (key_flags == LSB_KEYS_FLOAT && key_bits != 32 && key_bits != 64 && key_bits != 0)
    return LSB_EINVAL;
  return LSB_OK;
}

// Queues one checked job; the caller holds s->mu.
uint64_t enqueue_job(lsb_server* s, lsb_elem* A, size_t n, uint32_t key_bits,
                     uint32_t key_flags, double t) {
  Job j;
  j.id = s->next_id++;
  j.A = A;
  j.n = n;
  j.key_bits = key_bits ? key_bits : 64;
  j.key_flags = key_flags;
  j.submitted = t;
  if (s->first_submit == 0) s->first_submit = t;
  s->jobs[j.id] = j;
  s->pending.push_back(j.id);
  return j.id;
}

}  // namespace

extern "C" {

void lsb_server_config_init(lsb_server_config* c) {
  if (!c) return;
  memset(c, 0, sizeof *c);
  c->memory_budget = 200000000000ull;
  c->max_pending = 65536;
  c->batch_max_elems = 1u << 24;
  c->batch_max_total = 1ull << 28;
  c->streams = 8;
  c->use_graphs = 1;
  c->batch_window_us = 200;
}

int lsb_server_create(const lsb_server_config* cfg, lsb_server** out) {
  if (!out) return LSB_EINVAL;
  *out = nullptr;
  lsb_server_config c;
  lsb_server_config_init(&c);
  if (cfg) {
    if (cfg->memory_budget) c.memory_budget = cfg->memory_budget;
    if (cfg->max_pending) c.max_pending = cfg->max_pending;
    if (cfg->batch_max_elems) c.batch_max_elems = cfg->batch_max_elems;
    if (cfg->batch_max_total) c.batch_max_total = cfg->batch_max_total;
    if (cfg->streams) c.streams = cfg->streams;
    if (cfg->batch_window_us) c.batch_window_us = cfg->batch_window_us;
    c.use_graphs = cfg->use_graphs;
  }
  if (c.streams > 16 || c.batch_max_total < c.batch_max_elems ||
      c.batch_max_total > LSB_MAX_N)
    return LSB_EINVAL;
  lsb_server* s = new lsb_server();
  s->cfg = c;
  int rc = hip_rc(hipGetDevice(&s->device));
  s->slots.resize(c.streams);
  for (auto& sl : s->slots) {
    if (rc) break;
    rc = hip_rc(hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking));
    if (!rc) rc = hip_rc(hipMalloc(&sl.slab_b, c.batch_max_total * sizeof(lsb_elem)));
    if (!rc) rc = hip_rc(hipEventCreateWithFlags(&sl.slab_free, hipEventDisableTiming));
    if (!rc) rc = hip_rc(hipEventRecord(sl.slab_free, sl.s));
  }
  if (rc) {
    for (auto& sl : s->slots) {
      if (sl.slab_b) (void)hipFree(sl.slab_b);
      if (sl.slab_free) (void)hipEventDestroy(sl.slab_free);
      if (sl.s) (void)hipStreamDestroy(sl.s);
    }
    delete s;
    return rc == LSB_EHIP ? LSB_ENOMEM : rc;
  }
  s->worker = std::thread([s] { s->run(); });
  *out = s;
  return LSB_OK;
}


int lsb_server_submit(lsb_server* s, lsb_elem* A, size_t n, uint32_t key_bits,
                      uint32_t key_flags, uint64_t* job_id) {
  if (!s || !job_id) return LSB_EINVAL;
  int rc = check_job(A, n, key_bits, key_flags);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->stop) return LSB_EINVAL;
  if (s->pending.size() >= s->cfg.max_pending) return LSB_EBUSY;
  if (32ull * n > s->cfg.memory_budget) return LSB_ENOMEM;  // could never be admitted
  *job_id = enqueue_job(s, A, n, key_bits, key_flags, now_s());
  s->cv_submit.notify_one();
  return LSB_OK;
}

int lsb_server_submit_batch(lsb_server* s, lsb_elem* const* As, const size_t* ns,
                            const uint32_t* key_bits, const uint32_t* key_flags, size_t count,
                            uint64_t* job_ids) {
  if (!s || (count && (!As || !ns || !key_bits || !key_flags || !job_ids))) return LSB_EINVAL;
  for (size_t i = 0; i < count; ++i) {
    const int rc = check_job(As[i], ns[i], key_bits[i], key_flags[i]);
    if (rc) return rc;
  }
  std::lock_guard<std::mutex> lock(s->mu);
  if (s->stop) return LSB_EINVAL;
  if (s->pending.size() + count > s->cfg.max_pending) return LSB_EBUSY;
  for (size_t i = 0; i < count; ++i)
    if (32ull * ns[i] > s->cfg.memory_budget) return LSB_ENOMEM;
  const double t = now_s();
  for (size_t i = 0; i < count; ++i)
    job_ids[i] = enqueue_job(s, As[i], ns[i], key_bits[i], key_flags[i], t);
  if (count) s->cv_submit.notify_one();
  return LSB_OK;
}

int lsb_server_poll(lsb_server* s, uint64_t job_id) {
  if (!s) return LSB_EINVAL;
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return LSB_ENOTFOUND;
  if (it->second.state < 0) return it->second.state;
  return it->second.state == 2 ? 1 : 0;
}

int lsb_server_wait(lsb_server* s, uint64_t job_id, double timeout_s) {
  if (!s) return LSB_EINVAL;
  std::unique_lock<std::mutex> lock(s->mu);
  const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                           std::chrono::duration<double>(timeout_s));
  while (true) {
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return LSB_ENOTFOUND;
    if (it->second.state < 0) return it->second.state;
    if (it->second.state == 2) return LSB_OK;
    if (timeout_s >= 0 && Clock::now() >= deadline) return LSB_EBUSY;
    s->cv_done.wait_for(lock, std::chrono::milliseconds(2));
  }
}

int lsb_server_drain(lsb_server* s) {
  if (!s) return LSB_EINVAL;
  std::unique_lock<std::mutex> lock(s->mu);
  while (true) {
    bool busy = !s->pending.empty() || !s->inflight.empty();
    for (auto& kv : s->open_batch) busy |= !kv.second.empty();
    if (!busy) return LSB_OK;
    s->cv_done.wait_for(lock, std::chrono::milliseconds(2));
  }
}

int lsb_server_latency(lsb_server* s, uint64_t job_id, double* us) {
  if (!s || !us) return LSB_EINVAL;
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->jobs.find(job_id);
  if (it == s->jobs.end()) return LSB_ENOTFOUND;
  if (it->second.state != 2) return LSB_EBUSY;
  *us = (it->second.done_at - it->second.submitted) * 1e6;
  return LSB_OK;
}

int lsb_server_stats_get(lsb_server* s, lsb_server_stats* out) {
  if (!s || !out) return LSB_EINVAL;
  std::lock_guard<std::mutex> lock(s->mu);
  memset(out, 0, sizeof *out);
  out->jobs_done = s->done_jobs;
  out->elems_done = s->done_elems;
  out->batches = s->batches;
  out->batched_jobs = s->batched_jobs;
  out->plain_jobs = s->plain_jobs;
  out->pool_plain_jobs = s->pool_plain_jobs;
  out->admission_waits = s->admission_waits;
  out->memory_in_use = s->used;
  std::vector<double> v = s->lat_us;
  if (!v.empty()) {
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
      const size_t k = (size_t)std::min<double>(v.size() - 1, q * (v.size() - 1) + 0.5);
      return v[k];
    };
    out->p50_us = pct(0.50);
    out->p99_us = pct(0.99);
    out->max_us = v.back();
    double sum = 0;
    for (double x : v) sum += x;
    out->mean_us = sum / v.size();
  }
  out->span_s = s->last_done > s->first_submit ? s->last_done - s->first_submit : 0.0;
  return LSB_OK;
}

int lsb_server_reset_stats(lsb_server* s) {
  if (!s) return LSB_EINVAL;
  std::lock_guard<std::mutex> lock(s->mu);
  s->lat_us.clear();
  s->lat_next = 0;
  s->done_jobs = s->done_elems = s->batches = s->batched_jobs = s->plain_jobs = 0;
  s->pool_plain_jobs = s->admission_waits = 0;
  s->first_submit = s->last_done = 0;
  for (uint64_t id : s->finished) s->jobs.erase(id);
  s->finished.clear();
  return LSB_OK;
}

int lsb_server_destroy(lsb_server* s) {
  if (!s) return LSB_EINVAL;
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
    s->cv_submit.notify_all();
  }
  s->worker.join();
  for (auto& sl : s->slots) {
    (void)hipStreamSynchronize(sl.s);
    (void)hipFree(sl.slab_b);
    (void)hipEventDestroy(sl.slab_free);
    (void)hipStreamDestroy(sl.s);
  }
  delete s;
  return LSB_OK;
}

}  // extern "C"
