// host_selftest: drives the host side of liblsb_amd.so built with
// AddressSanitizer and UndefinedBehaviorSanitizer on its host code
// (make -C distributed-lsb_amd asan; make -C tests/cxx asan), so the
// library's planning and layout arithmetic, argument checks, tune-file
// parsing, workspace and graph caches and the serving scheduler's threads
// run under the sanitizers.
//
//   host_selftest          entry points that need no GPU (CPU container)
//   host_selftest --gpu    also sorts, batches and a two-thread serving run
//                          on device 0, each result proven with lsb_verify
//
// Run it with LSAN_OPTIONS=suppressions=tests/cxx/lsan_rocm.supp: the ROCm
// runtime keeps some allocations until the process ends.
//
// The reference guards the same kinds of invariants with asserts
// (mpi/mpi_lsbsort.cpp:364, 392, 487-488, 583).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <unistd.h>
#include <vector>

#include "lsb.h"

static int failures = 0;
#define CHECK(cond, what)                                                  \
  do {                                                                     \
    if (!(cond)) {                                                         \
      std::fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, what);  \
      ++failures;                                                          \
    }                                                                      \
  } while (0)

static void cpu_checks() {
  lsb_config c;
  lsb_config_init(&c);
  CHECK(c.key_bits == 64 && c.workspace == nullptr, "config defaults");
  // planning and workspace sizing across sizes, key widths and variants
  const size_t sizes[] = {2, 3, 100, 8192, 8193, 65537, (size_t)1 << 20, (size_t)1 << 24,
                          ((size_t)1 << 30) + 17, LSB_MAX_N};
  int nv = lsb_variants(nullptr, nullptr, nullptr, 0);
  CHECK(nv > 0, "variants");
  std::vector<uint32_t> vd(nv), vt(nv), vp(nv);
  lsb_variants(vd.data(), vt.data(), vp.data(), nv);
  for (size_t n : sizes) {
    for (uint32_t kb : {1u, 8u, 31u, 32u, 33u, 63u, 64u}) {
      c.key_bits = kb;
      c.variant = 0;
      uint32_t d = 0, t = 0, p = 0;
      CHECK(lsb_plan(n, &c, &d, &t, &p) == LSB_OK, "plan");
      CHECK(p == (kb + d - 1) / d, "pass count");
      CHECK(lsb_workspace_size(n, &c) > 0, "workspace size");
      for (int v = 1; v <= nv; ++v) {
        c.variant = (uint32_t)v;
        CHECK(lsb_plan(n, &c, &d, &t, &p) == LSB_OK && t == vt[v - 1], "plan per variant");
      }
      c.variant = 0;
      for (uint32_t nseg : {1u, 7u, 4096u}) {
        CHECK(lsb_segmented_workspace_size(n, nseg, &c) > 0, "segmented workspace size");
      }
    }
  }
  c.key_bits = 65;
  CHECK(lsb_plan(100, &c, nullptr, nullptr, nullptr) == LSB_EINVAL, "key_bits 65");
  lsb_config_init(&c);
  // argument errors that return before any HIP call
  lsb_elem* a = reinterpret_cast<lsb_elem*>(0x10000);
  lsb_elem* b = reinterpret_cast<lsb_elem*>((uintptr_t)1 << 40);
  CHECK(lsb_sort(nullptr, nullptr, 0, nullptr) == LSB_OK, "n = 0");
  CHECK(lsb_sort(a, b, 1, nullptr) == LSB_OK, "n = 1");
  CHECK(lsb_sort(nullptr, b, 5, nullptr) == LSB_EINVAL, "null A");
  CHECK(lsb_sort(a + 0, reinterpret_cast<lsb_elem*>(0x10008), 5, nullptr) == LSB_EINVAL,
        "misaligned / overlapping");
  CHECK(lsb_sort(a, b, LSB_MAX_N + 1, nullptr) == LSB_ERANGE, "above LSB_MAX_N");
  CHECK(lsb_sort_noalloc(a, b, LSB_NOALLOC_MAX_N + 1, nullptr) == LSB_ERANGE, "noalloc range");
  const uint64_t bad_off[] = {0, 50, 40, 100};
  CHECK(lsb_sort_segmented(a, b, bad_off, 3, nullptr) == LSB_EINVAL, "decreasing offsets");
  CHECK(lsb_sort_segmented(a, b, nullptr, 3, nullptr) == LSB_EINVAL, "null offsets");
  const uint64_t ns[] = {100000};
  lsb_elem* const As[] = {a};
  CHECK(lsb_sort_batch(As, nullptr, ns, 1, nullptr) == LSB_EINVAL, "batch without scratch");
  for (int code = 0; code >= -7; --code) CHECK(lsb_strerror(code) != nullptr, "strerror");
  // tune table: save, load (with malformed lines), clear
  const std::string path = std::string("/tmp/lsb_host_selftest_tune_") +
                           std::to_string((unsigned long)getpid()) + ".txt";
  {
    FILE* f = std::fopen(path.c_str(), "w");
    std::fprintf(f, "# comment\n20 64 8 8192 3 1.5 2.0\ngarbage line\n21 32 8 4096 3 0.5\n"
                    "22 64 99 1 1 1 1\n");
    std::fclose(f);
  }
  CHECK(lsb_tune_load(path.c_str()) == LSB_OK, "tune load");
  lsb_tune_entry e[8];
  const int ne = lsb_tune_table(e, 8);
  CHECK(ne == 1 && e[0].log2n == 20 && e[0].tile == 8192, "malformed tune lines dropped");
  CHECK(lsb_tune_save(path.c_str()) == LSB_OK, "tune save");
  CHECK(lsb_tune_load(path.c_str()) == LSB_OK && lsb_tune_table(e, 8) == 1, "tune reload");
  CHECK(lsb_tune_clear() == LSB_OK && lsb_tune_table(e, 8) == 0, "tune clear");
  std::remove(path.c_str());
  CHECK(lsb_tune_load("/nonexistent/lsb_tune.txt") == LSB_EINVAL, "missing tune file");
  // without a device the server cannot start, and says so without leaking
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    lsb_server* srv = nullptr;
    CHECK(lsb_server_create(nullptr, &srv) != LSB_OK && srv == nullptr, "server without GPU");
  }
}

// ---------------------------------------------------------------- GPU mode
static bool verify(lsb_elem* A, const uint64_t* keys, size_t n, uint32_t kb, hipStream_t s) {
  uint64_t bad = 0;
  int64_t first = 0;
  return lsb_verify(A, keys, n, 0, kb, 0, &bad, &first, s) == LSB_OK && bad == 0;
}

static lsb_elem* gen(size_t n, uint32_t kb, uint32_t kind, uint64_t seed, uint64_t** keys) {
  lsb_elem* A = nullptr;
  if (hipMalloc(&A, n * sizeof(lsb_elem)) != hipSuccess) return nullptr;
  lsb_gen_spec g;
  std::memset(&g, 0, sizeof g);
  g.kind = kind;
  g.ranks = 4;
  g.key_bits = kb;
  g.zipf_m = 1 << 12;
  g.zipf_s = 1.1;
  g.zipf_table_seed = 5;
  g.seed_base = seed;
  if (lsb_generate(A, n, &g, nullptr) != LSB_OK) return nullptr;
  if (hipMalloc(keys, n * sizeof(uint64_t)) != hipSuccess) return nullptr;
  if (lsb_extract_keys(A, n, *keys, nullptr) != LSB_OK) return nullptr;
  return A;
}

static void gpu_checks() {
  hipStream_t s = nullptr;
  CHECK(hipStreamCreate(&s) == hipSuccess, "stream");
  // plain sorts, growing and shrinking the cached workspace, eager and graph
  for (size_t n : {(size_t)1000, (size_t)300001, (size_t)1 << 22, (size_t)70000}) {
    uint64_t* keys = nullptr;
    lsb_elem* A = gen(n, 64, LSB_GEN_UNIFORM, n, &keys);
    lsb_elem* B = nullptr;
    CHECK(A && hipMalloc(&B, n * sizeof(lsb_elem)) == hipSuccess, "alloc");
    for (uint32_t flags : {0u, (uint32_t)LSB_GRAPH}) {
      lsb_config c;
      lsb_config_init(&c);
      c.stream = s;
      c.flags = flags | LSB_SYNC;
      CHECK(lsb_sort(A, B, n, &c) == LSB_OK, "lsb_sort");
      CHECK(verify(A, keys, n, 64, s), "lsb_sort result");
    }
    lsb_config c;
    lsb_config_init(&c);
    c.stream = s;
    c.flags = LSB_SYNC;
    CHECK(lsb_sort_noalloc(A, B, n, &c) == LSB_OK && verify(A, keys, n, 64, s), "noalloc");
    (void)hipFree(A);
    (void)hipFree(B);
    (void)hipFree(keys);
  }
  // a batch of separate buffers, MSD-first and LSD
  {
    const size_t ns[] = {5000, 9000, 200000, 1 << 20, 3, 70001};
    const int m = 6;
    std::vector<lsb_elem*> As(m), Bs(m);
    std::vector<uint64_t*> keys(m);
    std::vector<uint64_t> nn(ns, ns + m);
    for (int i = 0; i < m; ++i) {
      As[i] = gen(ns[i], 40, LSB_GEN_ZIPF, 100 + i, &keys[i]);
      CHECK(As[i] && hipMalloc(&Bs[i], ns[i] * sizeof(lsb_elem)) == hipSuccess, "alloc");
    }
    for (uint32_t flags : {0u, (uint32_t)LSB_GRAPH, (uint32_t)LSB_SEG_LSD}) {
      lsb_config c;
      lsb_config_init(&c);
      c.key_bits = 40;
      c.stream = s;
      c.flags = flags | LSB_SYNC;
      CHECK(lsb_sort_batch(As.data(), Bs.data(), nn.data(), m, &c) == LSB_OK, "batch");
      for (int i = 0; i < m; ++i) CHECK(verify(As[i], keys[i], ns[i], 40, s), "batch result");
    }
    for (int i = 0; i < m; ++i) {
      (void)hipFree(As[i]);
      (void)hipFree(Bs[i]);
      (void)hipFree(keys[i]);
    }
  }
  CHECK(lsb_release(s) == LSB_OK, "release");
  // serving: two submitter threads, a small budget so admission waits
  {
    const int jobs = 48;
    std::vector<lsb_elem*> A(jobs);
    std::vector<uint64_t*> keys(jobs);
    std::vector<size_t> n(jobs);
    std::vector<uint32_t> kb(jobs);
    for (int j = 0; j < jobs; ++j) {
      n[j] = (size_t)1 << (10 + (j * 7) % 13);
      kb[j] = j % 2 ? 32 : 64;
      A[j] = gen(n[j], kb[j], j % 3 ? LSB_GEN_UNIFORM : LSB_GEN_ZIPF, 500 + j, &keys[j]);
      CHECK(A[j] != nullptr, "alloc job");
    }
    lsb_server_config sc;
    lsb_server_config_init(&sc);
    sc.memory_budget = 32ull * ((size_t)1 << 23);
    sc.streams = 4;
    sc.batch_max_elems = 1 << 18;
    sc.batch_max_total = 1 << 21;
    lsb_server* srv = nullptr;
    CHECK(lsb_server_create(&sc, &srv) == LSB_OK, "server");
    std::atomic<int> errs{0};
    auto feed = [&](int first) {
      for (int j = first; j < jobs; j += 2) {
        uint64_t id = 0;
        if (lsb_server_submit(srv, A[j], n[j], kb[j], 0, &id) != LSB_OK ||
            lsb_server_wait(srv, id, 60.0) != LSB_OK)
          ++errs;
      }
    };
    std::thread t0(feed, 0), t1(feed, 1);
    t0.join();
    t1.join();
    CHECK(errs == 0, "serving submissions");
    CHECK(lsb_server_drain(srv) == LSB_OK, "drain");
    lsb_server_stats st;
    CHECK(lsb_server_stats_get(srv, &st) == LSB_OK && st.jobs_done == (uint64_t)jobs, "stats");
    CHECK(lsb_server_destroy(srv) == LSB_OK, "destroy");
    for (int j = 0; j < jobs; ++j) {
      CHECK(verify(A[j], keys[j], n[j], kb[j], nullptr), "served job result");
      (void)hipFree(A[j]);
      (void)hipFree(keys[j]);
    }
  }
  (void)hipStreamDestroy(s);
}

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && !std::strcmp(argv[1], "--gpu");
  cpu_checks();
  if (gpu) gpu_checks();
  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("host self-test passed (%s)\n", gpu ? "cpu and gpu checks" : "cpu checks");
  std::fflush(stdout);
  return 0;
}
