/*
 * oracle_selftest.c -- self-checks of the CPU oracle (lsb_oracle.c), built
 * with AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle asan)
 * so every index, shift and buffer the oracle touches is checked while it
 * runs.  TEST INFRASTRUCTURE ONLY, like the oracle itself.
 *
 * The checks are independent of the oracle's own arithmetic: published
 * known answers (FIPS 180-4 SHA-256 vectors, pcg64 outputs recorded from the
 * reference's generator), libc qsort on (masked key, original index) as the
 * stable order, brute-force digit counts, and mistakes planted in a sorted
 * output that the checker must catch.  The reference's verify is a
 * std::stable_sort comparison (mpi/mpi_lsbsort.cpp:755-784).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t key, val; } orc_elem;

void orc_pcg64_fill(uint64_t seed, uint64_t skip, uint64_t count, uint64_t* out);
void orc_gen_uniform(orc_elem* out, uint64_t n, uint32_t ranks, uint32_t key_bits,
                     uint64_t seed_base, uint64_t first, uint64_t count);
void orc_zipf_cdf(double* cdf, uint32_t m, double s);
void orc_histograms(const orc_elem* a, uint64_t n, uint32_t key_bits, uint32_t d,
                    uint32_t* counts);
void orc_lsd_sort(orc_elem* A, orc_elem* B, uint64_t n, uint32_t key_bits, uint32_t d);
void orc_stable_sort(orc_elem* a, uint64_t n, uint32_t key_bits);
int64_t orc_check(const orc_elem* out, const uint64_t* in_keys, uint64_t n, uint64_t val_base,
                  uint32_t key_bits, int32_t* why);
void orc_sha256(const void* data, uint64_t len, uint8_t out[32]);

static int failures = 0;
#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__); \
      fprintf(stderr, __VA_ARGS__);           \
      fprintf(stderr, "\n");                  \
      ++failures;                             \
    }                                         \
  } while (0)

static uint64_t g_mask;
static int by_masked_key_then_index(const void* x, const void* y) {
  const orc_elem* a = (const orc_elem*)x;
  const orc_elem* b = (const orc_elem*)y;
  const uint64_t ka = a->key & g_mask, kb = b->key & g_mask;
  if (ka != kb) return ka < kb ? -1 : 1;
  return a->val < b->val ? -1 : (a->val > b->val);
}

static void hex(const uint8_t* d, char* s) {
  for (int i = 0; i < 32; i++) sprintf(s + 2 * i, "%02x", d[i]);
}

static void check_sha256(void) {
  uint8_t d[32];
  char s[65];
  orc_sha256("abc", 3, d);
  hex(d, s);
  CHECK(!strcmp(s, "ba7816bf8f01cfea414140de5dae2223b00361a396177a9cb410ff61f20015ad"),
        "sha256(abc) = %s", s);
  orc_sha256("", 0, d);
  hex(d, s);
  CHECK(!strcmp(s, "e3b0c44298fc1c149afbf4c8996fb92427ae41e4649b934ca495991b7852b855"),
        "sha256('') = %s", s);
  const char* m = "abcdbcdecdefdefgefghfghighijhijkijkljklmklmnlmnomnopnopq";
  orc_sha256(m, strlen(m), d);
  hex(d, s);
  CHECK(!strcmp(s, "248d6a61d20638b8e5c026930c3e6039a33ce45964ff2167f6ecedd419db06c1"),
        "sha256(448-bit message) = %s", s);
}

static void check_pcg64(void) {
  /* outputs 0..2 of pcg64(0) and pcg64(1), recorded from pcg-cpp */
  const uint64_t want0[3] = {0x01070196e695f8f1ull, 0x703ec840c59f4493ull, 0xe54954914b3a44faull};
  const uint64_t want1[3] = {0xe175e32ed3507bfaull, 0xc0bf922a0b283109ull, 0x140bfa21e68785bbull};
  uint64_t out[3];
  orc_pcg64_fill(0, 0, 3, out);
  CHECK(!memcmp(out, want0, sizeof out), "pcg64(0) outputs");
  orc_pcg64_fill(1, 0, 3, out);
  CHECK(!memcmp(out, want1, sizeof out), "pcg64(1) outputs");
  /* jump-ahead equals stepping */
  uint64_t* all = malloc(5000 * sizeof(uint64_t));
  orc_pcg64_fill(7, 0, 5000, all);
  for (uint64_t skip = 0; skip < 5000; skip += 997) {
    orc_pcg64_fill(7, skip, 1, out);
    CHECK(out[0] == all[skip], "pcg64(7) skip %llu", (unsigned long long)skip);
  }
  free(all);
}

static void check_sorts(void) {
  const uint64_t sizes[] = {0, 1, 2, 3, 1000, 4097, 65537};
  const uint32_t widths[] = {1, 8, 13, 32, 57, 64};
  const uint32_t digits[] = {8, 11, 16};
  for (size_t si = 0; si < sizeof sizes / sizeof sizes[0]; si++) {
    const uint64_t n = sizes[si];
    orc_elem* in = malloc((n + 1) * sizeof(orc_elem));
    orc_elem* A = malloc((n + 1) * sizeof(orc_elem));
    orc_elem* B = malloc((n + 1) * sizeof(orc_elem));
    orc_elem* want = malloc((n + 1) * sizeof(orc_elem));
    uint64_t* keys = malloc((n + 1) * sizeof(uint64_t));
    orc_gen_uniform(in, n, 3, 64, 11 + si, 0, n);
    for (uint64_t i = 0; i < n; i++) {
      if (i % 3 == 0) in[i].key &= 0xFFFFull; /* duplicates of the low bits */
      keys[i] = in[i].key;
    }
    for (size_t wi = 0; wi < sizeof widths / sizeof widths[0]; wi++) {
      const uint32_t kb = widths[wi];
      g_mask = kb >= 64 ? ~0ull : ((1ull << kb) - 1);
      memcpy(want, in, n * sizeof(orc_elem));
      qsort(want, n, sizeof(orc_elem), by_masked_key_then_index);
      memcpy(A, in, n * sizeof(orc_elem));
      orc_stable_sort(A, n, kb);
      CHECK(!memcmp(A, want, n * sizeof(orc_elem)), "stable_sort n=%llu kb=%u",
            (unsigned long long)n, kb);
      int32_t first = -2;
      CHECK(orc_check(A, keys, n, 0, kb, &first) == -1, "check rejects a correct sort");
      for (size_t di = 0; di < sizeof digits / sizeof digits[0]; di++) {
        const uint32_t d = digits[di];
        memcpy(A, in, n * sizeof(orc_elem));
        orc_lsd_sort(A, B, n, kb, d);
        CHECK(!memcmp(A, want, n * sizeof(orc_elem)), "lsd_sort n=%llu kb=%u d=%u",
              (unsigned long long)n, kb, d);
        /* digit counts: brute force per digit */
        const uint32_t P = (kb + d - 1) / d, nb = 1u << d;
        uint32_t* c = calloc((size_t)P * nb, sizeof(uint32_t));
        orc_histograms(in, n, kb, d, c);
        for (uint32_t p = 0; p < P; p++) {
          uint64_t total = 0;
          for (uint32_t b = 0; b < nb; b++) total += c[(size_t)p * nb + b];
          CHECK(total == n, "histogram digit %u sums to n", p);
        }
        if (n) {
          const uint32_t sh = 0, w = kb < d ? kb : d;
          const uint64_t dk = (in[0].key >> sh) & ((1ull << w) - 1);
          uint64_t same = 0;
          for (uint64_t i = 0; i < n; i++) same += ((in[i].key >> sh) & ((1ull << w) - 1)) == dk;
          CHECK(c[dk] == same, "histogram digit 0 bucket of key 0");
        }
        free(c);
      }
      /* planted mistakes */
      if (n >= 3) {
        memcpy(A, want, n * sizeof(orc_elem));
        orc_elem t = A[0];
        A[0] = A[n - 1];
        A[n - 1] = t;
        int ne = (A[0].key & g_mask) != (A[n - 1].key & g_mask) || A[0].val != A[n - 1].val;
        if (ne) CHECK(orc_check(A, keys, n, 0, kb, &first) >= 0, "check misses a swap");
        memcpy(A, want, n * sizeof(orc_elem));
        A[1].val = A[0].val;
        CHECK(orc_check(A, keys, n, 0, kb, &first) >= 0, "check misses a duplicated index");
      }
    }
    free(in);
    free(A);
    free(B);
    free(want);
    free(keys);
  }
}

static void check_zipf(void) {
  const uint32_t m = 1000;
  double* cdf = malloc(m * sizeof(double));
  orc_zipf_cdf(cdf, m, 1.1);
  for (uint32_t k = 1; k < m; k++) CHECK(cdf[k] >= cdf[k - 1], "zipf cdf non-decreasing");
  CHECK(cdf[m - 1] == 1.0, "zipf cdf ends at 1");
  free(cdf);
}

int main(void) {
  check_sha256();
  check_pcg64();
  check_sorts();
  check_zipf();
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("oracle self-test passed\n");
  return 0;
}
