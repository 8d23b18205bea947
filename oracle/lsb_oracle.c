/*
 * lsb_oracle.c -- CPU restatement of the reference's one-rank LSD radix sort.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP product is
 * compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never links
 * or calls it, and nothing here is ever the thing being measured on the GPU.
 *
 * What each function restates (reference = mppf/distributed-lsb):
 *   orc_pcg64_*        pcg64 = PCG XSL-RR 128/64, setseq, as used for keygen
 *                      at mpi/mpi_lsbsort.cpp:686-690 (pcg-cpp, unpinned HEAD;
 *                      algorithm from O'Neill, "PCG: A Family of Simple Fast
 *                      Space-Efficient Statistically Good Algorithms", 2014).
 *                      Jump-ahead follows F. Brown, "Random Number Generation
 *                      with Arbitrary Strides", Trans. ANS 1994.
 *   orc_gen_uniform    per-rank keygen + block distribution,
 *                      mpi/mpi_lsbsort.cpp:146-171 (divCeil split) and 676-704.
 *   orc_histograms     count loop of localShuffle, mpi/mpi_lsbsort.cpp:237-241,
 *                      done for every digit at once.
 *   orc_counting_pass  one stable counting-sort pass (count, exclusive scan,
 *                      scatter), localShuffle mpi/mpi_lsbsort.cpp:225-259.
 *   orc_lsd_sort       mySort mpi/mpi_lsbsort.cpp:591-597 at one rank (the
 *                      count transpose / Exscan / Alltoallv are identities
 *                      when R = 1), generalised to digit width d and key
 *                      width key_bits like arkouda-radix-sort.chpl:69-78.
 *   orc_stable_sort    the verify oracle, std::stable_sort by key,
 *                      mpi/mpi_lsbsort.cpp:767-771 (restated as merge sort).
 *   orc_check          O(n) proof that out == stable_sort(in) when
 *                      val == original index (sorted, permutation, keys
 *                      match their source, equal keys keep index order).
 *   orc_sha256         FIPS 180-4 SHA-256 for the golden digests.
 *
 * Pinning (tests/test_oracle.py): pcg64 against the known-answer outputs of
 * pcg64(0..3); keygen + sort against the SHA-256 digests of the reference's
 * 2^24-element, 4-rank input and sorted output; histograms against
 * numpy.bincount; passes and full sort against numpy's stable argsort;
 * sha256 against hashlib; the checker against planted mistakes.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef struct { uint64_t key, val; } orc_elem;
typedef unsigned __int128 u128;

/* ---------------------------------------------------------------- pcg64 */

static const u128 PCG_MUL =
    ((u128)2549297995355413924ULL << 64) | (u128)4865540595714422341ULL;
static const u128 PCG_INC =
    ((u128)6364136223846793005ULL << 64) | (u128)1442695040888963407ULL;

static inline uint64_t pcg_out(u128 s) {
  uint64_t x = (uint64_t)(s >> 64) ^ (uint64_t)s;
  unsigned r = (unsigned)(s >> 122);
  return (x >> r) | (x << ((64 - r) & 63));
}

/* state of pcg64(seed) before its first output: bump(seed + inc) */
static inline u128 pcg_seed_state(uint64_t seed) {
  return ((u128)seed + PCG_INC) * PCG_MUL + PCG_INC;
}

/* advance an LCG state by delta steps in O(log delta) (Brown 1994) */
static u128 pcg_advance(u128 s, uint64_t delta) {
  u128 acc_mul = 1, acc_add = 0, cur_mul = PCG_MUL, cur_add = PCG_INC;
  while (delta) {
    if (delta & 1) { acc_mul *= cur_mul; acc_add = acc_add * cur_mul + cur_add; }
    cur_add = (cur_mul + 1) * cur_add;
    cur_mul *= cur_mul;
    delta >>= 1;
  }
  return acc_mul * s + acc_add;
}

/* out[k] = output number (skip + k) of pcg64(seed), 0-based */
void orc_pcg64_fill(uint64_t seed, uint64_t skip, uint64_t count, uint64_t* out) {
  u128 s = pcg_advance(pcg_seed_state(seed), skip);
  for (uint64_t k = 0; k < count; k++) {
    s = s * PCG_MUL + PCG_INC;           /* advance, then output */
    out[k] = pcg_out(s);
  }
}

static inline uint64_t key_mask(unsigned key_bits) {
  return key_bits >= 64 ? ~0ULL : ((1ULL << key_bits) - 1);
}

/* The reference's input: rank r of R draws ceil(n/R) keys from pcg64(r)
 * (the last rank's padding keys are drawn but never sorted) and sets
 * val = global index.  Element g of the global array is therefore output
 * number (g mod per) of pcg64(seed_base + g / per).  Keys are masked to the
 * low key_bits (64 = reference).  Only [first, first+count) is written. */
void orc_gen_uniform(orc_elem* out, uint64_t n, uint32_t ranks,
                     uint32_t key_bits, uint64_t seed_base, uint64_t first,
                     uint64_t count) {
  if (ranks == 0) ranks = 1;
  uint64_t per = (n + ranks - 1) / ranks;
  uint64_t m = key_mask(key_bits);
  uint64_t g = first, end = first + count;
  while (g < end) {
    uint64_t r = g / per, i = g % per;
    uint64_t run = per - i;
    if (run > end - g) run = end - g;
    u128 s = pcg_advance(pcg_seed_state(seed_base + r), i);
    for (uint64_t k = 0; k < run; k++) {
      s = s * PCG_MUL + PCG_INC;
      out[g - first + k].key = pcg_out(s) & m;
      out[g - first + k].val = g + k;
    }
    g += run;
  }
}

/* sorted (key = i) or reverse-sorted (key = n-1-i) keys, val = i */
void orc_gen_ordered(orc_elem* out, uint64_t n, int reverse) {
  for (uint64_t i = 0; i < n; i++) {
    out[i].key = reverse ? n - 1 - i : i;
    out[i].val = i;
  }
}

/* Zipf(s) over m distinct keys.  cdf[k] = sum_{j<=k} (j+1)^-s / sum_all,
 * summed in increasing k in double.  The rank of element g is the smallest
 * k with u < cdf[k], u = (x >> 11) * 2^-53 where x is the element's pcg64
 * draw laid out exactly like orc_gen_uniform; its key is output k of
 * pcg64(table_seed), masked to the low key_bits. */
void orc_zipf_cdf(double* cdf, uint32_t m, double s) {
  double acc = 0.0;
  for (uint32_t k = 0; k < m; k++) { acc += pow((double)(k + 1), -s); cdf[k] = acc; }
  for (uint32_t k = 0; k < m; k++) cdf[k] /= acc;
  cdf[m - 1] = 1.0;
}

static uint32_t zipf_rank(const double* cdf, uint32_t m, uint64_t x) {
  double u = (double)(x >> 11) * (1.0 / 9007199254740992.0);
  uint32_t lo = 0, hi = m - 1;          /* first k with u < cdf[k] */
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (u < cdf[mid]) hi = mid; else lo = mid + 1;
  }
  return lo;
}

void orc_gen_zipf(orc_elem* out, uint64_t n, uint32_t ranks, uint64_t seed_base,
                  const double* cdf, const uint64_t* table, uint32_t m,
                  uint64_t first, uint64_t count, uint32_t key_bits) {
  uint64_t mk = key_mask(key_bits);
  orc_gen_uniform(out, n, ranks, 64, seed_base, first, count);
  for (uint64_t k = 0; k < count; k++)
    out[k].key = table[zipf_rank(cdf, m, out[k].key)] & mk;
}

/* ------------------------------------------------------- radix sort core */

static inline unsigned n_digits(unsigned key_bits, unsigned d) {
  return (key_bits + d - 1) / d;
}

static inline uint64_t digit_of(uint64_t key, unsigned p, unsigned d,
                                unsigned key_bits) {
  unsigned shift = p * d;
  unsigned w = key_bits - shift < d ? key_bits - shift : d;
  return (key >> shift) & ((w >= 64) ? ~0ULL : ((1ULL << w) - 1));
}

/* counts[p * 2^d + b] for every digit p at once (uint32 counts). */
void orc_histograms(const orc_elem* a, uint64_t n, uint32_t key_bits,
                    uint32_t d, uint32_t* counts) {
  unsigned P = n_digits(key_bits, d);
  uint64_t nb = 1ULL << d;
  memset(counts, 0, sizeof(uint32_t) * P * nb);
  for (uint64_t i = 0; i < n; i++)
    for (unsigned p = 0; p < P; p++)
      counts[p * nb + digit_of(a[i].key, p, d, key_bits)]++;
}

/* One stable counting-sort pass on digit p: count, exclusive scan, scatter
 * in index order (localShuffle). */
void orc_counting_pass(const orc_elem* src, orc_elem* dst, uint64_t n,
                       uint32_t p, uint32_t d, uint32_t key_bits) {
  uint64_t nb = 1ULL << d;
  uint64_t* starts = (uint64_t*)calloc(nb, sizeof(uint64_t));
  for (uint64_t i = 0; i < n; i++) starts[digit_of(src[i].key, p, d, key_bits)]++;
  uint64_t sum = 0;
  for (uint64_t b = 0; b < nb; b++) { uint64_t c = starts[b]; starts[b] = sum; sum += c; }
  for (uint64_t i = 0; i < n; i++) dst[starts[digit_of(src[i].key, p, d, key_bits)]++] = src[i];
  free(starts);
}

/* mySort at one rank: every digit pass A->B, then the buffers swap roles;
 * the result ends in A (copied back if the pass count is odd). */
void orc_lsd_sort(orc_elem* A, orc_elem* B, uint64_t n, uint32_t key_bits,
                  uint32_t d) {
  unsigned P = n_digits(key_bits, d);
  orc_elem *src = A, *dst = B;
  for (unsigned p = 0; p < P; p++) {
    orc_counting_pass(src, dst, n, p, d, key_bits);
    orc_elem* t = src; src = dst; dst = t;
  }
  if (src != A) memcpy(A, src, n * sizeof(orc_elem));
}

/* std::stable_sort by (masked) key, as bottom-up merge sort. */
void orc_stable_sort(orc_elem* a, uint64_t n, uint32_t key_bits) {
  uint64_t m = key_mask(key_bits);
  orc_elem* tmp = (orc_elem*)malloc((n ? n : 1) * sizeof(orc_elem));
  orc_elem *src = a, *dst = tmp;
  for (uint64_t w = 1; w < n; w *= 2) {
    for (uint64_t lo = 0; lo < n; lo += 2 * w) {
      uint64_t mid = lo + w < n ? lo + w : n, hi = lo + 2 * w < n ? lo + 2 * w : n;
      uint64_t i = lo, j = mid, k = lo;
      while (i < mid && j < hi)
        dst[k++] = ((src[j].key & m) < (src[i].key & m)) ? src[j++] : src[i++];
      while (i < mid) dst[k++] = src[i++];
      while (j < hi) dst[k++] = src[j++];
    }
    orc_elem* t = src; src = dst; dst = t;
  }
  if (src != a) memcpy(a, src, n * sizeof(orc_elem));
  free(tmp);
}

/* O(n) check that out == stable_sort(input) by (masked) key, given that the
 * input's val fields are val_base + original index and in_keys[i] is the
 * input key of original index i.  Returns -1 when all hold, otherwise the
 * first failing position; *why gets 1 = val out of range, 2 = duplicate val
 * (not a permutation), 3 = key does not match its source, 4 = keys out of
 * order, 5 = equal keys out of index order (unstable). */
int64_t orc_check(const orc_elem* out, const uint64_t* in_keys, uint64_t n,
                  uint64_t val_base, uint32_t key_bits, int32_t* why) {
  uint64_t m = key_mask(key_bits);
  uint8_t* seen = (uint8_t*)calloc((n + 7) / 8 + 1, 1);
  int64_t bad = -1;
  *why = 0;
  for (uint64_t i = 0; i < n; i++) {
    uint64_t v = out[i].val - val_base;
    int code = 0;
    if (out[i].val < val_base || v >= n) code = 1;
    else if (seen[v >> 3] & (1u << (v & 7))) code = 2;
    else if (in_keys[v] != out[i].key) code = 3;
    else if (i > 0 && (out[i - 1].key & m) > (out[i].key & m)) code = 4;
    else if (i > 0 && (out[i - 1].key & m) == (out[i].key & m) &&
             out[i - 1].val >= out[i].val) code = 5;
    if (code) { bad = (int64_t)i; *why = code; break; }
    seen[v >> 3] |= (uint8_t)(1u << (v & 7));
  }
  free(seen);
  return bad;
}

/* ---------------------------------------------------------------- sha256 */

static const uint32_t SHA_K[64] = {
  0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
  0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
  0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
  0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
  0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
  0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
  0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
  0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};

static inline uint32_t rotr32(uint32_t x, unsigned r) { return (x >> r) | (x << (32 - r)); }

static void sha_block(uint32_t h[8], const uint8_t* p) {
  uint32_t w[64];
  for (int i = 0; i < 16; i++)
    w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 |
           (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
  for (int i = 16; i < 64; i++) {
    uint32_t s0 = rotr32(w[i - 15], 7) ^ rotr32(w[i - 15], 18) ^ (w[i - 15] >> 3);
    uint32_t s1 = rotr32(w[i - 2], 17) ^ rotr32(w[i - 2], 19) ^ (w[i - 2] >> 10);
    w[i] = w[i - 16] + s0 + w[i - 7] + s1;
  }
  uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], k = h[7];
  for (int i = 0; i < 64; i++) {
    uint32_t S1 = rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25);
    uint32_t ch = (e & f) ^ (~e & g);
    uint32_t t1 = k + S1 + ch + SHA_K[i] + w[i];
    uint32_t S0 = rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22);
    uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
    uint32_t t2 = S0 + mj;
    k = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
  }
  h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += k;
}

void orc_sha256(const void* data, uint64_t len, uint8_t out[32]) {
  uint32_t h[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                   0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
  const uint8_t* p = (const uint8_t*)data;
  uint64_t full = len / 64;
  for (uint64_t i = 0; i < full; i++) sha_block(h, p + 64 * i);
  uint8_t tail[128];
  uint64_t rem = len - 64 * full;
  memset(tail, 0, sizeof tail);
  memcpy(tail, p + 64 * full, rem);
  tail[rem] = 0x80;
  uint64_t tl = rem + 1 + 8 <= 64 ? 64 : 128;
  uint64_t bits = len * 8;
  for (int i = 0; i < 8; i++) tail[tl - 1 - i] = (uint8_t)(bits >> (8 * i));
  sha_block(h, tail);
  if (tl == 128) sha_block(h, tail + 64);
  for (int i = 0; i < 8; i++) {
    out[4 * i] = (uint8_t)(h[i] >> 24); out[4 * i + 1] = (uint8_t)(h[i] >> 16);
    out[4 * i + 2] = (uint8_t)(h[i] >> 8); out[4 * i + 3] = (uint8_t)h[i];
  }
}
