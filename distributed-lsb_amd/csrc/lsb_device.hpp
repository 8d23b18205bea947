// Device-side helpers shared by the sort, segmented-sort and workload kernels.
// CDNA4 (gfx950) only: 64-lane wavefronts, LDS, agent-scope atomics.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lsb_internal.hpp"

namespace lsb {

// One element, 16 bytes, key at offset 0 (SortElement, mpi/mpi_lsbsort.cpp:30-33).
struct __attribute__((aligned(16))) elem_t {
  uint64_t key;
  uint64_t val;
};

constexpr int kWave = 64;

// Look-back status word: 0 = tile not counted yet; otherwise bit 31 says
// the low 31 bits hold the INCLUSIVE prefix (all earlier tiles plus this
// one) rather than this tile's own AGGREGATE, and the low 31 bits are the
// count plus one (so a zero count still reads as published).
constexpr uint32_t kFlagInclusive = 0x80000000u;
constexpr uint32_t kCountMask = 0x7FFFFFFFu;

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// popcount(m & lanes below this one)
__device__ __forceinline__ uint32_t count_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint32_t ld_status(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Element loads and stores through the global address space.  Pointers the
// kernels read from memory (segment tables, range lists, runs) are generic
// to the compiler, which then emits flat instructions; those count against
// lgkmcnt as well as vmcnt, so every later wait for an LDS access would also
// wait for them to complete (a store's acknowledgement included).
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) u32x4_t g_u32x4_t;
typedef __attribute__((address_space(1))) uint64_t g_u64_t;

__device__ __forceinline__ elem_t gload(const elem_t* p) {
  const u32x4_t v = *(const g_u32x4_t*)p;
  elem_t x;
  x.key = ((uint64_t)v.y << 32) | v.x;
  x.val = ((uint64_t)v.w << 32) | v.z;
  return x;
}

__device__ __forceinline__ uint64_t gload_key(const elem_t* p) {
  return *(const g_u64_t*)&p->key;
}

__device__ __forceinline__ void gstore(elem_t* p, const elem_t& x) {
  const u32x4_t v = {(uint32_t)x.key, (uint32_t)(x.key >> 32), (uint32_t)x.val,
                     (uint32_t)(x.val >> 32)};
  *(g_u32x4_t*)p = v;
}

// Workgroup barrier that orders LDS only.  __syncthreads() also orders
// global memory, so the compiler drains every outstanding global load and
// store of the wave at it; the sort kernels communicate through LDS inside
// a block and through device-scope atomics (or kernel boundaries) across
// blocks, so they keep global traffic in flight across their barriers.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Inclusive prefix sum over the 64 lanes with DPP moves only (no LDS):
// row_shr 1, 2, 4, 8 scan each row of 16 lanes (lanes shifted in from
// outside the row read 0), then row_bcast:15 adds the last lane of rows 0
// and 2 into rows 1 and 3, and row_bcast:31 adds lane 31 into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, true);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, true);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}

// In-place exclusive scan of a[0..N) in LDS by a BLOCK-thread workgroup.
// Every thread of the block must call it; it begins and ends with a barrier
// so callers may write a[] right before and read it right after.  With
// ROWS > 0 the result is also added into ROWS rows of N words at `rows`
// (rows[r*N + i] += scan[i]) before the closing barrier.
//
// OWN: the caller wrote a[i] (and rows[r*N + i]) with thread i itself and
// N <= BLOCK, so each thread reads back only what it wrote and the opening
// barrier is skipped.
template <int N, int BLOCK, int ROWS = 0, bool OWN = false>
__device__ __forceinline__ void block_exclusive_scan(uint32_t* a, uint32_t* wave_sums,
                                                     uint32_t* rows = nullptr) {
  constexpr int PER = (N + BLOCK - 1) / BLOCK;
  constexpr int W = BLOCK / kWave;
  static_assert(!OWN || PER == 1, "OWN needs one element per thread");
  const int t = threadIdx.x;
  if (!OWN) lds_barrier();
  uint32_t v[PER];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = t * PER + k;
    v[k] = (i < N) ? a[i] : 0u;
    s += v[k];
  }
  const uint32_t inc = wave_inclusive_scan(s);
  if (lane_id() == kWave - 1) wave_sums[t / kWave] = inc;
  lds_barrier();
  uint32_t run = inc - s;
  const int w = t / kWave;
#pragma unroll
  for (int k = 0; k < W; ++k)
    if (k < w) run += wave_sums[k];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int i = t * PER + k;
    if (i < N) {
      a[i] = run;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) rows[r * N + i] += run;
    }
    run += v[k];
  }
  lds_barrier();
}

// Digit p of a key: bits [shift, shift + width) where the last digit is cut
// to the key width (Chapel's nBits, chpl/arkouda-radix-sort.chpl:69-78).
__device__ __forceinline__ uint32_t digit_of(uint64_t key, uint32_t shift, uint32_t mask) {
  return (uint32_t)(key >> shift) & mask;
}

// Key order.  Keys are unsigned integers of key_bits bits by default; a
// caller may declare them two's complement signed or IEEE-754 floats
// (binary32 in the low 32 bits, or binary64).  key_fwd maps a key to an
// unsigned integer of the same order, which is what every digit is taken
// from; key_inv maps it back before the element is written, so elements
// leave the sort unchanged.  sign = bit key_bits-1; fsign = sign for floats,
// else 0; full = the low key_bits bits.  Floats: a set sign bit flips every
// bit (negatives reverse), a clear one flips the sign bit alone, which puts
// -0 before +0 and orders NaNs by their bits at both ends.  All three are 0
// for unsigned keys, so both maps are the identity.  The reference carries
// an unused `negs` argument for this (chpl/arkouda-radix-sort.chpl:37-39).
__device__ __host__ __forceinline__ uint64_t key_fwd(uint64_t k, const KeyOrder& o) {
  return k ^ ((k & o.fsign) ? o.full : o.sign);
}
__device__ __host__ __forceinline__ uint64_t key_inv(uint64_t t, const KeyOrder& o) {
  return t ^ ((o.fsign && !(t & o.fsign)) ? o.full : o.sign);
}
// key_fwd on every element of a register tile (a no-op branch for unsigned).
template <int ITEMS>
__device__ __forceinline__ void keys_fwd(elem_t (&e)[ITEMS], const KeyOrder& o) {
  if (o.sign) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) e[i].key = key_fwd(e[i].key, o);
  }
}

// Adds one count for digit `d` of every valid lane.  When every active lane
// is valid and holds the same digit (sorted keys, zero high digits, heavy
// duplicates) one lane adds the whole count, which avoids a 64-way
// same-address LDS atomic; the test is one compare and one ballot.
__device__ __forceinline__ void hist_add(uint32_t* h, uint32_t d, bool valid) {
  // The active mask is taken before branching: inside the one-lane branch a
  // ballot would see that lane alone.
  const uint64_t active = __ballot(1);
  const uint32_t first = __builtin_amdgcn_readfirstlane(d);
  if (__ballot(!valid || d != first) == 0) {
    if (lane_id() == (uint32_t)__builtin_ctzll(active))
      atomicAdd(&h[first], (uint32_t)__popcll(active));
  } else if (valid) {
    atomicAdd(&h[d], 1u);
  }
}

// Counts digits 0 .. L-1 of one key per lane, all 64 lanes active and valid.
// A digit every lane of the wave shares (zero high digits, sorted or
// duplicate keys) is added once by lane 0 instead of by a 64-way
// same-address atomic.
template <int D, int L>
__device__ __forceinline__ void count_row(uint32_t* h, uint64_t key, const uint32_t (&mask)[L]) {
  constexpr int NB = 1 << D;
  if constexpr (D == 8) {
    // 8-bit digits never straddle the key's 32-bit halves: one field
    // extract each; the digits lane 0 shares with everyone are read from
    // the XOR with lane 0's key.
    const uint32_t lo = (uint32_t)key, hi = (uint32_t)(key >> 32);
    const uint32_t f_lo = __builtin_amdgcn_readfirstlane(lo);
    const uint32_t f_hi = __builtin_amdgcn_readfirstlane(hi);
    const uint32_t y_lo = lo ^ f_lo, y_hi = hi ^ f_hi;
#pragma unroll
    for (int p = 0; p < L; ++p) {
      const uint32_t sh = (uint32_t)(p * 8) & 31u;
      const uint32_t w = (uint32_t)__popc(mask[p]);
      if (__ballot(__builtin_amdgcn_ubfe(p < 4 ? y_lo : y_hi, sh, w) != 0u) == 0) {
        if (lane_id() == 0)
          atomicAdd(&h[p * NB + __builtin_amdgcn_ubfe(p < 4 ? f_lo : f_hi, sh, w)],
                    (uint32_t)kWave);
      } else {
        atomicAdd(&h[p * NB + __builtin_amdgcn_ubfe(p < 4 ? lo : hi, sh, w)], 1u);
      }
    }
  } else {
#pragma unroll
    for (int p = 0; p < L; ++p) {
      const uint32_t d = digit_of(key, p * D, mask[p]);
      const uint32_t first = __builtin_amdgcn_readfirstlane(d);
      if (__ballot(d != first) == 0) {
        if (lane_id() == 0) atomicAdd(&h[p * NB + first], (uint32_t)kWave);
      } else {
        atomicAdd(&h[p * NB + d], 1u);
      }
    }
  }
}

// ------------------------------------------------------------- pcg64 (device)
// PCG XSL-RR 128/64 with the pcg-cpp default multiplier and increment, the
// generator behind the reference keygen (mpi/mpi_lsbsort.cpp:686-690).
typedef unsigned __int128 u128;

__device__ __host__ __forceinline__ u128 pcg_mul() {
  return ((u128)2549297995355413924ULL << 64) | (u128)4865540595714422341ULL;
}
__device__ __host__ __forceinline__ u128 pcg_inc() {
  return ((u128)6364136223846793005ULL << 64) | (u128)1442695040888963407ULL;
}
__device__ __host__ __forceinline__ uint64_t pcg_output(u128 s) {
  const uint64_t x = (uint64_t)(s >> 64) ^ (uint64_t)s;
  const unsigned r = (unsigned)(s >> 122);
  return (x >> r) | (x << ((64 - r) & 63));
}
__device__ __host__ __forceinline__ u128 pcg_seed_state(uint64_t seed) {
  return ((u128)seed + pcg_inc()) * pcg_mul() + pcg_inc();
}
// Jump an LCG state ahead by delta steps in O(log delta) (F. Brown,
// "Random Number Generation with Arbitrary Strides", 1994).
__device__ __host__ inline u128 pcg_advance(u128 s, uint64_t delta) {
  u128 acc_mul = 1, acc_add = 0, cur_mul = pcg_mul(), cur_add = pcg_inc();
  while (delta) {
    if (delta & 1) {
      acc_mul *= cur_mul;
      acc_add = acc_add * cur_mul + cur_add;
    }
    cur_add = (cur_mul + 1) * cur_add;
    cur_mul *= cur_mul;
    delta >>= 1;
  }
  return acc_mul * s + acc_add;
}

}  // namespace lsb
