// Kernels of the stable LSD radix sort (one rank's mySort,
// mpi/mpi_lsbsort.cpp:591-597) for gfx950.
//
//   k_zero        clears the histogram counters before a sort.
//   k_histogram   ONE read pass over the keys that histograms every digit at
//                 once (the reference re-counts each digit inside every pass,
//                 localShuffle mpi/mpi_lsbsort.cpp:237-241).  LDS counters with
//                 64-lane atomics; a wave whose lanes share one digit adds once.
//   k_scan_plan   exclusive scan of every digit's counts (localShuffle 243-250 +
//                 exclusiveScan 397-426 at R = 1), detection of single-bucket
//                 digits, and the device-side pass plan (which passes run, in
//                 which direction) so the host never waits on the device.
//   k_onesweep    one stable scatter pass (localShuffle 252-258 + the global
//                 placement of globalShuffle 539-588): a tile of elements is
//                 ranked inside each wavefront with ballots, reordered by digit
//                 in LDS, given its global bucket offsets by decoupled
//                 look-back over earlier tiles, and written out in runs.
//   k_copy_back   moves the result from B to A when the executed pass count is
//                 odd (shmem/shmem_lsbsort.cpp:477-481 keeps it even instead).
#include "lsb_device.hpp"
#include "lsb_internal.hpp"
#include "lsb_onesweep.hpp"

namespace lsb {

// ----------------------------------------------------------------- k_zero
__global__ __launch_bounds__(256) void k_zero(uint32_t* __restrict__ p, size_t words) {
  for (size_t i = blockIdx.x * 256 + threadIdx.x; i < words; i += (size_t)gridDim.x * 256)
    p[i] = 0u;
}

void launch_zero(uint32_t* p, size_t words, hipStream_t s) {
  if (!words) return;
  size_t blocks = (words + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, s, p, words);
}

// ------------------------------------------------------------ k_histogram
// The array is cut into kRegions equal regions of `region_elems` elements
// (whole tiles); the scatter passes give each region its own look-back
// chain.  Each block owns one contiguous chunk inside one region and adds
// its counts for digits 0 .. L-1 into regcount[p][region][b].  P (the digit
// count of the key) and L are template parameters so the per-key digit loop
// unrolls.  Full rounds of 256 x U keys take count_row, the chunk's tail the
// checked hist_add.
//
// With orr given (L < P: the region passes count each later digit
// themselves) the kernel also ORs every key's XOR with key 0 into orr[0..1]:
// digit p is single-bucket exactly when its bits of that OR are zero.
template <int D, int P, int L>
__global__ __launch_bounds__(256) void k_histogram(const elem_t* __restrict__ A, uint32_t n,
                                                   uint32_t key_bits, KeyOrder ko,
                                                   uint32_t* __restrict__ regcount,
                                                   uint32_t region_elems,
                                                   uint32_t blocks_per_region,
                                                   uint32_t* __restrict__ zero,
                                                   size_t zero_words, uint32_t* __restrict__ orr) {
  static_assert(L >= 1 && L <= P, "digits counted");
  constexpr int NB = 1 << D;
  constexpr int U = 8;  // keys in flight per thread
  __shared__ uint32_t h[L * NB];
  __shared__ uint32_t s_or[2];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < (uint32_t)(L * NB); i += 256) h[i] = 0u;
  if (t < 2) s_or[t] = 0u;
  uint32_t mask[L];
#pragma unroll
  for (int p = 0; p < L; ++p) {
    const uint32_t w = key_bits - p * D < (uint32_t)D ? key_bits - p * D : (uint32_t)D;
    mask[p] = (1u << w) - 1u;
  }

  // Clear the first look-back slice of the scatter passes (16-byte stores).
  {
    uint4* z = reinterpret_cast<uint4*>(zero);
    const size_t z4 = zero_words / 4;
    for (size_t i = (size_t)blockIdx.x * 256 + t; i < z4; i += (size_t)gridDim.x * 256)
      z[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  __syncthreads();

  const uint32_t region = blockIdx.x / blocks_per_region;
  const uint32_t chunk = blockIdx.x % blocks_per_region;
  const size_t r0 = (size_t)region * region_elems;
  const size_t r1 = r0 + region_elems < n ? r0 + region_elems : n;
  const size_t per = (region_elems + blocks_per_region - 1) / blocks_per_region;
  const size_t c0 = r0 + (size_t)chunk * per;
  const size_t c1 = c0 + per < r1 ? c0 + per : r1;
  const uint64_t ref = orr ? key_fwd(A[0].key, ko) : 0ull;
  uint64_t acc = 0;
  size_t base = c0;
  // Double-buffered: the next round's loads are issued before this round is
  // counted, so the counting hides under them.
  uint64_t k[U], kn[U];
  if (base + 256 * U <= c1) {
#pragma unroll
    for (int u = 0; u < U; ++u) kn[u] = A[base + (size_t)u * 256 + t].key;
  }
  for (; base + 256 * U <= c1; base += 256 * U) {
#pragma unroll
    for (int u = 0; u < U; ++u) k[u] = kn[u];
    if (base + 2 * 256 * U <= c1) {
#pragma unroll
      for (int u = 0; u < U; ++u) kn[u] = A[base + 256 * U + (size_t)u * 256 + t].key;
    }
    if (ko.sign) {
#pragma unroll
      for (int u = 0; u < U; ++u) k[u] = key_fwd(k[u], ko);
    }
    if (orr) {
#pragma unroll
      for (int u = 0; u < U; ++u) acc |= k[u] ^ ref;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) count_row<D, L>(h, k[u], mask);
  }
  for (; base < c1; base += 256) {
    const size_t i = base + t;
    const bool v = i < c1;
    const uint64_t k = v ? key_fwd(A[i].key, ko) : 0ull;
    if (v) acc |= k ^ ref;
#pragma unroll
    for (int p = 0; p < L; ++p) hist_add(h + p * NB, digit_of(k, p * D, mask[p]), v);
  }
  if (orr) {
    if ((uint32_t)acc) atomicOr(&s_or[0], (uint32_t)acc);
    if ((uint32_t)(acc >> 32)) atomicOr(&s_or[1], (uint32_t)(acc >> 32));
  }
  __syncthreads();
  if (orr && t < 2 && s_or[t]) atomicOr(&orr[t], s_or[t]);
  if (c0 >= c1) return;
  for (uint32_t i = t; i < (uint32_t)(L * NB); i += 256) {
    const uint32_t c = h[i];
    const uint32_t p = i / NB, b = i % NB;
    if (c) atomicAdd(&regcount[((size_t)p * kRegions + region) * NB + b], c);
  }
}

template <int D, int P, int L>
static void launch_hist(const elem_t* a, uint32_t n, uint32_t key_bits, const KeyOrder& ko,
                        uint32_t* regcount, uint32_t region_elems, uint32_t per_region,
                        uint32_t regions, uint32_t* zero, size_t zero_words, uint32_t* orr,
                        hipStream_t s) {
  hipLaunchKernelGGL((k_histogram<D, P, L>), dim3(per_region * regions), dim3(256), 0, s, a, n,
                     key_bits, ko, regcount, region_elems, per_region, zero, zero_words, orr);
}

// Blocks per region of the histogram kernels: at least 8 rounds of 256 x 8
// keys per block, and at most two blocks per CU over the whole grid: the
// 2^30-key histogram on 256 CUs takes 2.82 ms with two long-running blocks
// per CU, 2.90 ms with four and 3.01 ms with eight (one per CU: 2.80 ms).
static uint32_t hist_blocks_per_region(uint32_t region_elems, uint32_t regions = kRegions) {
  static const uint32_t blocks = [] {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cus <= 0) {
      (void)hipGetLastError();
      cus = 256;
    }
    return (uint32_t)(2 * cus);
  }();
  const uint32_t cap = blocks / regions < 1 ? 1u : blocks / regions;
  size_t per_region = ((size_t)region_elems + 256 * 8 * 8 - 1) / (256 * 8 * 8);
  if (per_region < 1) per_region = 1;
  if (per_region > cap) per_region = cap;
  return (uint32_t)per_region;
}

hipError_t launch_histogram(uint32_t d, const void* A, uint32_t n, uint32_t key_bits,
                            const KeyOrder& ko, uint32_t passes, uint32_t counted,
                            uint32_t* regcount, uint32_t region_elems, uint32_t regions,
                            uint32_t* zero, size_t zero_words, uint32_t* orr, hipStream_t s) {
  if (regions < 1 || regions > (uint32_t)kRegions) return hipErrorInvalidValue;
  const uint32_t pr = hist_blocks_per_region(region_elems, regions);
  const elem_t* a = static_cast<const elem_t*>(A);
  // counted == passes: every digit (orr unused); counted == kLeanDigits with
  // orr: the lean form for the region passes.
  const bool lean = counted < passes;
  if (lean && (counted != (uint32_t)kLeanDigits || !orr)) return hipErrorInvalidValue;
#define LSB_HIST_CASE(DD, PP)                                                                    \
  case PP:                                                                                     \
    if (lean)                                                                                  \
      launch_hist<DD, PP, (PP > kLeanDigits ? kLeanDigits : PP)>(                              \
          a, n, key_bits, ko, regcount, region_elems, pr, regions, zero, zero_words, orr, s);  \
    else                                                                                       \
      launch_hist<DD, PP, PP>(a, n, key_bits, ko, regcount, region_elems, pr, regions, zero,   \
                              zero_words, orr, s);                                             \
    break;
  if (d == 8) {
    switch (passes) {
      LSB_HIST_CASE(8, 1) LSB_HIST_CASE(8, 2) LSB_HIST_CASE(8, 3) LSB_HIST_CASE(8, 4)
      LSB_HIST_CASE(8, 5) LSB_HIST_CASE(8, 6) LSB_HIST_CASE(8, 7) LSB_HIST_CASE(8, 8)
      default: return hipErrorInvalidValue;
    }
  } else if (d == 11) {
    switch (passes) {
      LSB_HIST_CASE(11, 1) LSB_HIST_CASE(11, 2) LSB_HIST_CASE(11, 3)
      LSB_HIST_CASE(11, 4) LSB_HIST_CASE(11, 5) LSB_HIST_CASE(11, 6)
      default: return hipErrorInvalidValue;
    }
  } else {
    return hipErrorInvalidValue;
  }
#undef LSB_HIST_CASE
  return hipGetLastError();
}

// ----------------------------------------------------------- k_plan_first
// Lean form, after the histogram: the pass plan and, when the digits counted
// up front are all single-bucket, the first executed digit's per-region
// counts (one more read of the keys).  Every block derives the plan itself
// from the two OR words (digit p runs unless its bits of the OR are zero,
// or always with LSB_NO_SKIP); block 0 writes it, all blocks clear the
// tickets and the per-region counters the scatter passes fill, and the
// extra count runs only when it is needed.  One launch does what k_scan_plan
// and a separate count would.
template <int D>
__global__ __launch_bounds__(256) void k_plan_first(const elem_t* __restrict__ A, uint32_t n,
                                                    uint32_t key_bits, KeyOrder ko,
                                                    uint32_t* __restrict__ regcount,
                                                    uint32_t region_elems,
                                                    uint32_t blocks_per_region,
                                                    const uint32_t* __restrict__ orr,
                                                    uint32_t* __restrict__ plan,
                                                    uint32_t* __restrict__ tickets,
                                                    uint32_t* __restrict__ regnext,
                                                    uint32_t passes, uint32_t counted,
                                                    uint32_t flags) {
  constexpr int NB = 1 << D;
  constexpr int U = 8;
  const uint32_t t = threadIdx.x;
  const uint64_t x = (uint64_t)orr[0] | ((uint64_t)orr[1] << 32);
  uint32_t f = passes;
  for (uint32_t p = passes; p-- > 0;) {
    const uint32_t shift = p * D;
    const uint32_t w = key_bits - shift < (uint32_t)D ? key_bits - shift : (uint32_t)D;
    if ((flags & 2u) || ((x >> shift) & ((1ull << w) - 1ull)) != 0) f = p;
  }
  if (blockIdx.x == 0 && t == 0) {
    uint32_t ord = 0;
    for (uint32_t p = 0; p < passes; ++p) {
      const uint32_t shift = p * D;
      const uint32_t w = key_bits - shift < (uint32_t)D ? key_bits - shift : (uint32_t)D;
      const uint32_t run =
          ((flags & 2u) || ((x >> shift) & ((1ull << w) - 1ull)) != 0) ? 1u : 0u;
      plan[4 * p + 0] = run;
      plan[4 * p + 1] = ord;
      plan[4 * p + 2] = ord & 1u;
      plan[4 * p + 3] = 0u;
      ord += run;
    }
    plan[4 * passes] = ord & 1u;
  }
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t i = (size_t)blockIdx.x * 256 + t; i < (size_t)passes * kRegions * NB; i += stride)
    regnext[i] = 0u;
  for (size_t i = (size_t)blockIdx.x * 256 + t; i < (size_t)passes * kRegions; i += stride)
    tickets[i] = 0u;
  if (f >= passes || f < counted) return;

  __shared__ uint32_t h[NB];
  for (uint32_t i = t; i < (uint32_t)NB; i += 256) h[i] = 0u;
  __syncthreads();
  const uint32_t shift = f * D;
  const uint32_t w = key_bits - shift < (uint32_t)D ? key_bits - shift : (uint32_t)D;
  const uint32_t mask = (1u << w) - 1u;
  const uint32_t region = blockIdx.x / blocks_per_region;
  const uint32_t chunk = blockIdx.x % blocks_per_region;
  const size_t r0 = (size_t)region * region_elems;
  const size_t r1 = r0 + region_elems < n ? r0 + region_elems : n;
  const size_t per = (region_elems + blocks_per_region - 1) / blocks_per_region;
  const size_t c0 = r0 + (size_t)chunk * per;
  const size_t c1 = c0 + per < r1 ? c0 + per : r1;
  size_t base = c0;
  for (; base + 256 * U <= c1; base += 256 * U) {
    uint64_t k[U];
#pragma unroll
    for (int u = 0; u < U; ++u) k[u] = key_fwd(A[base + (size_t)u * 256 + t].key, ko);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t d = digit_of(k[u], shift, mask);
      const uint32_t first = __builtin_amdgcn_readfirstlane(d);
      if (__ballot(d != first) == 0) {
        if (lane_id() == 0) atomicAdd(&h[first], (uint32_t)kWave);
      } else {
        atomicAdd(&h[d], 1u);
      }
    }
  }
  for (; base < c1; base += 256) {
    const size_t i = base + t;
    const bool v = i < c1;
    const uint64_t k = v ? key_fwd(A[i].key, ko) : 0ull;
    hist_add(h, digit_of(k, shift, mask), v);
  }
  __syncthreads();
  if (c0 >= c1) return;
  for (uint32_t i = t; i < (uint32_t)NB; i += 256) {
    const uint32_t c = h[i];
    if (c) atomicAdd(&regcount[((size_t)f * kRegions + region) * NB + i], c);
  }
}

hipError_t launch_plan_first(uint32_t d, const void* A, uint32_t n, uint32_t key_bits,
                             const KeyOrder& ko, uint32_t* regcount, uint32_t region_elems,
                             const uint32_t* orr, uint32_t* plan, uint32_t* tickets,
                             uint32_t* regnext, uint32_t passes, uint32_t counted,
                             uint32_t flags, hipStream_t s) {
  if (d != 8) return hipErrorInvalidValue;
  const uint32_t pr = hist_blocks_per_region(region_elems);
  hipLaunchKernelGGL(k_plan_first<8>, dim3(pr * kRegions), dim3(256), 0, s,
                     static_cast<const elem_t*>(A), n, key_bits, ko, regcount, region_elems, pr,
                     orr, plan, tickets, regnext, passes, counted, flags);
  return hipGetLastError();
}

// ------------------------------------------------------------ k_scan_plan
// Global count of bucket b = sum over regions; exclusive scan per digit;
// single-bucket detection; plan; clears the tickets and the per-region
// counters the scatter passes fill for their successors.  The lean form
// (region passes) plans with k_plan_first instead.
template <int NB>
__global__ __launch_bounds__(1024) void k_scan_plan(const uint32_t* __restrict__ regcount,
                                                    uint32_t* __restrict__ starts,
                                                    uint32_t* __restrict__ plan,
                                                    uint32_t* __restrict__ tickets,
                                                    uint32_t* __restrict__ regnext, uint32_t n,
                                                    uint32_t passes, uint32_t flags) {
  __shared__ uint32_t a[NB];
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t trivial[16];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < passes * kRegions * NB; i += 1024) regnext[i] = 0u;
  for (uint32_t i = t; i < passes * kRegions; i += 1024) tickets[i] = 0u;
  for (uint32_t p = 0; p < passes; ++p) {
    int triv = 0;
    for (uint32_t i = t; i < NB; i += 1024) {
      uint32_t c = 0;
#pragma unroll
      for (int r = 0; r < kRegions; ++r) c += regcount[((size_t)p * kRegions + r) * NB + i];
      a[i] = c;
      triv |= (c == n);
    }
    triv = __syncthreads_or(triv);
    if (t == 0) trivial[p] = (uint32_t)triv;
    block_exclusive_scan<NB, 1024>(a, wsum);
    for (uint32_t i = t; i < NB; i += 1024) starts[p * NB + i] = a[i];
    __syncthreads();
  }
  if (t == 0) {
    uint32_t ord = 0;
    for (uint32_t p = 0; p < passes; ++p) {
      const uint32_t run = (flags & 2u) ? 1u : (trivial[p] ? 0u : 1u);
      plan[4 * p + 0] = run;
      plan[4 * p + 1] = ord;
      plan[4 * p + 2] = ord & 1u;
      plan[4 * p + 3] = 0u;
      ord += run;
    }
    plan[4 * passes] = ord & 1u;
  }
}

hipError_t launch_scan_plan(uint32_t d, const uint32_t* regcount, uint32_t* starts,
                            uint32_t* plan, uint32_t* tickets, uint32_t* regnext, uint32_t n,
                            uint32_t passes, uint32_t flags, hipStream_t s) {
  if (d == 8)
    hipLaunchKernelGGL(k_scan_plan<256>, dim3(1), dim3(1024), 0, s, regcount, starts, plan,
                       tickets, regnext, n, passes, flags);
  else if (d == 11)
    hipLaunchKernelGGL(k_scan_plan<2048>, dim3(1), dim3(1024), 0, s, regcount, starts, plan,
                       tickets, regnext, n, passes, flags);
  else
    return hipErrorInvalidValue;
  return hipSuccess;
}

// -------------------------------------------------------------- k_onesweep
// One launch per digit pass, one tile per block.  Tile ids come from a
// ticket, so a tile only ever waits on tiles whose blocks are already
// running: look-back cannot deadlock whatever order the hardware dispatches
// blocks in.
template <int D, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_onesweep(PassArgs args) {
  constexpr int NB = 1 << D;
  __shared__ TileSmem<D, BLOCK, ITEMS> sm;
  const uint32_t* plan = args.plan;
  const uint32_t pass = args.pass;
  if (!plan[4 * pass]) return;  // single-bucket digit: identity pass
  const uint32_t ord = plan[4 * pass + 1];
  const bool src_b = plan[4 * pass + 2] != 0;
  const elem_t* src = static_cast<const elem_t*>(src_b ? args.B : args.A);
  elem_t* dst = static_cast<elem_t*>(src_b ? args.A : args.B);
  const size_t slice = (size_t)args.ntiles * NB;
  uint32_t* st_cur = args.status + (size_t)(ord & 1u) * slice;
  uint32_t* st_next = args.status + (size_t)((ord + 1u) & 1u) * slice;
  if (threadIdx.x == 0) sm.scalar[0] = atomicAdd(&args.tickets[pass], 1u);
  __syncthreads();
  const uint32_t tile = sm.scalar[0];
  onesweep_tile<D, BLOCK, ITEMS>(sm, src, dst, args.n, tile, st_cur, st_next, args.starts,
                                 args.shift, args.mask, args.ko);
}

// Region form: a persistent grid of kRegions groups of blocks; group r takes
// tickets for the tiles of region r only and runs that region's look-back
// chain, so kRegions chains advance side by side.  Each region's bucket
// bases come from the per-region counts of the previous executed pass (or of
// the histogram kernel for the first one); with NEXT the pass itself counts,
// per region of its output, the digits of the next executed pass.  Without
// NEXT every pass's per-region counts come from the histogram kernel, which
// then counts every digit up front: the form for wide digits, whose
// per-region next-pass counters (kRegions x 2^d words) would not fit in LDS
// beside the tile.  That is exact only while the per-region counts do not
// depend on the order a pass leaves behind: with one region (R = 1, one
// look-back chain over the whole array).
template <int D, int BLOCK, int ITEMS, bool NEXT = true, int R = kRegions>
__global__ __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(ITEMS <= 8 ? 3 : 2))) void k_onesweep_rg(PassArgs args) {
  constexpr int NB = 1 << D;
  static_assert(kRegions == kRegionsDev, "region count");
  static_assert(NEXT || R == 1, "counts from the histogram hold for one region only");
  __shared__ TileSmem<D, BLOCK, ITEMS> sm;
  __shared__ uint32_t s_rbase[NB];
  __shared__ uint32_t s_hnext[NEXT ? kRegions * NB : 1];
  // Per-wave counters get their own LDS when a block fills the CU alone
  // (512+ threads) and 8-bit digits leave room; otherwise they live in the
  // staging buffer (so two small blocks still fit).
  constexpr bool kOwnCounters = BLOCK >= 512 && D <= 8;
  __shared__ uint32_t s_whist[kOwnCounters ? (BLOCK / kWave) * NB : 1];
  const uint32_t* plan = args.plan;
  const uint32_t pass = args.pass;
  if (!plan[4 * pass]) return;
  const uint32_t ord = plan[4 * pass + 1];
  const bool src_b = plan[4 * pass + 2] != 0;
  const elem_t* src = static_cast<const elem_t*>(src_b ? args.B : args.A);
  elem_t* dst = static_cast<elem_t*>(src_b ? args.A : args.B);
  const size_t slice = (size_t)args.ntiles * NB;
  uint32_t* st_cur = args.status + (size_t)(ord & 1u) * slice;
  uint32_t* st_next = args.status + (size_t)((ord + 1u) & 1u) * slice;
  const uint32_t region = blockIdx.x % (uint32_t)R;
  const uint32_t t_begin = region * args.region_tiles;
  if (t_begin >= args.ntiles) return;
  const uint32_t t_count =
      args.ntiles - t_begin < args.region_tiles ? args.ntiles - t_begin : args.region_tiles;
  const uint32_t region_elems = args.region_tiles * (uint32_t)(BLOCK * ITEMS);

  // the next executed pass, whose per-region counts this pass produces
  int q = -1;
  if (NEXT)
    for (uint32_t p2 = pass + 1; p2 < args.passes; ++p2)
      if (plan[4 * p2]) {
        q = (int)p2;
        break;
      }
  uint32_t nshift = 0, nmask = 0;
  if (q >= 0) {
    nshift = (uint32_t)q * D;
    const uint32_t w = args.key_bits - nshift < (uint32_t)D ? args.key_bits - nshift : (uint32_t)D;
    nmask = (1u << w) - 1u;
  }
  // Bucket b of region r starts at the global start of b (exclusive scan of
  // the counts summed over regions) plus b's count in regions before r.
  const uint32_t* C =
      (!NEXT || ord == 0 ? args.regcount : args.regnext) + (size_t)pass * kRegions * NB;
  for (uint32_t b = threadIdx.x; b < (uint32_t)NB; b += BLOCK) {
    uint32_t before = 0, all = 0;
#pragma unroll
    for (uint32_t r = 0; r < (uint32_t)kRegions; ++r) {
      const uint32_t c = C[(size_t)r * NB + b];
      before += r < region ? c : 0u;
      all += c;
    }
    s_rbase[b] = before;
    sm.start[b] = all;
  }
  block_exclusive_scan<NB, BLOCK>(sm.start, sm.wsum);
  for (uint32_t b = threadIdx.x; b < (uint32_t)NB; b += BLOCK) s_rbase[b] += sm.start[b];
  if (NEXT)
    for (uint32_t i = threadIdx.x; i < (uint32_t)(kRegions * NB); i += BLOCK) s_hnext[i] = 0u;
  if (kOwnCounters)
    for (uint32_t i = threadIdx.x; i < (uint32_t)((BLOCK / kWave) * NB); i += BLOCK) s_whist[i] = 0u;
  uint32_t* ticket = &args.tickets[pass * kRegions + region];
  if (threadIdx.x == 0) sm.scalar[0] = atomicAdd(ticket, 1u);
  lds_barrier();
  // With its own counters a block leaves the previous tile in LDS until the
  // next tile's rank is done, so it counts that tile for the next pass while
  // the next tile's loads are in flight instead of in the store phase.
  constexpr bool kCountLater = kOwnCounters;
  const bool count_later = kCountLater && q >= 0;
  uint32_t prev_n = 0;  // elements of the tile this block finished last
  while (true) {
    const uint32_t local = sm.scalar[0];
    if (local >= t_count) break;
    const uint32_t tile = t_begin + local;
    elem_t e[ITEMS];
    tile_load<BLOCK, ITEMS>(e, src, args.n, tile);
    if (count_later && prev_n)
      count_run<BLOCK, ITEMS, NB>(sm.elems, sm.grs, s_hnext, prev_n, args.shift, args.mask,
                                  nshift, nmask);
    keys_fwd<ITEMS>(e, args.ko);
    onesweep_tile_region<D, BLOCK, ITEMS>(sm, e, dst, args.n, tile, local == 0, st_cur, st_next,
                                          s_rbase, args.shift, args.mask, ticket,
                                          q >= 0 ? s_hnext : nullptr, nshift, nmask,
                                          region_elems, args.ko,
                                          kOwnCounters ? s_whist : nullptr, count_later);
    const uint32_t tb = tile * (uint32_t)(BLOCK * ITEMS);
    prev_n = args.n - tb < (uint32_t)(BLOCK * ITEMS) ? args.n - tb : (uint32_t)(BLOCK * ITEMS);
    lds_barrier();
  }
  if (count_later && prev_n) {
    count_run<BLOCK, ITEMS, NB>(sm.elems, sm.grs, s_hnext, prev_n, args.shift, args.mask, nshift,
                                nmask);
    lds_barrier();
  }
  if (NEXT && q >= 0) {
    uint32_t* out = args.regnext + (size_t)q * kRegions * NB;
    for (uint32_t i = threadIdx.x; i < (uint32_t)(kRegions * NB); i += BLOCK) {
      const uint32_t c = s_hnext[i];
      if (c) atomicAdd(&out[i], c);
    }
  }
}

template <int D, int BLOCK, int ITEMS, bool NEXT = true, int R = kRegions>
static void launch_onesweep_rg(const PassArgs& a, hipStream_t s) {
  // resident blocks per region, found once (thread-safe static init)
  static const uint32_t per_region = [] {
    int dev = 0, cus = 0, occ = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ,
                                                     k_onesweep_rg<D, BLOCK, ITEMS, NEXT, R>,
                                                     BLOCK, 0) != hipSuccess) {
      (void)hipGetLastError();
      cus = 256;
      occ = 1;
    }
    const uint32_t pr = (uint32_t)((cus * (occ > 0 ? occ : 1)) / R);
    return pr < 1 ? 1u : pr;
  }();
  const uint32_t pr = a.region_tiles < per_region ? a.region_tiles : per_region;
  hipLaunchKernelGGL((k_onesweep_rg<D, BLOCK, ITEMS, NEXT, R>), dim3(pr * R), dim3(BLOCK), 0, s,
                     a);
}

template <int D, int BLOCK, int ITEMS>
static void launch_onesweep(const PassArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_onesweep<D, BLOCK, ITEMS>), dim3(a.ntiles), dim3(BLOCK), 0, s, a);
}

// Variants are (digit width, tile, block, form, launcher); form 3 is the
// region form, 1 the persistent form with one look-back chain (every pass's
// counts from the histogram kernel), 0 the one-tile-per-block form.  The first one per digit
// width is the fallback for that width when no tuned choice applies.
static const SortVariant kVariants[] = {
    {8, 1024 * 8, 1024, 3, &launch_onesweep_rg<8, 1024, 8>},
    {8, 256 * 16, 256, 3, &launch_onesweep_rg<8, 256, 16>},
    {8, 256 * 8, 256, 3, &launch_onesweep_rg<8, 256, 8>},
    {8, 256 * 16, 256, 0, &launch_onesweep<8, 256, 16>},
    {8, 512 * 16, 512, 0, &launch_onesweep<8, 512, 16>},
    {8, 256 * 8, 256, 0, &launch_onesweep<8, 256, 8>},
    {11, 512 * 12, 512, 1, &launch_onesweep_rg<11, 512, 12, false, 1>},
    {11, 512 * 12, 512, 0, &launch_onesweep<11, 512, 12>},
};

const SortVariant* sort_variants(int* count) {
  *count = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
  return kVariants;
}

// ----------------------------------------------------------- k_small_sort
// Whole sort of one array of n <= BLOCK*ITEMS elements in one block (LDS).
template <int D, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_small_sort(elem_t* A, uint32_t n, uint32_t key_bits,
                                                      KeyOrder ko) {
  __shared__ TileSmem<D, BLOCK, ITEMS> sm;
  small_sort<D, BLOCK, ITEMS>(sm, A, n, key_bits, ko);
}

uint32_t small_sort_capacity() { return 512 * 16; }

hipError_t launch_small_sort(void* A, uint32_t n, uint32_t key_bits, const KeyOrder& ko,
                             hipStream_t s) {
  if (n > small_sort_capacity()) return hipErrorInvalidValue;
  if (n <= 256 * 8)
    hipLaunchKernelGGL((k_small_sort<8, 256, 8>), dim3(1), dim3(256), 0, s,
                       static_cast<elem_t*>(A), n, key_bits, ko);
  else
    hipLaunchKernelGGL((k_small_sort<8, 512, 16>), dim3(1), dim3(512), 0, s,
                       static_cast<elem_t*>(A), n, key_bits, ko);
  return hipGetLastError();
}

// ------------------------------------------------------------ k_copy_back
__global__ __launch_bounds__(256) void k_copy_back(elem_t* __restrict__ A,
                                                   const elem_t* __restrict__ B, uint32_t n,
                                                   const uint32_t* __restrict__ flag) {
  if (!*flag) return;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    A[i] = B[i];
}

void launch_copy_back(void* A, const void* B, uint32_t n, const uint32_t* flag,
                      hipStream_t s) {
  size_t blocks = ((size_t)n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_copy_back, dim3((unsigned)blocks), dim3(256), 0, s,
                     static_cast<elem_t*>(A), static_cast<const elem_t*>(B), n, flag);
}

}  // namespace lsb
