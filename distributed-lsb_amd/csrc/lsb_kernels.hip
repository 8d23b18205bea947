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
// Adds one count for digit `d` of every valid lane.  When every valid lane
// of the wave holds the same digit (sorted keys, zero high digits, heavy
// duplicates) a single lane adds the whole count, which avoids a 64-way
// same-address LDS atomic.
__device__ __forceinline__ void hist_add(uint32_t* h, uint32_t d, bool valid) {
  const uint64_t act = __ballot(valid);
  if (act == 0) return;
  const uint32_t first = __builtin_amdgcn_readfirstlane(d);
  const uint64_t same = __ballot(valid && d == first);
  if (same == act) {
    if (valid && count_below(act) == 0) atomicAdd(&h[first], (uint32_t)__popcll(act));
  } else if (valid) {
    atomicAdd(&h[d], 1u);
  }
}

template <int D>
__global__ __launch_bounds__(256) void k_histogram(const elem_t* __restrict__ A, uint32_t n,
                                                   uint32_t key_bits, uint32_t passes,
                                                   uint32_t* __restrict__ counts,
                                                   uint32_t* __restrict__ zero,
                                                   size_t zero_words) {
  constexpr int NB = 1 << D;
  constexpr int MAXP = (64 + D - 1) / D;
  constexpr int U = 4;  // keys in flight per thread
  __shared__ uint32_t h[MAXP * NB];
  const uint32_t t = threadIdx.x;
  for (uint32_t i = t; i < passes * NB; i += 256) h[i] = 0u;

  // Clear the first look-back slice of the scatter passes (16-byte stores).
  {
    uint4* z = reinterpret_cast<uint4*>(zero);
    const size_t z4 = zero_words / 4;
    for (size_t i = (size_t)blockIdx.x * 256 + t; i < z4; i += (size_t)gridDim.x * 256)
      z[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  __syncthreads();

  const size_t step = (size_t)gridDim.x * 256 * U;
  for (size_t base = (size_t)blockIdx.x * 256 * U; base < n; base += step) {
    uint64_t k[U];
    bool v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = base + (size_t)u * 256 + t;
      v[u] = i < n;
      k[u] = v[u] ? A[i].key : 0ull;
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t shift = p * D;
        const uint32_t w = key_bits - shift < (uint32_t)D ? key_bits - shift : (uint32_t)D;
        const uint32_t dg = digit_of(k[u], shift, (1u << w) - 1u);
        hist_add(h + p * NB, dg, v[u]);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = t; i < passes * NB; i += 256) {
    const uint32_t c = h[i];
    if (c) atomicAdd(&counts[i], c);
  }
}

hipError_t launch_histogram(uint32_t d, const void* A, uint32_t n, uint32_t key_bits,
                            uint32_t passes, uint32_t* counts, uint32_t* zero,
                            size_t zero_words, hipStream_t s) {
  // Enough blocks to keep every CU streaming, few enough that the final
  // flush of LDS counters to HBM stays small.
  size_t blocks = ((size_t)n + 256 * 4 * 8 - 1) / (256 * 4 * 8);
  if (blocks < 1) blocks = 1;
  if (blocks > 2048) blocks = 2048;
  const size_t zblocks = (zero_words / 4 + 255) / 256;
  if (blocks < zblocks && zblocks <= 2048) blocks = zblocks;
  if (blocks < zblocks) blocks = 2048;
  const elem_t* a = static_cast<const elem_t*>(A);
  if (d == 8)
    hipLaunchKernelGGL(k_histogram<8>, dim3((unsigned)blocks), dim3(256), 0, s, a, n, key_bits,
                       passes, counts, zero, zero_words);
  else if (d == 11)
    hipLaunchKernelGGL(k_histogram<11>, dim3((unsigned)blocks), dim3(256), 0, s, a, n,
                       key_bits, passes, counts, zero, zero_words);
  else
    return hipErrorInvalidValue;
  return hipSuccess;
}

// ------------------------------------------------------------ k_scan_plan
template <int NB>
__global__ __launch_bounds__(1024) void k_scan_plan(const uint32_t* __restrict__ counts,
                                                    uint32_t* __restrict__ starts,
                                                    uint32_t* __restrict__ plan,
                                                    uint32_t* __restrict__ tickets, uint32_t n,
                                                    uint32_t passes, uint32_t flags) {
  __shared__ uint32_t a[NB];
  __shared__ uint32_t wsum[16];
  __shared__ uint32_t trivial[16];
  const uint32_t t = threadIdx.x;
  for (uint32_t p = 0; p < passes; ++p) {
    int triv = 0;
    for (uint32_t i = t; i < NB; i += 1024) {
      const uint32_t c = counts[p * NB + i];
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
      tickets[p] = 0u;
      ord += run;
    }
    plan[4 * passes] = ord & 1u;
  }
}

hipError_t launch_scan_plan(uint32_t d, const uint32_t* counts, uint32_t* starts,
                            uint32_t* plan, uint32_t* tickets, uint32_t n, uint32_t passes,
                            uint32_t flags, hipStream_t s) {
  if (d == 8)
    hipLaunchKernelGGL(k_scan_plan<256>, dim3(1), dim3(1024), 0, s, counts, starts, plan,
                       tickets, n, passes, flags);
  else if (d == 11)
    hipLaunchKernelGGL(k_scan_plan<2048>, dim3(1), dim3(1024), 0, s, counts, starts, plan,
                       tickets, n, passes, flags);
  else
    return hipErrorInvalidValue;
  return hipSuccess;
}

// -------------------------------------------------------------- k_onesweep
// Tile layout: wave w of the block owns the contiguous run of 64*ITEMS
// elements starting at w*64*ITEMS; its item i, lane l is element
// i*64 + l of that run.  Every global load of a wave is one contiguous 1 KiB.
// Ranking visits (wave, item, lane) in element order, so the position an
// element gets inside its bucket follows input order: the pass is stable.
template <int D, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_onesweep(PassArgs args) {
  constexpr int NB = 1 << D;
  constexpr int W = BLOCK / kWave;
  constexpr int T = BLOCK * ITEMS;
  constexpr int DPT = (NB + BLOCK - 1) / BLOCK;  // digits per thread
  static_assert(W * NB * 4 <= T * 16, "per-wave counters must fit in the tile buffer");

  // The tile buffer first holds the per-wave digit counters, then the
  // elements in digit order.
  __shared__ elem_t s_elems[T];
  __shared__ uint32_t s_start[NB];  // tile-local exclusive bucket starts
  __shared__ uint32_t s_gbase[NB];  // global index of tile-local position 0, per bucket
  __shared__ uint32_t s_wsum[W];
  __shared__ uint32_t s_tile;

  const uint32_t* plan = args.plan;
  const uint32_t pass = args.pass;
  if (!plan[4 * pass]) return;  // single-bucket digit: identity pass
  const uint32_t ord = plan[4 * pass + 1];
  const bool src_b = plan[4 * pass + 2] != 0;
  const elem_t* __restrict__ src = static_cast<const elem_t*>(src_b ? args.B : args.A);
  elem_t* __restrict__ dst = static_cast<elem_t*>(src_b ? args.A : args.B);
  const size_t slice = (size_t)args.ntiles * NB;
  uint32_t* st_cur = args.status + (size_t)(ord & 1u) * slice;
  uint32_t* st_next = args.status + (size_t)((ord + 1u) & 1u) * slice;
  const uint32_t n = args.n, shift = args.shift, mask = args.mask;

  const uint32_t t = threadIdx.x;
  const uint32_t lane = t & (kWave - 1);
  const uint32_t wave = t / kWave;
  uint32_t* whist = reinterpret_cast<uint32_t*>(s_elems);

  // Tile ids come from a ticket, so a tile only ever waits on tiles whose
  // blocks are already running: look-back cannot deadlock whatever order the
  // hardware dispatches blocks in.
  if (t == 0) s_tile = atomicAdd(&args.tickets[pass], 1u);
  for (uint32_t i = t; i < (uint32_t)(W * NB); i += BLOCK) whist[i] = 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  for (uint32_t b = t; b < (uint32_t)NB; b += BLOCK) st_next[(size_t)tile * NB + b] = 0u;

  const size_t tile_base = (size_t)tile * T;
  const size_t wbase = tile_base + (size_t)wave * (kWave * ITEMS) + lane;
  elem_t e[ITEMS];
  uint32_t dg[ITEMS], rk[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const size_t idx = wbase + (size_t)i * kWave;
    if (idx < n) e[i] = src[idx];
    else { e[i].key = 0; e[i].val = 0; }
  }

  // Wave-level ranking: peers = lanes holding the same digit, found with D
  // ballots.  The lowest peer bumps the wave's counter for the digit; every
  // peer's rank is the old counter plus the peers below it.
  uint32_t* wh = whist + wave * NB;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const bool valid = wbase + (size_t)i * kWave < n;
    const uint32_t d = digit_of(e[i].key, shift, mask);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < D; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t below = count_below(peers);
    uint32_t old = 0;
    if (valid) old = wh[d];
    if (valid && below == 0) wh[d] = old + (uint32_t)__popcll(peers);
    dg[i] = d;
    rk[i] = old + below;
  }
  __syncthreads();

  // Per digit: exclusive prefix over waves (in place) and the tile total,
  // published at once as this tile's aggregate.
  uint32_t tot[DPT];
#pragma unroll
  for (int k = 0; k < DPT; ++k) {
    const uint32_t b = t + k * BLOCK;
    tot[k] = 0;
    if (b < (uint32_t)NB) {
      uint32_t s = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t c = whist[w * NB + b];
        whist[w * NB + b] = s;
        s += c;
      }
      tot[k] = s;
      s_start[b] = s;
      st_status(st_cur + (size_t)tile * NB + b, tile == 0 ? ((s + 1u) | kFlagInclusive) : (s + 1u));
    }
  }
  block_exclusive_scan<NB, BLOCK>(s_start, s_wsum);

  uint32_t pos[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) pos[i] = s_start[dg[i]] + wh[dg[i]] + rk[i];
  __syncthreads();  // counters are dead from here: the buffer takes elements
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if (wbase + (size_t)i * kWave < n) s_elems[pos[i]] = e[i];

  // Decoupled look-back, one thread per digit: sum earlier tiles' aggregates
  // until one publishes an inclusive prefix.
#pragma unroll
  for (int k = 0; k < DPT; ++k) {
    const uint32_t b = t + k * BLOCK;
    if (b < (uint32_t)NB) {
      uint32_t prefix = 0;
      if (tile > 0) {
        int64_t j = (int64_t)tile - 1;
        while (true) {
          const uint32_t w = ld_status(st_cur + (size_t)j * NB + b);
          if (w == 0u) {
            __builtin_amdgcn_s_sleep(1);
            continue;
          }
          prefix += (w & kCountMask) - 1u;
          if (w & kFlagInclusive) break;
          --j;
        }
        st_status(st_cur + (size_t)tile * NB + b, (prefix + tot[k] + 1u) | kFlagInclusive);
      }
      s_gbase[b] = args.starts[b] + prefix - s_start[b];
    }
  }
  __syncthreads();

  // Write out in digit order: consecutive lanes write consecutive slots of
  // one bucket run.
  const uint32_t tile_n = (n - tile_base) < (size_t)T ? (uint32_t)(n - tile_base) : (uint32_t)T;
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t j = k * BLOCK + t;
    if (j < tile_n) {
      const elem_t x = s_elems[j];
      const uint32_t d = digit_of(x.key, shift, mask);
      dst[(size_t)(uint32_t)(s_gbase[d] + j)] = x;
    }
  }
}

template <int D, int BLOCK, int ITEMS>
static void launch_onesweep(const PassArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_onesweep<D, BLOCK, ITEMS>), dim3(a.ntiles), dim3(BLOCK), 0, s, a);
}

static const SortVariant kVariants[] = {
    {8, 256 * 16, 256, &launch_onesweep<8, 256, 16>},
    {8, 512 * 16, 512, &launch_onesweep<8, 512, 16>},
    {8, 256 * 8, 256, &launch_onesweep<8, 256, 8>},
    {11, 512 * 12, 512, &launch_onesweep<11, 512, 12>},
};

const SortVariant* sort_variants(int* count) {
  *count = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
  return kVariants;
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
