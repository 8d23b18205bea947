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
                                 args.shift, args.mask);
}

// Persistent form: a grid that fits the device at once; each block keeps
// taking tickets, and takes its NEXT ticket and issues that tile's loads
// before it ranks, looks back and stores the current tile, so the loads of
// one tile are in flight under the work of the previous one.  A block waits
// only on tiles with smaller tickets, which running blocks hold and finish
// without waiting on later tiles, so the loop always drains.
template <int D, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_onesweep_pf(PassArgs args) {
  constexpr int NB = 1 << D;
  __shared__ TileSmem<D, BLOCK, ITEMS> sm;
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
  const uint32_t ntiles = args.ntiles;
  uint32_t* ticket = &args.tickets[pass];

  if (threadIdx.x == 0) sm.scalar[0] = atomicAdd(ticket, 1u);
  __syncthreads();
  uint32_t tile = sm.scalar[0];
  if (tile >= ntiles) return;
  elem_t e[ITEMS];
  tile_load<BLOCK, ITEMS>(e, src, args.n, tile);
  while (true) {
    __syncthreads();  // every thread has read scalar[] and finished with sm
    if (threadIdx.x == 0) sm.scalar[1] = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t next = sm.scalar[1];
    elem_t en[ITEMS];
    if (next < ntiles) tile_load<BLOCK, ITEMS>(en, src, args.n, next);
    onesweep_tile_regs<D, BLOCK, ITEMS>(sm, e, dst, args.n, tile, st_cur, st_next, args.starts,
                                        args.shift, args.mask);
    if (next >= ntiles) break;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) e[i] = en[i];
    tile = next;
  }
}

template <int D, int BLOCK, int ITEMS>
static void launch_onesweep_pf(const PassArgs& a, hipStream_t s) {
  static uint32_t grid = 0;
  if (!grid) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_onesweep_pf<D, BLOCK, ITEMS>, BLOCK,
                                                     0) != hipSuccess) {
      (void)hipGetLastError();
      cus = 256;
      per = 1;
    }
    grid = (uint32_t)(cus * (per > 0 ? per : 1));
  }
  const uint32_t g = a.ntiles < grid ? a.ntiles : grid;
  hipLaunchKernelGGL((k_onesweep_pf<D, BLOCK, ITEMS>), dim3(g), dim3(BLOCK), 0, s, a);
}

// Split-staging form (keys, then values, through an 8-byte buffer).
template <int D, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_onesweep_split(PassArgs args) {
  constexpr int NB = 1 << D;
  __shared__ TileSmemSplit<D, BLOCK, ITEMS> sm;
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
  if (threadIdx.x == 0) sm.scalar[0] = atomicAdd(&args.tickets[pass], 1u);
  __syncthreads();
  const uint32_t tile = sm.scalar[0];
  elem_t e[ITEMS];
  tile_load<BLOCK, ITEMS>(e, src, args.n, tile);
  onesweep_tile_split<D, BLOCK, ITEMS>(sm, e, dst, args.n, tile, st_cur, st_next, args.starts,
                                       args.shift, args.mask);
}

template <int D, int BLOCK, int ITEMS>
static void launch_onesweep_split(const PassArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_onesweep_split<D, BLOCK, ITEMS>), dim3(a.ntiles), dim3(BLOCK), 0, s, a);
}

template <int D, int BLOCK, int ITEMS>
static void launch_onesweep(const PassArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((k_onesweep<D, BLOCK, ITEMS>), dim3(a.ntiles), dim3(BLOCK), 0, s, a);
}

// Variants are (digit width, tile, block, persistent-prefetch?).  The
// first one per digit width is the default for that width.
static const SortVariant kVariants[] = {
    {8, 256 * 16, 256, 0, &launch_onesweep<8, 256, 16>},
    {8, 256 * 12, 256, 0, &launch_onesweep<8, 256, 12>},
    {8, 512 * 16, 512, 0, &launch_onesweep<8, 512, 16>},
    {8, 256 * 8, 256, 0, &launch_onesweep<8, 256, 8>},
    {11, 512 * 12, 512, 0, &launch_onesweep<11, 512, 12>},
    {8, 256 * 16, 256, 2, &launch_onesweep_split<8, 256, 16>},
    {8, 512 * 12, 512, 2, &launch_onesweep_split<8, 512, 12>},
    {8, 256 * 12, 256, 2, &launch_onesweep_split<8, 256, 12>},
    {8, 256 * 12, 256, 1, &launch_onesweep_pf<8, 256, 12>},
    {8, 256 * 8, 256, 1, &launch_onesweep_pf<8, 256, 8>},
    {11, 512 * 8, 512, 1, &launch_onesweep_pf<11, 512, 8>},
};

const SortVariant* sort_variants(int* count) {
  *count = (int)(sizeof(kVariants) / sizeof(kVariants[0]));
  return kVariants;
}

// ----------------------------------------------------------- k_small_sort
// Whole sort of one array of n <= BLOCK*ITEMS elements in one block (LDS).
template <int D, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_small_sort(elem_t* A, uint32_t n, uint32_t key_bits) {
  __shared__ TileSmem<D, BLOCK, ITEMS> sm;
  small_sort<D, BLOCK, ITEMS>(sm, A, n, key_bits);
}

uint32_t small_sort_capacity() { return 512 * 16; }

hipError_t launch_small_sort(void* A, uint32_t n, uint32_t key_bits, hipStream_t s) {
  if (n > small_sort_capacity()) return hipErrorInvalidValue;
  if (n <= 256 * 8)
    hipLaunchKernelGGL((k_small_sort<8, 256, 8>), dim3(1), dim3(256), 0, s,
                       static_cast<elem_t*>(A), n, key_bits);
  else
    hipLaunchKernelGGL((k_small_sort<8, 512, 16>), dim3(1), dim3(512), 0, s,
                       static_cast<elem_t*>(A), n, key_bits);
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
