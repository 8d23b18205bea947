// Batched segmented sort: thousands of independent arrays ("segments") of
// uneven sizes, each sorted stably by key on its own, in one launch sequence.
// The reference sorts one array per rank (mySort, mpi/mpi_lsbsort.cpp:
// 591-597); a segment here is one such array.
//
// Every kernel reads the batch shape from a device header and runs a grid
// sized to the device, not to the batch, so one captured hipGraph replays
// for any batch that fits its workspace:
//   k_seg_small     segments of at most one workgroup's capacity, sorted
//                   entirely in LDS (small_sort): one HBM read + write.
//   k_seg_zero      clears the per-segment digit counters.
//   k_seg_hist      per-segment all-digit histograms over the tiles of the
//                   large segments (+ clears look-back slice 0).
//   k_seg_scan      per-segment exclusive scans, skip mask, pass parity.
//   k_seg_onesweep  one launch per digit pass, persistent: the large
//                   segments are split whole over kRegions groups of blocks,
//                   each with its own tile tickets; look-back stays inside a
//                   segment (its first tile publishes inclusive).
//   k_seg_copy      moves a segment from B to A when it ran an odd number
//                   of passes.
#include "lsb_device.hpp"
#include "lsb_internal.hpp"
#include "lsb_onesweep.hpp"

namespace lsb {

// ----------------------------------------------------------- small segments
template <int D, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_seg_small(const SegHeader* __restrict__ hdr,
                                                     const SegSmall* __restrict__ list,
                                                     int which, uint32_t key_bits, KeyOrder ko) {
  __shared__ TileSmem<D, BLOCK, ITEMS> sm;
  // The metadata array holds the large segments, then the tiny, then the
  // small ones.
  const uint32_t count = which == 0 ? hdr->n_tiny : hdr->n_small;
  const SegSmall* l = reinterpret_cast<const SegSmall*>(
                          reinterpret_cast<const SegLarge*>(list) + hdr->n_large) +
                      (which == 0 ? 0u : hdr->n_tiny);
  for (uint32_t i = blockIdx.x; i < count; i += gridDim.x) {
    const SegSmall s = l[i];
    small_sort<D, BLOCK, ITEMS>(sm, reinterpret_cast<elem_t*>(s.a), s.n, key_bits, ko);
    __syncthreads();
  }
}

// ------------------------------------------------------------ tile map
// tile_seg[t] = the large segment that global tile t belongs to; perm lists
// the global tiles in the groups' ticket order (each segment's tiles
// consecutive and in order from its perm_pos).
__global__ __launch_bounds__(256) void k_seg_tilemap(const SegHeader* __restrict__ hdr,
                                                     const SegLarge* __restrict__ large,
                                                     uint32_t tile, uint32_t* __restrict__ tile_seg,
                                                     uint32_t* __restrict__ perm) {
  const uint32_t nl = hdr->n_large;
  for (uint32_t s = blockIdx.x; s < nl; s += gridDim.x) {
    const SegLarge L = large[s];
    const uint32_t nt = (L.n + tile - 1) / tile;
    for (uint32_t i = threadIdx.x; i < nt; i += 256) {
      tile_seg[L.tile_begin + i] = s;
      perm[L.perm_pos + i] = L.tile_begin + i;
    }
  }
}

// ------------------------------------------------------------ counters
__global__ __launch_bounds__(256) void k_seg_zero(const SegHeader* __restrict__ hdr,
                                                  uint32_t* __restrict__ counts,
                                                  uint32_t words_per_seg) {
  const size_t words = (size_t)hdr->n_large * words_per_seg;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < words; i += (size_t)gridDim.x * 256)
    counts[i] = 0u;
}

// Each block owns a contiguous range of tiles, accumulates the counts of
// digits 0 .. P-1 over consecutive tiles of one segment in LDS and flushes
// them when the segment changes (tile_seg: the tile-to-segment map).  P is a
// template parameter so the digit loop unrolls; full tiles are read in
// unchecked rounds of 256 x 8 keys counted by count_row, a segment's last
// tile key by key.
template <int D, int P>
__global__ __launch_bounds__(256) void k_seg_hist(const SegHeader* __restrict__ hdr,
                                                  const SegLarge* __restrict__ large,
                                                  uint32_t tile, uint32_t key_bits, KeyOrder ko,
                                                  uint32_t* __restrict__ counts,
                                                  uint32_t* __restrict__ status0,
                                                  const uint32_t* __restrict__ tile_seg) {
  constexpr int NB = 1 << D;
  constexpr int MAXP = (64 + D - 1) / D;
  constexpr int U = 8;
  __shared__ uint32_t h[P * NB];
  const uint32_t t = threadIdx.x;
  const uint32_t total = hdr->total_tiles;
  const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
  const uint32_t t0 = blockIdx.x * per;
  const uint32_t t1 = t0 + per < total ? t0 + per : total;
  if (t0 >= t1) return;

  // clear look-back slice 0 for these tiles
  {
    uint4* z = reinterpret_cast<uint4*>(status0 + (size_t)t0 * NB);
    const size_t z4 = (size_t)(t1 - t0) * NB / 4;
    for (size_t i = t; i < z4; i += 256) z[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  for (uint32_t i = t; i < (uint32_t)(P * NB); i += 256) h[i] = 0u;
  uint32_t mask[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const uint32_t w = key_bits - p * D < (uint32_t)D ? key_bits - p * D : (uint32_t)D;
    mask[p] = (1u << w) - 1u;
  }
  __syncthreads();

  uint32_t seg = tile_seg[t0];
  for (uint32_t tl = t0; tl < t1; ++tl) {
    const uint32_t next = tile_seg[tl];
    if (next != seg) {
      // flush the finished segment
      __syncthreads();
      for (uint32_t i = t; i < (uint32_t)(P * NB); i += 256) {
        const uint32_t c = h[i];
        if (c) atomicAdd(&counts[(size_t)seg * MAXP * NB + i], c);
        h[i] = 0u;
      }
      __syncthreads();
      seg = next;
    }
    const SegLarge L = large[seg];
    const uint32_t first = (tl - L.tile_begin) * tile;
    const uint32_t cnt = L.n - first < tile ? L.n - first : tile;
    const elem_t* a = reinterpret_cast<const elem_t*>(L.a) + first;
    if (cnt == tile && tile % (256 * U) == 0) {
      for (uint32_t i0 = 0; i0 < cnt; i0 += 256 * U) {
        uint64_t k[U];
#pragma unroll
        for (int u = 0; u < U; ++u) k[u] = gload_key(a + i0 + u * 256 + t);
        if (ko.sign) {
#pragma unroll
          for (int u = 0; u < U; ++u) k[u] = key_fwd(k[u], ko);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) count_row<D, P>(h, k[u], mask);
      }
    } else {
      for (uint32_t i0 = 0; i0 < cnt; i0 += 256) {
        const uint32_t i = i0 + t;
        const bool v = i < cnt;
        const uint64_t k = v ? key_fwd(gload_key(a + i), ko) : 0ull;
#pragma unroll
        for (int p = 0; p < P; ++p) hist_add(h + p * NB, digit_of(k, p * D, mask[p]), v);
      }
    }
  }
  __syncthreads();
  for (uint32_t i = t; i < (uint32_t)(P * NB); i += 256) {
    const uint32_t c = h[i];
    if (c) atomicAdd(&counts[(size_t)seg * MAXP * NB + i], c);
  }
}

// One block per large segment (grid-stride): in-place exclusive scan of
// every pass's counts, then the segment's pass mask and final parity.
template <int D>
__global__ __launch_bounds__(1024) void k_seg_scan(const SegHeader* __restrict__ hdr,
                                                   const SegLarge* __restrict__ large,
                                                   uint32_t* __restrict__ counts,
                                                   uint32_t* __restrict__ seg_plan,
                                                   uint32_t* __restrict__ tickets,
                                                   uint32_t passes, uint32_t flags) {
  constexpr int NB = 1 << D;
  constexpr int MAXP = (64 + D - 1) / D;
  __shared__ uint32_t a[NB];
  __shared__ uint32_t wsum[16];
  const uint32_t t = threadIdx.x;
  if (blockIdx.x == 0 && t < 8 * kRegions) tickets[t] = 0u;
  const uint32_t nlarge = hdr->n_large;
  for (uint32_t s = blockIdx.x; s < nlarge; s += gridDim.x) {
    const uint32_t n = large[s].n;
    uint32_t mask = 0;
    for (uint32_t p = 0; p < passes; ++p) {
      uint32_t* c = counts + (size_t)s * MAXP * NB + (size_t)p * NB;
      int triv = 0;
      for (uint32_t i = t; i < (uint32_t)NB; i += 1024) {
        const uint32_t v = c[i];
        a[i] = v;
        triv |= (v == n);
      }
      triv = __syncthreads_or(triv);
      if (!triv || (flags & 2u)) mask |= 1u << p;
      block_exclusive_scan<NB, 1024>(a, wsum);
      for (uint32_t i = t; i < (uint32_t)NB; i += 1024) c[i] = a[i];
      __syncthreads();
    }
    if (t == 0) {
      seg_plan[2 * s] = mask;
      seg_plan[2 * s + 1] = (uint32_t)__popc(mask) & 1u;
    }
  }
}

// --------------------------------------------------------- k_seg_onesweep
template <int D, int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_seg_onesweep(SegPassArgs args) {
  constexpr int NB = 1 << D;
  constexpr int MAXP = (64 + D - 1) / D;
  __shared__ TileSmem<D, BLOCK, ITEMS> sm;
  // Blocks that fill a CU alone (512+ threads) keep their per-wave counters
  // in LDS of their own, so a tile starts ranking without a barrier (8-bit
  // digits: the 11-bit counters would not fit beside the tile).
  constexpr bool kOwn = BLOCK >= 512 && D <= 8;
  __shared__ uint32_t s_whist[kOwn ? (BLOCK / kWave) * NB : 1];
  const uint32_t pass = args.pass;
  const size_t slice = (size_t)args.max_tiles * NB;
  // Group g = blockIdx % groups (with kRegions groups, one per XCD when
  // blocks are dispatched round-robin) works through its own segments, so a
  // segment's look-back chain and the partial lines between its tiles' runs
  // stay in one L2.
  const uint32_t group = blockIdx.x % args.hdr->groups;
  const uint32_t g_off = args.hdr->group_off[group];
  const uint32_t g_tiles = args.hdr->group_tiles[group];
  uint32_t* ticket = &args.tickets[pass * kRegions + group];
  if (threadIdx.x == 0) sm.scalar[3] = atomicAdd(ticket, 1u);
  __syncthreads();
  while (true) {
    // The next tile comes from the previous tile's early ticket; its segment
    // from the tile map (uniform loads).
    const uint32_t t = sm.scalar[3];
    if (t >= g_tiles) return;
    const uint32_t tile = args.perm[g_off + t];
    const uint32_t seg = args.tile_seg[tile];
    const uint32_t mask = args.seg_plan[2 * seg];
    if (mask & (1u << pass)) {
      const SegLarge L = args.large[seg];
      const uint32_t ord = (uint32_t)__popc(mask & ((1u << pass) - 1u));
      elem_t* a = reinterpret_cast<elem_t*>(L.a);
      elem_t* b = reinterpret_cast<elem_t*>(L.b);
      const elem_t* src = (ord & 1u) ? b : a;
      elem_t* dst = (ord & 1u) ? a : b;
      uint32_t* st_cur = args.status + (size_t)(ord & 1u) * slice + (size_t)L.tile_begin * NB;
      uint32_t* st_next =
          args.status + (size_t)((ord + 1u) & 1u) * slice + (size_t)L.tile_begin * NB;
      const uint32_t* starts = args.counts + (size_t)seg * MAXP * NB + (size_t)pass * NB;
      onesweep_tile<D, BLOCK, ITEMS>(sm, src, dst, L.n, tile - L.tile_begin, st_cur, st_next,
                                     starts, args.shift, args.mask, args.ko, ticket,
                                     kOwn ? s_whist : nullptr);
    } else if (threadIdx.x == 0) {
      sm.scalar[3] = atomicAdd(ticket, 1u);  // this segment skips the pass
    }
    __syncthreads();
  }
}

// -------------------------------------------------------------- k_seg_copy
__global__ __launch_bounds__(256) void k_seg_copy(const SegHeader* __restrict__ hdr,
                                                  const SegLarge* __restrict__ large,
                                                  const uint32_t* __restrict__ seg_plan,
                                                  const uint32_t* __restrict__ tile_seg,
                                                  uint32_t tile) {
  const uint32_t total = hdr->total_tiles;
  for (uint32_t tl = blockIdx.x; tl < total; tl += gridDim.x) {
    const uint32_t seg = tile_seg[tl];
    if (!seg_plan[2 * seg + 1]) continue;
    const SegLarge L = large[seg];
    const uint32_t first = (tl - L.tile_begin) * tile;
    const uint32_t cnt = L.n - first < tile ? L.n - first : tile;
    for (uint32_t i = threadIdx.x; i < cnt; i += 256)
      gstore(reinterpret_cast<elem_t*>(L.a) + first + i, gload(reinterpret_cast<const elem_t*>(L.b) + first + i));
  }
}

// ----------------------------------------------------------------- launch
template <int D, int BLOCK, int ITEMS>
static hipError_t seg_onesweep_launch(const SegPassArgs& a, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((k_seg_onesweep<D, BLOCK, ITEMS>), dim3(grid), dim3(BLOCK), 0, s, a);
  return hipGetLastError();
}

template <int D, int BLOCK, int ITEMS>
static uint32_t seg_onesweep_occupancy() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_seg_onesweep<D, BLOCK, ITEMS>,
                                                   BLOCK, 0) != hipSuccess) {
    (void)hipGetLastError();
    blocks = 1;
  }
  return blocks > 0 ? (uint32_t)blocks : 1u;
}

static const SegVariant kSegVariants[] = {
    {8, 1024 * 8, 1024, &seg_onesweep_launch<8, 1024, 8>, &seg_onesweep_occupancy<8, 1024, 8>},
    {8, 256 * 16, 256, &seg_onesweep_launch<8, 256, 16>, &seg_onesweep_occupancy<8, 256, 16>},
    {8, 512 * 16, 512, &seg_onesweep_launch<8, 512, 16>, &seg_onesweep_occupancy<8, 512, 16>},
    {8, 256 * 8, 256, &seg_onesweep_launch<8, 256, 8>, &seg_onesweep_occupancy<8, 256, 8>},
    {11, 512 * 12, 512, &seg_onesweep_launch<11, 512, 12>, &seg_onesweep_occupancy<11, 512, 12>},
};

const SegVariant* seg_variants(int* count) {
  *count = (int)(sizeof(kSegVariants) / sizeof(kSegVariants[0]));
  return kSegVariants;
}

uint32_t seg_tiny_capacity() { return 256 * 8; }
uint32_t seg_small_capacity() { return 512 * 16; }

hipError_t launch_seg_small(const SegHeader* hdr, const SegSmall* list,
                            uint32_t key_bits, const KeyOrder& ko, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL((k_seg_small<8, 256, 8>), dim3(grid), dim3(256), 0, s, hdr, list, 0,
                     key_bits, ko);
  hipLaunchKernelGGL((k_seg_small<8, 512, 16>), dim3(grid), dim3(512), 0, s, hdr, list, 1,
                     key_bits, ko);
  return hipGetLastError();
}

hipError_t launch_seg_prepare(uint32_t d, const SegHeader* hdr,
                              const SegLarge* large, uint32_t tile, uint32_t key_bits,
                              const KeyOrder& ko, uint32_t passes, uint32_t* counts, uint32_t* seg_plan,
                              uint32_t* tickets, uint32_t* status0, uint32_t* tile_seg,
                              uint32_t* perm, uint32_t flags, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_tilemap, dim3(grid), dim3(256), 0, s, hdr, large, tile, tile_seg,
                     perm);
  if (d == 8) {
    constexpr int MAXP = 8;
    hipLaunchKernelGGL(k_seg_zero, dim3(grid), dim3(256), 0, s, hdr, counts, MAXP * 256u);
#define LSB_SEG_HIST(PP)                                                                   \
  case PP:                                                                                 \
    hipLaunchKernelGGL((k_seg_hist<8, PP>), dim3(grid), dim3(256), 0, s, hdr, large, tile, \
                       key_bits, ko, counts, status0, tile_seg);                           \
    break;
    switch (passes) {
      LSB_SEG_HIST(1) LSB_SEG_HIST(2) LSB_SEG_HIST(3) LSB_SEG_HIST(4)
      LSB_SEG_HIST(5) LSB_SEG_HIST(6) LSB_SEG_HIST(7) LSB_SEG_HIST(8)
      default: return hipErrorInvalidValue;
    }
#undef LSB_SEG_HIST
    hipLaunchKernelGGL(k_seg_scan<8>, dim3(grid), dim3(1024), 0, s, hdr, large, counts,
                       seg_plan, tickets, passes, flags);
  } else if (d == 11) {
    constexpr int MAXP = 6;
    hipLaunchKernelGGL(k_seg_zero, dim3(grid), dim3(256), 0, s, hdr, counts, MAXP * 2048u);
#define LSB_SEG_HIST(PP)                                                                    \
  case PP:                                                                                  \
    hipLaunchKernelGGL((k_seg_hist<11, PP>), dim3(grid), dim3(256), 0, s, hdr, large, tile, \
                       key_bits, ko, counts, status0, tile_seg);                            \
    break;
    switch (passes) {
      LSB_SEG_HIST(1) LSB_SEG_HIST(2) LSB_SEG_HIST(3)
      LSB_SEG_HIST(4) LSB_SEG_HIST(5) LSB_SEG_HIST(6)
      default: return hipErrorInvalidValue;
    }
#undef LSB_SEG_HIST
    hipLaunchKernelGGL(k_seg_scan<11>, dim3(grid), dim3(1024), 0, s, hdr, large, counts,
                       seg_plan, tickets, passes, flags);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_seg_copy(const SegHeader* hdr, const SegLarge* large,
                           const uint32_t* seg_plan, const uint32_t* tile_seg, uint32_t tile,
                           uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(k_seg_copy, dim3(grid), dim3(256), 0, s, hdr, large, seg_plan, tile_seg,
                     tile);
  return hipGetLastError();
}

}  // namespace lsb
