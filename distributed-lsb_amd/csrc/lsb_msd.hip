// MSD-first sort of the large segments of a batch.
//
// Each segment is one rank's mySort (mpi/mpi_lsbsort.cpp:591-597); its
// result must be the stable sort by key, which any stable sorting order
// produces.  Instead of the reference's fixed sequence of low-to-high digit
// passes over the whole array, a range is partitioned on its highest digit
// that still varies, with the same stable onesweep tile pass as the LSD
// sort, and each resulting bucket is either small enough to be finished
// inside one workgroup (all its lower digits sorted in LDS, one HBM read and
// write) or partitioned again on its next varying digit.  Uniform 64-bit
// keys in arrays of up to ~2^29 elements take two partition passes and the
// in-LDS finish instead of eight passes.
//
// Per level k (launch_msd_level), over that level's list of open ranges:
//   k_msd_hist     counts each range's partition digit and takes the OR and
//                  the AND of its keys (their XOR: the bits that vary), tile
//                  by tile; clears the look-back words of the level's tiles.
//   k_msd_plan     one block per range: a range whose keys are all equal is
//                  done (or copied to A when it sits in B); one whose digit
//                  is constant moves on to the next level with its next
//                  varying digit; any other is partitioned (bucket starts =
//                  exclusive scan of its counts) and its buckets become
//                  next-level ranges (> kMsdFinishCap elements) or runs of
//                  consecutive small buckets for the finish.
//   k_msd_scatter  the partition: persistent onesweep over the level's tiles
//                  with look-back inside each range; copies the tiles of
//                  ranges marked for copying.
// Then k_msd_finish sorts every run in LDS by the digits it still needs.
#include "lsb_device.hpp"
#include "lsb_internal.hpp"
#include "lsb_onesweep.hpp"

namespace lsb {

namespace {

constexpr int kNB = 256;  // 8-bit digits

__device__ __forceinline__ uint64_t low_bits(uint32_t bits) {
  return bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
}

// Mask of digit p (8 bits, the last one cut to the key width).
__device__ __forceinline__ uint32_t digit_mask(uint32_t p, uint32_t key_bits) {
  const uint32_t sh = p * 8u;
  const uint32_t w = key_bits - sh < 8u ? key_bits - sh : 8u;
  return (1u << w) - 1u;
}

// Highest digit below `below` with a set bit in x, or -1.
__device__ __forceinline__ int top_varying_digit(uint64_t x, int below) {
  for (int p = below - 1; p >= 0; --p)
    if ((x >> (8 * p)) & 0xFFull) return p;
  return -1;
}

__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }

// The thread index through an empty asm: every use site computes its own
// copy, so the finish kernel's loop over runs does not hoist the many values
// derived from it (slot offsets, row limits) out of the loop and spill them.
__device__ __forceinline__ uint32_t opaque_tid() {
  uint32_t t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// LDS passes for a run of n elements over nb non-empty buckets: the partition digit
// (it keeps the buckets in order) and enough digits below it that a bucket of
// the average size m = n / nb has few keys left tied for the fix-up: with
// uniform keys about m^2 / 2^(8d+1) tied pairs per bucket, so d digits with
// 2^(8d) >= 2m leave about one tied pair per four elements, which the fix-up's
// transposition rounds sort.
__device__ __forceinline__ uint32_t finish_passes(uint32_t n, uint32_t nb) {
  const uint32_t m = n / (nb ? nb : 1u);
  uint32_t d = 1;
  while (d < 3 && (1ull << (8 * d)) < 2ull * m) ++d;
  return 1u + d;
}

// LDS passes for a run of the wide partition, over its digits `digits` from
// the top: digit p0 (holding the wide digit's lowest bits) sorts only its
// bits below the wide digit (fshift - 8 p0 of them) inside a bucket.  As
// many digits as bring the bits sorted inside a bucket of the average size
// m to 2^bits >= 2m (about one tied pair per four elements for uniform
// keys; the fix-up's transposition rounds take them).
__device__ __forceinline__ uint32_t wide_finish_passes(uint32_t digits, uint32_t fshift,
                                                       uint32_t m) {
  uint32_t want = 0, bits = 0;
  for (int p = 7; p >= 0; --p) {
    if (!((digits >> p) & 1u)) continue;
    ++want;
    const uint32_t lo = 8u * (uint32_t)p, hi = umin(lo + 8u, fshift);
    bits += hi > lo ? hi - lo : 0u;
    if (bits >= 40u || (1ull << bits) >= 2ull * m) break;
  }
  return want;
}

// Appends one range to level k1's list (a single thread): its entry, its
// zeroed counters and its tiles in the level's tile map.
__device__ void msd_emit(const MsdArgs& a, uint32_t k1, uint64_t ea, uint64_t eb, uint32_t n,
                         uint32_t dig, uint32_t flags) {
  if (k1 >= (uint32_t)kMsdMaxLevels) {
    atomicAdd(&a.run_ctr[2], 1u);
    return;
  }
  const uint32_t N = k1 & 1u;
  MsdLevel* lv = &a.lvl[k1];
  const uint32_t nt = (n + kMsdTile - 1) / kMsdTile;
  const uint32_t id = atomicAdd(&lv->entries, 1u);
  const uint32_t tb = atomicAdd(&lv->tiles, nt);
  if (id >= a.cap_entries || (uint64_t)tb + nt > a.cap_tiles) {
    atomicAdd(&a.run_ctr[2], 1u);  // cannot happen within the workspace's capacities
    if (id < a.cap_entries) {
      MsdEntry dead = {ea, eb, 0u, 0xFFFFFFFFu, 0u, 0u};
      a.ent[N][id] = dead;
    }
    return;
  }
  MsdEntry e = {ea, eb, n, tb, dig, flags};
  a.ent[N][id] = e;
  uint4* c = reinterpret_cast<uint4*>(a.counts[N] + (size_t)id * kNB);
  for (int i = 0; i < kNB / 4; ++i) c[i] = make_uint4(0u, 0u, 0u, 0u);
  reinterpret_cast<uint4*>(a.orr[N])[id] = make_uint4(0u, 0u, ~0u, ~0u);
  for (uint32_t i = 0; i < nt; ++i) a.tilemap[N][tb + i] = id;
}

}  // namespace

// ----------------------------------------------------------- k_msd_init
// Level 0 = the batch's large segments (host table), partition digit = the
// key's top digit; clears every level's counters and the run counters.
__global__ __launch_bounds__(256) void k_msd_init(MsdArgs a) {
  const uint32_t t = threadIdx.x;
  const uint32_t nl = umin(a.hdr->n_large, a.cap_entries);
  const uint32_t P = (a.key_bits + 7u) / 8u;
  if (blockIdx.x == 0) {
    if (t < (uint32_t)kMsdMaxLevels) {
      MsdLevel z = {0u, 0u, 0u, 0u};
      if (t == 0) {
        z.entries = nl;
        z.tiles = umin(a.hdr->total_tiles, a.cap_tiles);
      }
      a.lvl[t] = z;
    }
    if (t < 4) a.run_ctr[t] = 0u;
    for (uint32_t i = t; i < (uint32_t)(kRegions * kNB + kRegions); i += 256) a.rg[i] = 0u;
  }
  if (blockIdx.x == 0 && t == 0) {
    MsdLevel w = {umin(a.hdr->n_wide, a.cap_wide), umin(a.hdr->wide_tiles, a.cap_wide_tiles),
                  0u, 0u};
    a.wlvl[0] = w;
    MsdLevel z = {0u, 0u, 0u, 0u};
    for (int i = 1; i < kMsdWideSteps; ++i) a.wlvl[i] = z;
  }
  // A segment marked wide goes to the wide partition; its level-0 entry is
  // left empty and its level-0 tiles map to no entry.
  for (uint32_t s = blockIdx.x; s < nl; s += gridDim.x) {
    const SegLarge L = a.large[s];
    const uint32_t nt = (L.n + kMsdTile - 1) / kMsdTile;
    const bool wide = L.wide < a.cap_wide;
    if (t == 0) {
      MsdEntry e = {L.a, L.b, wide ? 0u : L.n, wide ? 0xFFFFFFFFu : L.tile_begin, P - 1u, 0u};
      a.ent[0][s] = e;
      if (wide) {
        // the key's top kMsdWideBits bits, elements in A
        MsdEntry w = {L.a, L.b, L.n, L.perm_pos, a.key_bits - kMsdWideBits, 0u};
        a.went[L.wide] = w;
      }
    }
    a.counts[0][(size_t)s * kNB + t] = 0u;
    if (t < 4) a.orr[0][4 * (size_t)s + t] = t < 2 ? 0u : ~0u;
    for (uint32_t i = t; i < nt; i += 256)
      if (L.tile_begin + i < a.cap_tiles) a.tilemap[0][L.tile_begin + i] = wide ? 0xFFFFFFFFu : s;
    if (wide) {
      constexpr uint32_t WNB = 1u << kMsdWideBits;
      for (uint32_t i = t; i < WNB; i += 256) a.wcounts[(size_t)L.wide * WNB + i] = 0u;
      if (t < 4) a.worr[4 * (size_t)L.wide + t] = t < 2 ? 0u : ~0u;
      for (uint32_t i = t; i < nt; i += 256)
        if (L.perm_pos + i < a.cap_wide_tiles) a.wtile[L.perm_pos + i] = L.wide;
    }
  }
}

// ------------------------------------------------------------ region form
// Tiles per region when a range of n elements is cut into kRegions slices of
// whole tiles (the last ones may hold fewer or none).
__device__ __forceinline__ uint32_t msd_region_tiles(uint32_t n) {
  const uint32_t nt = (n + kMsdTile - 1) / kMsdTile;
  return (nt + kRegions - 1) / kRegions;
}

// Whether level k is partitioned in the region form: level 0 holding one
// range of more than kMsdRegionMin elements (a plain sort's array; a wide
// segment's emptied entry holds none).  Histogram, scatter and region
// scatter decide it alike, from the device's own level table.
__device__ __forceinline__ bool msd_region_level(const MsdArgs& a, uint32_t k) {
  return k == 0 && umin(a.lvl[0].entries, a.cap_entries) == 1u &&
         a.ent[0][0].n > kMsdRegionMin;
}

// ----------------------------------------------------------- k_msd_hist
// Each block owns a contiguous range of the level's tiles, counts each
// range's partition digit in LDS and accumulates the OR and the AND of its
// keys (their XOR is the set of bits in which the keys differ), flushing
// when the range changes.  A wave whose 64 keys share one digit adds them
// from one lane.  The block's part of the tile map is staged in LDS and a
// range's entry is read once for its consecutive tiles, so a tile's key
// loads wait on no other load.
// NB = 256: each range's own 8-bit digit E.dig; NB = 2^kMsdWideBits (a
// wide partition): kMsdWideBits bits from bit E.dig up.
// With `reg` (a level 0 partitioned in the region form: one range) the
// counts are also added per region of the range's tiles into reg[r][NB].
template <int NB>
__device__ __forceinline__ void msd_hist(const MsdEntry* __restrict__ ent,
                                         uint32_t* __restrict__ counts,
                                         uint32_t* __restrict__ orr,
                                         const uint32_t* __restrict__ tilemap,
                                         uint32_t* __restrict__ status, uint32_t tiles,
                                         uint32_t nent, uint32_t key_bits, const KeyOrder& ko,
                                         uint32_t* __restrict__ reg = nullptr) {
  constexpr int U = 8;
  constexpr int PER = NB / 256;  // counters per thread
  const uint32_t t = threadIdx.x;
  const uint32_t per = (tiles + gridDim.x - 1) / gridDim.x;
  const uint32_t t0 = blockIdx.x * per;
  const uint32_t t1 = umin(t0 + per, tiles);
  if (t0 >= t1) return;
  __shared__ uint32_t h[NB];
  __shared__ uint32_t s_or[4];  // OR lo, OR hi, AND lo, AND hi
  __shared__ uint32_t s_tm[256];
  {  // the level's scatter starts from clear look-back words
    uint4* z = reinterpret_cast<uint4*>(status + (size_t)t0 * NB);
    const size_t z4 = (size_t)(t1 - t0) * NB / 4;
    for (size_t i = t; i < z4; i += 256) z[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  for (int i = 0; i < PER; ++i) h[t + 256 * i] = 0u;
  if (t < 4) s_or[t] = t < 2 ? 0u : ~0u;
  const uint64_t kbm = low_bits(key_bits);
  uint32_t cur = 0xFFFFFFFFu;     // range of the counts in h
  uint32_t cur_rg = 0xFFFFFFFFu;  // and their region (region form)
  uint64_t acc_or = 0, acc_and = ~0ull;
  MsdEntry E = {0ull, 0ull, 0u, 0xFFFFFFFFu, 0u, kMsdCopy};
  auto flush = [&](uint32_t e, uint32_t rg) {
    acc_or &= kbm;
    acc_and |= ~kbm;
    if ((uint32_t)acc_or) atomicOr(&s_or[0], (uint32_t)acc_or);
    if ((uint32_t)(acc_or >> 32)) atomicOr(&s_or[1], (uint32_t)(acc_or >> 32));
    if (~(uint32_t)acc_and) atomicAnd(&s_or[2], (uint32_t)acc_and);
    if (~(uint32_t)(acc_and >> 32)) atomicAnd(&s_or[3], (uint32_t)(acc_and >> 32));
    acc_or = 0;
    acc_and = ~0ull;
    __syncthreads();
    for (int i = 0; i < PER; ++i) {
      const uint32_t c = h[t + 256 * i];
      if (c && e < nent) atomicAdd(&counts[(size_t)e * NB + t + 256 * i], c);
      if (c && reg && rg < (uint32_t)kRegions) atomicAdd(&reg[rg * NB + t + 256 * i], c);
      h[t + 256 * i] = 0u;
    }
    if (t < 4) {
      const uint32_t v = s_or[t], idle = t < 2 ? 0u : ~0u;
      if (v != idle && e < nent) {
        if (t < 2) atomicOr(&orr[4 * (size_t)e + t], v);
        else atomicAnd(&orr[4 * (size_t)e + t], v);
      }
      s_or[t] = idle;
    }
    __syncthreads();
  };
  for (uint32_t c0 = t0; c0 < t1; c0 += 256) {
    __syncthreads();  // the previous chunk of the map is consumed
    if (c0 + t < t1) s_tm[t] = tilemap[c0 + t];
    __syncthreads();
    const uint32_t c1 = umin(c0 + 256, t1);
    for (uint32_t tl = c0; tl < c1; ++tl) {
      const uint32_t e = s_tm[tl - c0];
      if (e != cur) {
        if (cur != 0xFFFFFFFFu) flush(cur, cur_rg);
        cur = e;
        cur_rg = 0xFFFFFFFFu;
        if (e < nent) E = ent[e];
      }
      if (e >= nent || (E.flags & kMsdCopy)) continue;
      if (tl < E.tile_begin || (uint64_t)(tl - E.tile_begin) * kMsdTile >= E.n) continue;
      if (reg) {
        const uint32_t rt = msd_region_tiles(E.n);
        const uint32_t rg = (tl - E.tile_begin) / rt;
        if (rg != cur_rg) {
          if (cur_rg != 0xFFFFFFFFu) flush(cur, cur_rg);
          cur_rg = rg;
        }
      }
      const elem_t* src = reinterpret_cast<const elem_t*>((E.flags & kMsdInB) ? E.b : E.a);
      const uint32_t shift = NB == kNB ? E.dig * 8u : E.dig;
      const uint32_t mask = NB == kNB ? digit_mask(E.dig, key_bits) : (uint32_t)NB - 1u;
      const uint32_t first = (tl - E.tile_begin) * kMsdTile;
      const uint32_t cnt = umin(kMsdTile, E.n - first);
      const elem_t* p = src + first;
      if (cnt == kMsdTile) {
        for (uint32_t i0 = 0; i0 < cnt; i0 += 256 * U) {
          uint64_t kk[U];
#pragma unroll
          for (int u = 0; u < U; ++u) kk[u] = key_fwd(gload_key(p + i0 + u * 256 + t), ko);
#pragma unroll
          for (int u = 0; u < U; ++u) {
            acc_or |= kk[u];
            acc_and &= kk[u];
            const uint32_t d = digit_of(kk[u], shift, mask);
            const uint32_t f = __builtin_amdgcn_readfirstlane(d);
            if (__ballot(d != f) == 0) {
              if (lane_id() == 0) atomicAdd(&h[f], (uint32_t)kWave);
            } else {
              atomicAdd(&h[d], 1u);
            }
          }
        }
      } else {
        for (uint32_t i0 = 0; i0 < cnt; i0 += 256) {
          const uint32_t i = i0 + t;
          const bool v = i < cnt;
          const uint64_t kk = v ? key_fwd(gload_key(p + i), ko) : 0ull;
          if (v) {
            acc_or |= kk;
            acc_and &= kk;
          }
          hist_add(h, digit_of(kk, shift, mask), v);
        }
      }
    }
  }
  if (cur != 0xFFFFFFFFu) flush(cur, cur_rg);
}

__global__ __launch_bounds__(256) void k_msd_hist(MsdArgs a) {
  const uint32_t k = a.level, L = k & 1u;
  msd_hist<kNB>(a.ent[L], a.counts[L], a.orr[L], a.tilemap[L], a.status,
                umin(a.lvl[k].tiles, a.cap_tiles), umin(a.lvl[k].entries, a.cap_entries),
                a.key_bits, a.ko, msd_region_level(a, k) ? a.rg : nullptr);
}

// A wide step's histogram: each range's kMsdWideBits-bit digit.
__global__ __launch_bounds__(256) void k_msd_whist(MsdArgs a) {
  const MsdLevel* wl = &a.wlvl[a.wstep];
  msd_hist<(1 << kMsdWideBits)>(a.went, a.wcounts, a.worr, a.wtile, a.wstatus,
                                umin(wl->tiles, a.cap_wide_tiles),
                                umin(wl->entries, a.cap_wide), a.key_bits, a.ko);
}

// ----------------------------------------------------------- k_msd_plan
// One block per range; thread b owns bucket b of a partitioned range.
__global__ __launch_bounds__(256) void k_msd_plan(MsdArgs a) {
  constexpr uint32_t C = kMsdFinishCap;
  const uint32_t k = a.level, L = k & 1u;
  const uint32_t t = threadIdx.x;
  const uint32_t nent = umin(a.lvl[k].entries, a.cap_entries);
  const uint64_t kbm = low_bits(a.key_bits);
  __shared__ uint32_t s_start[kNB];   // element offset of each bucket
  __shared__ uint32_t s_q[kNB + 1];   // exclusive prefix of min(count, C + 1)
  __shared__ uint32_t s_nl[kNB + 1];  // exclusive prefix of buckets of <= C elements (low 16
                                      // bits) and of non-empty buckets (high 16 bits)
  __shared__ uint32_t s_emit[kNB];    // exclusive prefix of (tiles << 9 | 1) of next-level buckets
  __shared__ uint32_t s_next[kNB];    // end of a run that starts at b; for a large b, the next
                                      // small bucket | 0x80000000
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_rlo[kNB], s_rn[kNB], s_rb[kNB];
  __shared__ uint32_t s_misc[4];
  for (uint32_t e = blockIdx.x; e < nent; e += gridDim.x) {
    const MsdEntry E = a.ent[L][e];
    if ((E.flags & kMsdCopy) || E.n == 0) {
      // a copy (emitted by the level before) is moved by this level's scatter
      if (t == 0 && E.n) atomicAdd(&a.lvl[k].moved, E.n);
      continue;
    }
    const uint32_t rem = umin(8u * (E.dig + 1u), a.key_bits);
    // bits in which the range's keys differ: OR of the keys ^ AND of the keys
    const uint4 oa = reinterpret_cast<const uint4*>(a.orr[L])[e];
    const uint64_t x = ((((uint64_t)oa.y << 32) | oa.x) ^ (((uint64_t)oa.w << 32) | oa.z)) & kbm &
                       low_bits(rem);
    const bool in_b = (E.flags & kMsdInB) != 0;
    if (x == 0) {
      // every key equal: done where it is, or copied to A by this level's scatter
      if (in_b && t == 0) {
        a.ent[L][e].flags = E.flags | kMsdCopy;
        atomicAdd(&a.lvl[k].moved, E.n);
      }
      continue;
    }
    if (((x >> (8u * E.dig)) & 0xFFull) == 0) {
      // this digit is constant: the next level partitions on the next varying one
      if (t == 0) msd_emit(a, k + 1, E.a, E.b, E.n, (uint32_t)top_varying_digit(x, (int)E.dig),
                           E.flags & kMsdInB);
      continue;
    }
    // Partition on E.dig.  Buckets of more than C elements become ranges of
    // the next level (with the next varying digit, or, when no digit below
    // varies and they land in B, as copies); the others are finished in runs.
    const int cd = top_varying_digit(x, (int)E.dig);
    const bool out_b = !in_b;
    const uint32_t c = a.counts[L][(size_t)e * kNB + t];
    const bool large = c > C;
    // At level k < kMsdWideSteps - 1 a bucket of the wide size range whose
    // next varying digit leaves room for a wide digit goes to wide step k+1
    // instead of level k+1: its buckets then finish after one more partition.
    const bool to_wide = k + 1 < (uint32_t)kMsdWideSteps && large && c > kMsdWideMin &&
                         c <= kMsdWideMax && cd >= 1 && a.key_bits >= kMsdWideKeyBits;
    const bool emit = large && !to_wide && (cd >= 0 || out_b);
    const uint32_t nt = (c + kMsdTile - 1) / kMsdTile;
    const uint32_t my_q = large ? C + 1 : c;
    const uint32_t my_nl = (large ? 0u : 1u) | (c ? 0x10000u : 0u);
    const uint32_t my_emit = emit ? ((nt << 9) | 1u) : 0u;
    s_start[t] = c;
    s_q[t] = my_q;
    s_nl[t] = my_nl;
    s_emit[t] = my_emit;
    block_exclusive_scan<kNB, 256, 0, true>(s_start, s_wsum);
    block_exclusive_scan<kNB, 256, 0, true>(s_q, s_wsum);
    block_exclusive_scan<kNB, 256, 0, true>(s_nl, s_wsum);
    block_exclusive_scan<kNB, 256, 0, true>(s_emit, s_wsum);
    const uint32_t start = s_start[t];
    a.counts[L][(size_t)e * kNB + t] = start;
    int any_small = (!large && c > 0) ? 1 : 0;
    if (t == kNB - 1) {
      s_q[kNB] = s_q[t] + my_q;
      s_nl[kNB] = s_nl[t] + my_nl;
      const uint32_t tot = s_emit[t] + my_emit;
      uint32_t be = 0, bt = 0;
      if (tot && k + 1 < (uint32_t)kMsdMaxLevels) {
        be = atomicAdd(&a.lvl[k + 1].entries, tot & 511u);
        bt = atomicAdd(&a.lvl[k + 1].tiles, tot >> 9);
      } else if (tot) {
        atomicAdd(&a.run_ctr[2], 1u);
      }
      s_misc[0] = be;
      s_misc[1] = bt;
      a.ent[L][e].flags = E.flags | kMsdMove;
      atomicAdd(&a.lvl[k].moved, E.n);
    }
    any_small = __syncthreads_or(any_small);
    if (to_wide) {
      constexpr uint32_t WNB = 1u << kMsdWideBits;
      const uint32_t nt = (c + kMsdTile - 1) / kMsdTile;
      const uint64_t ca = E.a + 16ull * start, cb = E.b + 16ull * start;
      const uint32_t id = atomicAdd(&a.wlvl[k + 1].entries, 1u);
      const uint32_t tb = atomicAdd(&a.wlvl[k + 1].tiles, nt);
      if (id >= a.cap_wide || (uint64_t)tb + nt > a.cap_wide_tiles) {
        atomicAdd(&a.run_ctr[2], 1u);  // cannot happen within the workspace's capacities
        if (id < a.cap_wide) {
          MsdEntry dead = {ca, cb, 0u, 0xFFFFFFFFu, 0u, 0u};
          a.went[id] = dead;
        }
      } else {
        MsdEntry w = {ca, cb, c, tb, 8u * (uint32_t)cd + 8u - kMsdWideBits,
                      out_b ? kMsdInB : 0u};
        a.went[id] = w;
        uint4* z = reinterpret_cast<uint4*>(a.wcounts + (size_t)id * WNB);
        for (uint32_t i = 0; i < WNB / 4; ++i) z[i] = make_uint4(0u, 0u, 0u, 0u);
        reinterpret_cast<uint4*>(a.worr)[id] = make_uint4(0u, 0u, ~0u, ~0u);
        for (uint32_t i = 0; i < nt; ++i) a.wtile[tb + i] = id;
      }
    }
    if (emit && k + 1 < (uint32_t)kMsdMaxLevels) {
      const uint32_t N = (k + 1) & 1u;
      const uint32_t id = s_misc[0] + (s_emit[t] & 511u);
      const uint32_t tb = s_misc[1] + (s_emit[t] >> 9);
      const uint64_t ca = E.a + 16ull * start, cb = E.b + 16ull * start;
      if (id >= a.cap_entries || (uint64_t)tb + nt > a.cap_tiles) {
        atomicAdd(&a.run_ctr[2], 1u);  // cannot happen within the workspace's capacities
        if (id < a.cap_entries) {
          MsdEntry dead = {ca, cb, 0u, 0xFFFFFFFFu, 0u, 0u};
          a.ent[N][id] = dead;
        }
      } else {
        MsdEntry ch = {ca, cb, c, tb, cd >= 0 ? (uint32_t)cd : 0u,
                       (out_b ? kMsdInB : 0u) | (cd >= 0 ? 0u : kMsdCopy)};
        a.ent[N][id] = ch;
        uint4* z = reinterpret_cast<uint4*>(a.counts[N] + (size_t)id * kNB);
        for (int i = 0; i < kNB / 4; ++i) z[i] = make_uint4(0u, 0u, 0u, 0u);
        reinterpret_cast<uint4*>(a.orr[N])[id] = make_uint4(0u, 0u, ~0u, ~0u);
        for (uint32_t i = 0; i < nt; ++i) a.tilemap[N][tb + i] = id;
      }
    }
    // Runs: consecutive buckets of at most C elements packed greedily up to
    // C elements; a larger bucket ends a run.  Each run is sorted by E.dig
    // (so its buckets keep their order; not needed in a run of one bucket)
    // and the digits below it that vary in the range; with none varying its
    // buckets are already sorted, and it is a copy when they sit in B.
    if (!any_small || (cd < 0 && !out_b)) {
      __syncthreads();  // the shared arrays are rewritten for the next range
      continue;
    }
    uint32_t digits = 0;
    if (cd >= 0) {
      digits = 1u << E.dig;
      for (int p = 0; p < (int)E.dig; ++p)
        if ((x >> (8 * p)) & 0xFFull) digits |= 1u << p;
    }
    // Where a run starting at bucket t ends: the last bucket boundary j with
    // at most C elements in [t, j) and no large bucket among them (a large
    // bucket counts C + 1); for a large t, the next small bucket.
    if (large) {
      // smallest m in (t + 1, kNB] with s_nl[m] > s_nl[t + 1]: bucket m - 1
      const uint32_t want = (s_nl[t + 1] & 0xFFFFu) + 1u;
      uint32_t lo = t + 2, hi = kNB + 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) / 2;
        if ((s_nl[mid] & 0xFFFFu) >= want) hi = mid;
        else lo = mid + 1;
      }
      s_next[t] = 0x80000000u | (lo <= (uint32_t)kNB ? lo - 1 : (uint32_t)kNB);
    } else {
      const uint32_t base = s_q[t];
      uint32_t lo = t + 1, hi = kNB;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (s_q[mid] - base <= C) lo = mid;
        else hi = mid - 1;
      }
      s_next[t] = lo;
    }
    __syncthreads();
    if (t == 0) {
      uint32_t R = 0;
      for (uint32_t b = 0; b < (uint32_t)kNB;) {
        const uint32_t v = s_next[b];
        if (v & 0x80000000u) {
          b = v & 0x1FFu;
          continue;
        }
        const uint32_t rn = s_q[v] - s_q[b];
        if (rn) {
          s_rlo[R] = s_start[b];
          s_rn[R] = rn;
          s_rb[R] = (s_nl[v] >> 16) - (s_nl[b] >> 16);  // its non-empty buckets
          ++R;
        }
        b = v;
      }
      uint32_t base = R ? atomicAdd(&a.run_ctr[0], R) : 0u;
      if (R && (uint64_t)base + R > a.cap_runs) {
        atomicAdd(&a.run_ctr[2], 1u);  // cannot happen within the workspace's capacities
        R = 0;
      }
      s_misc[2] = R;
      s_misc[3] = base;
    }
    __syncthreads();
    const uint32_t R = s_misc[2], rbase = s_misc[3];
    const uint64_t from = out_b ? E.b : E.a;
    for (uint32_t r = t; r < R; r += 256) {
      // A run of one bucket has a constant E.dig: no pass over it.
      const bool one = s_rb[r] == 1u;
      const uint32_t dr = one ? digits & ~(1u << E.dig) : digits;
      const uint32_t passes = finish_passes(s_rn[r], s_rb[r]) - (one ? 1u : 0u);
      MsdRun run = {from + 16ull * s_rlo[r], E.a + 16ull * s_rlo[r], s_rn[r],
                    dr | (passes << 16)};
      a.runs[rbase + r] = run;
    }
    __syncthreads();
  }
}

// -------------------------------------------------------- k_msd_scatter
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_msd_scatter(MsdArgs a) {
  static_assert(BLOCK * ITEMS == (int)kMsdTile, "one onesweep tile per level tile");
  __shared__ TileSmem<8, BLOCK, ITEMS> sm;
  constexpr bool kOwn = BLOCK >= 512;
  __shared__ uint32_t s_whist[kOwn ? (BLOCK / kWave) * kNB : 1];
  const uint32_t k = a.level, L = k & 1u;
  if (msd_region_level(a, k)) return;  // k_msd_scatter_rg partitions it
  const uint32_t tiles = umin(a.lvl[k].tiles, a.cap_tiles);
  const uint32_t nent = umin(a.lvl[k].entries, a.cap_entries);
  uint32_t* ticket = &a.lvl[k].ticket;
  if (threadIdx.x == 0) sm.scalar[3] = atomicAdd(ticket, 1u);
  __syncthreads();
  while (true) {
    const uint32_t tl = sm.scalar[3];
    if (tl >= tiles) return;
    const uint32_t e = a.tilemap[L][tl];
    bool took = false;
    if (e < nent) {
      const MsdEntry E = a.ent[L][e];
      const uint32_t local = tl - E.tile_begin;
      if (tl >= E.tile_begin && (uint64_t)local * kMsdTile < E.n) {
        if (E.flags & kMsdMove) {
          const bool in_b = (E.flags & kMsdInB) != 0;
          const elem_t* src = reinterpret_cast<const elem_t*>(in_b ? E.b : E.a);
          elem_t* dst = reinterpret_cast<elem_t*>(in_b ? E.a : E.b);
          onesweep_tile<8, BLOCK, ITEMS>(sm, src, dst, E.n, local,
                                         a.status + (size_t)E.tile_begin * kNB, nullptr,
                                         a.counts[L] + (size_t)e * kNB, E.dig * 8u,
                                         digit_mask(E.dig, a.key_bits), a.ko, ticket,
                                         kOwn ? s_whist : nullptr);
          took = true;
        } else if (E.flags & kMsdCopy) {
          const uint32_t first = local * kMsdTile;
          const uint32_t cnt = umin(kMsdTile, E.n - first);
          const elem_t* src = reinterpret_cast<const elem_t*>(E.b) + first;
          elem_t* dst = reinterpret_cast<elem_t*>(E.a) + first;
          for (uint32_t i = threadIdx.x; i < cnt; i += BLOCK) gstore(dst + i, gload(src + i));
        }
      }
    }
    if (!took && threadIdx.x == 0) sm.scalar[3] = atomicAdd(ticket, 1u);
    __syncthreads();
  }
}

// ------------------------------------------------------ k_msd_scatter_rg
// The region form of a level-0 range (msd_region_level): the range's tiles
// are cut into kRegions slices; block b serves slice b % kRegions with its
// own ticket counter and look-back chain (the slice's first tile publishes
// an inclusive word), so kRegions chains advance side by side instead of
// one, as in the LSD passes' region form.  A slice's bucket bases are the
// range's bucket starts (the plan's) plus the bucket's counts in earlier
// slices (the histogram's per-region counts).  The output is the same
// stable partition k_msd_scatter writes.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK)
__attribute__((amdgpu_waves_per_eu(3))) void k_msd_scatter_rg(MsdArgs a) {
  static_assert(BLOCK * ITEMS == (int)kMsdTile, "one onesweep tile per level tile");
  constexpr int W = BLOCK / kWave;
  __shared__ TileSmem<8, BLOCK, ITEMS> sm;
  __shared__ uint32_t s_rbase[kNB];
  __shared__ uint32_t s_whist[W * kNB];
  if (!msd_region_level(a, 0)) return;
  const MsdEntry E = a.ent[0][0];
  if (!(E.flags & kMsdMove)) return;
  const uint32_t region = blockIdx.x % (uint32_t)kRegions;
  const uint32_t nt = (E.n + kMsdTile - 1) / kMsdTile;
  const uint32_t rt = msd_region_tiles(E.n);
  const uint32_t t_begin = region * rt;
  if (t_begin >= nt) return;
  const uint32_t t_count = umin(rt, nt - t_begin);
  const uint32_t* starts = a.counts[0];  // entry 0: bucket starts after the plan
  for (uint32_t b = threadIdx.x; b < (uint32_t)kNB; b += BLOCK) {
    uint32_t before = 0;
    for (uint32_t r = 0; r < region; ++r) before += a.rg[r * kNB + b];
    s_rbase[b] = starts[b] + before;
  }
  for (uint32_t i = threadIdx.x; i < (uint32_t)(W * kNB); i += BLOCK) s_whist[i] = 0u;
  uint32_t* ticket = a.rg + kRegions * kNB + region;
  if (threadIdx.x == 0) sm.scalar[0] = atomicAdd(ticket, 1u);
  lds_barrier();
  const bool in_b = (E.flags & kMsdInB) != 0;
  const elem_t* src = reinterpret_cast<const elem_t*>(in_b ? E.b : E.a);
  elem_t* dst = reinterpret_cast<elem_t*>(in_b ? E.a : E.b);
  const uint32_t shift = E.dig * 8u, mask = digit_mask(E.dig, a.key_bits);
  uint32_t* st = a.status + (size_t)E.tile_begin * kNB;  // cleared by the histogram
  while (true) {
    const uint32_t local = sm.scalar[0];
    if (local >= t_count) break;
    const uint32_t tile = t_begin + local;
    elem_t e[ITEMS];
    tile_load<BLOCK, ITEMS>(e, src, E.n, tile);
    keys_fwd<ITEMS>(e, a.ko);
    onesweep_tile_region<8, BLOCK, ITEMS>(sm, e, dst, E.n, tile, local == 0, st, nullptr,
                                          s_rbase, shift, mask, ticket, nullptr, 0u, 0u,
                                          rt * kMsdTile, a.ko, s_whist, false);
    lds_barrier();
  }
}

// --------------------------------------------------------- k_msd_finish
// A run is sorted in LDS by LSD passes over the highest few of its digits
// (as many as its buckets' size needs: finish_passes).  Keys that still tie
// on everything from the lowest sorted digit up then sit next to each other
// in input order, and fix_groups sorts each such group that is out of order
// by insertion.  A group of many distinct keys sends the run through LDS
// passes over every one of its digits instead; groups of equal keys are in
// order and cost nothing.

// One stable LDS pass: ranks e[] by the digit at shift/mask and writes the
// elements to sm.elems in digit order.  Returns false (nothing written) when
// all n elements share one bucket.  whist: the waves' own counter rows.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ bool lds_pass(TileSmem<8, BLOCK, ITEMS>& sm, uint32_t* whist,
                                         const elem_t (&e)[ITEMS], uint32_t lim, uint32_t n,
                                         uint32_t shift, uint32_t mask, uint32_t* one_flag,
                                         uint32_t pass_id) {
  constexpr int W = BLOCK / kWave;
  static_assert(BLOCK >= kNB, "one bucket per thread");
  const uint32_t t = opaque_tid();
  const uint32_t lane = t & (kWave - 1);
  const uint32_t wave = t / kWave;
  uint32_t* wh = whist + wave * kNB;
  // Only this wave has touched its row since the last pass's scan.
  for (uint32_t i = lane; i < (uint32_t)kNB; i += kWave) wh[i] = 0u;
  uint32_t dg[ITEMS], rk[ITEMS];
  wave_rank<8, ITEMS, true>(e, lim, shift, mask, wh, dg, rk);
  lds_barrier();
  if (t < (uint32_t)kNB) {
    uint32_t s = 0;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const uint32_t c = whist[w * kNB + t];
      whist[w * kNB + t] = s;
      s += c;
    }
    sm.start[t] = s;
    // the bucket holding all n elements marks the pass (no word to clear:
    // each pass of a block has its own id)
    if (s == n) *one_flag = pass_id;
  }
  // adds each bucket's start into every wave's row: slot = wh[digit] + rank
  // (its barriers also publish the mark)
  block_exclusive_scan<kNB, BLOCK, W, true>(sm.start, sm.wsum, whist);
  if (*one_flag == pass_id) return false;
  uint32_t pos[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) pos[i] = wh[dg[i]] + rk[i];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if ((uint32_t)i * kWave + lane < lim) sm.elems[pos[i]] = e[i];
  return true;
}

// LSD passes over the digits in `digits`, lowest first, from e[]; on return
// sm.elems[0..n) holds the elements in the resulting order.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ void lds_sort(TileSmem<8, BLOCK, ITEMS>& sm, uint32_t* whist,
                                         elem_t (&e)[ITEMS], uint32_t lim, uint32_t n,
                                         uint32_t key_bits, uint32_t digits, uint32_t* one_flag,
                                         uint32_t& pass_id) {
  const uint32_t tid = opaque_tid();
  const uint32_t lane = tid & (kWave - 1);
  const uint32_t wrun = (tid / kWave) * (kWave * ITEMS);
  bool in_lds = false;  // sm.elems holds the current order (else e[] does)
  for (uint32_t p = 0; p < 8; ++p) {
    if (!((digits >> p) & 1u)) continue;
    if (in_lds) {
      lds_barrier();
#pragma unroll
      for (int i = 0; i < ITEMS; ++i)
        if ((uint32_t)i * kWave + lane < lim) e[i] = sm.elems[wrun + (uint32_t)i * kWave + lane];
      in_lds = false;
    }
    if (lds_pass<BLOCK, ITEMS>(sm, whist, e, lim, n, p * 8u, digit_mask(p, key_bits), one_flag,
                               ++pass_id))
      in_lds = true;
  }
  if (!in_lds) {
    lds_barrier();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if ((uint32_t)i * kWave + lane < lim) sm.elems[wrun + (uint32_t)i * kWave + lane] = e[i];
  }
  lds_barrier();
}

constexpr int kRawBufferFlags = 0x00020000;  // dword 3 of a raw gfx9 buffer resource

// Issues the loads of the run at src (n elements, the same in every lane)
// into e[] in the wave-contiguous layout: raw buffer loads over exactly the
// run, so slots past its end read zero with no branch and the loads go out
// back to back.  The resource lives in scalar registers (readfirstlane
// returns int: each word goes through uint32_t so that no sign extension
// reaches the address's high word).  Each load keeps its whole offset in
// the vector operand: the range check covers that one (and the
// instruction's offset), not the scalar offset.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ void run_load(elem_t (&e)[ITEMS], const elem_t* src, uint32_t n) {
  const uint32_t tid = opaque_tid();
  const uint32_t lane = tid & (kWave - 1);
  const uint32_t wrun = (tid / kWave) * (kWave * ITEMS);
  const uint64_t sa = reinterpret_cast<uint64_t>(src);
  const uint64_t su =
      ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(sa >> 32)) << 32) |
      (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)sa);
  const uint32_t nu = (uint32_t)__builtin_amdgcn_readfirstlane(n);
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      reinterpret_cast<void*>(su), (short)0, (int)(nu * (uint32_t)sizeof(elem_t)),
      kRawBufferFlags);
  const uint32_t off = (wrun + lane) * (uint32_t)sizeof(elem_t);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const auto x = __builtin_amdgcn_raw_buffer_load_b128(
        rs, (int)(off + (uint32_t)(i * kWave) * (uint32_t)sizeof(elem_t)), 0, 0);
    e[i].key = ((uint64_t)x[1] << 32) | x[0];
    e[i].val = ((uint64_t)x[3] << 32) | x[2];
  }
}

// Ties left by the LDS passes.  The passes sorted the run on its key bits
// from `cut` up (within kbm), so elements that tie there sit next to each
// other in input order.  Bit p of tmask: elements p and p + 1 tie from `cut`
// up; bit p of dmask: they tie and p's key is the larger (out of order).  A
// group (a maximal run of tmask bits and the element after it) holding an
// out-of-order pair is sorted by insertion, by the one thread whose positions
// [PER t, PER t + PER) hold the group's first element; only strictly larger
// keys move, so equal keys keep their order.  With uniform keys the groups
// are pairs and triples.  Returns false when a group that needs sorting is
// longer than kFixMaxGroup (the caller then sorts the whole run on all its
// digits; what was sorted here was sorted stably, so that stays correct).
// flags: 2 * BLOCK / 64 words of LDS (the two block-wide ORs).
constexpr uint32_t kFixMaxGroup = 16;

template <int BLOCK, int ITEMS>
__device__ __forceinline__ bool fix_groups(TileSmem<8, BLOCK, ITEMS>& sm, uint64_t* tmask,
                                           uint64_t* dmask, uint32_t* flags, uint32_t n,
                                           uint64_t kbm, uint32_t cut) {
  constexpr int W = BLOCK / kWave;
  constexpr uint32_t PER = (uint32_t)ITEMS;  // positions per thread: BLOCK * PER = the run cap
  static_assert(PER < 32 && kWave % PER == 0, "a thread's positions lie in one mask word");
  const uint32_t t = opaque_tid();
  const uint32_t lane = t & (kWave - 1);
  const uint32_t wave = t / kWave;
  bool any = false;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t p = (uint32_t)i * BLOCK + t;
    bool tie = false, desc = false;
    if (p + 1 < n) {
      const uint64_t k0 = sm.elems[p].key & kbm, k1 = sm.elems[p + 1].key & kbm;
      tie = (k0 >> cut) == (k1 >> cut);
      desc = tie && k0 > k1;
    }
    const uint64_t bt = __ballot(tie), bd = __ballot(desc);
    if (lane == 0) {
      tmask[p / kWave] = bt;
      dmask[p / kWave] = bd;
    }
    any |= bd != 0;
  }
  if (lane == 0) flags[wave] = any ? 1u : 0u;
  lds_barrier();
  uint32_t need = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) need |= flags[w];
  if (!need) return true;

  uint32_t too_long = 0;
  const uint32_t p0 = t * PER;
  if (p0 < n) {
    const uint32_t bits = (uint32_t)(tmask[p0 / kWave] >> (p0 % kWave)) & ((1u << PER) - 1u);
    const uint32_t prev = p0 ? (uint32_t)(tmask[(p0 - 1) / kWave] >> ((p0 - 1) % kWave)) & 1u : 0u;
    uint32_t starts = bits & ~((bits << 1) | prev);
    while (starts) {
      const uint32_t s = p0 + (uint32_t)__builtin_ctz(starts);
      starts &= starts - 1u;
      // the group's last element: the first position from s on whose bit is
      // clear (position n - 1 has a clear bit, so the walk stops inside the run)
      uint32_t wi = s / kWave;
      uint64_t x = ~tmask[wi] & (~0ull << (s % kWave));
      while (!x) x = ~tmask[++wi];
      const uint32_t e = wi * kWave + (uint32_t)__builtin_ctzll(x);
      // an out-of-order pair among [s, e)?
      bool bad = false;
      for (uint32_t w2 = s / kWave; w2 <= (e - 1) / kWave && !bad; ++w2) {
        uint64_t m = dmask[w2];
        if (w2 == s / kWave) m &= ~0ull << (s % kWave);
        if (w2 == e / kWave) m &= (1ull << (e % kWave)) - 1ull;
        bad = m != 0;
      }
      if (!bad) continue;
      if (e - s + 1u > kFixMaxGroup) {
        too_long = 1;
        continue;
      }
      for (uint32_t j = s + 1; j <= e; ++j) {
        const elem_t v = sm.elems[j];
        const uint64_t kv = v.key & kbm;
        uint32_t i = j;
        while (i > s && (sm.elems[i - 1].key & kbm) > kv) {
          sm.elems[i] = sm.elems[i - 1];
          --i;
        }
        if (i != j) sm.elems[i] = v;
      }
    }
  }
  const bool lw = __ballot(too_long != 0) != 0;
  if (lane == 0) flags[W + wave] = lw ? 1u : 0u;
  lds_barrier();  // also orders the groups' writes before the stores' reads
  uint32_t fail = 0;
#pragma unroll
  for (int w = 0; w < W; ++w) fail |= flags[W + w];
  return fail == 0;
}

// A run's descriptor as the finish kernel keeps it: the run and its index.
struct RunDesc {
  uint64_t src, dst;
  uint32_t n, digits, idx;
};

// The descriptor that ends a block's loop: no run has every digits bit set
// (a run's word holds an 8-bit digit mask and its LDS pass count << 16).
constexpr uint32_t kNoRun = 0xFFFFFFFFu;

// Run `r`'s descriptor (thread 0), through the global address space so the
// load counts against vmcnt only; the end marker past the last run.
__device__ __forceinline__ RunDesc fetch_run(const MsdRun* runs, uint32_t r, uint32_t nr) {
  RunDesc x = {0ull, 0ull, 0u, kNoRun, r};
  if (r < nr) {
    const g_u64_t* q = (const g_u64_t*)&runs[r];
    x.src = q[0];
    x.dst = q[1];
    const uint64_t w = q[2];
    x.n = (uint32_t)w;
    x.digits = (uint32_t)(w >> 32);
  }
  return x;
}

// A 64-bit value every lane holds the same copy of, in scalar registers.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  return ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
         (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
}

// Takes a ticket (thread 0): a returning atomic through the global address
// space, so LDS waits do not wait for its round trip.
__device__ __forceinline__ uint32_t take_ticket(uint32_t* ticket) {
  typedef __attribute__((address_space(1))) uint32_t g_u32_t;
  return __hip_atomic_fetch_add((g_u32_t*)ticket, 1u, __ATOMIC_RELAXED,
                                __HIP_MEMORY_SCOPE_AGENT);
}

// Per-block LDS of the finish kernel beside the tile buffer.
template <int BLOCK>
struct FinishSmem {
  uint32_t whist[(BLOCK / kWave) * kNB];  // the waves' counter rows (LDS passes; proxy_sort:
                                          // the waves' ordinal tables)
  uint64_t tmask[kMsdFinishCap / kWave];  // fix_groups: tie bits
  uint64_t dmask[kMsdFinishCap / kWave];  // fix_groups: out-of-order bits
  uint64_t red[2 * (BLOCK / kWave)];      // proxy_sort: the waves' OR and AND of keys
  uint32_t ord[257];                      // proxy_sort: top-window values present
                                          // ([256]: rows past the run)
  RunDesc run[3];                         // descriptors of this run, the next, the one after
  uint32_t flags[2 * (BLOCK / kWave)];
  uint32_t pf[2 * (BLOCK / kWave)];       // proxy_sort: large buckets, ties
  uint32_t one;
};

// ------------------------------------------------------------ proxy_sort
// The finish's main path.  Each element of the run gets a 32-bit proxy,
// unique within the run: key bits below the counting digit (see below) over
// its position in the run (kOriginBits bits).  Sorting the elements by
// (counting digit, proxy) sorts the run by those key bits, ties in input
// order, so no step has to be stable: elements are counted by their
// counting digit with LDS atomics (in whatever order the atomics land), the
// proxies are scattered by bucket, and each element's final slot is its
// bucket's start plus the number of smaller proxies in its bucket (none or
// one compare for buckets of one or two elements, the common case for
// spread keys; larger buckets through a per-wave list).
//   The key bits used: the run's varying part (from the highest bit that
// differs within the run down), with its top 8 bits replaced by their
// ordinal among the values that occur (a run of many small buckets of a
// partition holds few values of its top digit; the ordinals pack them so
// the counting buckets stay small); the counting digit is the top
// kCountBits of that, the proxy holds the next kProxyKeyBits, and key bits
// below those (`cut`) are left out.
//   Returns kProxyEqual (all keys equal: the run is sorted as it is),
// kProxyPasses (some counting bucket exceeds kMaxBucket elements, e.g.
// heavy duplicates: the caller leaves the run to the LDS passes) or
// kProxyDone: pos[] holds each element's slot, and fs.pf[W + w] says
// whether wave w saw two elements tie on all the key bits used (they may
// differ below `cut`: fix_groups sorts those groups) -- read after the
// caller's next barrier.  The caller writes the elements to their slots
// after that barrier.
constexpr uint32_t kOriginBits = 13;                  // positions in a run
constexpr uint32_t kCountBits = 13;                   // counting digit
constexpr uint32_t kProxyKeyBits = 32 - kOriginBits;  // key bits a proxy holds below it
constexpr uint32_t kMaxBucket = 16;
static_assert(kMsdFinishCap == (1u << kOriginBits), "origins are the positions of a run");
enum : uint32_t { kProxyDone = 0, kProxyPasses = 1, kProxyEqual = 2 };

// OR (AND) of x over the 64 lanes, complete in lane 63: the DPP pattern of
// wave_inclusive_scan, with lanes shifted in from outside a row reading the
// identity (0 for OR, ~0 for AND).
template <bool AND>
__device__ __forceinline__ uint32_t wave_reduce_bits(uint32_t x) {
  const int id = AND ? -1 : 0;
  const auto op = [](uint32_t p, uint32_t q) { return AND ? (p & q) : (p | q); };
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x111, 0xf, 0xf, false));  // row_shr:1
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x112, 0xf, 0xf, false));  // row_shr:2
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x114, 0xf, 0xf, false));  // row_shr:4
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x118, 0xf, 0xf, false));  // row_shr:8
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x142, 0xa, 0xf, false));  // row_bcast:15
  x = op(x, (uint32_t)__builtin_amdgcn_update_dpp(id, (int)x, 0x143, 0xc, 0xf, false));  // row_bcast:31
  return x;
}

template <int BLOCK, int ITEMS>
__device__ __forceinline__ uint32_t proxy_sort(TileSmem<8, BLOCK, ITEMS>& sm,
                                               FinishSmem<BLOCK>& fs,
                                               const elem_t (&e)[ITEMS], uint32_t n,
                                               uint32_t lim, uint64_t kbm, uint32_t& cut,
                                               uint32_t& lowvary, uint32_t (&pos)[ITEMS]) {
  constexpr int W = BLOCK / kWave;
  constexpr uint32_t C = 1u << kCountBits;
  constexpr uint32_t CPT = C / BLOCK;  // counters per thread in the scan
  static_assert(BLOCK * ITEMS == (int)kMsdFinishCap && CPT * BLOCK == C && CPT % 2 == 0,
                "one run per block, whole counters per thread");
  // proxies [kMsdFinishCap] u32, then counters [C + 2] (C: the sentinel, C + 1:
  // where slots past the run count, so no atomic needs a branch)
  static_assert(4 * kMsdFinishCap + 4 * (C + 2) <= sizeof(sm.elems), "fits the tile buffer");
  uint32_t* P = reinterpret_cast<uint32_t*>(sm.elems);
  uint32_t* cnt = reinterpret_cast<uint32_t*>(P + kMsdFinishCap);
  const uint32_t t = opaque_tid();
  const uint32_t lane = t & (kWave - 1);
  const uint32_t wave = t / kWave;
  const uint32_t wrun = wave * (kWave * ITEMS);
  // bit i: row i of this lane holds an element of the run
  uint32_t vm = 0;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) vm |= ((uint32_t)i * kWave + lane < lim ? 1u : 0u) << i;

  // Counters and presence flags start clear (the caller's loop ends each
  // run with a barrier, after which the tile buffer is free).
  for (uint32_t i = t; i < C / 2 + 1; i += BLOCK) reinterpret_cast<uint2*>(cnt)[i] = make_uint2(0u, 0u);
  if (t < 256) fs.ord[t] = 0u;

  // The bits that vary within the run: OR xor AND of its keys.
  uint64_t o = 0, a = ~0ull;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint64_t k = e[i].key & kbm;
    const bool ok = (vm >> i) & 1u;
    o |= ok ? k : 0ull;
    a &= ok ? k : ~0ull;
  }
  const uint32_t olo = wave_reduce_bits<false>((uint32_t)o);
  const uint32_t ohi = wave_reduce_bits<false>((uint32_t)(o >> 32));
  const uint32_t alo = wave_reduce_bits<true>((uint32_t)a);
  const uint32_t ahi = wave_reduce_bits<true>((uint32_t)(a >> 32));
  if (lane == kWave - 1) {
    fs.red[wave] = ((uint64_t)ohi << 32) | olo;
    fs.red[W + wave] = ((uint64_t)ahi << 32) | alo;
  }
  lds_barrier();
  uint64_t O = 0, A = ~0ull;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    O |= fs.red[w];
    A &= fs.red[W + w];
  }
  const uint64_t vary = (O ^ A) & kbm;
  if (vary == 0) return kProxyEqual;
  const uint32_t hb = 63u - (uint32_t)__builtin_clzll(vary);
  const uint32_t lb = (uint32_t)__builtin_ctzll(vary);
  lowvary = lb;
  const uint32_t w0 = hb >= 7u ? hb - 7u : 0u;  // the top window: bits [w0, w0 + 8)

  // Ordinals of the top window's values that occur (lanes past the run in
  // its last row mark fs.ord[256], which nothing reads).  Every wave turns the flags into its
  // own table of ordinals (in its counter row, unused on this path), so no
  // further barrier is needed: lane l covers values 4l .. 4l + 3.
  // Rows wholly past the run are skipped by a branch on the wave-uniform
  // limit: their 64 lanes would all hit one LDS word, which the LDS takes
  // one lane at a time.
  const uint32_t ulim = (uint32_t)__builtin_amdgcn_readfirstlane(lim);
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if ((uint32_t)i * kWave < ulim)
      fs.ord[((vm >> i) & 1u) ? (uint32_t)((e[i].key & kbm) >> w0) & 0xFFu : 256u] = 1u;
  lds_barrier();
  uint32_t* const tab = fs.whist + wave * kNB;
  uint32_t np;  // values present (>= 2: bit hb varies)
  {
    const uint32_t f0 = fs.ord[4u * lane], f1 = fs.ord[4u * lane + 1u];
    const uint32_t f2 = fs.ord[4u * lane + 2u], f3 = fs.ord[4u * lane + 3u];
    const uint32_t s4 = f0 + f1 + f2 + f3;
    const uint32_t inc = wave_inclusive_scan(s4);
    const uint32_t ex = inc - s4;
    tab[4u * lane] = ex;
    tab[4u * lane + 1u] = ex + f0;
    tab[4u * lane + 2u] = ex + f0 + f1;
    tab[4u * lane + 3u] = ex + f0 + f1 + f2;
    np = (uint32_t)__builtin_amdgcn_readlane((int)inc, kWave - 1);
  }
  const uint32_t ob = 32u - (uint32_t)__builtin_clz(np - 1u);  // bits of the largest ordinal
  const uint32_t L = w0;                                      // key bits below the window
  const uint32_t T = ob + L;                                  // bits of the varying part
  // the counting digit: the top cb bits of the varying part; a proxy: the
  // next pb - cb bits and the origin; key bits below `cut` are left out
  const uint32_t cb = T < kCountBits ? T : kCountBits;
  const uint32_t pb = T < kCountBits + kProxyKeyBits ? T : kCountBits + kProxyKeyBits;
  cut = T - pb;  // < L: only key bits below the window are left out
  const uint64_t lowm = L ? (1ull << L) - 1ull : 0ull;
  const uint32_t pmask = (1u << (pb - cb)) - 1u;

  // Counting: lanes past the run in its last row count in cnt[C + 1], which
  // nothing reads; rows wholly past it skip the atomic (the branch is
  // wave-uniform).  Every LDS load in these loops is unconditional (rows past
  // the run hold zero keys, whose lookups stay in bounds), so a row's load
  // never sits behind a branch that waits for it alone.
  uint32_t c[ITEMS], old[ITEMS], px[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint64_t k = e[i].key & kbm;
    const uint64_t v = ((uint64_t)tab[(uint32_t)(k >> w0) & 0xFFu] << L) | (k & lowm);
    c[i] = ((vm >> i) & 1u) ? (uint32_t)(v >> (T - cb)) : C + 1u;
    px[i] = (((uint32_t)(v >> cut) & pmask) << kOriginBits) | (wrun + (uint32_t)i * kWave + lane);
  }
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    old[i] = 0u;
    if ((uint32_t)i * kWave < ulim) old[i] = atomicAdd(&cnt[c[i]], 1u);
  }
  lds_barrier();

  // Bucket starts (CPT consecutive counters per thread); any bucket above
  // kMaxBucket sends the run to the LDS passes.
  uint32_t v[CPT];
#pragma unroll
  for (uint32_t k = 0; k < CPT; k += 2) {
    const uint2 x = reinterpret_cast<const uint2*>(cnt)[(t * CPT + k) / 2];
    v[k] = x.x;
    v[k + 1] = x.y;
  }
  uint32_t sum = 0;
  bool big = false;
#pragma unroll
  for (uint32_t k = 0; k < CPT; ++k) {
    sum += v[k];
    big |= v[k] > kMaxBucket;
  }
  const uint32_t inc = wave_inclusive_scan(sum);
  if (lane == kWave - 1) sm.wsum[wave] = inc;
  const bool wbig = __ballot(big) != 0;
  if (lane == 0) fs.pf[wave] = wbig ? 1u : 0u;
  lds_barrier();
  uint32_t anybig = 0, run = inc - sum;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    anybig |= fs.pf[w];
    run += (uint32_t)w < wave ? sm.wsum[w] : 0u;
  }
  if (anybig) return kProxyPasses;
#pragma unroll
  for (uint32_t k = 0; k < CPT; k += 2) {
    reinterpret_cast<uint2*>(cnt)[(t * CPT + k) / 2] = make_uint2(run, run + v[k]);
    run += v[k] + v[k + 1];
  }
  if (t == 0) cnt[C] = n;
  lds_barrier();

  // Proxies in bucket order (in whatever order within a bucket).  From
  // here c[] packs each element's bucket start (13 bits), its bucket's size
  // (5 bits: at most kMaxBucket) and its arrival index in the bucket (4
  // bits); rows past the run hold the empty bucket at 0.
  static_assert(kMaxBucket <= 16u && kMsdFinishCap <= (1u << 13), "the packing below");
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const bool ok = (vm >> i) & 1u;
    const uint32_t cc = ok ? c[i] : 0u;
    const uint32_t b0 = cnt[cc], b1 = cnt[cc + 1u];
    if (ok) P[b0 + old[i]] = px[i];
    c[i] = ok ? b0 | ((b1 - b0) << 13) | (old[i] << 18) : 0u;
  }
  lds_barrier();
  // Each element's final slot: its bucket's start plus the proxies of its
  // bucket below its own.  A bucket of one element needs nothing and one of
  // two a single compare (most elements: the buckets hold about 0.7
  // elements on average); the elements of larger buckets go to a list of
  // the wave's own in LDS that its lanes then work through, so the wave
  // loops over those buckets only instead of over its largest bucket for
  // every row.  One wave's LDS operations complete in order, so the list
  // and its results need no barrier.  Two proxies with the same key bits
  // (different origins) mean the window left out key bits that may differ.
  static_assert(4 * (kMsdFinishCap + C + 4) + 4 * W * kWave * ITEMS <= sizeof(sm.elems),
                "the lists fit the tile buffer beside the proxies and counters");
  uint32_t* lst = P + kMsdFinishCap + (C + 4u) + wave * (uint32_t)(kWave * ITEMS);
  bool tie = false;
  uint32_t nl = 0;  // entries of the wave's list (the same in every lane)
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t b0 = c[i] & 0x1FFFu, sz = (c[i] >> 13) & 31u, a = c[i] >> 18;
    const bool two = sz == 2u;
    const uint32_t q = P[b0 + (two ? 1u - a : 0u)];  // the other element of a pair
    pos[i] = b0 + (two && q < px[i] ? 1u : 0u);
    tie |= two && (q >> kOriginBits) == (px[i] >> kOriginBits);
    const bool big = sz > 2u;
    const uint64_t bm = __ballot(big);
    if (big) {
      const uint32_t k = nl + count_below(bm);
      lst[k] = c[i] | (((uint32_t)i * kWave + lane) << 22);
      pos[i] = k;  // the list entry, replaced by its result below
    }
    nl += (uint32_t)__popcll(bm);
  }
  for (uint32_t k0 = 0; k0 < nl; k0 += kWave) {
    const uint32_t k = k0 + lane;
    if (k < nl) {
      const uint32_t v = lst[k];
      const uint32_t b0 = v & 0x1FFFu, sz = (v >> 13) & 31u, a = (v >> 18) & 15u;
      const uint32_t me = P[b0 + a];
      uint32_t rk = 0;
      for (uint32_t m = 0; m < sz; ++m) {
        const uint32_t q = P[b0 + m];
        rk += q < me ? 1u : 0u;
        tie |= m != a && (q >> kOriginBits) == (me >> kOriginBits);
      }
      lst[k] = b0 + rk;
    }
  }
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if (((c[i] >> 13) & 31u) > 2u) pos[i] = lst[pos[i]];
  const bool wtie = __ballot(tie) != 0;
  if (lane == 0) fs.pf[W + wave] = wtie ? 1u : 0u;
  return kProxyDone;
}

// Stores sm.elems[0, n) (key order mapped back) to dst, coalesced.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ void store_sorted(const TileSmem<8, BLOCK, ITEMS>& sm, elem_t* dst,
                                             uint32_t n, const KeyOrder& ko) {
  const uint32_t t = opaque_tid();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint32_t p = t + (uint32_t)i * BLOCK;
    if (p < n) {
      elem_t x = sm.elems[p];
      x.key = key_inv(x.key, ko);
      gstore(dst + p, x);
    }
  }
}

// LDS passes over the highest few of the run's digits (as many as its
// buckets' size needs: finish_passes), then fix_groups; if a group is too
// long for it, passes over every digit of the run.  e[] holds the run in
// input order on entry; on return sm.elems[0, n) holds it sorted.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ void passes_sort(TileSmem<8, BLOCK, ITEMS>& sm, FinishSmem<BLOCK>& fs,
                                            elem_t (&e)[ITEMS], uint32_t n, uint32_t lim,
                                            uint32_t key_bits, uint32_t digits,
                                            uint32_t& pass_id) {
  const uint32_t t = opaque_tid();
  const uint32_t lane = t & (kWave - 1);
  const uint32_t wrun = (t / kWave) * (kWave * ITEMS);
  const uint32_t want = digits >> 16;
  digits &= 0xFFu;
  uint32_t top = 0, ntop = 0;
  for (int p = 7; p >= 0 && ntop < want; --p)
    if ((digits >> p) & 1u) {
      top |= 1u << p;
      ++ntop;
    }
  lds_sort<BLOCK, ITEMS>(sm, fs.whist, e, lim, n, key_bits, top, &fs.one, pass_id);
  if ((digits & ~top) &&
      !fix_groups<BLOCK, ITEMS>(sm, fs.tmask, fs.dmask, fs.flags, n, low_bits(key_bits),
                                8u * (uint32_t)__builtin_ctz(top))) {
    // every slot is read (those past the run hold stale elements that the
    // passes ignore)
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) e[i] = sm.elems[wrun + (uint32_t)i * kWave + lane];
    lds_sort<BLOCK, ITEMS>(sm, fs.whist, e, lim, n, key_bits, digits, &fs.one, pass_id);
  }
}

// The run in sm.elems (its equal keys in input order) sorted again by LDS
// passes over all of `digits`: for the rare runs whose tied groups are too
// long for fix_groups.  Out of line, so its registers do not weigh on the
// common path.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ void resort_all(TileSmem<8, BLOCK, ITEMS>& sm, FinishSmem<BLOCK>& fs,
                                           uint32_t n, uint32_t lim, uint32_t key_bits,
                                           uint32_t digits, uint32_t& pass_id) {
  const uint32_t t = opaque_tid();
  const uint32_t lane = t & (kWave - 1);
  const uint32_t wrun = (t / kWave) * (kWave * ITEMS);
  // every slot is read (those past the run hold stale elements that the
  // passes ignore)
  elem_t e[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) e[i] = sm.elems[wrun + (uint32_t)i * kWave + lane];
  lds_sort<BLOCK, ITEMS>(sm, fs.whist, e, lim, n, key_bits, digits, &fs.one, pass_id);
}

// Sorts one run (n <= BLOCK * ITEMS) from src into dst: by proxy_sort (and
// fix_groups where ties remain), or by passes_sort for the runs proxy_sort
// leaves (a counting bucket of more than kMaxBucket elements: heavy
// duplicates, few varying bits) and those whose tied groups are too long for
// fix_groups.  Thread 0 stores the descriptor of the block's run after next
// (ticket `nt`, taken while the previous run was stored) into fs.run[nxt]
// right after this run's loads go out: waiting for it waits for them, which
// the run needs at once anyway.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ void finish_run(TileSmem<8, BLOCK, ITEMS>& sm, FinishSmem<BLOCK>& fs,
                                           const elem_t* src, elem_t* dst, uint32_t n,
                                           uint32_t key_bits, uint32_t digits,
                                           const KeyOrder& ko, const MsdRun* runs, uint32_t nt,
                                           uint32_t nr, uint32_t nxt, uint32_t& pass_id) {
  constexpr int W = BLOCK / kWave;
  const uint32_t t = opaque_tid();
  const uint32_t lane = t & (kWave - 1);
  const uint32_t wrun = (t / kWave) * (kWave * ITEMS);
  const uint32_t lim = n > wrun ? n - wrun : 0u;
  const uint64_t kbm = low_bits(key_bits);
  elem_t e[ITEMS];
  run_load<BLOCK, ITEMS>(e, src, n);
  if (t == 0) fs.run[nxt] = fetch_run(runs, nt, nr);
  keys_fwd<ITEMS>(e, ko);
  uint32_t cut = 0, lb = 0, pos[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) pos[i] = wrun + (uint32_t)i * kWave + lane;  // input order
  const uint32_t how =
      n > 1u ? proxy_sort<BLOCK, ITEMS>(sm, fs, e, n, lim, kbm, cut, lb, pos) : kProxyEqual;
  lds_barrier();  // the proxies and counters are dead: the buffer takes the elements
  if (how == kProxyPasses) {
    passes_sort<BLOCK, ITEMS>(sm, fs, e, n, lim, key_bits, digits, pass_id);
  } else {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if ((uint32_t)i * kWave + lane < lim) sm.elems[pos[i]] = e[i];
    bool ties = false;
    if (how == kProxyDone && cut > lb) {
      uint32_t any = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) any |= fs.pf[W + w];
      ties = any != 0;
    }
    lds_barrier();
    if (ties && !fix_groups<BLOCK, ITEMS>(sm, fs.tmask, fs.dmask, fs.flags, n, kbm, cut))
      resort_all<BLOCK, ITEMS>(sm, fs, n, lim, key_bits, digits & 0xFFu, pass_id);
  }
  store_sorted<BLOCK, ITEMS>(sm, dst, n, ko);
}

// Sorts every run inside one workgroup (runs handed out by a ticket, so
// uneven runs balance across the grid).  A run's loads were consumed before
// anything is stored, so a run sorted in place (src == dst) needs no further
// wait; runs are disjoint.  The descriptors of a block's current, next and
// after-next runs rotate through fs.run[0..2]; thread 0 takes each ticket
// two runs ahead, after the stores of a run have gone out.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_msd_finish(MsdArgs a) {
  static_assert(BLOCK * ITEMS >= (int)kMsdFinishCap, "a run fits one workgroup");
  __shared__ TileSmem<8, BLOCK, ITEMS> sm;
  __shared__ FinishSmem<BLOCK> fs;
  const uint32_t nr = umin(a.run_ctr[0], a.cap_runs);
  uint32_t* ticket = &a.run_ctr[1];
  uint32_t pass_id = 0;  // LDS passes of this block so far
  uint32_t nt = 0;       // thread 0: the ticket of the run after next
  if (threadIdx.x == 0) {
    fs.one = 0u;
    fs.run[0] = fetch_run(a.runs, take_ticket(ticket), nr);
    fs.run[1] = fetch_run(a.runs, take_ticket(ticket), nr);
    nt = take_ticket(ticket);
  }
  lds_barrier();
  for (uint32_t cur = 0;; cur = cur == 2u ? 0u : cur + 1u) {
    const uint32_t digits = (uint32_t)__builtin_amdgcn_readfirstlane(fs.run[cur].digits);
    if (digits == kNoRun) return;
    const uint32_t n = (uint32_t)__builtin_amdgcn_readfirstlane(fs.run[cur].n);
    const elem_t* src = reinterpret_cast<const elem_t*>(uniform_u64(fs.run[cur].src));
    elem_t* dst = reinterpret_cast<elem_t*>(uniform_u64(fs.run[cur].dst));
    const uint32_t after = cur == 0u ? 2u : cur - 1u;  // slot of the run after next
    if ((digits & 0xFFu) == 0) {
      // a copy
      elem_t e[ITEMS];
      run_load<BLOCK, ITEMS>(e, src, n);
      if (threadIdx.x == 0) fs.run[after] = fetch_run(a.runs, nt, nr);
      const uint32_t tid = opaque_tid();
      const uint32_t lane = tid & (kWave - 1);
      const uint32_t wrun = (tid / kWave) * (kWave * ITEMS);
#pragma unroll
      for (int i = 0; i < ITEMS; ++i)
        if (wrun + (uint32_t)i * kWave + lane < n) gstore(dst + wrun + (uint32_t)i * kWave + lane, e[i]);
    } else {
      finish_run<BLOCK, ITEMS>(sm, fs, src, dst, n, a.key_bits, digits, a.ko, a.runs, nt, nr,
                               after, pass_id);
    }
    if (threadIdx.x == 0) nt = take_ticket(ticket);
    lds_barrier();  // fs.run[after] is written; LDS is free again
  }
}

// ------------------------------------------------------- wide partition
// The segments the host marked (kMsdWideMin < n <= kMsdWideMax) are first
// partitioned on the top kMsdWideBits bits of the key, so their buckets
// average at most kMsdFinishCap elements and most finish in LDS at once.
// k_msd_wplan is k_msd_plan for one such segment with 2^kMsdWideBits buckets
// (several per thread); its large buckets join level 0, its small ones form
// runs.
__global__ __launch_bounds__(256) void k_msd_wplan(MsdArgs a) {
  constexpr uint32_t C = kMsdFinishCap;
  constexpr int WNB = 1 << kMsdWideBits, PER = WNB / 256;
  const uint32_t t = threadIdx.x;
  const uint32_t nw = umin(a.wlvl[a.wstep].entries, a.cap_wide);
  const uint32_t kt = a.wstep;  // the level its large buckets join
  const uint32_t kb = a.key_bits;
  const uint64_t kbm = low_bits(kb);
  __shared__ uint32_t s_start[WNB];
  __shared__ uint32_t s_q[WNB + 1];   // exclusive prefix of min(count, C + 1)
  __shared__ uint32_t s_nl[WNB + 1];  // small buckets (low 16 bits), non-empty ones (high 16)
  __shared__ uint32_t s_emit[WNB];    // (tiles << 12 | 1) of buckets that join level 0
  __shared__ uint32_t s_next[WNB];
  __shared__ uint32_t s_rlo[WNB], s_rn[WNB], s_rb[WNB];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_misc[4];
  for (uint32_t w = blockIdx.x; w < nw; w += gridDim.x) {
    const MsdEntry E = a.went[w];
    const uint32_t fshift = E.dig;  // the wide digit's low bit
    const bool in_b = (E.flags & kMsdInB) != 0;
    const bool out_b = !in_b;       // where the partition puts the elements
    const uint4 oa = reinterpret_cast<const uint4*>(a.worr)[w];
    const uint64_t x = ((((uint64_t)oa.y << 32) | oa.x) ^ (((uint64_t)oa.w << 32) | oa.z)) & kbm &
                       low_bits(fshift + kMsdWideBits);
    if (x == 0) {
      // every key equal: done in A, or copied there by level kt's scatter
      if (in_b && t == 0) msd_emit(a, kt, E.a, E.b, E.n, 0u, kMsdInB | kMsdCopy);
      continue;
    }
    if (((x >> fshift) & (uint64_t)(WNB - 1)) == 0) {
      // the wide digit is constant: level kt partitions on the next varying one
      if (t == 0)
        msd_emit(a, kt, E.a, E.b, E.n, (uint32_t)top_varying_digit(x, 8), E.flags & kMsdInB);
      continue;
    }
    const int cd = top_varying_digit(x & low_bits(fshift), 8);
    // Buckets of more than C elements join level kt (with the next varying
    // digit, or, when no digit below varies and they land in B, as copies).
    const bool emit_large = cd >= 0 || out_b;
    uint32_t c[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const uint32_t b = t * PER + i;
      c[i] = a.wcounts[(size_t)w * WNB + b];
      const bool large = c[i] > C;
      const uint32_t nt = (c[i] + kMsdTile - 1) / kMsdTile;
      s_start[b] = c[i];
      s_q[b] = large ? C + 1 : c[i];
      s_nl[b] = (large ? 0u : 1u) | (c[i] ? 0x10000u : 0u);
      s_emit[b] = large && emit_large ? ((nt << 12) | 1u) : 0u;
    }
    block_exclusive_scan<WNB, 256>(s_start, s_wsum);
    block_exclusive_scan<WNB, 256>(s_q, s_wsum);
    block_exclusive_scan<WNB, 256>(s_nl, s_wsum);
    block_exclusive_scan<WNB, 256>(s_emit, s_wsum);
    int any_small = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const uint32_t b = t * PER + i;
      a.wcounts[(size_t)w * WNB + b] = s_start[b];
      any_small |= (c[i] && c[i] <= C) ? 1 : 0;
    }
    if (t == 255) {
      const uint32_t lb = WNB - 1, lc = c[PER - 1];
      s_q[WNB] = s_q[lb] + (lc > C ? C + 1 : lc);
      s_nl[WNB] = s_nl[lb] + ((lc > C ? 0u : 1u) | (lc ? 0x10000u : 0u));
      const uint32_t tot =
          s_emit[lb] +
          (lc > C && emit_large ? (((lc + kMsdTile - 1) / kMsdTile) << 12) | 1u : 0u);
      s_misc[0] = tot ? atomicAdd(&a.lvl[kt].entries, tot & 4095u) : 0u;
      s_misc[1] = tot ? atomicAdd(&a.lvl[kt].tiles, tot >> 12) : 0u;
      a.went[w].flags = E.flags | kMsdMove;
      atomicAdd(&a.wlvl[kt].moved, E.n);
    }
    any_small = __syncthreads_or(any_small);
    const uint32_t NL = kt & 1u;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const uint32_t b = t * PER + i;
      if (c[i] <= C || !emit_large) continue;
      const uint32_t nt = (c[i] + kMsdTile - 1) / kMsdTile;
      const uint32_t id = s_misc[0] + (s_emit[b] & 4095u);
      const uint32_t tb = s_misc[1] + (s_emit[b] >> 12);
      const uint64_t ca = E.a + 16ull * s_start[b], cb = E.b + 16ull * s_start[b];
      if (id >= a.cap_entries || (uint64_t)tb + nt > a.cap_tiles) {
        atomicAdd(&a.run_ctr[2], 1u);  // cannot happen within the workspace's capacities
        if (id < a.cap_entries) {
          MsdEntry dead = {ca, cb, 0u, 0xFFFFFFFFu, 0u, 0u};
          a.ent[NL][id] = dead;
        }
        continue;
      }
      MsdEntry ch = {ca, cb, c[i], tb, cd >= 0 ? (uint32_t)cd : 0u,
                     (out_b ? kMsdInB : 0u) | (cd >= 0 ? 0u : kMsdCopy)};
      a.ent[NL][id] = ch;
      uint4* z = reinterpret_cast<uint4*>(a.counts[NL] + (size_t)id * kNB);
      for (int j = 0; j < kNB / 4; ++j) z[j] = make_uint4(0u, 0u, 0u, 0u);
      reinterpret_cast<uint4*>(a.orr[NL])[id] = make_uint4(0u, 0u, ~0u, ~0u);
      for (uint32_t j = 0; j < nt; ++j) a.tilemap[NL][tb + j] = id;
    }
    if (!any_small || (cd < 0 && !out_b)) {
      __syncthreads();  // the shared arrays are rewritten for the next range
      continue;
    }
    // Runs of consecutive small buckets, as in k_msd_plan, that also stay
    // inside one aligned group of G buckets: the 8-bit digit p0 holding the
    // wide digit's lowest bits then tells a run's buckets apart, so one LDS
    // pass on it keeps them in order.  With no varying bit below the wide
    // digit a run's buckets are already in order and it is a copy to A.
    const uint32_t p0 = fshift / 8u;
    const uint32_t G = 1u << (8u * p0 + 8u - fshift);
    uint32_t lower = 0;
    for (uint32_t p = 0; p < 8; ++p)
      if (((x & low_bits(fshift)) >> (8 * p)) & 0xFFull) lower |= 1u << p;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const uint32_t b = t * PER + i;
      if (c[i] > C) {
        // smallest m in (b + 1, WNB] with more small buckets before it: bucket m - 1
        const uint32_t want = (s_nl[b + 1] & 0xFFFFu) + 1u;
        uint32_t lo = b + 2, hi = WNB + 1;
        while (lo < hi) {
          const uint32_t mid = (lo + hi) / 2;
          if ((s_nl[mid] & 0xFFFFu) >= want) hi = mid;
          else lo = mid + 1;
        }
        s_next[b] = 0x80000000u | (lo <= (uint32_t)WNB ? lo - 1 : (uint32_t)WNB);
      } else {
        const uint32_t base = s_q[b];
        uint32_t lo = b + 1, hi = umin((b / G + 1u) * G, (uint32_t)WNB);
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) / 2;
          if (s_q[mid] - base <= C) lo = mid;
          else hi = mid - 1;
        }
        s_next[b] = lo;
      }
    }
    __syncthreads();
    if (t == 0) {
      uint32_t R = 0;
      for (uint32_t b = 0; b < (uint32_t)WNB;) {
        const uint32_t v = s_next[b];
        if (v & 0x80000000u) {
          b = v & 0xFFFFu;
          continue;
        }
        const uint32_t rn = s_q[v] - s_q[b];
        if (rn) {
          s_rlo[R] = s_start[b];
          s_rn[R] = rn;
          s_rb[R] = (s_nl[v] >> 16) - (s_nl[b] >> 16);
          ++R;
        }
        b = v;
      }
      uint32_t base = R ? atomicAdd(&a.run_ctr[0], R) : 0u;
      if (R && (uint64_t)base + R > a.cap_runs) {
        atomicAdd(&a.run_ctr[2], 1u);  // cannot happen within the workspace's capacities
        R = 0;
      }
      s_misc[2] = R;
      s_misc[3] = base;
    }
    __syncthreads();
    const uint32_t R = s_misc[2], rbase = s_misc[3];
    for (uint32_t r = t; r < R; r += 256) {
      const bool one = s_rb[r] == 1u;
      uint32_t digits = 0, passes = 0;
      if (cd >= 0) {
        digits = one ? lower : (lower | (1u << p0));
        passes = wide_finish_passes(digits, fshift, s_rn[r] / s_rb[r]);
      }
      MsdRun run = {(out_b ? E.b : E.a) + 16ull * s_rlo[r], E.a + 16ull * s_rlo[r], s_rn[r],
                    digits | (passes << 16)};
      a.runs[rbase + r] = run;
    }
    __syncthreads();
  }
}

template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_msd_wscatter(MsdArgs a) {
  static_assert(BLOCK * ITEMS == (int)kMsdTile, "one onesweep tile per level tile");
  constexpr int WNB = 1 << kMsdWideBits;
  // per-wave counter rows of 2^kMsdWideBits buckets do not fit beside the
  // tile: they share the staging buffer
  __shared__ TileSmem<kMsdWideBits, BLOCK, ITEMS> sm;
  MsdLevel* wl = &a.wlvl[a.wstep];
  const uint32_t tiles = umin(wl->tiles, a.cap_wide_tiles);
  const uint32_t nw = umin(wl->entries, a.cap_wide);
  uint32_t* ticket = &wl->ticket;
  if (threadIdx.x == 0) sm.scalar[3] = atomicAdd(ticket, 1u);
  __syncthreads();
  while (true) {
    const uint32_t tl = sm.scalar[3];
    if (tl >= tiles) return;
    const uint32_t e = a.wtile[tl];
    bool took = false;
    if (e < nw) {
      const MsdEntry E = a.went[e];
      const uint32_t local = tl - E.tile_begin;
      if ((E.flags & kMsdMove) && tl >= E.tile_begin && (uint64_t)local * kMsdTile < E.n) {
        const bool in_b = (E.flags & kMsdInB) != 0;
        onesweep_tile<kMsdWideBits, BLOCK, ITEMS>(
            sm, reinterpret_cast<const elem_t*>(in_b ? E.b : E.a),
            reinterpret_cast<elem_t*>(in_b ? E.a : E.b), E.n, local,
            a.wstatus + (size_t)E.tile_begin * WNB, nullptr, a.wcounts + (size_t)e * WNB, E.dig,
            (uint32_t)WNB - 1u, a.ko, ticket, nullptr);
        took = true;
      }
    }
    if (!took && threadIdx.x == 0) sm.scalar[3] = atomicAdd(ticket, 1u);
    __syncthreads();
  }
}

// ----------------------------------------------------------------- launch
namespace {

template <class K>
uint32_t resident_blocks(K kernel, int block) {
  int occ = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, 0) != hipSuccess) {
    (void)hipGetLastError();
    occ = 1;
  }
  return occ > 0 ? (uint32_t)occ : 1u;
}

constexpr int kScatterBlock = 1024, kScatterItems = 8;
constexpr int kFinishBlock = 1024, kFinishItems = 8;

}  // namespace

hipError_t launch_msd_init(const MsdArgs& a, uint32_t grid, hipStream_t s) {
  hipLaunchKernelGGL(k_msd_init, dim3(grid), dim3(256), 0, s, a);
  return hipGetLastError();
}

namespace {
hipError_t mark(hipEvent_t* after, int i, hipStream_t s) {
  return after ? hipEventRecord(after[i], s) : hipGetLastError();
}
}  // namespace

hipError_t launch_msd_wide(const MsdArgs& a, uint32_t cus, hipStream_t s, hipEvent_t* after) {
  hipError_t e;
  hipLaunchKernelGGL(k_msd_whist, dim3(2 * cus), dim3(256), 0, s, a);
  if ((e = mark(after, 0, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_msd_wplan, dim3(cus), dim3(256), 0, s, a);
  if ((e = mark(after, 1, s)) != hipSuccess) return e;
  hipLaunchKernelGGL((k_msd_wscatter<kScatterBlock, kScatterItems>), dim3(cus),
                     dim3(kScatterBlock), 0, s, a);
  if ((e = mark(after, 2, s)) != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t launch_msd_level(const MsdArgs& a, uint32_t cus, hipStream_t s, hipEvent_t* after) {
  static const uint32_t scatter_occ =
      resident_blocks(k_msd_scatter<kScatterBlock, kScatterItems>, kScatterBlock);
  hipError_t e;
  hipLaunchKernelGGL(k_msd_hist, dim3(2 * cus), dim3(256), 0, s, a);
  if ((e = mark(after, 0, s)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_msd_plan, dim3(4 * cus), dim3(256), 0, s, a);
  if ((e = mark(after, 1, s)) != hipSuccess) return e;
  hipLaunchKernelGGL((k_msd_scatter<kScatterBlock, kScatterItems>), dim3(scatter_occ * cus),
                     dim3(kScatterBlock), 0, s, a);
  if (a.level == 0) {
    // the region form of a level 0 of one large range (no-op otherwise)
    static const uint32_t rg_occ =
        resident_blocks(k_msd_scatter_rg<kScatterBlock, kScatterItems>, kScatterBlock);
    hipLaunchKernelGGL((k_msd_scatter_rg<kScatterBlock, kScatterItems>), dim3(rg_occ * cus),
                       dim3(kScatterBlock), 0, s, a);
  }
  if ((e = mark(after, 2, s)) != hipSuccess) return e;
  return hipGetLastError();
}

hipError_t launch_msd_finish(const MsdArgs& a, uint32_t cus, hipStream_t s) {
  static const uint32_t occ =
      resident_blocks(k_msd_finish<kFinishBlock, kFinishItems>, kFinishBlock);
  hipLaunchKernelGGL((k_msd_finish<kFinishBlock, kFinishItems>), dim3(occ * cus),
                     dim3(kFinishBlock), 0, s, a);
  return hipGetLastError();
}

}  // namespace lsb
