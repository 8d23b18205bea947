// Device building blocks of one stable counting-sort pass over 16-byte
// elements, shared by the plain sort (k_onesweep), the segmented sort
// (k_seg_onesweep) and the LDS-resident small-array sort (k_small_sort).
//
// A tile is BLOCK*ITEMS elements.  Wave w owns the contiguous run of
// 64*ITEMS elements starting at w*64*ITEMS; its item i, lane l is element
// i*64 + l of that run, so every global load of a wave reads one contiguous
// 1 KiB, and visiting (wave, item, lane) in order visits the tile in input
// order -- which is what makes every pass stable.
#pragma once

#include "lsb_device.hpp"

namespace lsb {

// Regions per scatter pass (kRegions in lsb_internal.hpp).
constexpr int kRegionsDev = 8;

// ------------------------------------------------------------- wave_rank
// Ranks this wave's ITEMS rows into its LDS counters wh[0 .. 2^D).  Row i,
// lane l is valid when i*64 + l < lim.  For every valid element: dg = its
// digit, rk = number of elements of the same digit earlier in the wave's
// run.  Peers (lanes holding the same digit) are found with D ballots.  The
// lowest peer of each group adds the group's size to the counter with one
// returning LDS atomic; the rows' atomics go out back to back (LDS
// operations of one wave complete in order, so row i+1 sees row i's adds),
// and each lane then fetches its group leader's old count with
// ds_bpermute.  No row waits on the previous row's LDS round trip.
// (wave_rank_digits takes the rows' digits already computed in dg.)  With
// SKIP, the ballots of row groups holding no valid element are skipped (a
// branch on the wave-uniform lim): worth it where runs often fill only part
// of the block.
template <int D, int ITEMS, bool SKIP = false>
__device__ __forceinline__ void wave_rank_digits(const uint32_t (&dg)[ITEMS], uint32_t lim,
                                                 uint32_t* wh, uint32_t (&rk)[ITEMS]) {
  const uint32_t lane = lane_id();
  const uint32_t ulim = SKIP ? (uint32_t)__builtin_amdgcn_readfirstlane(lim) : lim;
  uint32_t old[ITEMS], lead[ITEMS], plo[ITEMS], phi[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    const uint64_t vmask = __ballot((uint32_t)i * kWave + lane < lim);
    plo[i] = (uint32_t)vmask;
    phi[i] = (uint32_t)(vmask >> 32);
  }
  // Per bit: c = the lane's bit broadcast to 0 / ~0 (v_bfe_i32), the wave's
  // ballot of it, and peers &= ~(ballot ^ c) on both halves with one v_bitop3
  // each (truth table 0x90 = a & (b == c)).  Bits outside, a group of rows
  // inside: the rows are independent, so one row's instructions fill the
  // wait states between the other's ballot and its use.
  // Rows in pairs: with four or more at once the wider live set spills.
  constexpr int G = ITEMS < 2 ? ITEMS : 2;
  static_assert(ITEMS % G == 0, "rows come in whole groups");
#pragma unroll
  for (int i0 = 0; i0 < ITEMS; i0 += G) {
    if (SKIP && (uint32_t)i0 * kWave >= ulim) {
      // no valid element from here on: peers stay empty, no counter changes
#pragma unroll
      for (int i = i0; i < ITEMS; ++i) {
        old[i] = 0;
        lead[i] = lane;
        rk[i] = 0;
      }
      break;
    }
#pragma unroll
    for (int b = 0; b < D; ++b) {
#pragma unroll
      for (int i = i0; i < i0 + G; ++i) {
        // A pure VALU op in asm: the compiler would otherwise rewrite the
        // field extract as two shifts and test the sign of the first.
        uint32_t c;
        asm("v_bfe_i32 %0, %1, %2, 1" : "=v"(c) : "v"(dg[i]), "i"(b));
        const uint64_t bal = __ballot(c != 0u);
        plo[i] = __builtin_amdgcn_bitop3_b32(plo[i], (uint32_t)bal, c, 0x90);
        phi[i] = __builtin_amdgcn_bitop3_b32(phi[i], (uint32_t)(bal >> 32), c, 0x90);
      }
    }
#pragma unroll
    for (int i = i0; i < i0 + G; ++i) {
      const bool valid = (uint32_t)i * kWave + lane < lim;
      const uint64_t peers = ((uint64_t)phi[i] << 32) | plo[i];
      const uint32_t below = count_below(peers);
      old[i] = 0;
      if (valid && below == 0) old[i] = atomicAdd(&wh[dg[i]], (uint32_t)__popcll(peers));
      lead[i] = peers ? (uint32_t)__builtin_ctzll(peers) : lane;
      rk[i] = below;
    }
  }
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    rk[i] += (uint32_t)__builtin_amdgcn_ds_bpermute((int)(lead[i] << 2), (int)old[i]);
}

template <int D, int ITEMS, bool SKIP = false>
__device__ __forceinline__ void wave_rank(const elem_t (&e)[ITEMS], uint32_t lim, uint32_t shift,
                                          uint32_t mask, uint32_t* wh, uint32_t (&dg)[ITEMS],
                                          uint32_t (&rk)[ITEMS]) {
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) dg[i] = digit_of(e[i].key, shift, mask);
  wave_rank_digits<D, ITEMS, SKIP>(dg, lim, wh, rk);
}

// ------------------------------------------------------------- look-back
// Exclusive prefix of digit b over tiles [0, tile) of one pass: walks back
// from tile-1 adding AGGREGATE words until an INCLUSIVE word ends the walk.
// Each round issues kLookback independent status loads, so a run of
// aggregates costs one memory round trip per kLookback tiles.  Tile 0 always
// publishes an inclusive word, so the walk never passes it; slots before
// tile 0 read as "not ready" and are never consumed.
constexpr int kLookback = 4;

// Issues the first round of a look-back (tiles tile-1 .. tile-kLookback)
// without waiting for it, so the caller can do other work meanwhile.
__device__ __forceinline__ void lookback_issue(const uint32_t* __restrict__ st, uint32_t tile,
                                               uint32_t nb, uint32_t b,
                                               uint32_t (&w)[kLookback]) {
  const int64_t j = (int64_t)tile - 1;
#pragma unroll
  for (int k = 0; k < kLookback; ++k)
    w[k] = (j - k >= 0) ? ld_status(st + (size_t)(j - k) * nb + b) : 0u;
}

// Finishes a look-back whose first round lookback_issue read into w.
__device__ __forceinline__ uint32_t lookback_finish(const uint32_t* __restrict__ st,
                                                    uint32_t tile, uint32_t nb, uint32_t b,
                                                    uint32_t (&w)[kLookback]) {
  uint32_t prefix = 0;
  int64_t j = (int64_t)tile - 1;
  bool first = true;
  while (true) {
    if (!first) {
#pragma unroll
      for (int k = 0; k < kLookback; ++k)
        w[k] = (j - k >= 0) ? ld_status(st + (size_t)(j - k) * nb + b) : 0u;
    }
    first = false;
    uint32_t used = 0;
    bool stop = false, done = false;
#pragma unroll
    for (int k = 0; k < kLookback; ++k) {
      if (!stop) {
        if (w[k] == 0u) {
          stop = true;
        } else {
          prefix += (w[k] & kCountMask) - 1u;
          ++used;
          if (w[k] & kFlagInclusive) stop = done = true;
        }
      }
    }
    if (done) return prefix;
    j -= used;
    if (used == 0) __builtin_amdgcn_s_sleep(1);
  }
}

__device__ __forceinline__ uint32_t lookback(const uint32_t* __restrict__ st, uint32_t tile,
                                             uint32_t nb, uint32_t b) {
  uint32_t w[kLookback];
  lookback_issue(st, tile, nb, b, w);
  return lookback_finish(st, tile, nb, b, w);
}

// Shared memory of one tile pass.  The element buffer first holds the
// per-wave digit counters (while the elements sit in registers), then the
// elements in digit order.
template <int D, int BLOCK, int ITEMS>
struct TileSmem {
  static constexpr int NB = 1 << D;
  static constexpr int W = BLOCK / kWave;
  static constexpr int T = BLOCK * ITEMS;
  static_assert(W * NB * 4 <= T * 16, "per-wave counters must fit in the tile buffer");
  elem_t elems[T];
  uint32_t start[NB];  // tile-local exclusive bucket starts
  // per bucket: x = destination index of tile-local slot 0; y (region form
  // only) = (first slot in the next region << 3) | region.  Interleaved so
  // the store phase reads both with one 8-byte LDS load.
  uint2 grs[NB];
  uint32_t wsum[W];
  uint32_t scalar[4];
};

// ------------------------------------------------------------- tile_load
// Loads tile `tile` of an array of n elements into registers in the
// wave-contiguous layout; slots past the end read as zero.
template <int BLOCK, int ITEMS>
__device__ __forceinline__ void tile_load(elem_t (&e)[ITEMS], const elem_t* __restrict__ src,
                                          uint32_t n, uint32_t tile) {
  constexpr uint32_t T = BLOCK * ITEMS;
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t tile_base = tile * T;
  const uint32_t tile_n = n - tile_base < T ? n - tile_base : T;
  const uint32_t wrun = wave * (kWave * ITEMS);
  const uint32_t lim = tile_n > wrun ? tile_n - wrun : 0u;
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if ((uint32_t)i * kWave + lane < lim) {
      e[i] = gload(src + (size_t)tile_base + wrun + (uint32_t)i * kWave + lane);
    } else {
      e[i].key = 0;
      e[i].val = 0;
    }
  }
}

// ------------------------------------------------------------ store_run
// Writes a tile held in LDS in digit order (slot j of bucket d goes to
// dst[grs[d].x + j]).  With COUNT, every element is also counted in hnext
// by the region its destination falls in (grs[d].y = (first slot of the
// next region << 3) | region) and by its digit for the next executed pass.
// A full tile takes a branch-free path, so the LDS reads of all ITEMS rows
// are in flight together instead of one row's round trips at a time.
template <int BLOCK, int ITEMS, int NB, bool COUNT>
__device__ __forceinline__ void store_run(const elem_t* elems, const uint2* grs, uint32_t* hnext,
                                          uint32_t tile_n, elem_t* __restrict__ dst,
                                          uint32_t shift, uint32_t mask, uint32_t nshift,
                                          uint32_t nmask, const KeyOrder& ko) {
  const uint32_t t = threadIdx.x;
  if (tile_n == (uint32_t)(BLOCK * ITEMS)) {
    elem_t x[ITEMS];
    uint32_t d[ITEMS];
    uint2 g[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) x[k] = elems[k * BLOCK + t];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) d[k] = digit_of(x[k].key, shift, mask);
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) g[k] = grs[d[k]];
    if (COUNT) {
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) {
        const uint32_t j = k * BLOCK + t;
        const uint32_t r = (g[k].y & 7u) + (j >= (g[k].y >> 3) ? 1u : 0u);
        atomicAdd(&hnext[r * NB + digit_of(x[k].key, nshift, nmask)], 1u);
      }
    }
    if (ko.sign) {
#pragma unroll
      for (int k = 0; k < ITEMS; ++k) x[k].key = key_inv(x[k].key, ko);
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) gstore(dst + (size_t)(g[k].x + (uint32_t)(k * BLOCK) + t), x[k]);
    return;
  }
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t j = k * BLOCK + t;
    if (j < tile_n) {
      elem_t x = elems[j];
      const uint32_t d = digit_of(x.key, shift, mask);
      const uint2 g = grs[d];
      if (COUNT) {
        const uint32_t r = (g.y & 7u) + (j >= (g.y >> 3) ? 1u : 0u);
        atomicAdd(&hnext[r * NB + digit_of(x.key, nshift, nmask)], 1u);
      }
      x.key = key_inv(x.key, ko);
      gstore(dst + (size_t)(g.x + j), x);
    }
  }
}

// ------------------------------------------------------------ count_run
// The counting half of store_run<..., true> on its own: counts the tile
// still held in LDS (elements in digit order, grs as store_run reads it) into
// hnext by destination region and next digit.  Run while the block waits
// for its next tile's loads, it takes that work off the store phase.
template <int BLOCK, int ITEMS, int NB>
__device__ __forceinline__ void count_run(const elem_t* elems, const uint2* grs, uint32_t* hnext,
                                          uint32_t tile_n, uint32_t shift, uint32_t mask,
                                          uint32_t nshift, uint32_t nmask) {
  const uint32_t t = threadIdx.x;
  if (tile_n == (uint32_t)(BLOCK * ITEMS)) {
    uint64_t key[ITEMS];
    uint2 g[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) key[k] = elems[k * BLOCK + t].key;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) g[k] = grs[digit_of(key[k], shift, mask)];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const uint32_t j = k * BLOCK + t;
      const uint32_t r = (g[k].y & 7u) + (j >= (g[k].y >> 3) ? 1u : 0u);
      atomicAdd(&hnext[r * NB + digit_of(key[k], nshift, nmask)], 1u);
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const uint32_t j = k * BLOCK + t;
    if (j < tile_n) {
      const uint64_t key = elems[j].key;
      const uint2 g = grs[digit_of(key, shift, mask)];
      const uint32_t r = (g.y & 7u) + (j >= (g.y >> 3) ? 1u : 0u);
      atomicAdd(&hnext[r * NB + digit_of(key, nshift, nmask)], 1u);
    }
  }
}

// --------------------------------------------------------- onesweep_tile
// One tile of one pass over an array (or one segment) of n elements whose
// elements are already in registers (tile_load): dst points at element 0 of
// that array; `tile` counts from its first tile; st_cur/st_next are the
// array's rows of the two look-back slices; starts[b] is the global
// exclusive start of bucket b for this digit.  Must be entered by the whole
// block after a barrier that ends every earlier use of sm.
template <int D, int BLOCK, int ITEMS>
__device__ __forceinline__ void onesweep_tile_regs(TileSmem<D, BLOCK, ITEMS>& sm,
                                                   const elem_t (&e)[ITEMS],
                                                   elem_t* __restrict__ dst, uint32_t n,
                                                   uint32_t tile, uint32_t* st_cur,
                                                   uint32_t* st_next,
                                                   const uint32_t* __restrict__ starts,
                                                   uint32_t shift, uint32_t mask,
                                                   const KeyOrder& ko,
                                                   uint32_t* next_ticket = nullptr,
                                                   uint32_t* own = nullptr) {
  using S = TileSmem<D, BLOCK, ITEMS>;
  constexpr int NB = S::NB, W = S::W, T = S::T;
  constexpr int DPT = (NB + BLOCK - 1) / BLOCK;  // digits per thread
  const uint32_t t = threadIdx.x;
  const uint32_t lane = t & (kWave - 1);
  const uint32_t wave = t / kWave;
  // Per-wave counters: in `own` (W * NB words of their own) each wave clears
  // its row and ranks at once -- since the last tile's staging barrier only
  // that wave has touched the row; else in the staging buffer, cleared by
  // all threads behind a barrier.
  uint32_t* whist = own ? own : reinterpret_cast<uint32_t*>(sm.elems);
  if (own) {
    for (uint32_t i = lane; i < (uint32_t)NB; i += kWave) whist[wave * NB + i] = 0u;
  } else {
    for (uint32_t i = t; i < (uint32_t)(W * NB); i += BLOCK) whist[i] = 0u;
  }
  // The next pass's row of this tile starts clear (null: the caller clears
  // its status rows itself).
  if (st_next)
    for (uint32_t b = t; b < (uint32_t)NB; b += BLOCK) st_next[(size_t)tile * NB + b] = 0u;

  const uint32_t tile_base = tile * (uint32_t)T;
  const uint32_t tile_n = n - tile_base < (uint32_t)T ? n - tile_base : (uint32_t)T;
  const uint32_t wrun = wave * (kWave * ITEMS);
  const uint32_t lim = tile_n > wrun ? tile_n - wrun : 0u;
  if (!own) lds_barrier();

  uint32_t dg[ITEMS], rk[ITEMS];
  uint32_t* wh = whist + wave * NB;
  wave_rank<D, ITEMS>(e, lim, shift, mask, wh, dg, rk);
  lds_barrier();

  // Per digit: exclusive prefix over waves (in place) and the tile total,
  // published at once as this tile's aggregate.  The first look-back round
  // is issued right away and consumed only after the tile-local scan and
  // the reorder through LDS, so its memory latency hides under that work.
  uint32_t tot[DPT], gstart[DPT];
  uint32_t lbw[DPT][kLookback];
#pragma unroll
  for (int k = 0; k < DPT; ++k) {
    const uint32_t b = t + k * BLOCK;
    tot[k] = 0;
    gstart[k] = 0;
    if (b < (uint32_t)NB) {
      gstart[k] = starts[b];
      uint32_t s = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t c = whist[w * NB + b];
        whist[w * NB + b] = s;
        s += c;
      }
      tot[k] = s;
      sm.start[b] = s;
      uint32_t* mine = st_cur + (size_t)tile * NB + b;
      if (tile == 0) {
        st_status(mine, (s + 1u) | kFlagInclusive);
      } else {
        st_status(mine, s + 1u);
        lookback_issue(st_cur, tile, NB, b, lbw[k]);
      }
    }
  }
  // The scan also adds each bucket's tile start into every wave's row, so
  // an element's slot is one LDS read: wh[digit] + rank.
  block_exclusive_scan<NB, BLOCK, W, DPT == 1>(sm.start, sm.wsum, whist);

  uint32_t pos[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) pos[i] = wh[dg[i]] + rk[i];
  lds_barrier();  // counters are dead from here: the buffer takes elements
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if ((uint32_t)i * kWave + lane < lim) sm.elems[pos[i]] = e[i];

#pragma unroll
  for (int k = 0; k < DPT; ++k) {
    const uint32_t b = t + k * BLOCK;
    if (b < (uint32_t)NB) {
      uint32_t prefix = 0;
      if (tile > 0) {
        prefix = lookback_finish(st_cur, tile, NB, b, lbw[k]);
        st_status(st_cur + (size_t)tile * NB + b, (prefix + tot[k] + 1u) | kFlagInclusive);
      }
      sm.grs[b].x = gstart[k] + prefix - sm.start[b];
    }
  }
  lds_barrier();

  // A persistent caller's next ticket is requested before the store phase
  // so its latency hides under the stores; thread 0 leaves it in
  // sm.scalar[3].
  uint32_t nt;
  if (next_ticket && t == 0) nt = atomicAdd(next_ticket, 1u);

  // Write out in digit order: consecutive lanes write consecutive slots of
  // one bucket run.
  store_run<BLOCK, ITEMS, NB, false>(sm.elems, sm.grs, nullptr, tile_n, dst, shift, mask, 0u, 0u,
                                     ko);
  if (next_ticket && t == 0) sm.scalar[3] = nt;
}

// ------------------------------------------------------- region form
// One tile of a pass whose array is cut into kRegions regions, each with its
// own ticket counter and look-back chain (`chain_first` = the region's first
// tile, which publishes an inclusive word at once).  rbase[b] (LDS) is the
// destination of the region's first element of bucket b: the global bucket
// start plus the bucket's elements in earlier regions.  Two more duties:
//  * the block's next ticket is requested before the store phase, so its
//    latency hides under the stores (thread 0 leaves it in sm.scalar[0]).
//    It is not requested earlier: a ticket held by a block that cannot yet
//    start its tile delays every later tile's look-back in that region;
//  * when hnext is given, every element is counted by the region its
//    destination falls in and by its digit for the next executed pass
//    (nshift, nmask) -- the per-region counts that pass needs.  With
//    count_later the store phase leaves that to the caller, who runs
//    count_run over the tile left in LDS before the block's next tile
//    overwrites it.
template <int D, int BLOCK, int ITEMS>
__device__ __forceinline__ void onesweep_tile_region(
    TileSmem<D, BLOCK, ITEMS>& sm, const elem_t (&e)[ITEMS], elem_t* __restrict__ dst,
    uint32_t n, uint32_t tile, bool chain_first, uint32_t* st_cur, uint32_t* st_next,
    const uint32_t* rbase, uint32_t shift, uint32_t mask, uint32_t* ticket, uint32_t* hnext,
    uint32_t nshift, uint32_t nmask, uint32_t region_elems, const KeyOrder& ko,
    uint32_t* whist, bool count_later) {
  using S = TileSmem<D, BLOCK, ITEMS>;
  constexpr int NB = S::NB, W = S::W, T = S::T;
  constexpr int DPT = (NB + BLOCK - 1) / BLOCK;
  const uint32_t t = threadIdx.x;
  const uint32_t lane = t & (kWave - 1);
  const uint32_t wave = t / kWave;
  // whist: the per-wave counters.  Given their own LDS (W * NB words) they
  // are zero on entry and cleared again once this tile's slots are known,
  // so a tile starts ranking without a barrier of its own; passed as null
  // they share the staging buffer (blocks whose LDS has no room to spare)
  // and are cleared here, behind a barrier.
  const bool own_counters = whist != nullptr;
  if (!own_counters) {
    whist = reinterpret_cast<uint32_t*>(sm.elems);
    for (uint32_t i = t; i < (uint32_t)(W * NB); i += BLOCK) whist[i] = 0u;
  }
  // (null st_next: no later pass reads this chain's other slice)
  if (st_next)
    for (uint32_t b = t; b < (uint32_t)NB; b += BLOCK) st_next[(size_t)tile * NB + b] = 0u;
  const uint32_t tile_base = tile * (uint32_t)T;
  const uint32_t tile_n = n - tile_base < (uint32_t)T ? n - tile_base : (uint32_t)T;
  const uint32_t wrun = wave * (kWave * ITEMS);
  const uint32_t lim = tile_n > wrun ? tile_n - wrun : 0u;
  if (!own_counters) lds_barrier();

  uint32_t dg[ITEMS], rk[ITEMS];
  uint32_t* wh = whist + wave * NB;
  wave_rank<D, ITEMS>(e, lim, shift, mask, wh, dg, rk);
  lds_barrier();

  uint32_t tot[DPT], gstart[DPT];
  uint32_t lbw[DPT][kLookback] = {};
#pragma unroll
  for (int k = 0; k < DPT; ++k) {
    const uint32_t b = t + k * BLOCK;
    tot[k] = 0;
    gstart[k] = 0;
    if (b < (uint32_t)NB) {
      gstart[k] = rbase[b];
      uint32_t s = 0;
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const uint32_t c = whist[w * NB + b];
        whist[w * NB + b] = s;
        s += c;
      }
      tot[k] = s;
      sm.start[b] = s;
      uint32_t* mine = st_cur + (size_t)tile * NB + b;
      if (chain_first) {
        st_status(mine, (s + 1u) | kFlagInclusive);
      } else {
        st_status(mine, s + 1u);
        lookback_issue(st_cur, tile, NB, b, lbw[k]);
      }
    }
  }
  // The scan also adds each bucket's tile start into every wave's row, so
  // an element's slot is one LDS read: wh[digit] + rank.
  block_exclusive_scan<NB, BLOCK, W, DPT == 1>(sm.start, sm.wsum, whist);

  uint32_t pos[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) pos[i] = wh[dg[i]] + rk[i];
  lds_barrier();
  if (own_counters)
    for (uint32_t i = t; i < (uint32_t)(W * NB); i += BLOCK) whist[i] = 0u;  // for the next tile
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if ((uint32_t)i * kWave + lane < lim) sm.elems[pos[i]] = e[i];

#pragma unroll
  for (int k = 0; k < DPT; ++k) {
    const uint32_t b = t + k * BLOCK;
    if (b < (uint32_t)NB) {
      uint32_t prefix = 0;
      if (!chain_first) {
        prefix = lookback_finish(st_cur, tile, NB, b, lbw[k]);
        st_status(st_cur + (size_t)tile * NB + b, (prefix + tot[k] + 1u) | kFlagInclusive);
      }
      // The bucket's run is contiguous in dst and no longer than a region,
      // so it crosses at most one region boundary: note its first region and
      // the tile slot where the next region begins.
      const uint32_t first = gstart[k] + prefix;
      const uint32_t st = sm.start[b];
      uint32_t info = 0;
      if (hnext) {
        uint32_t r0 = 0;
#pragma unroll
        for (int q = 1; q < kRegionsDev; ++q) r0 += first >= (uint32_t)q * region_elems ? 1u : 0u;
        // slots at or past `split` fall in region r0 + 1 (none when the
        // boundary lies beyond this tile; clamped so the shift cannot
        // overflow)
        uint32_t split = st + (r0 + 1u) * region_elems - first;
        if (split > (uint32_t)T) split = (uint32_t)T;
        info = (split << 3) | r0;
      }
      sm.grs[b] = make_uint2(first - st, info);
    }
  }
  // Take every look-back load, used or not, on every path: one the compiler
  // cannot prove retired makes it wait for all outstanding memory operations
  // (the next tile's loads included) before it next reuses that register.
#pragma unroll
  for (int k = 0; k < DPT; ++k)
#pragma unroll
    for (int j = 0; j < kLookback; ++j) asm volatile("" ::"v"(lbw[k][j]));
  lds_barrier();

  uint32_t next;
  if (t == 0) next = atomicAdd(ticket, 1u);
  if (hnext && !count_later)
    store_run<BLOCK, ITEMS, NB, true>(sm.elems, sm.grs, hnext, tile_n, dst, shift, mask, nshift,
                                      nmask, ko);
  else
    store_run<BLOCK, ITEMS, NB, false>(sm.elems, sm.grs, hnext, tile_n, dst, shift, mask, nshift,
                                       nmask, ko);
  if (t == 0) sm.scalar[0] = next;
}

// Loads and processes one tile (the whole block, after a barrier).
template <int D, int BLOCK, int ITEMS>
__device__ __forceinline__ void onesweep_tile(TileSmem<D, BLOCK, ITEMS>& sm,
                                              const elem_t* __restrict__ src,
                                              elem_t* __restrict__ dst, uint32_t n,
                                              uint32_t tile, uint32_t* st_cur,
                                              uint32_t* st_next,
                                              const uint32_t* __restrict__ starts,
                                              uint32_t shift, uint32_t mask,
                                              const KeyOrder& ko,
                                              uint32_t* next_ticket = nullptr,
                                              uint32_t* own = nullptr) {
  elem_t e[ITEMS];
  tile_load<BLOCK, ITEMS>(e, src, n, tile);
  keys_fwd<ITEMS>(e, ko);
  onesweep_tile_regs<D, BLOCK, ITEMS>(sm, e, dst, n, tile, st_cur, st_next, starts, shift, mask,
                                      ko, next_ticket, own);
}

// ----------------------------------------------------------- small_sort
// Sorts n <= BLOCK*ITEMS elements of one array entirely inside the block:
// every digit pass ranks the elements held in registers, writes them to
// LDS in digit order and reloads them in the wave-contiguous layout.  Only
// one read of src and one write of dst per element, whatever the pass count
// (src may equal dst: every element is loaded before the first barrier and
// stored after the last).  Digits p with bit p of `digits` set are sorted,
// lowest first; digits whose elements all share one bucket are skipped.
template <int D, int BLOCK, int ITEMS>
__device__ __forceinline__ void small_sort(TileSmem<D, BLOCK, ITEMS>& sm, const elem_t* src,
                                           elem_t* dst, uint32_t n, uint32_t key_bits,
                                           uint32_t digits, const KeyOrder& ko) {
  using S = TileSmem<D, BLOCK, ITEMS>;
  constexpr int NB = S::NB, W = S::W;
  const uint32_t t = threadIdx.x;
  const uint32_t lane = t & (kWave - 1);
  const uint32_t wave = t / kWave;
  uint32_t* whist = reinterpret_cast<uint32_t*>(sm.elems);
  const uint32_t wrun = wave * (kWave * ITEMS);
  const uint32_t lim = n > wrun ? n - wrun : 0u;

  elem_t e[ITEMS];
#pragma unroll
  for (int i = 0; i < ITEMS; ++i) {
    if ((uint32_t)i * kWave + lane < lim) {
      e[i] = gload(src + wrun + (uint32_t)i * kWave + lane);
    } else {
      e[i].key = 0;
      e[i].val = 0;
    }
  }
  keys_fwd<ITEMS>(e, ko);
  const uint32_t passes = (key_bits + D - 1) / D;
  for (uint32_t p = 0; p < passes; ++p) {
    if (!((digits >> p) & 1u)) continue;
    const uint32_t shift = p * D;
    const uint32_t w = key_bits - shift < (uint32_t)D ? key_bits - shift : (uint32_t)D;
    const uint32_t mask = (1u << w) - 1u;
    lds_barrier();  // previous pass's reload is done with the buffer
    for (uint32_t i = t; i < (uint32_t)(W * NB); i += BLOCK) whist[i] = 0u;
    if (t == 0) sm.scalar[0] = 0u;
    lds_barrier();
    uint32_t dg[ITEMS], rk[ITEMS];
    uint32_t* wh = whist + wave * NB;
    wave_rank<D, ITEMS>(e, lim, shift, mask, wh, dg, rk);
    lds_barrier();
    for (uint32_t b = t; b < (uint32_t)NB; b += BLOCK) {
      uint32_t s = 0;
#pragma unroll
      for (int k = 0; k < W; ++k) {
        const uint32_t c = whist[k * NB + b];
        whist[k * NB + b] = s;
        s += c;
      }
      sm.start[b] = s;
      if (s == n) sm.scalar[0] = 1u;  // one bucket holds everything
    }
    block_exclusive_scan<NB, BLOCK>(sm.start, sm.wsum);
    if (sm.scalar[0]) continue;  // identity pass (uniform across the block)
    uint32_t pos[ITEMS];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) pos[i] = sm.start[dg[i]] + wh[dg[i]] + rk[i];
    lds_barrier();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if ((uint32_t)i * kWave + lane < lim) sm.elems[pos[i]] = e[i];
    lds_barrier();
#pragma unroll
    for (int i = 0; i < ITEMS; ++i)
      if ((uint32_t)i * kWave + lane < lim) e[i] = sm.elems[wrun + (uint32_t)i * kWave + lane];
  }
  if (ko.sign) {
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) e[i].key = key_inv(e[i].key, ko);
  }
  // With src == dst no element may be stored before every thread's loads of
  // it have completed: a full barrier (waits for this thread's loads too).
  __syncthreads();
#pragma unroll
  for (int i = 0; i < ITEMS; ++i)
    if ((uint32_t)i * kWave + lane < lim) gstore(dst + wrun + (uint32_t)i * kWave + lane, e[i]);
}

// Whole sort of one array in place: every digit of the key.
template <int D, int BLOCK, int ITEMS>
__device__ __forceinline__ void small_sort(TileSmem<D, BLOCK, ITEMS>& sm, elem_t* a, uint32_t n,
                                           uint32_t key_bits, const KeyOrder& ko) {
  small_sort<D, BLOCK, ITEMS>(sm, a, a, n, key_bits, ~0u, ko);
}

}  // namespace lsb
