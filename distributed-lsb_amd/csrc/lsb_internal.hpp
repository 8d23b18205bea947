// Host-side interface between the C ABI (lsb_api.cpp) and the kernel
// launchers (lsb_kernels.hip, lsb_workload.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lsb {

// Key order (key_fwd / key_inv in lsb_device.hpp): all zero for unsigned
// keys; sign = bit key_bits-1 for signed and float keys; fsign = sign for
// float keys; full = the low key_bits bits.
struct KeyOrder {
  uint64_t sign;
  uint64_t fsign;
  uint64_t full;
};

// Each scatter pass cuts the array into this many regions of whole tiles;
// every region has its own tile tickets and look-back chain, so eight
// chains advance side by side instead of one.
constexpr int kRegions = 8;

// Per-pass plan written by the scan kernel, read by the scatter kernels:
// plan[4*p + 0] = pass p executes, [4*p+1] = its ordinal among executed
// passes, [4*p+2] = its source is B (else A).  plan[4*P] = 1 when the
// executed-pass count is odd, so the result sits in B and is copied to A.
struct PassArgs {
  void* A;
  void* B;
  uint32_t n;
  uint32_t pass;
  uint32_t shift;
  uint32_t mask;
  const uint32_t* plan;
  const uint32_t* starts;  // starts of this pass: global exclusive bucket offsets
                           // (the region form makes its own from the counts)
  uint32_t* tickets;       // tile ticket counters [pass][region]
  uint32_t* status;        // two look-back slices of tiles * 2^d words
  uint32_t ntiles;
  // region form
  uint32_t passes;
  uint32_t key_bits;
  uint32_t region_tiles;        // tiles per region (the last region may hold fewer)
  const uint32_t* regcount;     // histogram kernel's per-region counts [P][R][2^d]
  uint32_t* regnext;            // per-region counts filled by the pass before [P][R][2^d]
  KeyOrder ko;                  // key order (identity for unsigned keys)
};

typedef void (*ScatterLaunch)(const PassArgs&, hipStream_t);

struct SortVariant {
  uint32_t digit_bits;
  uint32_t tile;  // elements per tile (block * items)
  uint32_t block;
  uint32_t persistent;  // 0 = one tile per block, 3 = region form (persistent),
                        // 1 = persistent with one look-back chain
  ScatterLaunch scatter;
};

// Compiled variants (lsb_kernels.hip).
const SortVariant* sort_variants(int* count);

// Stream-ordered allocation from the library's own memory pool on the
// current device (freed with hipFreeAsync); the pool keeps freed memory for
// reuse (lsb_api.cpp).
hipError_t pool_alloc(void** p, size_t bytes, hipStream_t s);
// Returns the unused memory of device dev's library pool to the device.
hipError_t pool_trim(int dev);
// lsb_sort's host-side argument checks (null, alignment, overlap, n).
int check_sort_args(const void* A, const void* B, size_t n);
// Records e as this thread's lsb_last_hip_error() and clears HIP's sticky
// last error.
void note_hip_error(hipError_t e);

// Zeroes p[0..words), launched before the histogram.
void launch_zero(uint32_t* p, size_t words, hipStream_t s);
// Digits the lean histogram counts up front (the region passes count each
// later digit for their successor while they scatter).
constexpr int kLeanDigits = 2;
// Histogram per region of digits 0 .. counted-1: regcount[p][region][b] +=
// ..., with regions of region_elems elements; zeroes zero[0..zero_words).
// counted == passes counts every digit; counted == kLeanDigits (d = 8 only)
// is the lean form, which also ORs every key's XOR with key 0 into orr[0..1]
// (zeroed beforehand; n >= 1).
// `regions` (1 .. kRegions) regions of region_elems elements are counted.
hipError_t launch_histogram(uint32_t d, const void* A, uint32_t n, uint32_t key_bits,
                            const KeyOrder& ko, uint32_t passes, uint32_t counted,
                            uint32_t* regcount, uint32_t region_elems, uint32_t regions,
                            uint32_t* zero, size_t zero_words, uint32_t* orr, hipStream_t s);
// Lean form, after the histogram: skip detection from orr, the plan, the
// cleared tickets[P][R] and regnext[P][R][2^d], and the per-region counts of
// the first executed pass when it is not one of the `counted` digits (d = 8).
hipError_t launch_plan_first(uint32_t d, const void* A, uint32_t n, uint32_t key_bits,
                             const KeyOrder& ko, uint32_t* regcount, uint32_t region_elems,
                             const uint32_t* orr, uint32_t* plan, uint32_t* tickets,
                             uint32_t* regnext, uint32_t passes, uint32_t counted,
                             uint32_t flags, hipStream_t s);
// Global counts (sum of regions), exclusive scans, skip detection, plan;
// clears tickets[P][R] and regnext[P][R][2^d] (all-digit histogram).
hipError_t launch_scan_plan(uint32_t d, const uint32_t* regcount, uint32_t* starts,
                            uint32_t* plan, uint32_t* tickets, uint32_t* regnext, uint32_t n,
                            uint32_t passes, uint32_t flags, hipStream_t s);
// Whole sort of one array of n <= small_sort_capacity() elements inside one
// workgroup (one HBM read and write per element).
uint32_t small_sort_capacity();
hipError_t launch_small_sort(void* A, uint32_t n, uint32_t key_bits, const KeyOrder& ko,
                             hipStream_t s);
// Copies B to A when plan says the result is in B.
void launch_copy_back(void* A, const void* B, uint32_t n, const uint32_t* flag,
                      hipStream_t s);

// ------------------------------------------------------ segmented sort
// Batch shape, uploaded per call; every segmented kernel reads it from the
// device so one captured graph serves any batch that fits its workspace.
struct SegHeader {
  uint32_t n_large;      // segments sorted by the tile pipeline
  uint32_t n_tiny;       // segments <= seg_tiny_capacity()
  uint32_t n_small;      // segments in (tiny, seg_small_capacity()]
  uint32_t total_tiles;  // tiles over all large segments
  // The scatter passes split the large segments, whole, over `groups`
  // groups of blocks (kRegions, one per XCD, when that balances; else 1);
  // block b serves group b % groups, which takes the tiles
  // perm[group_off[g] .. group_off[g] + group_tiles[g]) in ticket order.
  uint32_t groups;
  uint32_t group_off[kRegions];
  uint32_t group_tiles[kRegions];
  // MSD-first form: large segments partitioned first on an 11-bit digit
  uint32_t n_wide;
  uint32_t wide_tiles;
};
// Segment tables hold device addresses, so one launch sequence serves
// segments carved out of one array (lsb_sort_segmented) or scattered over
// separate buffers (lsb_sort_batch) alike.  Large entries come first, then
// the tiny ones, then the small ones.
struct SegLarge {
  uint64_t a;           // the segment's elements (sorted in place)
  uint64_t b;           // its scratch of n elements
  uint32_t n;
  uint32_t tile_begin;  // global index of the segment's first tile
  uint32_t perm_pos;    // LSD form: where its tiles start in its group's tile order;
                        // MSD form: its first tile among the wide segments' tiles
  uint32_t wide;        // MSD form: its index among the wide segments, or kNotWide
};
constexpr uint32_t kNotWide = 0xFFFFFFFFu;
struct SegSmall {
  uint64_t a;
  uint32_t n;
  uint32_t pad;
};
struct SegPassArgs {
  const SegHeader* hdr;
  const SegLarge* large;
  const uint32_t* seg_plan;  // [2*s] executed-pass mask, [2*s+1] result-in-B flag
  const uint32_t* tile_seg;  // large segment of every global tile
  const uint32_t* perm;      // global tiles in the groups' ticket order
  const uint32_t* counts;    // per segment [MAXP][2^d] starts (after the scan)
  uint32_t* tickets;         // [pass][group]
  uint32_t* status;          // two slices of max_tiles * 2^d words
  uint32_t max_tiles;
  uint32_t pass;
  uint32_t shift;
  uint32_t mask;
  KeyOrder ko;
};
typedef hipError_t (*SegScatterLaunch)(const SegPassArgs&, uint32_t grid, hipStream_t);
struct SegVariant {
  uint32_t digit_bits;
  uint32_t tile;
  uint32_t block;
  SegScatterLaunch scatter;
  uint32_t (*blocks_per_cu)();
};
const SegVariant* seg_variants(int* count);
uint32_t seg_tiny_capacity();
uint32_t seg_small_capacity();
hipError_t launch_seg_small(const SegHeader* hdr, const SegSmall* list,
                            uint32_t key_bits, const KeyOrder& ko, uint32_t grid, hipStream_t s);
hipError_t launch_seg_prepare(uint32_t d, const SegHeader* hdr,
                              const SegLarge* large, uint32_t tile, uint32_t key_bits,
                              const KeyOrder& ko, uint32_t passes, uint32_t* counts, uint32_t* seg_plan,
                              uint32_t* tickets, uint32_t* status0, uint32_t* tile_seg,
                              uint32_t* perm, uint32_t flags, uint32_t grid, hipStream_t s);
hipError_t launch_seg_copy(const SegHeader* hdr, const SegLarge* large,
                           const uint32_t* seg_plan, const uint32_t* tile_seg, uint32_t tile,
                           uint32_t grid,
                           hipStream_t s);

// ------------------------------------------- MSD-first batched sort
// The large segments of a batch are sorted most-significant digit first:
// level k partitions every open range ("entry") on its highest digit that
// still varies, with the same stable onesweep tile pass; children of at
// most kMsdFinishCap elements are grouped into runs that one workgroup
// sorts in LDS by their remaining digits (k_msd_finish), larger children
// become the entries of level k+1.  Every list and count lives on the
// device, so the launch sequence is fixed and can be captured.
constexpr uint32_t kMsdTile = 8192;       // elements per onesweep tile
constexpr uint32_t kMsdFinishCap = 8192;  // largest run one workgroup sorts in LDS
constexpr int kMsdMaxLevels = 9;          // 8 digits of 8 bits + a copy-only level
// Segments of more than kMsdWideMin and at most kMsdWideMax elements, with
// keys of at least kMsdWideKeyBits bits, are first partitioned on their top
// kMsdWideBits bits: their buckets then average at most kMsdFinishCap
// elements and finish in LDS, where 8-bit digits would take two partitions.
constexpr uint32_t kMsdWideBits = 10;
constexpr uint32_t kMsdWideMin = 256 * kMsdFinishCap;
constexpr uint32_t kMsdWideMax = 1024 * kMsdFinishCap;
constexpr uint32_t kMsdWideKeyBits = 16;
// Wide steps: step 0 before level 0, step s > 0 after level s - 1.
constexpr int kMsdWideSteps = 3;

// flags of an entry
constexpr uint32_t kMsdInB = 1u;   // its elements are in the scratch buffer
constexpr uint32_t kMsdCopy = 2u;  // all keys equal and in B: copy to A at this level
constexpr uint32_t kMsdMove = 4u;  // partitioned on `dig` at this level (set by the plan)
// A level 0 of one range of more than kMsdRegionMin elements (a plain sort's
// array) is partitioned in the region form: kRegions look-back chains side
// by side, each over a contiguous slice of the range's tiles, with bucket
// bases from per-region counts the level's histogram takes.
constexpr uint32_t kMsdRegionMin = kMsdWideMax;

struct MsdEntry {
  uint64_t a;           // the range in the final buffer
  uint64_t b;           // the same range in the scratch buffer
  uint32_t n;
  uint32_t tile_begin;  // first tile in its level's tile space
  uint32_t dig;         // digit to partition on
  uint32_t flags;
};
// A run: consecutive children of one partitioned entry, sorted in LDS by
// the digits in the low byte of `digits` (bit p = digit p), read from src,
// written to dst; bits 16+ give how many of the highest of them the LDS
// passes sort before the tie fix-up.
struct MsdRun {
  uint64_t src;
  uint64_t dst;
  uint32_t n;
  uint32_t digits;
};
struct MsdLevel {
  uint32_t entries, tiles, ticket;
  uint32_t moved;  // elements the level's scatter partitions or copies
};
struct MsdArgs {
  const SegHeader* hdr;
  const SegLarge* large;
  MsdEntry* ent[2];        // entries of even / odd levels
  uint32_t* counts[2];     // [entry][256] digit counts, then bucket starts
  uint32_t* orr[2];        // [entry][4] OR (lo, hi) and AND (lo, hi) of the entry's keys
  uint32_t* tilemap[2];    // tile -> entry
  uint32_t* status;        // [tile][256] look-back words of the current level
  MsdLevel* lvl;           // [kMsdMaxLevels]
  MsdRun* runs;
  uint32_t* run_ctr;       // [0] runs emitted, [1] finish ticket, [2] overflow
  uint32_t* rg;            // region form of level 0: [kRegions][256] counts, then
                           // [kRegions] tickets
  uint32_t cap_entries, cap_tiles, cap_runs;
  // Wide partitions (kMsdWideBits-bit digit) of ranges of kMsdWideMin ..
  // kMsdWideMax elements, in kMsdWideSteps steps: step 0 takes the segments
  // the host marked and sends its large buckets to level 0; step s > 0 takes
  // level s-1's children of that size and sends its large buckets to level s
  // (a range's first partition is at level 0, or at level 1 after a level
  // that found its top digit constant).  An entry's `dig` is the wide
  // digit's lowest bit; kMsdInB says where its elements are.  The steps run
  // one after another and share the wide arrays.
  MsdEntry* went;          // [wide] (tile_begin: first tile among wide tiles)
  uint32_t* wcounts;       // [wide][2^kMsdWideBits] digit counts, then bucket starts
  uint32_t* worr;          // [wide][4] OR and AND of its keys
  uint32_t* wtile;         // wide tile -> wide range
  uint32_t* wstatus;       // [wide tile][2^kMsdWideBits] look-back words
  MsdLevel* wlvl;          // [kMsdWideSteps] entries, tiles, ticket of each step
  uint32_t wstep;          // the step a wide launch runs
  uint32_t cap_wide, cap_wide_tiles;
  uint32_t key_bits;
  KeyOrder ko;
  uint32_t level;
};
// Every launch reads its work counts from the device.  A wide step and a
// level are three launches (histogram, plan, scatter); with `after` set, an
// event is recorded after each of them into after[0..2].
hipError_t launch_msd_init(const MsdArgs& a, uint32_t grid, hipStream_t s);
hipError_t launch_msd_wide(const MsdArgs& a, uint32_t cus, hipStream_t s,
                           hipEvent_t* after = nullptr);
hipError_t launch_msd_level(const MsdArgs& a, uint32_t cus, hipStream_t s,
                            hipEvent_t* after = nullptr);
hipError_t launch_msd_finish(const MsdArgs& a, uint32_t cus, hipStream_t s);

// Workload generation and verification (lsb_workload.hip).
hipError_t launch_generate(void* A, uint64_t n, uint32_t kind, uint64_t per,
                           uint64_t seed_base, uint64_t key_mask, uint64_t val_base,
                           const double* cdf, const uint64_t* table, uint32_t m,
                           hipStream_t s);
hipError_t launch_verify(const void* out, const uint64_t* in_keys, uint64_t n,
                         uint64_t val_base, uint64_t key_mask, const KeyOrder& ko,
                         uint32_t* bitmap,
                         unsigned long long* result, hipStream_t s);
hipError_t launch_extract_keys(const void* A, uint64_t n, uint64_t* keys, hipStream_t s);
// One scatter pass's memory traffic alone (lsb_debug_scatter_floor): n a
// multiple of scatter_floor_tile(), runs a power of two dividing it.
uint32_t scatter_floor_tile();
hipError_t launch_scatter_floor(const void* src, void* dst, uint64_t n, uint32_t runs,
                                uint32_t grid, hipStream_t s);

}  // namespace lsb
