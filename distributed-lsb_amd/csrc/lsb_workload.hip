// Synthetic inputs and the on-device result check.
//
//   k_generate      the reference's keygen (mpi/mpi_lsbsort.cpp:676-704) for
//                   a block distribution over virtual ranks: element g takes
//                   output number (g mod per) of pcg64(seed_base + g/per).
//                   Each thread jumps its stream ahead once and then steps.
//   k_verify        O(n) proof of out == stable_sort(in) given val = index,
//                   under the caller's key order (key_fwd)
//                   (replaces the gather + std::stable_sort comparison of
//                   mpi/mpi_lsbsort.cpp:755-784, which is O(n log n) on one
//                   host core).
#include "lsb_device.hpp"
#include "lsb_internal.hpp"

namespace lsb {

constexpr uint32_t kGenChunk = 64;  // consecutive elements per thread

__global__ __launch_bounds__(256) void k_generate(elem_t* __restrict__ A, uint64_t n,
                                                  uint32_t kind, uint64_t per,
                                                  uint64_t seed_base, uint64_t key_mask,
                                                  uint64_t val_base,
                                                  const double* __restrict__ cdf,
                                                  const uint64_t* __restrict__ table,
                                                  uint32_t m) {
  const uint64_t g0 = ((uint64_t)blockIdx.x * 256 + threadIdx.x) * kGenChunk;
  if (g0 >= n) return;
  const uint64_t g1 = g0 + kGenChunk < n ? g0 + kGenChunk : n;
  if (kind == 1u || kind == 2u) {  // sorted / reverse-sorted
    for (uint64_t g = g0; g < g1; ++g) {
      elem_t x;
      x.key = kind == 1u ? g : n - 1 - g;
      x.val = val_base + g;
      A[g] = x;
    }
    return;
  }
  uint64_t r = g0 / per, i = g0 % per;
  u128 s = pcg_advance(pcg_seed_state(seed_base + r), i);
  for (uint64_t g = g0; g < g1; ++g) {
    if (i == per) {
      ++r;
      i = 0;
      s = pcg_seed_state(seed_base + r);
    }
    s = s * pcg_mul() + pcg_inc();  // advance, then output
    ++i;
    uint64_t key = pcg_output(s);
    if (kind == 3u) {
      // Zipf: smallest k with u < cdf[k], u uniform with 53 bits.
      const double u = (double)(key >> 11) * (1.0 / 9007199254740992.0);
      uint32_t lo = 0, hi = m - 1;
      while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (u < cdf[mid]) hi = mid;
        else lo = mid + 1;
      }
      key = table[lo] & key_mask;
    } else {
      key &= key_mask;
    }
    elem_t x;
    x.key = key;
    x.val = val_base + g;
    A[g] = x;
  }
}

hipError_t launch_generate(void* A, uint64_t n, uint32_t kind, uint64_t per,
                           uint64_t seed_base, uint64_t key_mask, uint64_t val_base,
                           const double* cdf, const uint64_t* table, uint32_t m,
                           hipStream_t s) {
  const uint64_t threads = (n + kGenChunk - 1) / kGenChunk;
  const uint64_t blocks = (threads + 255) / 256;
  if (blocks == 0) return hipSuccess;
  if (blocks > 0x7FFFFFFFull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_generate, dim3((unsigned)blocks), dim3(256), 0, s,
                     static_cast<elem_t*>(A), n, kind, per, seed_base, key_mask, val_base, cdf,
                     table, m);
  return hipGetLastError();
}

// result[0] = failing positions, result[1] = lowest failing position
// (initialised to ~0 by the host).  bitmap holds one bit per val, zeroed.
__global__ __launch_bounds__(256) void k_verify(const elem_t* __restrict__ out,
                                                const uint64_t* __restrict__ in_keys,
                                                uint64_t n, uint64_t val_base,
                                                uint64_t key_mask, KeyOrder ko,
                                                uint32_t* __restrict__ bitmap,
                                                unsigned long long* __restrict__ result) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256) {
    const elem_t e = out[i];
    bool ok = e.val >= val_base && e.val - val_base < n;
    if (ok) {
      const uint64_t v = e.val - val_base;
      const uint32_t bit = 1u << (v & 31);
      const uint32_t old = atomicOr(&bitmap[v >> 5], bit);
      ok = !(old & bit) && in_keys[v] == e.key;
    }
    if (ok && i > 0) {
      const elem_t p = out[i - 1];
      const uint64_t pk = key_fwd(p.key, ko) & key_mask, k = key_fwd(e.key, ko) & key_mask;
      ok = pk < k || (pk == k && p.val < e.val);
    }
    if (!ok) {
      atomicAdd(&result[0], 1ull);
      atomicMin(&result[1], (unsigned long long)i);
    }
  }
}

__global__ void k_zero_words(uint32_t* p, uint64_t words) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < words;
       i += (uint64_t)gridDim.x * 256)
    p[i] = 0u;
}

__global__ void k_init_result(unsigned long long* r) {
  r[0] = 0ull;
  r[1] = ~0ull;
}

hipError_t launch_verify(const void* out, const uint64_t* in_keys, uint64_t n,
                         uint64_t val_base, uint64_t key_mask, const KeyOrder& ko,
                         uint32_t* bitmap,
                         unsigned long long* result, hipStream_t s) {
  const uint64_t words = (n + 31) / 32;
  hipLaunchKernelGGL(k_zero_words, dim3(2048), dim3(256), 0, s, bitmap, words);
  hipLaunchKernelGGL(k_init_result, dim3(1), dim3(1), 0, s, result);
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_verify, dim3((unsigned)blocks), dim3(256), 0, s,
                     static_cast<const elem_t*>(out), in_keys, n, val_base, key_mask, ko, bitmap,
                     result);
  return hipGetLastError();
}

__global__ void k_extract_keys(const elem_t* __restrict__ A, uint64_t n,
                               uint64_t* __restrict__ keys) {
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * 256)
    keys[i] = A[i].key;
}

hipError_t launch_extract_keys(const void* A, uint64_t n, uint64_t* keys, hipStream_t s) {
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_extract_keys, dim3((unsigned)blocks), dim3(256), 0, s,
                     static_cast<const elem_t*>(A), n, keys);
  return hipGetLastError();
}

// ------------------------------------------------------- k_scatter_floor
// The memory traffic of one scatter pass without its sorting: tile t of
// BLOCK * ITEMS elements is read in order (each wave its 64 * ITEMS
// contiguous elements, as the pass loads them) and written as `runs` runs
// of consecutive elements, run r of tile t at dst[r * region + t * run_len]
// -- the layout one pass over uniform keys produces with runs = 256
// buckets, consecutive lanes writing consecutive slots of a run.  One block
// per CU walks the tiles in order, like the pass's persistent grid.
template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void k_scatter_floor(const elem_t* __restrict__ src,
                                                         elem_t* __restrict__ dst,
                                                         uint32_t ntiles, uint32_t run_len,
                                                         uint64_t region) {
  constexpr uint32_t T = BLOCK * ITEMS;
  const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    elem_t e[ITEMS];
    const uint32_t w0 = wave * kWave * ITEMS;
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) e[i] = src[(size_t)t * T + w0 + (uint32_t)i * kWave + lane];
#pragma unroll
    for (int i = 0; i < ITEMS; ++i) {
      const uint32_t j = w0 + (uint32_t)i * kWave + lane;
      const uint32_t r = j / run_len, q = j % run_len;
      dst[(size_t)r * region + (size_t)t * run_len + q] = e[i];
    }
  }
}

uint32_t scatter_floor_tile() { return 1024 * 8; }

hipError_t launch_scatter_floor(const void* src, void* dst, uint64_t n, uint32_t runs,
                                uint32_t grid, hipStream_t s) {
  const uint32_t T = scatter_floor_tile();
  hipLaunchKernelGGL((k_scatter_floor<1024, 8>), dim3(grid), dim3(1024), 0, s,
                     static_cast<const elem_t*>(src), static_cast<elem_t*>(dst),
                     (uint32_t)(n / T), T / runs, n / runs);
  return hipGetLastError();
}

}  // namespace lsb
