// Host-side interface between the C ABI (lsb_api.cpp) and the kernel
// launchers (lsb_kernels.hip, lsb_workload.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lsb {

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
  uint32_t* tickets;       // one tile ticket counter per pass
  uint32_t* status;        // two look-back slices of tiles * 2^d words
  uint32_t ntiles;
};

typedef void (*ScatterLaunch)(const PassArgs&, hipStream_t);

struct SortVariant {
  uint32_t digit_bits;
  uint32_t tile;  // elements per tile (block * items)
  uint32_t block;
  ScatterLaunch scatter;
};

// Compiled variants (lsb_kernels.hip).
const SortVariant* sort_variants(int* count);

// Zeroes counts[0..words) (and tickets), launched before the histogram.
void launch_zero(uint32_t* p, size_t words, hipStream_t s);
// All-digit histogram: counts[p * 2^d + b] += ..., zeroes zero[0..zero_words).
hipError_t launch_histogram(uint32_t d, const void* A, uint32_t n, uint32_t key_bits,
                            uint32_t passes, uint32_t* counts, uint32_t* zero,
                            size_t zero_words, hipStream_t s);
// Exclusive scan of every pass's counts, skip detection, plan, tickets reset.
hipError_t launch_scan_plan(uint32_t d, const uint32_t* counts, uint32_t* starts,
                            uint32_t* plan, uint32_t* tickets, uint32_t n, uint32_t passes,
                            uint32_t flags, hipStream_t s);
// Copies B to A when plan says the result is in B.
void launch_copy_back(void* A, const void* B, uint32_t n, const uint32_t* flag,
                      hipStream_t s);

// Workload generation and verification (lsb_workload.hip).
hipError_t launch_generate(void* A, uint64_t n, uint32_t kind, uint64_t per,
                           uint64_t seed_base, uint64_t key_mask, uint64_t val_base,
                           const double* cdf, const uint64_t* table, uint32_t m,
                           hipStream_t s);
hipError_t launch_verify(const void* out, const uint64_t* in_keys, uint64_t n,
                         uint64_t val_base, uint64_t key_mask, uint32_t* bitmap,
                         unsigned long long* result, hipStream_t s);
hipError_t launch_extract_keys(const void* A, uint64_t n, uint64_t* keys, hipStream_t s);

}  // namespace lsb
