// lsb_sort_noalloc: the plain sort with no device allocation at call time.
//
// Its workspace (histogram counters, pass plan, tile tickets and the two
// look-back slices of status words) lives in `scratch`, a fixed device array
// that the loader places on every GPU the library runs on when the code
// object is loaded there, and that stays until the process ends.  The first
// call on a device looks the array's address up once (hipGetSymbolAddress,
// with that device current) and keeps it; every call on that device then
// passes that address to the launch sequence lsb_sort enqueues in its LSD
// form (the MSD form stages a range table in pinned host memory), so the
// results are identical.  The kernels clear the counters, tickets and status
// words they use themselves, so what the last call left in the array does
// not matter; two calls that overlap on one device share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/lsb.h"
#include "lsb_internal.hpp"

namespace {

// Workspace of the largest array accepted: LSB_NOALLOC_MAX_N elements with
// 8-bit digits and 8192-element tiles need 2 x 16384 tiles x 256 status
// words (32 MiB) plus 140 KiB of counters, plan and tickets.
constexpr size_t kScratchBytes = size_t(34) << 20;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices];

// The current device's copy of `scratch`: its address, looked up on the
// first call on that device.  Host code uses this address only; the array's
// own name is a host-side stand-in.
int scratch_on_current_device(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    lsb::note_hip_error(e);
    return LSB_EHIP;
  }
  if (dev < 0 || dev >= kMaxDevices) return LSB_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      lsb::note_hip_error(e != hipSuccess ? e : hipErrorInvalidSymbol);
      return LSB_EHIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return LSB_OK;
}

}  // namespace

extern "C" {

size_t lsb_noalloc_scratch_bytes(void) { return kScratchBytes; }

int lsb_sort_noalloc(lsb_elem* A, lsb_elem* B, size_t n, const lsb_config* cfg) {
  lsb_config c;
  lsb_config_init(&c);
  if (cfg) c = *cfg;
  // Events (profiling) and graph instantiation allocate; a caller workspace
  // would make the scratch array pointless.
  if (c.flags & (LSB_GRAPH | LSB_PROFILE)) return LSB_EINVAL;
  if (c.workspace || c.workspace_bytes) return LSB_EINVAL;
  // The MSD form stages its range table in pinned host memory: LSD only.
  if ((c.flags & LSB_FORM_MSD) && !(c.flags & LSB_FORM_LSD)) return LSB_EINVAL;
  c.flags |= LSB_FORM_LSD;
  if (n > LSB_NOALLOC_MAX_N) return LSB_ERANGE;
  int rc = lsb::check_sort_args(A, B, n);
  if (rc || n <= 1) return rc;
  size_t need = lsb_workspace_size(n, &c);
  if (need == 0) return LSB_EINVAL;
  if (need > kScratchBytes) {
    // the automatic choice for this size has small tiles: use the largest
    if (c.digit_bits || c.tile || c.variant) return LSB_ERANGE;
    c.digit_bits = 8;
    c.tile = 8192;
    need = lsb_workspace_size(n, &c);
    if (need == 0 || need > kScratchBytes) return LSB_ERANGE;
  }
  void* p = nullptr;
  if ((rc = scratch_on_current_device(&p))) return rc;
  c.workspace = p;
  c.workspace_bytes = kScratchBytes;
  return lsb_sort(A, B, n, &c);
}

}  // extern "C"
