// C++ host-side mirror of the reference's per-rank sort interface, above the
// C ABI of lsb.h.  Header-only; a caller compiles it with the HIP runtime
// headers (-D__HIP_PLATFORM_AMD__ -I/opt/rocm/include, or hipcc) and links
// liblsb_amd.so and libamdhip64.
//
//   void mySort(DistributedArray<SortElement>& A,
//               DistributedArray<SortElement>& B)      mpi/mpi_lsbsort.cpp:591-597
//
// becomes, for one rank whose block lives in host memory,
//
//   lsb::mySort(std::vector<lsb::SortElement>& A, std::vector<lsb::SortElement>& B)
//
// with the same contract: on entry A holds the elements and B is scratch of
// at least the same size; on exit A holds the stable sort of A by key.  B's
// host contents are left as they were (the reference leaves its B
// clobbered, which this satisfies).  The block is copied to HBM, sorted there
// by lsb_sort and copied back, so a timer around the call (as the reference's,
// mpi_lsbsort.cpp:731-745) includes both PCIe copies.  Where the reference
// asserts or aborts (mpi_lsbsort.cpp:18-20, 364, 392, 487-488, 583), this
// throws lsb::Error carrying the LSB_E* code.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "lsb.h"

namespace lsb {

// The reference's element, mpi/mpi_lsbsort.cpp:30-33: key, then value.
struct SortElement {
  uint64_t key;
  uint64_t val;
};
static_assert(sizeof(SortElement) == sizeof(lsb_elem), "16-byte element, key first");

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

namespace detail {

inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(LSB_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer released on every path out of mySort.
struct DeviceBuffer {
  lsb_elem* p = nullptr;
  explicit DeviceBuffer(size_t n) {
    check_hip(hipMalloc(&p, n * sizeof(lsb_elem)), "hipMalloc");
  }
  ~DeviceBuffer() {
    if (p) (void)hipFree(p);
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

}  // namespace detail

// key_bits: sort by the low key_bits of the key (64 = the reference).
inline void mySort(std::vector<SortElement>& A, std::vector<SortElement>& B,
                   uint32_t key_bits = 64) {
  const size_t n = A.size();
  if (B.size() < n) throw Error(LSB_EINVAL, "mySort: B is smaller than A");
  if (n > LSB_MAX_N) throw Error(LSB_ERANGE, "mySort: more than LSB_MAX_N elements");
  if (n < 2) return;
  detail::DeviceBuffer dA(n), dB(n);
  detail::check_hip(hipMemcpy(dA.p, A.data(), n * sizeof(lsb_elem), hipMemcpyHostToDevice),
                    "hipMemcpy to device");
  lsb_config c;
  lsb_config_init(&c);
  c.key_bits = key_bits;
  c.flags = LSB_SYNC;
  const int rc = lsb_sort(dA.p, dB.p, n, &c);
  if (rc != LSB_OK) throw Error(rc, std::string("lsb_sort: ") + lsb_strerror(rc));
  detail::check_hip(hipMemcpy(A.data(), dA.p, n * sizeof(lsb_elem), hipMemcpyDeviceToHost),
                    "hipMemcpy to host");
}

}  // namespace lsb
