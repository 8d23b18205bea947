// A C++ caller of the library through lsb::mySort (include/lsb_mysort.hpp),
// checked the way the reference checks itself: the input is the reference's
// keygen (pcg64(rank) keys, val = global index; mpi/mpi_lsbsort.cpp:676-704,
// restated by the oracle), the expected output is std::stable_sort by key of
// that input (mpi_lsbsort.cpp:767-771), compared element by element on key
// and value (operator== of SortElement, mpi_lsbsort.cpp:49-51).  The timer
// brackets mySort as the reference's does (mpi_lsbsort.cpp:731-745), so its
// figure includes the copies to and from HBM.
//
//   test_mysort N:R [N:R ...]     one case per argument: n elements drawn
//                                 over R virtual ranks
//
// For each case it prints one line
//   case n=N ranks=R verify=ok|FAILED sha256=<hex of the sorted image>
// after the reference's own "Sorted N values in T" / "That's X M elements
// sorted / s" lines, and exits 1 if any case failed.  Test infrastructure:
// it links the oracle for keygen and SHA-256.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <vector>

#include "lsb_mysort.hpp"

extern "C" {
typedef struct {
  uint64_t key, val;
} orc_elem;
void orc_gen_uniform(orc_elem* out, uint64_t n, uint32_t ranks, uint32_t key_bits,
                     uint64_t seed_base, uint64_t first, uint64_t count);
void orc_sha256(const void* data, uint64_t len, uint8_t out[32]);
}

static bool run_case(uint64_t n, uint32_t ranks) {
  std::vector<lsb::SortElement> A(n), B(n);
  if (n) orc_gen_uniform(reinterpret_cast<orc_elem*>(A.data()), n, ranks, 64, 0, 0, n);
  std::vector<lsb::SortElement> expected = A;
  std::stable_sort(expected.begin(), expected.end(),
                   [](const lsb::SortElement& a, const lsb::SortElement& b) {
                     return a.key < b.key;
                   });

  const auto start = std::chrono::steady_clock::now();
  lsb::mySort(A, B);
  const auto end = std::chrono::steady_clock::now();
  const std::chrono::duration<double> elapsed = end - start;
  std::cout << "Sorted " << n << " values in " << elapsed.count() << "\n";
  std::cout << "That's " << n / elapsed.count() / 1000.0 / 1000.0 << " M elements sorted / s\n";

  int64_t first_bad = -1;
  for (uint64_t i = 0; i < n; ++i) {
    if (A[i].key != expected[i].key || A[i].val != expected[i].val) {
      first_bad = (int64_t)i;
      break;
    }
  }
  uint8_t h[32];
  orc_sha256(A.data(), n * sizeof(lsb::SortElement), h);
  char hex[65];
  for (int i = 0; i < 32; ++i) snprintf(hex + 2 * i, 3, "%02x", h[i]);
  std::cout << "case n=" << n << " ranks=" << ranks
            << " verify=" << (first_bad < 0 ? "ok" : "FAILED") << " sha256=" << hex;
  if (first_bad >= 0) std::cout << " first_bad=" << first_bad;
  std::cout << std::endl;
  return first_bad < 0;
}

int main(int argc, char** argv) {
  bool all = true;
  for (int a = 1; a < argc; ++a) {
    unsigned long long n = 0;
    unsigned ranks = 0;
    if (sscanf(argv[a], "%llu:%u", &n, &ranks) != 2 || ranks == 0) {
      std::cerr << "bad case '" << argv[a] << "' (want N:R)\n";
      return 2;
    }
    std::cout << "running n=" << n << " ranks=" << ranks << std::endl;
    try {
      all &= run_case(n, ranks);
    } catch (const lsb::Error& e) {
      std::cout << "case n=" << n << " ranks=" << ranks << " error=" << e.code() << " "
                << e.what() << std::endl;
      all = false;
    }
  }
  return all ? 0 : 1;
}
