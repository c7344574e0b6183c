// Compiled and run by tests/test_stress.py: counts the floats in [-1, 1]
// for which the product's acosf_fdlibm (pmmc_amd/csrc/acosf_fdlibm.hpp, the
// kernels' arc cosine) and the C library's acosf (the reference's) differ.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../pmmc_amd/csrc/acosf_fdlibm.hpp"

int main() {
  const unsigned T = std::thread::hardware_concurrency() ? std::thread::hardware_concurrency() : 4;
  std::vector<uint64_t> bad(T, 0);
  std::vector<std::thread> th;
  for (unsigned t = 0; t < T; ++t)
    th.emplace_back([&, t] {
      for (uint32_t sign = 0; sign < 2; ++sign)
        for (uint32_t u = t; u <= 0x3f800000u; u += T) {
          const uint32_t b = u | (sign << 31);
          float x;
          std::memcpy(&x, &b, 4);
          const float a = acosf(x), m = pmmc::acosf_fdlibm(x);
          if (std::memcmp(&a, &m, 4) != 0) ++bad[t];
        }
    });
  for (auto &x : th) x.join();
  uint64_t n = 0;
  for (uint64_t b : bad) n += b;
  std::printf("mismatches %llu of %llu\n", static_cast<unsigned long long>(n),
              2ull * (0x3f800000ull + 1));
  return n == 0 ? 0 : 1;
}
