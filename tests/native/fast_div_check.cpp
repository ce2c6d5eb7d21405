// Checks FastDiv (common/dev_types.h) against integer division: for every
// divisor of a set (all of 1..4096, the pixel counts and tile-row widths of
// common image sizes, values next to powers of two and to 2^31) and, for
// each, dividends next to 0, next to 2^31 - 1, next to multiples of the
// divisor and random ones below 2^31, (n * m) >> p must equal n / d. Prints
// the number of cases and exits non-zero on the first mismatch.
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "common/dev_types.h"

static bool check(uint32_t n, uint32_t d, const xnh::FastDiv& f) {
  const uint32_t q = static_cast<uint32_t>((static_cast<uint64_t>(n) * f.m) >> f.p);
  if (q != n / d) {
    std::printf("mismatch: %u / %u = %u, multiply-shift gives %u (m %u, p %u)\n", n, d, n / d, q, f.m, f.p);
    return false;
  }
  return true;
}

int main() {
  std::vector<uint32_t> ds;
  for (uint32_t d = 1; d <= 4096; ++d) ds.push_back(d);
  const uint32_t sizes[][2] = {{512, 512}, {1920, 1080}, {1280, 720}, {3840, 2160}, {2048, 2048}, {256, 256},
                               {24, 14}, {160, 90}, {7, 5}};
  for (const auto& wh : sizes) {
    ds.push_back(wh[0] * wh[1]);
    ds.push_back(wh[0] * wh[1] / 8);
    ds.push_back(8 * wh[0]);
  }
  for (uint32_t k = 12; k < 31; ++k) {
    ds.push_back((1u << k) - 1);
    ds.push_back(1u << k);
    ds.push_back((1u << k) + 1);
  }
  ds.push_back(0x7FFFFFFFu);
  std::mt19937 rng(12345);
  uint64_t cases = 0;
  for (uint32_t d : ds) {
    const xnh::FastDiv f = xnh::fast_div(d);
    std::vector<uint32_t> ns;
    for (uint32_t i = 0; i < 64; ++i) {
      ns.push_back(i);
      ns.push_back(0x7FFFFFFFu - i);
    }
    for (uint32_t i = 0; i < 64; ++i) {
      const uint32_t base = (rng() & 0x7FFFFFFFu) / d * d;
      ns.push_back(base);
      if (base > 0) ns.push_back(base - 1);
      if (base + d - 1 <= 0x7FFFFFFFu) ns.push_back(base + d - 1);
      ns.push_back(rng() & 0x7FFFFFFFu);
    }
    const uint32_t top = 0x7FFFFFFFu / d * d;
    ns.push_back(top);
    if (top > 0) ns.push_back(top - 1);
    for (uint32_t n : ns) {
      ++cases;
      if (!check(n, d, f)) return 1;
    }
  }
  std::printf("fast_div: %llu cases over %zu divisors agree\n", static_cast<unsigned long long>(cases), ds.size());
  return 0;
}
