// Checker for the MT19937 jump-ahead of rrtmg_sw_amd/csrc/rrtmgsw_mtjump.cpp (compiled with it by
// tests/test_mtjump.py): for jumps of 2^5, 2^12 and 2^21 words and three seeds, the window of 624
// raw words formed from the coefficients equals the sequentially generated window (all bits that
// are part of the state), and 5000 words generated from it equal the sequential stream. Prints the
// mismatch count per case and exits non-zero on any mismatch.
#include <cstdio>
#include <cstdint>
#include <vector>
#include <chrono>
namespace rrtmgsw { std::vector<uint32_t> mt_jump_coefficients(int log2_jump);
                    std::vector<uint32_t> mt_jump_multiples(int log2_jump, int count); }
static std::vector<uint32_t> raw(uint32_t seed, size_t n) {
  std::vector<uint32_t> x(n); x[0] = seed;
  for (size_t i = 1; i < 624; i++) x[i] = 1812433253u * (x[i-1] ^ (x[i-1] >> 30)) + (uint32_t)i;
  for (size_t r = 624; r < n; r++) { uint32_t y = (x[r-624] & 0x80000000u) | (x[r-623] & 0x7fffffffu);
    x[r] = x[r-227] ^ (y >> 1) ^ ((x[r-623] & 1u) ? 0x9908b0dfu : 0u); }
  return x;
}
int main() {
  for (int lj : {5, 12, 21}) {
    auto t0 = std::chrono::steady_clock::now();
    auto c = rrtmgsw::mt_jump_coefficients(lj);
    double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (c.empty()) { printf("no polynomial\n"); return 1; }
    size_t J = (size_t)1 << lj;
    for (uint32_t seed : {112u, 5489u, 4357u}) {
      auto x = raw(seed, J + 624 + 5000 + 20000);
      // window at J from windows i < 19937
      std::vector<uint32_t> w(624, 0);
      for (int i = 0; i < 19937; i++) if ((c[i >> 5] >> (i & 31)) & 1u) for (int k = 0; k < 624; k++) w[k] ^= x[i + k];
      int bad = 0;
      for (int k = 1; k < 624; k++) bad += w[k] != x[J + k];
      bad += (w[0] & 0x80000000u) != (x[J] & 0x80000000u);
      // generate forward from the jumped window
      std::vector<uint32_t> y(w); y.resize(624 + 5000);
      for (size_t r = 624; r < y.size(); r++) { uint32_t v = (y[r-624] & 0x80000000u) | (y[r-623] & 0x7fffffffu);
        y[r] = y[r-227] ^ (v >> 1) ^ ((y[r-623] & 1u) ? 0x9908b0dfu : 0u); }
      for (size_t r = 624; r < y.size(); r++) bad += y[r] != x[J + r];
      printf("log2J=%d seed=%u mismatches=%d (%.2f s for the coefficients)\n", lj, seed, bad, dt);
      if (bad) return 2;
    }
  }
  // x^(k J) mod P for k = 1..4 (J = 2^12) and k = 1..3 (J = 2^19): the window k J ahead
  for (auto [lj, cnt] : {std::pair<int, int>{12, 4}, std::pair<int, int>{19, 3}}) {
    auto t0 = std::chrono::steady_clock::now();
    auto m = rrtmgsw::mt_jump_multiples(lj, cnt);
    double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (m.size() != (size_t)cnt * 624) { printf("no multiples\n"); return 1; }
    size_t J = (size_t)1 << lj;
    auto x = raw(5489u, cnt * J + 624 + 20000);
    for (int k = 1; k <= cnt; k++) {
      const uint32_t *c = m.data() + (size_t)(k - 1) * 624;
      std::vector<uint32_t> w(624, 0);
      for (int i = 0; i < 19937; i++) if ((c[i >> 5] >> (i & 31)) & 1u) for (int q = 0; q < 624; q++) w[q] ^= x[i + q];
      int bad = 0;
      for (int q = 1; q < 624; q++) bad += w[q] != x[k * J + q];
      bad += (w[0] & 0x80000000u) != (x[k * J] & 0x80000000u);
      printf("multiple k=%d log2J=%d mismatches=%d (%.2f s for all %d)\n", k, lj, bad, dt, cnt);
      if (bad) return 3;
    }
  }
  return 0;
}
