// rrtmgsw_mtjump.cpp -- jump-ahead of the MT19937 stream (host side).
//
// MT19937 (Matsumoto and Nishimura 1998) is linear over GF(2): its raw words satisfy
// x[r] = x[r - 227] ^ twist(x[r - 624], x[r - 623]), and the 19937-bit state T advances by a
// fixed linear map whose characteristic polynomial P has degree 19937. With
// x^J = sum_i c_i x^i (mod P), i < 19937, the window of 624 raw words J positions ahead is the
// XOR of the windows i positions ahead over the set bits c_i (Haramoto, Matsumoto, Nishimura,
// Panneton and L'Ecuyer 2008, "Efficient jump ahead for F2-linear random number generators").
// P is found once by the Berlekamp-Massey algorithm on one output bit of the stream (P is
// primitive, so that bit's minimal polynomial is P itself), and x^J mod P by repeated squaring.
#include <immintrin.h>

#include <cstdint>
#include <cstring>
#include <vector>

namespace rrtmgsw {

std::vector<uint32_t> mt_jump_coefficients(int log2_jump);

namespace {

constexpr int MT_N = 624, MT_DEG = 19937;

using Bits = std::vector<uint64_t>;

inline int bit(const Bits &b, size_t i) { return (int)((b[i >> 6] >> (i & 63)) & 1u); }
inline void flip(Bits &b, size_t i) { b[i >> 6] ^= (uint64_t)1 << (i & 63); }

// dst ^= src << sh (bit shift towards higher degree), within dst's length
void xor_shifted(Bits &dst, const Bits &src, size_t sh) {
  const size_t ws = sh >> 6, bs = sh & 63;
  const size_t top = ws + src.size() + 1 < dst.size() ? ws + src.size() + 1 : dst.size();
  for (size_t i = top; i-- > ws;) {
    const size_t j = i - ws;
    uint64_t v = j < src.size() ? src[j] << bs : 0;
    if (bs && j >= 1 && j - 1 < src.size()) v |= src[j - 1] >> (64 - bs);
    dst[i] ^= v;
  }
}

// Raw MT19937 words x[0 .. n) from a seed (init_genrand, then the recurrence).
std::vector<uint32_t> raw_words(uint32_t seed, size_t n) {
  std::vector<uint32_t> x(n);
  x[0] = seed;
  for (size_t i = 1; i < (size_t)MT_N && i < n; i++) x[i] = 1812433253u * (x[i - 1] ^ (x[i - 1] >> 30)) + (uint32_t)i;
  for (size_t r = MT_N; r < n; r++) {
    const uint32_t y = (x[r - 624] & 0x80000000u) | (x[r - 623] & 0x7fffffffu);
    x[r] = x[r - 227] ^ (y >> 1) ^ ((x[r - 623] & 1u) ? 0x9908b0dfu : 0u);
  }
  return x;
}

// Characteristic polynomial P of the MT19937 transition, bit j = coefficient of x^j.
Bits charpoly() {
  const size_t N = 2 * MT_DEG + 64, W = (N + 63) / 64 + 1;
  std::vector<uint32_t> x = raw_words(5489u, MT_N + N);
  // Berlekamp-Massey on s_n = lowest bit of x[624 + n]; R holds s_(n-i) at bit i
  Bits C(W, 0), B(W, 0), R(W, 0), T;
  C[0] = B[0] = 1;
  size_t L = 0, m = 1;
  for (size_t n = 0; n < N; n++) {
    for (size_t i = W; i-- > 1;) R[i] = (R[i] << 1) | (R[i - 1] >> 63);   // R <<= 1
    R[0] = (R[0] << 1) | (x[MT_N + n] & 1u);
    uint64_t acc = 0;
    for (size_t i = 0; i < W; i++) acc ^= C[i] & R[i];
    const int d = __builtin_parityll(acc);
    if (!d) {
      m++;
    } else if (2 * L <= n) {
      T = C;
      xor_shifted(C, B, m);
      L = n + 1 - L;
      B = T;
      m = 1;
    } else {
      xor_shifted(C, B, m);
      m++;
    }
  }
  // the recurrence s_n = sum_(i=1..L) C_i s_(n-i) has characteristic polynomial x^L + sum C_i x^(L-i)
  Bits P((MT_DEG + 64) / 64 + 1, 0);
  if (L != (size_t)MT_DEG) return Bits();
  for (size_t i = 0; i <= L; i++)
    if (bit(C, i)) flip(P, L - i);
  return P;
}

// a mod P in place (a has room for degree < 2 * MT_DEG)
void reduce(Bits &a, const Bits &P) {
  for (size_t p = (size_t)2 * MT_DEG; p-- > (size_t)MT_DEG;)
    if (bit(a, p)) xor_shifted(a, P, p - MT_DEG);
}

// The characteristic polynomial, found once per process (the callers hold the library's lock).
const Bits &charpoly_cached() {
  static Bits P = charpoly();
  return P;
}

constexpr size_t PW = (MT_DEG + 63) / 64;   // words of a reduced polynomial (degree < 19937)

// Carry-less 64 x 64 -> 128-bit products: the CPU's carry-less multiply where it has one, else
// 4-bit windows of b.
__attribute__((target("pclmul,sse2"))) inline void clmul64_hw(uint64_t a, uint64_t b, uint64_t &lo,
                                                             uint64_t &hi) {
  const __m128i p = _mm_clmulepi64_si128(_mm_set_epi64x(0, (long long)a), _mm_set_epi64x(0, (long long)b), 0);
  lo = (uint64_t)_mm_cvtsi128_si64(p);
  hi = (uint64_t)_mm_cvtsi128_si64(_mm_unpackhi_epi64(p, p));
}
inline void clmul64_sw(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
  uint64_t tl[16], th[16];
  tl[0] = th[0] = 0;
  for (int v = 1; v < 16; v++) {   // a * v for a 4-bit v: 67 bits
    uint64_t l = 0, h = 0;
    for (int b2 = 0; b2 < 4; b2++)
      if ((v >> b2) & 1) l ^= a << b2, h ^= b2 ? a >> (64 - b2) : 0;
    tl[v] = l, th[v] = h;
  }
  lo = hi = 0;
  for (int s4 = 0; s4 < 64; s4 += 4) {
    const int v = (int)((b >> s4) & 15);
    lo ^= s4 ? tl[v] << s4 : tl[v];
    hi ^= s4 ? (tl[v] >> (64 - s4)) ^ (th[v] << s4) : th[v];
  }
}

// Reduction modulo P of a product of two reduced polynomials, from the top down a byte at a
// time: the byte v of bits [19937 + 8i, 19937 + 8i + 8) is x^(8i) v(x) x^19937, and
// v(x) x^19937 mod P = sum_b v_b (x^(19937 + b) mod P) has degree below 19937, so its shifted
// copy lies below the byte.
struct Reducer {
  std::vector<Bits> tab;   // tab[v] = v(x) x^19937 mod P, for the 256 bytes v
  explicit Reducer(const Bits &P) : tab(256, Bits(PW + 1, 0)) {
    std::vector<Bits> r(8, Bits(2 * PW + 2, 0));
    for (int b = 0; b < 8; b++) {   // x^(19937 + b) mod P
      flip(r[b], (size_t)MT_DEG + b);
      for (size_t p = (size_t)MT_DEG + 8; p-- > (size_t)MT_DEG;)
        if (bit(r[b], p)) xor_shifted(r[b], P, p - MT_DEG);
      r[b].resize(PW + 1);
    }
    for (int v = 1; v < 256; v++)
      for (int b = 0; b < 8; b++)
        if ((v >> b) & 1)
          for (size_t i = 0; i <= PW; i++) tab[v][i] ^= r[b][i];
  }
  void reduce(Bits &a) const {   // a: degree < 2 * 19937 - 1
    for (size_t i = (size_t)(MT_DEG - 1) / 8 + 1; i-- > 0;) {
      const size_t pos = (size_t)MT_DEG + 8 * i;
      uint64_t v = 0;
      for (int b = 0; b < 8; b++)
        if (pos + b < a.size() * 64 && bit(a, pos + b)) v |= (uint64_t)1 << b, flip(a, pos + b);
      if (v) xor_shifted(a, tab[v], 8 * i);
    }
  }
};

// a * b mod P for reduced a, b
template <bool HW>
void mul_words(const Bits &a, const Bits &b, Bits &pr) {
  for (size_t i = 0; i < PW; i++) {
    if (!a[i]) continue;
    for (size_t j = 0; j < PW; j++) {
      if (!b[j]) continue;
      uint64_t lo, hi;
      if (HW) clmul64_hw(a[i], b[j], lo, hi);
      else clmul64_sw(a[i], b[j], lo, hi);
      pr[i + j] ^= lo, pr[i + j + 1] ^= hi;
    }
  }
}
Bits mulmod(const Bits &a, const Bits &b, const Reducer &R) {
  static const bool hw = __builtin_cpu_supports("pclmul");
  Bits pr(2 * PW + 2, 0);
  if (hw) mul_words<true>(a, b, pr);
  else mul_words<false>(a, b, pr);
  R.reduce(pr);
  pr.resize(PW);
  return pr;
}

}  // namespace

// Coefficients of x^(k J) mod P for k = 1 .. count, J = 2^log2_jump, each packed little-endian into
// 624 words (bit i = the coefficient of x^i), one after the other; empty if the characteristic
// polynomial could not be found. The segment k of a jumped stream starts from the window that
// x^(k J) selects, so every segment's start can be formed at once from the stream's own start.
std::vector<uint32_t> mt_jump_multiples(int log2_jump, int count) {
  std::vector<uint32_t> one = mt_jump_coefficients(log2_jump);
  if (one.empty() || count < 1) return {};
  const Bits &P = charpoly_cached();
  const Reducer R(P);
  Bits a(PW, 0);
  for (size_t i = 0; i < (size_t)MT_DEG; i++)
    if ((one[i >> 5] >> (i & 31)) & 1u) flip(a, i);
  std::vector<uint32_t> out((size_t)count * MT_N, 0);
  Bits m = a;
  for (int k = 1; k <= count; k++) {
    if (k > 1) m = mulmod(m, a, R);
    for (size_t i = 0; i < (size_t)MT_DEG; i++)
      if (bit(m, i)) out[(size_t)(k - 1) * MT_N + (i >> 5)] |= 1u << (i & 31);
  }
  return out;
}

// Coefficients c_0 .. c_19936 of x^(2^log2_jump) mod P, packed little-endian into 624 words
// (bit i of the result = c_i); empty if the characteristic polynomial could not be found.
std::vector<uint32_t> mt_jump_coefficients(int log2_jump) {
  const Bits &P = charpoly_cached();
  if (P.empty()) return {};
  const size_t W = (2 * MT_DEG + 64) / 64 + 1;
  Bits a(W, 0);
  flip(a, 1);   // x
  for (int k = 0; k < log2_jump; k++) {
    Bits sq(W, 0);   // squaring over GF(2) spreads the bits
    for (size_t i = 0; i < (size_t)MT_DEG; i++)
      if (bit(a, i)) flip(sq, 2 * i);
    reduce(sq, P);
    a.swap(sq);
  }
  std::vector<uint32_t> out(MT_N, 0);
  for (size_t i = 0; i < (size_t)MT_DEG; i++)
    if (bit(a, i)) out[i >> 5] |= 1u << (i & 31);
  return out;
}

}  // namespace rrtmgsw
