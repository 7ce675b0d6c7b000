// Exact grouped float sums: the integer arithmetic and the one rounding step.
//
// What is computed (the reference's contract, Sources/ArrowMetal/Kernels/GroupSumExact.swift:11-60,
// Sources/ArrowMetalC/ArrowMetalC_GroupByExtra.swift:185-196): a group's sum is the exact sum of its
// non-null values rounded once to nearest-even, and its mean is that exact sum divided by the number
// of non-null values, rounded once.
//
// How (this file's own method). A group's finite non-zero values are added as signed 128-bit
// fixed-point integers whose least significant bit weighs 2^(lsb - 1074); `lsb` comes from the
// group's largest exponent and its row count so that the sum cannot overflow. Each value is rounded
// toward minus infinity onto that grid, so for the accumulated W and `drops`, the number of values
// that lost bits, the exact sum S (in grid units) satisfies W <= S < W + drops, and S == W exactly
// when drops == 0. One rounding routine turns a positive 128-bit magnitude, plus the knowledge of
// whether the true value lies a little above or below it, into a binary64 or binary32 pattern:
//   - `rne`     normalises the magnitude so its leading one is bit 127, keeps the top 64 bits, folds
//               whatever lies below into a sticky flag, and rounds at the format's precision (fewer
//               bits for a subnormal result);
//   - `rne_div` first divides the normalised magnitude by the count with two 64-by-32-bit divisions
//               (the quotient's top word already holds at least 63 significant bits), the remainder
//               becoming the sticky flag.
// When drops > 0, rounding is monotonic, so if the lowest and the highest value of the open interval
// (W, W + drops) round to the same pattern, that pattern is the answer; otherwise the group is summed
// again with a full-width accumulator (groupby.hip) and rounded by the same routine.
#pragma once

// Included after the HIP runtime header (for __host__ / __device__); everything here is integer
// arithmetic that runs on either side.
#include <cstdint>

namespace am {
namespace fs {

struct U128 {
  unsigned long long lo, hi;
};

__host__ __device__ inline bool is_zero(U128 x) { return (x.lo | x.hi) == 0ull; }
__host__ __device__ inline bool is_negative(U128 x) { return (long long)x.hi < 0; }

__host__ __device__ inline U128 negate(U128 x) {
  U128 r;
  r.lo = ~x.lo + 1ull;
  r.hi = ~x.hi + (r.lo == 0ull ? 1ull : 0ull);
  return r;
}

__host__ __device__ inline U128 add_small(U128 x, unsigned long long d) {
  U128 r;
  r.lo = x.lo + d;
  r.hi = x.hi + (r.lo < d ? 1ull : 0ull);
  return r;
}

// x << s for 0 <= s < 128.
__host__ __device__ inline U128 shl(U128 x, int s) {
  if (s == 0) return x;
  if (s >= 64) return U128{0ull, x.lo << (s - 64)};
  return U128{x.lo << s, (x.hi << s) | (x.lo >> (64 - s))};
}

__host__ __device__ inline int clz128(U128 x) { return x.hi ? __builtin_clzll(x.hi) : 64 + __builtin_clzll(x.lo); }

// The binary64 pattern of a float32 pattern (exact: every float32 is a float64). Integer-only, so that
// float32 subnormals survive whatever denormal mode the floating-point unit runs in. NaN payloads move
// up with the fraction bits, as a hardware conversion moves them.
__host__ __device__ inline unsigned long long widen_f32(uint32_t x) {
  const unsigned long long sign = (unsigned long long)(x >> 31) << 63;
  const uint32_t e = (x >> 23) & 0xFFu, f = x & 0x7FFFFFu;
  if (e == 0xFFu) return sign | 0x7FF0000000000000ull | ((unsigned long long)f << 29);
  if (e == 0u) {
    if (f == 0u) return sign;
    // A float32 subnormal f * 2^-149 is a normal float64: shift its leading one to the hidden bit.
    const int lead = 31 - __builtin_clz(f);   // 0 .. 22
    const unsigned long long frac = ((unsigned long long)f << (52 - lead)) & 0xFFFFFFFFFFFFFull;
    return sign | ((unsigned long long)(lead - 149 + 1023) << 52) | frac;
  }
  return sign | ((unsigned long long)(e - 127 + 1023) << 52) | ((unsigned long long)f << 29);
}

// The grid of a group: the bit offset `lsb` above 2^-1074 of the fixed-point unit, from the group's
// largest biased exponent (1 for subnormals) and its count of values. cnt values below
// 2^(emax + 52) * 2^-1074 each sum to less than 2^(bitlen(cnt) + emax + 52 - lsb) units, which is at
// most 2^127 for lsb = emax + bitlen(cnt) - 75; a grid finer than 2^-1074 gains nothing.
__host__ __device__ inline int grid_lsb(uint32_t emax, unsigned long long cnt) {
  const int bitlen = 64 - __builtin_clzll(cnt);
  const int lsb = int(emax) + bitlen - 75;
  return lsb > 0 ? lsb : 0;
}

// floor(v / 2^(lsb - 1074)) as a two's-complement 128-bit integer, for a finite non-zero binary64
// pattern v whose exponent is at most the one `lsb` was made for; `lost` when the floor dropped bits.
__host__ __device__ inline U128 to_grid(unsigned long long v, int lsb, bool& lost) {
  const uint32_t be = uint32_t(v >> 52) & 0x7FFu;
  const unsigned long long m = (v & 0xFFFFFFFFFFFFFull) | (be ? (1ull << 52) : 0ull);
  const int sh = int(be ? be : 1u) - 1 - lsb;   // m's bit 0 sits at grid bit sh (sh <= 73)
  U128 x{0ull, 0ull};
  lost = false;
  if (sh >= 0) {
    x = shl(U128{m, 0ull}, sh);
  } else if (sh > -64) {
    x.lo = m >> -sh;
    lost = (m << (64 + sh)) != 0ull;
  } else {
    lost = true;
  }
  if (v >> 63) x = negate(add_small(x, lost ? 1ull : 0ull));   // floor(-a) = -ceil(a)
  return x;
}

// Round-to-nearest-even of (m + t) * 2^scale, m > 0, to a format with P significand bits and minimum
// normal exponent EMIN (binary64: 53, -1022; binary32: 24, -126). t is a fraction of one unit of m:
// `tail` says 0 < t < 1 (bits known to lie below m); otherwise t is 0 (bias 0), an infinitesimal
// above (bias > 0) or an infinitesimal below (bias < 0). Returns the magnitude's pattern in the
// format (binary32 in the low 32 bits); overflow gives infinity.
template <int P, int EMIN>
__host__ __device__ unsigned long long rne(U128 m, bool tail, int bias, int scale) {
  constexpr unsigned long long INF = (unsigned long long)(2 * (1 - EMIN) + 1) << (P - 1);
  const int lz = clz128(m);
  const U128 n = shl(m, lz);
  const unsigned long long top = n.hi;
  const bool more = tail || n.lo != 0ull;        // a positive fraction below `top`
  const bool above = more || bias > 0;
  const bool below = !more && bias < 0;
  const int e = scale + 127 - lz;                // the value lies in [2^e, 2^(e+1))
  const int cut = 64 - P + (e < EMIN ? EMIN - e : 0);   // bits of `top` below the last kept bit
  if (cut > 64) return 0ull;                     // under half the smallest subnormal
  unsigned long long keep, rem, half;
  if (cut == 64) {
    keep = 0ull;
    rem = top;
    half = 1ull << 63;
  } else {
    keep = top >> cut;
    rem = top & ((1ull << cut) - 1ull);
    half = 1ull << (cut - 1);
  }
  const bool up = rem > half || (rem == half && (above || (!below && (keep & 1ull))));
  keep += up ? 1ull : 0ull;
  // A normal `keep` carries the hidden bit, so adding it to the exponent field minus one gives the
  // pattern, and a carry out of the significand moves the exponent up by itself.
  const unsigned long long bits = e >= EMIN ? ((unsigned long long)(e - EMIN) << (P - 1)) + keep : keep;
  return bits >= INF ? INF : bits;
}

// Round-to-nearest-even of (m + t) / c * 2^scale, m > 0, 0 < c < 2^32; t as for `rne`.
template <int P, int EMIN>
__host__ __device__ unsigned long long rne_div(U128 m, bool tail, int bias, unsigned long long c, int scale) {
  const int lz = clz128(m);
  const U128 n = shl(m, lz);                     // n.hi >= 2^63
  const unsigned long long q1 = n.hi / c, r1 = n.hi % c;
  const unsigned long long x = (r1 << 32) | (n.lo >> 32);   // r1 < c < 2^32
  const unsigned long long q2 = x / c, r2 = x % c;          // q2 < 2^32
  // n / c = (q1 * 2^32 + q2) * 2^32 + f with 0 <= f < 2^32; f and t only decide the sticky bit,
  // because the quotient q1 * 2^32 + q2 >= 2^63 already reaches past the rounding position.
  const bool frac = tail || r2 != 0ull || (n.lo & 0xFFFFFFFFull) != 0ull;
  const U128 q{(q1 << 32) | q2, q1 >> 32};
  return rne<P, EMIN>(q, frac, frac ? 0 : bias, scale + 32 - lz);
}

// The sum and mean of a group from its fixed-point accumulator: W (two's complement) in units of
// 2^(lsb - 1074), `drops` values rounded down onto the grid, `cnt` values. Writes the patterns
// (binary32 in the low 32 bits when P == 24) and returns true, or returns false when the interval
// (W, W + drops) does not pin both roundings.
template <int P, int EMIN>
__host__ __device__ bool settle(U128 w, unsigned long long drops, int lsb, unsigned long long cnt,
                       unsigned long long& sum, unsigned long long& mean) {
  constexpr unsigned long long SIGN = 1ull << (P == 53 ? 63 : 31);
  const int scale = lsb - 1074;
  if (drops == 0ull) {
    if (is_zero(w)) {   // an exact cancellation of non-zero values: +0.0
      sum = 0ull;
      mean = 0ull;
      return true;
    }
    const bool neg = is_negative(w);
    const U128 mag = neg ? negate(w) : w;
    const unsigned long long s = neg ? SIGN : 0ull;
    sum = s | rne<P, EMIN>(mag, false, 0, scale);
    mean = s | rne_div<P, EMIN>(mag, false, 0, cnt, scale);
    return true;
  }
  const U128 end = add_small(w, drops);          // S lies strictly between w and end
  U128 lo_mag, hi_mag;                            // |S| lies strictly between these
  unsigned long long s;
  if (!is_negative(w) && !is_zero(w)) {
    lo_mag = w;
    hi_mag = end;
    s = 0ull;
  } else if (is_negative(end)) {
    lo_mag = negate(end);
    hi_mag = negate(w);
    s = SIGN;
  } else {
    return false;                                 // the interval reaches zero
  }
  const unsigned long long s0 = rne<P, EMIN>(lo_mag, false, +1, scale);
  const unsigned long long s1 = rne<P, EMIN>(hi_mag, false, -1, scale);
  if (s0 != s1) return false;
  const unsigned long long m0 = rne_div<P, EMIN>(lo_mag, false, +1, cnt, scale);
  const unsigned long long m1 = rne_div<P, EMIN>(hi_mag, false, -1, cnt, scale);
  if (m0 != m1) return false;
  sum = s | s0;
  mean = s | m0;
  return true;
}

// The 128 bits of the non-negative multi-word integer w[0..nw) that start at its leading one (bit
// position `lead`), the position of their lowest bit, and whether any bit below them is set.
__host__ __device__ inline U128 top128(const unsigned long long* w, int nw, int lead, int& low, bool& tail) {
  low = lead - 127;
  U128 r{0ull, 0ull};
  tail = false;
  for (int j = 0; j < nw; j++) {
    const unsigned long long x = w[j];
    if (x == 0ull) continue;
    // Word j covers bits [64 j, 64 j + 64); place them relative to `low`.
    const int off = j * 64 - low;
    if (off >= 128) continue;
    if (off <= -64) {
      tail = true;
      continue;
    }
    if (off < 0) {
      const int d = -off;
      r.lo |= x >> d;
      if (x << (64 - d)) tail = true;
    } else if (off < 64) {
      r.lo |= x << off;
      if (off) r.hi |= x >> (64 - off);
    } else {
      r.hi |= x << (off - 64);
    }
  }
  return r;
}

}  // namespace fs
}  // namespace am
