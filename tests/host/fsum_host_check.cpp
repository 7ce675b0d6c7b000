// Host check of the exact float-sum arithmetic of the HIP library (arrowmetal_amd/csrc/fsum.hpp),
// compiled for the CPU: random groups of float64 values go through the same grid, floor, settle and
// full-width steps the kernels run, and the sum and mean bits are compared with the oracle's
// (oracle/oracle.c or_group_agg ops 0 and 3, a 2,304-bit exact accumulator). Float32 results (the
// dense group-by's float32 sum) are compared with a plain 128-bit integer rounding of the exact sum of
// float32 values whose exponents are kept in a range where that integer cannot overflow.
//
// Usage: fsum_host_check SEED TRIALS; prints one summary line, exits 1 on any mismatch.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <vector>
#include <random>
#include "../../arrowmetal_amd/csrc/fsum.hpp"
extern "C" int or_group_agg(const int32_t* ids, int64_t n, int64_t K, const void* vals, int t,
                 const uint8_t* valid, int op, uint64_t* out_bits, uint8_t* out_valid);
using namespace am::fs;
static const int OR_F64 = 9;

static void full_width(const std::vector<uint64_t>& v, uint64_t cnt, unsigned long long& sum, unsigned long long& mean, int P) {
  // 36 words two's complement exact accumulator
  const int NW = 36; unsigned long long w[NW] = {0};
  for (uint64_t b : v) {
    uint32_t e = (b >> 52) & 0x7FF; if (e == 0x7FF || (b << 1) == 0) continue;
    uint64_t m = (b & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0);
    uint32_t off = (e ? e : 1) - 1; int j = off / 64, s = off % 64;
    uint64_t lo = m << s, hi = s ? m >> (64 - s) : 0;
    uint64_t add[NW] = {0}; add[j] = lo; if (j + 1 < NW) add[j + 1] = hi;
    if (b >> 63) { uint64_t c = 1; for (int k = 0; k < NW; k++) { uint64_t x = ~add[k] + c; c = (c && x == 0); add[k] = x; } }
    uint64_t c = 0; for (int k = 0; k < NW; k++) { unsigned __int128 t = (unsigned __int128)w[k] + add[k] + c; w[k] = (uint64_t)t; c = (uint64_t)(t >> 64); }
  }
  bool neg = (int64_t)w[NW - 1] < 0;
  if (neg) { uint64_t c = 1; for (int k = 0; k < NW; k++) { uint64_t x = ~w[k] + c; c = (c && x == 0); w[k] = x; } }
  int lead = -1; for (int j = NW - 1; j >= 0 && lead < 0; j--) if (w[j]) lead = j * 64 + 63 - __builtin_clzll(w[j]);
  if (lead < 0) { sum = mean = 0; return; }
  int low; bool tail; U128 top = top128(w, NW, lead, low, tail);
  uint64_t S = neg ? (P == 53 ? 1ull << 63 : 1ull << 31) : 0;
  if (P == 53) { sum = S | rne<53, -1022>(top, tail, 0, low - 1074); mean = S | rne_div<53, -1022>(top, tail, 0, cnt, low - 1074); }
  else { sum = S | rne<24, -126>(top, tail, 0, low - 1074); mean = S | rne_div<24, -126>(top, tail, 0, cnt, low - 1074); }
}

// The product pipeline for one group of finite non-zero/zero values (no specials).
static bool product(const std::vector<uint64_t>& v, unsigned long long& sum, unsigned long long& mean, bool& settled) {
  uint32_t emax = 0; for (uint64_t b : v) { uint32_t e = (b >> 52) & 0x7FF; if ((b << 1) == 0) continue; uint32_t ee = e ? e : 1; if (ee > emax) emax = ee; }
  uint64_t cnt = v.size();
  if (emax == 0) return false;
  int lsb = grid_lsb(emax, cnt);
  U128 W{0, 0}; uint64_t drops = 0;
  for (uint64_t b : v) {
    if ((b << 1) == 0) continue;
    bool lost; U128 x = to_grid(b, lsb, lost);
    unsigned __int128 a = ((unsigned __int128)W.hi << 64) | W.lo, y = ((unsigned __int128)x.hi << 64) | x.lo;
    a += y; W.lo = (uint64_t)a; W.hi = (uint64_t)(a >> 64); drops += lost;
  }
  settled = settle<53, -1022>(W, drops, lsb, cnt, sum, mean);
  if (!settled) full_width(v, cnt, sum, mean, 53);
  return true;
}


// Round-to-nearest-even of the positive integer x * 2^-S (x < 2^120) to binary32, as a float32
// pattern; a plain reading of the definition (no subnormal results in the ranges used).
static uint32_t rn32_of(unsigned __int128 x, int S, bool sticky) {
  int L = 0; for (unsigned __int128 t = x; t; t >>= 1) L++;
  int sh = L - 24; unsigned __int128 q; bool up = false;
  if (sh <= 0) { q = x << -sh; }
  else {
    q = x >> sh; unsigned __int128 rem = x & ((((unsigned __int128)1) << sh) - 1), half = ((unsigned __int128)1) << (sh - 1);
    up = rem > half || (rem == half && (sticky || (q & 1)));
    if (rem == half && !sticky && !(q & 1)) up = false;
  }
  if (sh <= 0 && sticky) { /* exact bits below the kept ones are zero; sticky alone never rounds up */ }
  q += up;
  int e = L - 1 - S;                       // exponent of the leading bit
  if (q >> 24) { q >>= 1; e++; }
  return (uint32_t)((e + 127) << 23) | (uint32_t)(q & 0x7FFFFF);
}

static long check_f32(std::mt19937_64& rng, long trials, long& unsettled) {
  long bad = 0;
  for (long t = 0; t < trials; t++) {
    int n = 1 + rng() % 60;
    std::vector<uint64_t> v(n);
    // float32 values with exponents in [2^-20, 2^20): exact sums fit x * 2^-44 with x < 2^90.
    unsigned __int128 pos = 0, negs = 0;
    int base = 107 + (int)(rng() % 30);
    for (int i = 0; i < n; i++) {
      uint32_t e = (uint32_t)(base + (int)(rng() % 11)) , f = (uint32_t)(rng() & 0x7FFFFF);
      if (rng() % 4 == 0) f &= 0x7F0000;
      uint32_t x = ((uint32_t)(rng() & 1) << 31) | (e << 23) | f;
      v[i] = widen_f32(x);
      unsigned __int128 m = (unsigned __int128)(f | 0x800000) << (e - 107);   // units of 2^-150... scaled below
      if (x >> 31) negs += m; else pos += m;
    }
    // x * 2^-(150 - 107 + 23 - ...) : value = m * 2^(e - 150) = (m << (e - 107)) * 2^-43
    const int S = 43;
    bool neg = negs > pos; unsigned __int128 mag = neg ? negs - pos : pos - negs;
    uint64_t cnt = n;
    uint32_t ws = 0, wm = 0;
    if (mag != 0) {
      ws = rn32_of(mag, S, false) | (neg ? 0x80000000u : 0);
      // mean: (mag << 40) / cnt with the remainder as sticky
      unsigned __int128 num = mag << 30, q = num / cnt; bool st = (num % cnt) != 0;
      wm = rn32_of(q, S + 30, st) | (neg ? 0x80000000u : 0);
    }
    uint32_t emax = 0; for (uint64_t b : v) { uint32_t e = (b >> 52) & 0x7FF; if (e > emax) emax = e; }
    int lsb = grid_lsb(emax, cnt);
    U128 W{0, 0}; uint64_t drops = 0;
    for (uint64_t b : v) {
      bool lost; U128 x = to_grid(b, lsb, lost);
      unsigned __int128 a = ((unsigned __int128)W.hi << 64) | W.lo, y = ((unsigned __int128)x.hi << 64) | x.lo;
      a += y; W.lo = (uint64_t)a; W.hi = (uint64_t)(a >> 64); drops += lost;
    }
    unsigned long long s, m;
    if (!settle<24, -126>(W, drops, lsb, cnt, s, m)) { unsettled++; full_width(v, cnt, s, m, 24); }
    if (s != ws || m != wm) {
      if (bad < 10) printf("F32 MISMATCH t=%ld n=%d sum %08x vs %08x mean %08x vs %08x\n", t, n, (unsigned)s, ws, (unsigned)m, wm);
      bad++;
    }
  }
  return bad;
}

static uint64_t dbits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

int main(int argc, char** argv) {
  std::mt19937_64 rng(argc > 1 ? atoll(argv[1]) : 1);
  long trials = argc > 2 ? atol(argv[2]) : 200000;
  long bad = 0, unsettled = 0, checked = 0;
  for (long t = 0; t < trials; t++) {
    int n = 1 + rng() % (t % 7 == 0 ? 3000 : 40);
    int mode = rng() % 8;
    std::vector<uint64_t> v(n);
    int base = 1 + rng() % 2040;
    for (int i = 0; i < n; i++) {
      uint64_t b;
      switch (mode) {
        case 0: b = rng(); break;                                   // random bits (exclude specials below)
        case 1: b = dbits(((double)(int64_t)(rng() % 2000001) - 1e6)); break;   // integers
        case 2: { int e = base + (int)(rng() % 120) - 60; if (e < 0) e = 0; if (e > 2046) e = 2046;
                  b = ((uint64_t)(rng() & 1) << 63) | ((uint64_t)e << 52) | (rng() & 0xFFFFFFFFFFFFFull); break; }
        case 3: b = ((uint64_t)(rng() & 1) << 63) | (rng() & 0xFFFFFFFFFFFFFull); break;   // subnormals
        case 4: { int e = (int)(rng() % 2047); b = ((uint64_t)(rng() & 1) << 63) | ((uint64_t)e << 52) | (rng() & 0xFFFFFFFFFFFFFull); break; }
        case 5: { // near-cancellation pairs
                  if (i & 1) b = v[i - 1] ^ (1ull << 63) ^ (rng() % 4); else { int e = base; b = ((uint64_t)e << 52) | (rng() & 0xFFFFFFFFFFFFFull) | ((uint64_t)(rng() & 1) << 63); }
                  break; }
        case 6: { // big + tiny values (drops)
                  int e = (i == 0) ? 2000 : (int)(rng() % 1900); b = ((uint64_t)(rng() & 1) << 63) | ((uint64_t)e << 52) | (rng() & 0xFFFFFFFFFFFFFull); break; }
        default: { // ties: values with few bits, near 2^53 boundary
                  int e = 1075 + (int)(rng() % 3); b = ((uint64_t)e << 52) | ((rng() % 8) << 49) | ((uint64_t)(rng() & 1) << 63); if (rng() % 3 == 0) b = dbits((double)(rng() % 16) * 0.5); break; }
      }
      uint32_t e = (b >> 52) & 0x7FF; if (e == 0x7FF) b &= ~(1ull << 62);   // keep finite
      v[i] = b;
    }
    unsigned long long s, m; bool settled = true;
    if (!product(v, s, m, settled)) continue;
    if (!settled) unsettled++;
    std::vector<int32_t> ids(n, 0);
    uint64_t os, om; uint8_t ov;
    or_group_agg(ids.data(), n, 1, v.data(), OR_F64, nullptr, 0, &os, &ov);
    or_group_agg(ids.data(), n, 1, v.data(), OR_F64, nullptr, 3, &om, &ov);
    checked++;
    if (s != os || m != om) {
      if (bad < 10) printf("MISMATCH t=%ld mode=%d n=%d settled=%d sum %016llx vs %016llx mean %016llx vs %016llx\n", t, mode, n, settled,
                           (unsigned long long)s, (unsigned long long)os, (unsigned long long)m, (unsigned long long)om);
      bad++;
    }
  }
  long unsettled32 = 0;
  long bad32 = check_f32(rng, trials / 4, unsettled32);
  printf("float64: checked %ld, unsettled %ld, mismatches %ld; float32: checked %ld, unsettled %ld, mismatches %ld\n",
         checked, unsettled, bad, trials / 4, unsettled32, bad32);
  return (bad != 0 || bad32 != 0) ? 1 : 0;
}
