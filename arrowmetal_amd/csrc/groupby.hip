// Hash group-by for CDNA4 (gfx950).
//
// Semantics follow the reference's GroupByKeys / GroupBy / GroupSumExact / GroupByExtrema
// (Sources/ArrowMetal/Kernels/GroupByKeys.swift:49-58, 80-149, 387-413; HashTable.swift:290-302;
// GroupBy.swift:46-182; GroupSumExactSource.swift:3-32; GroupByExtremaSource.swift:3-55):
//   * one group per distinct key; -0.0 == 0.0 and every NaN is one key; a null key is its own group;
//   * group order ascending by key with the null group last (lexicographic for several columns);
//   * the representative row of a group is its lowest row;
//   * integer sums wrap in 64 bits; float64 sums and means are correctly rounded;
//   * min / max compare on an order-preserving key (float64: the raw sign flip, so -0.0 < +0.0) and
//     skip NaN; a group with no value to answer with is null.
//
// How it runs here (this build's own design; CDNA4 has native 64-bit global and LDS atomics, so none
// of the reference's 32-bit carry work-arounds are needed):
//   gb_build     one read of the key column into an open-addressing table of 16-byte slots
//                (64-bit key, lowest row) with linear probing and a 64-bit CAS insert; every row
//                records its slot. The key 0x8000000000000000 is the empty marker and has a slot of
//                its own past the end of the table.
//   gb_compact   the occupied slots, in any order
//   (argsort)    the distinct keys' order keys, sorted by the radix argsort of sort.hip
//   gb_relabel   slot -> ascending group id, and the representative rows
//   gb_ids       each row's slot -> its group id, in place
//   gb_agg<P>    one pass over (ids, values) for a family of aggregates: a private LDS table per
//                workgroup when the groups fit, otherwise an LDS cache of the groups a workgroup sees
//                (claimed by CAS) in front of global 64-bit atomics; flushed once per workgroup.
#include "am_internal.hpp"
#include "fsum.hpp"

#include <algorithm>
#include <type_traits>
#include <map>

namespace am {

namespace {

constexpr unsigned long long EMPTY = 0x8000000000000000ull;
constexpr uint32_t NULL_SLOT = 0xFFFFFFFFu;

__device__ __forceinline__ bool bit_get(const uint8_t* v, int64_t i) { return (v[i >> 3] >> (i & 7)) & 1; }

// fmix64, the MurmurHash3 finaliser (a published mixing function), after a golden-ratio xor.
__device__ __forceinline__ unsigned long long mix64(unsigned long long k) {
  k ^= 0x9E3779B97F4A7C15ull;
  k ^= k >> 33;
  k *= 0xFF51AFD7ED558CCDull;
  k ^= k >> 33;
  k *= 0xC4CEB9FE1A85EC53ull;
  k ^= k >> 33;
  return k;
}

// The 64-bit grouping key of a non-null element: integers widen (sign-extended when signed), floats
// are normalised so that bit equality is Arrow value equality (HashTable.swift:290-302).
// Column reads of the row passes are streaming (non-temporal): each row is read once, and the hash
// and aggregate tables stay resident in L2 instead.
template <typename T>
__device__ __forceinline__ T ld_stream(const void* p, int64_t i) {
  return __builtin_nontemporal_load(static_cast<const T*>(p) + i);
}

// The row passes issue all their loads before using any: raw bits of a W-byte element, loaded
// unconditionally (a type switch around the loads would make the compiler wait on each at the branch
// join), and a validity byte from a one-byte all-valid constant when a column has no bitmap.
__device__ const uint8_t kAllValid = 0xFFu;

template <int W>
__device__ __forceinline__ unsigned long long ld_raw(const void* p, int64_t i) {
  if constexpr (W == 8) return ld_stream<unsigned long long>(p, i);
  else if constexpr (W == 4) return ld_stream<uint32_t>(p, i);
  else if constexpr (W == 2) return ld_stream<uint16_t>(p, i);
  else return ld_stream<uint8_t>(p, i);
}

// Whether row i is valid, from an unconditional load.
__device__ __forceinline__ bool ld_valid(const uint8_t* valid, int64_t voff, int64_t i) {
  const int64_t b = voff + i;
  const uint8_t* p = valid ? valid + (b >> 3) : &kAllValid;
  return (load_global(p) >> (valid ? (b & 7) : 0)) & 1;
}

// A row's validity byte, loaded (a column without a bitmap loads the all-valid constant).
__device__ __forceinline__ uint32_t ld_vbyte(const uint8_t* valid, int64_t voff, int64_t i) {
  return load_global(valid ? valid + ((voff + i) >> 3) : &kAllValid);
}

// Row i's bit of its loaded validity byte (a column without a bitmap loads 0xFF: every bit set). No
// branch: the byte is used on every path, so that its register is never left with a load in flight.
__device__ __forceinline__ uint32_t vbit(int64_t voff, int64_t i, uint32_t byte) {
  return (byte >> uint32_t((voff + i) & 7)) & 1u;
}

// The group key of raw element bits: integers widened with their sign, float NaN made canonical and
// -0.0 made +0.0 (registers only).
__device__ __forceinline__ unsigned long long key_of_raw(unsigned long long r, int type) {
  switch (type) {
    case T_I32: return (unsigned long long)(long long)int32_t(uint32_t(r));
    case T_I16: return (unsigned long long)(long long)int16_t(uint16_t(r));
    case T_I8: return (unsigned long long)(long long)int8_t(uint8_t(r));
    case T_F64:
      if ((r & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) return 0x7FF8000000000000ull;
      if ((r & 0x7FFFFFFFFFFFFFFFull) == 0ull) return 0ull;
      return r;
    case T_F32: {
      uint32_t b = uint32_t(r);
      if ((b & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC00000u;
      if ((b & 0x7FFFFFFFu) == 0u) return 0u;
      return b;
    }
    default: return r;   // unsigned integers and int64: the bits themselves
  }
}

__device__ __forceinline__ unsigned long long group_key(const void* vals, int64_t i, int type) {
  switch (type) {
    case T_I64: case T_U64: return ld_stream<unsigned long long>(vals, i);
    case T_I32: return (unsigned long long)(long long)ld_stream<int32_t>(vals, i);
    case T_I16: return (unsigned long long)(long long)ld_stream<int16_t>(vals, i);
    case T_I8: return (unsigned long long)(long long)ld_stream<int8_t>(vals, i);
    case T_U32: return ld_stream<uint32_t>(vals, i);
    case T_U16: return ld_stream<uint16_t>(vals, i);
    case T_U8: return ld_stream<uint8_t>(vals, i);
    case T_F64: {
      unsigned long long b = ld_stream<unsigned long long>(vals, i);
      if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) return 0x7FF8000000000000ull;
      if ((b & 0x7FFFFFFFFFFFFFFFull) == 0ull) return 0ull;
      return b;
    }
    default: {   // T_F32
      uint32_t b = ld_stream<uint32_t>(vals, i);
      if ((b & 0x7FFFFFFFu) > 0x7F800000u) return 0x7FC00000u;
      if ((b & 0x7FFFFFFFu) == 0u) return 0u;
      return b;
    }
  }
}

// Ascending order of the groups: signed, unsigned or float value order (the one NaN after +inf).
__device__ __forceinline__ unsigned long long order_key(unsigned long long gk, int type) {
  if (type == T_F64) return (gk & 0x8000000000000000ull) ? ~gk : (gk | 0x8000000000000000ull);
  if (type == T_F32) {
    uint32_t x = uint32_t(gk);
    return (x & 0x80000000u) ? uint32_t(~x) : (x | 0x80000000u);
  }
  if (type == T_I8 || type == T_I16 || type == T_I32 || type == T_I64) return gk ^ 0x8000000000000000ull;
  return gk;
}

// The hash table is two arrays of cap + 1 entries: the keys (EMPTY when free; entry cap holds the
// EMPTY key itself) and, for builds that track them, the lowest row of each key.
struct Table {
  unsigned long long* key;
  uint32_t* first;   // null when not tracked
};

struct BuildStatus {
  uint32_t overflow;
  uint32_t null_first;
  uint32_t distinct;
  uint32_t special_used;   // the EMPTY key has been seen: slot C holds it
  unsigned long long nulls;
  uint32_t xspecial[NXCD]; // the same, per XCD-private table
};

__global__ void gb_table_init(Table table, size_t slots, BuildStatus* st) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < slots; i += stride) {
    table.key[i] = EMPTY;
    if (table.first) table.first[i] = 0xFFFFFFFFu;
  }
  if (st && blockIdx.x == 0 && threadIdx.x == 0) {
    st->overflow = 0;
    st->null_first = 0xFFFFFFFFu;
    st->distinct = 0;
    st->special_used = 0;
    for (int x = 0; x < NXCD; x++) st->xspecial[x] = 0;
    st->nulls = 0;
  }
}

// A slot's lowest row only decreases, so a plain (possibly stale) read is a safe filter: a stale
// value is never below the true one, and the atomic settles every row that could lower it.
__device__ __forceinline__ void lower_first(uint32_t* p, uint32_t row) {
  if (row < *p) atomicMin(p, row);
}

// A probe's read of a slot key, through this XCD's L2. A stale EMPTY only sends the row to the CAS,
// which returns the settled key (device-scope reads measured slower on MI355X).
__device__ __forceinline__ unsigned long long probe_load(const unsigned long long* p) { return *p; }

// Probes (and inserts) key k, starting at its hash slot. Returns its slot, or NULL_SLOT when the probe
// budget ran out (the table is too full: the caller grows it and builds again).
__device__ __forceinline__ uint32_t probe_insert(unsigned long long* tkey, unsigned long long k, uint32_t s,
                                                 unsigned long long cur, uint32_t cap_mask, uint32_t budget) {
  uint32_t probes = 0;
  while (true) {
    if (cur == k) return s;
    if (cur == EMPTY) {
      unsigned long long prev = atomicCAS(&tkey[s], EMPTY, k);
      if (prev == EMPTY || prev == k) return s;
    }
    if (++probes >= budget) return NULL_SLOT;
    s = (s + 1u) & cap_mask;
    // A slot's key goes from EMPTY to its key once and never changes again, so a plain read is
    // enough: a stale EMPTY only sends the row to the CAS, which returns the settled key.
    cur = probe_load(&tkey[s]);
  }
}

// Statistics of a key column, read back once before the column is mapped:
//  - a HyperLogLog estimate of its distinct keys (2^12 registers), which sizes the hash table at about
//    5 slots per distinct key rather than two per row, so that the table stays small enough for L2;
//  - for integer keys, the smallest and largest order key, the nulls and the lowest null row, which
//    decide whether the range path maps the column.
constexpr int HLL_BITS = 12;
constexpr uint32_t HLL_M = 1u << HLL_BITS;

// The register hash: the high and low halves of a Fibonacci multiplicative hash of the key, folded
// and mixed by the first steps of the MurmurHash3 32-bit finaliser (published functions). 32 bits are
// enough: the estimate only sizes a table of at most 2^32 slots.
__device__ __forceinline__ uint32_t hll_hash(unsigned long long k) {
  const unsigned long long h = k * 0x9E3779B97F4A7C15ull;
  uint32_t x = uint32_t(h >> 32) ^ uint32_t(h);
  x ^= x >> 16;
  x *= 0x85EBCA6Bu;
  x ^= x >> 13;
  return x;
}

struct RangeStat {
  unsigned long long kmin, kmax, nulls;
  uint32_t null_first, pad;
};

struct KeyStat {
  RangeStat r;
  unsigned long long est;
};

__global__ __launch_bounds__(256) void gb_key_init(uint32_t* regs, RangeStat* st) {
  for (uint32_t j = threadIdx.x; j < HLL_M; j += blockDim.x) regs[j] = 0u;
  if (threadIdx.x == 0) {
    st->kmin = ~0ull;
    st->kmax = 0ull;
    st->nulls = 0ull;
    st->null_first = 0xFFFFFFFFu;
    st->pad = 0u;
  }
}

// One pass over the keys: each workgroup keeps its own registers in LDS, then raises the global ones.
// A global register is read first and raised only when the workgroup's is larger: registers only
// grow, so a stale read can only send a value to the atomic, which decides. Integer keys also give
// the smallest and largest order key, the nulls and the lowest null row.
//
// The key type is a template argument (no type switch per row) and each thread has two steps of
// SCAN_ROWS rows in flight: the next step's keys and validity bytes are loaded before this step's are
// hashed, and every loaded register is taken at the top of a step, so that no load is moved into a
// branch where it would be waited for alone.
constexpr int SCAN_ROWS = 8;
// Threads per workgroup of the key scan, 1024 threads per CU: one workgroup of 1024, or, for a column
// with a validity bitmap (whose bytes need more registers than the 128 a 1024-thread workgroup allows),
// two of 512.
template <bool HAS_VALID>
constexpr uint32_t scan_threads() {
  return HAS_VALID ? 512u : 1024u;
}
#ifndef AM_OCC_BLOCKS_PER_CU
#define AM_OCC_BLOCKS_PER_CU 2
#endif

struct ScanRows {
  unsigned long long raw[SCAN_ROWS];
  uint32_t vb[SCAN_ROWS];
};

template <int W, bool HAS_VALID>
__device__ __forceinline__ void scan_load(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n,
                                          uint64_t row0, uint32_t gsize, ScanRows& x) {
#pragma unroll
  for (int r = 0; r < SCAN_ROWS; r++) {
    const uint64_t i = row0 + uint64_t(r) * gsize;
    const int64_t ii = int64_t(i < n ? i : uint64_t(n) - 1u);
    x.raw[r] = ld_raw<W>(vals, ii);
    if (HAS_VALID) x.vb[r] = load_global(valid + ((voff + ii) >> 3));
  }
}

// Runs rows(row0, step) over every step of the grid-stride row loop, one step's rows loaded while the
// step before is processed (the step count is the same for every thread).
template <int W, bool HAS_VALID, class F>
__device__ __forceinline__ void scan_steps(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n, F&& rows) {
  const uint32_t gsize = gridDim.x * blockDim.x;
  const uint64_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = uint64_t(gsize) * SCAN_ROWS;
  const int steps = int((uint64_t(n) + stride - 1) / stride);
  if (steps == 0) return;
  ScanRows xa, xb;
  scan_load<W, HAS_VALID>(vals, valid, voff, n, tid, gsize, xa);
  for (int t = 0; t < steps; t += 2) {
    const uint64_t base = uint64_t(t) * stride + tid;
    scan_load<W, HAS_VALID>(vals, valid, voff, n, base + stride, gsize, xb);
    rows(base, xa);
    if (t + 1 >= steps) break;
    scan_load<W, HAS_VALID>(vals, valid, voff, n, base + 2 * stride, gsize, xa);
    rows(base + stride, xb);
  }
}

// Takes every loaded register of a step at once (before any branch uses one).
template <bool HAS_VALID>
__device__ __forceinline__ void scan_take(const ScanRows& x) {
#pragma unroll
  for (int r = 0; r < SCAN_ROWS; r++) {
    asm volatile("" ::"v"(x.raw[r]));
    if (HAS_VALID) asm volatile("" ::"v"(x.vb[r]));
  }
}

constexpr int width_of(int t) {
  return (t == T_I8 || t == T_U8) ? 1 : (t == T_I16 || t == T_U16) ? 2 : (t == T_I64 || t == T_U64 || t == T_F64) ? 8 : 4;
}

// Calls f(std::integral_constant<int, TYPE>) for a primitive key type.
template <class F>
void with_key_type(int type, F&& f) {
  switch (type) {
    case T_I8: f(std::integral_constant<int, T_I8>{}); break;
    case T_U8: f(std::integral_constant<int, T_U8>{}); break;
    case T_I16: f(std::integral_constant<int, T_I16>{}); break;
    case T_U16: f(std::integral_constant<int, T_U16>{}); break;
    case T_I32: f(std::integral_constant<int, T_I32>{}); break;
    case T_U32: f(std::integral_constant<int, T_U32>{}); break;
    case T_F32: f(std::integral_constant<int, T_F32>{}); break;
    case T_I64: f(std::integral_constant<int, T_I64>{}); break;
    case T_U64: f(std::integral_constant<int, T_U64>{}); break;
    case T_F64: f(std::integral_constant<int, T_F64>{}); break;
    default: fail("group-by: key type " + std::to_string(type) + " is not a primitive type");
  }
}

struct ScanAcc {
  unsigned long long lo = ~0ull, hi = 0ull, nz = 0ull;
  uint32_t nf = 0xFFFFFFFFu;
};

template <int TYPE, bool HAS_VALID>
__device__ __forceinline__ void scan_rows(int64_t voff, uint32_t n, uint64_t row0, uint32_t gsize, const ScanRows& x,
                                          uint32_t* r_lds, ScanAcc& acc) {
  constexpr bool RANGE = TYPE >= T_I8 && TYPE <= T_U64;
  scan_take<HAS_VALID>(x);
  uint32_t idx[SCAN_ROWS], rank[SCAN_ROWS];
#pragma unroll
  for (int r = 0; r < SCAN_ROWS; r++) {
    const uint64_t i = row0 + uint64_t(r) * gsize;
    const bool in = i < n;
    const int64_t ii = int64_t(in ? i : uint64_t(n) - 1u);
    const bool v = HAS_VALID ? vbit(voff, ii, x.vb[r]) != 0u : true;
    const bool ok = in && v;
    const unsigned long long gk = key_of_raw(x.raw[r], TYPE);
    if (RANGE) {
      const unsigned long long k = order_key(gk, TYPE);
      acc.lo = ok && k < acc.lo ? k : acc.lo;
      acc.hi = ok && k > acc.hi ? k : acc.hi;
      if (HAS_VALID) {
        const bool null_row = in && !v;
        acc.nz += null_row ? 1u : 0u;
        acc.nf = null_row ? min(acc.nf, uint32_t(i)) : acc.nf;
      }
    }
    const uint32_t h = hll_hash(gk);
    idx[r] = h >> (32 - HLL_BITS);
    const uint32_t w = h << HLL_BITS;
    rank[r] = !ok ? 0u : w ? uint32_t(__clz(int(w))) + 1u : uint32_t(32 - HLL_BITS + 1);
  }
  // The step's register reads are issued together; a register is raised only when the row's rank is
  // larger (a null row or a row past the end has rank 0).
  uint32_t cur[SCAN_ROWS];
#pragma unroll
  for (int r = 0; r < SCAN_ROWS; r++) cur[r] = r_lds[idx[r]];
#pragma unroll
  for (int r = 0; r < SCAN_ROWS; r++)
    if (rank[r] > cur[r]) atomicMax(&r_lds[idx[r]], rank[r]);
}

template <int W, int TYPE, bool HAS_VALID>
__global__ __launch_bounds__(scan_threads<HAS_VALID>()) void gb_key_scan(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n,
                                                   uint32_t* regs, RangeStat* st) {
  constexpr bool RANGE = TYPE >= T_I8 && TYPE <= T_U64;
  __shared__ uint32_t r[HLL_M];
  constexpr uint32_t SCAN_WAVES = scan_threads<HAS_VALID>() / 64;
  __shared__ unsigned long long s_lo[SCAN_WAVES], s_hi[SCAN_WAVES], s_nz[SCAN_WAVES];
  __shared__ uint32_t s_nf[SCAN_WAVES];
  for (uint32_t j = threadIdx.x; j < HLL_M; j += blockDim.x) r[j] = 0u;
  __syncthreads();
  const uint32_t gsize = gridDim.x * blockDim.x;
  ScanAcc acc;
  scan_steps<W, HAS_VALID>(vals, valid, voff, n, [&](uint64_t row0, const ScanRows& x) {
    scan_rows<TYPE, HAS_VALID>(voff, n, row0, gsize, x, r, acc);
  });
  if (RANGE) {
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long l2 = __shfl_xor(acc.lo, off, 64), h2 = __shfl_xor(acc.hi, off, 64);
      acc.lo = l2 < acc.lo ? l2 : acc.lo;
      acc.hi = h2 > acc.hi ? h2 : acc.hi;
      acc.nz += __shfl_xor(acc.nz, off, 64);
      acc.nf = min(acc.nf, uint32_t(__shfl_xor(int(acc.nf), off, 64)));
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
      s_lo[wv] = acc.lo;
      s_hi[wv] = acc.hi;
      s_nz[wv] = acc.nz;
      s_nf[wv] = acc.nf;
    }
  }
  __syncthreads();
  if (RANGE && threadIdx.x == 0) {
    for (uint32_t k = 1; k < SCAN_WAVES; k++) {
      acc.lo = s_lo[k] < acc.lo ? s_lo[k] : acc.lo;
      acc.hi = s_hi[k] > acc.hi ? s_hi[k] : acc.hi;
      acc.nz += s_nz[k];
      acc.nf = min(acc.nf, s_nf[k]);
    }
    if (acc.lo != ~0ull) atomicMin(&st->kmin, acc.lo);
    if (acc.hi != 0ull) atomicMax(&st->kmax, acc.hi);
    if (acc.nz) {
      atomicAdd(&st->nulls, acc.nz);
      atomicMin(&st->null_first, acc.nf);
    }
  }
  for (uint32_t j = threadIdx.x; j < HLL_M; j += blockDim.x) {
    const uint32_t mine = r[j];
    if (mine > load_global(&regs[j])) atomicMax(&regs[j], mine);
  }
}

void key_scan(int type, hipStream_t s, const void* vals, const uint8_t* valid, int64_t voff, uint32_t n,
              uint32_t* regs, RangeStat* st) {
  const int64_t threads = valid ? scan_threads<true>() : scan_threads<false>();
  const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((int64_t(n) + threads * SCAN_ROWS - 1) / (threads * SCAN_ROWS),
                                                                int64_t(device().cu_count) * (1024 / threads)));
  const dim3 grid{uint32_t(blocks)}, block{uint32_t(threads)};
  with_key_type(type, [&](auto tc) {
    constexpr int TT = decltype(tc)::value;
    if (valid)
      hipLaunchKernelGGL((gb_key_scan<width_of(TT), TT, true>), grid, block, 0, s, vals, valid, voff, n, regs, st);
    else
      hipLaunchKernelGGL((gb_key_scan<width_of(TT), TT, false>), grid, block, 0, s, vals, valid, voff, n, regs, st);
  });
}

// One block: the HyperLogLog estimate (with the linear-counting correction for small counts), written
// with the range statistics straight into pinned host memory.
__global__ __launch_bounds__(256) void gb_key_estimate(const uint32_t* regs, const RangeStat* st, KeyStat* out) {
  __shared__ float s_sum[256];
  __shared__ uint32_t s_zero[256];
  float sum = 0.f;
  uint32_t zeros = 0;
  for (uint32_t j = threadIdx.x; j < HLL_M; j += blockDim.x) {
    sum += __builtin_ldexpf(1.0f, -int(regs[j]));
    zeros += regs[j] == 0u ? 1u : 0u;
  }
  s_sum[threadIdx.x] = sum;
  s_zero[threadIdx.x] = zeros;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (int(threadIdx.x) < off) {
      s_sum[threadIdx.x] += s_sum[threadIdx.x + off];
      s_zero[threadIdx.x] += s_zero[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double m = double(HLL_M);
    double e = (0.7213 / (1.0 + 1.079 / m)) * m * m / double(s_sum[0]);
    if (e <= 2.5 * m && s_zero[0] > 0) e = m * log(m / double(s_zero[0]));
    out->r = *st;
    out->est = (unsigned long long)(e + 0.5);
  }
}

// Inserts every row's key; writes the slot of each row (NULL_SLOT for a null key). Each thread keeps
// BUILD_ROWS rows in flight: their keys are loaded, hashed and their first slots read together, so a
// wave waits on one memory latency per BUILD_ROWS rows rather than per row. TRACK_FIRST also keeps
// every slot's lowest row (the representative rows of float and multi-column keys).
//
// Each workgroup keeps the slots of the keys it has seen in an LDS cache (BUILD_CACHE entries, direct
// mapped by high hash bits, each written once: the key by CAS, then its slot). A row whose key is
// cached never reads the table: with skewed keys, the few hottest slot lines would otherwise take a
// large share of all table reads, on one L2 channel each.
constexpr int BUILD_ROWS = 8;
#ifndef AM_BUILD_CACHE
#define AM_BUILD_CACHE 8192
#endif
constexpr uint32_t BUILD_CACHE = AM_BUILD_CACHE;   // a power of two
// Threads per workgroup of the build and of gb_ids (the same geometry, so that a row's id lookup runs
// on the XCD whose table the row went into), and workgroups per CU. One workgroup of 16 waves per CU
// shares one 96 KiB cache: a key is missed once per CU rather than once per 256 threads, and the 8192
// entries hold more of a skewed key set (50M rows: 0.45 -> 0.41 ms with 100k Zipf keys, 0.29 -> 0.20
// ms with 200 keys).
#ifndef AM_BUILD_THREADS
#define AM_BUILD_THREADS 1024
#endif
#ifndef AM_BUILD_BLOCKS_PER_CU
#define AM_BUILD_BLOCKS_PER_CU 1
#endif
constexpr uint32_t BUILD_THREADS = AM_BUILD_THREADS;

// The cache entry of a key from its hash (bits above the table slot's).
__device__ __forceinline__ uint32_t cache_entry(unsigned long long h) { return uint32_t(h >> 40) & (BUILD_CACHE - 1u); }

// Caches key -> slot in the workgroup's LDS (entry c, the key's cache_entry) if the entry is still
// free. The slot is written after the key; a reader that finds the key with no slot yet reads the
// table instead.
__device__ __forceinline__ void cache_slot(unsigned long long* ckey, uint32_t* cslot, uint32_t c, unsigned long long k,
                                           uint32_t slot) {
  if (ckey[c] == EMPTY && atomicCAS(&ckey[c], EMPTY, k) == EMPTY) cslot[c] = slot;
}

// The keys and validity bytes of one build step's rows, as loaded.
struct BuildRows {
  unsigned long long k[BUILD_ROWS];
  uint32_t vb[BUILD_ROWS];
};

// KT: the key type when fixed at compile time (the common 64-bit integer key), or -1 for the
// column's type at run time; NV: the column has no validity bitmap (no byte is loaded).
template <int W, bool NV>
__device__ __forceinline__ void build_load(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n,
                                           uint64_t base, uint32_t gsize, uint32_t tid, BuildRows& x) {
#pragma unroll
  for (int r = 0; r < BUILD_ROWS; r++) {   // rows past the end read the last row
    const uint64_t iu = base + uint64_t(r) * gsize + tid;
    const int64_t ii = int64_t(iu < uint64_t(n) ? iu : uint64_t(n) - 1u);
    x.k[r] = ld_raw<W>(vals, ii);
    x.vb[r] = NV ? 0xFFu : ld_vbyte(valid, voff, ii);
  }
}

// What a thread of gb_build keeps across its steps.
struct BuildState {
  unsigned long long* tkey;
  uint32_t* tfirst;
  uint32_t* special_flag;
  uint32_t tbase, special, cap_mask, budget;
  unsigned long long* ckey;
  uint32_t* cslot;
  unsigned long long nulls;
  bool overflow;
};

// One step: the rows at `base` (loaded in `cur`). Their first table slots are read first, then the
// next step's rows are loaded (into `nxt`, at `nbase`), so that those loads are in flight while this
// step waits for its slots, and waiting for the slots never waits for them (memory operations
// complete in issue order).
template <bool TRACK_FIRST, int W, int KT, bool NV>
__device__ __forceinline__ void build_step(const void* vals, int type, const uint8_t* valid, int64_t voff, uint32_t n,
                                           uint64_t base, uint64_t nbase, uint32_t gsize, uint32_t tid,
                                           const BuildRows& cur, BuildRows& nxt, uint32_t* slot_of, BuildStatus* st,
                                           BuildState& b) {
#pragma unroll
  for (int r = 0; r < BUILD_ROWS; r++) asm volatile("" ::"v"(cur.k[r]), "v"(cur.vb[r]));   // every loaded register taken here
  unsigned long long k[BUILD_ROWS], cv[BUILD_ROWS];
  uint32_t s[BUILD_ROWS], ce[BUILD_ROWS];
  int cat[BUILD_ROWS];   // 0 key, 1 null, 2 the EMPTY key, 3 past the end, 4 cached
#pragma unroll
  for (int r = 0; r < BUILD_ROWS; r++) {
    const uint64_t i = base + uint64_t(r) * gsize + tid;
    const bool vok = NV || vbit(voff, int64_t(i < n ? i : uint64_t(n) - 1u), cur.vb[r]) != 0u;
    cat[r] = 3;
    k[r] = 0;
    if (i < n) {
      if (!vok) {
        cat[r] = 1;
      } else {
        k[r] = key_of_raw(cur.k[r], KT >= 0 ? KT : type);
        cat[r] = k[r] == EMPTY ? 2 : 0;
      }
    }
    const unsigned long long h = mix64(k[r]);
    s[r] = uint32_t(h) & b.cap_mask;
    ce[r] = cache_entry(h);
    // A cached key takes its slot from LDS (cat 4).
    if (cat[r] == 0) {
      const uint32_t c = ce[r];
      const uint32_t cs = b.cslot[c];
      if (b.ckey[c] == k[r] && cs != NULL_SLOT) {
        cat[r] = 4;
        s[r] = cs;
      }
    }
  }
#pragma unroll
  for (int r = 0; r < BUILD_ROWS; r++) cv[r] = cat[r] == 0 ? probe_load(&b.tkey[s[r]]) : 0ull;
  build_load<W, NV>(vals, valid, voff, n, nbase, gsize, tid, nxt);
  // Rows whose key already sits in its home slot (most rows of a repeated key) finish here; the
  // others (a first insert or a collision) go through the probe loop one at a time below, with the
  // key read again, so that the loop is not inlined per row.
  uint32_t slow = 0;
#pragma unroll
  for (int r = 0; r < BUILD_ROWS; r++) {
    const uint32_t i = uint32_t(base + uint64_t(r) * gsize + tid);
    if (cat[r] == 3) continue;
    uint32_t slot;
    if (cat[r] == 1) {
      slot = NULL_SLOT;
      b.nulls++;
      if (TRACK_FIRST) lower_first(&st->null_first, i);
    } else if (cat[r] == 2) {
      slot = b.special;
      if (!*b.special_flag) atomicExch(b.special_flag, 1u);
    } else if (cat[r] == 4) {
      slot = s[r];
    } else if (cv[r] == k[r]) {
      slot = s[r];
      cache_slot(b.ckey, b.cslot, ce[r], k[r], slot);
    } else {
      slow |= 1u << r;
      continue;
    }
    __builtin_nontemporal_store(slot == NULL_SLOT ? NULL_SLOT : b.tbase + slot, slot_of + i);
    if (TRACK_FIRST && slot != NULL_SLOT) lower_first(&b.tfirst[slot], i);
  }
  while (slow) {
    const int r = __builtin_ctz(slow);
    slow &= slow - 1u;
    const uint32_t i = uint32_t(base + uint64_t(r) * gsize + tid);
    const unsigned long long kk = key_of_raw(ld_raw<W>(vals, int64_t(i)), KT >= 0 ? KT : type);
    const unsigned long long hh = mix64(kk);
    const uint32_t ss = uint32_t(hh) & b.cap_mask;
    const uint32_t slot = probe_insert(b.tkey, kk, ss, probe_load(&b.tkey[ss]), b.cap_mask, b.budget);
    if (slot == NULL_SLOT) b.overflow = true;
    else cache_slot(b.ckey, b.cslot, cache_entry(hh), kk, slot);
    __builtin_nontemporal_store(slot == NULL_SLOT ? NULL_SLOT : b.tbase + slot, slot_of + i);
    if (TRACK_FIRST && slot != NULL_SLOT) lower_first(&b.tfirst[slot], i);
  }
}

template <bool TRACK_FIRST, int W, int KT, bool NV>
__global__ __launch_bounds__(BUILD_THREADS) void gb_build(const void* vals, int type, const uint8_t* valid, int64_t voff,
                                                uint32_t n, Table table, uint32_t cap_mask, uint32_t budget,
                                                uint32_t* slot_of, BuildStatus* st, int private_tables) {
  const uint32_t gsize = gridDim.x * blockDim.x;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  // With private tables each XCD inserts into its own copy (cap + 1 entries each), so a slot line in
  // an XCD's L2 is only ever changed by that XCD's own atomics; rows store the slot's index over all
  // the copies.
  const uint32_t tbl = private_tables ? xcd_id() : 0u;
  BuildState b;
  b.tbase = tbl * (cap_mask + 2u);
  b.tkey = table.key + b.tbase;
  b.tfirst = TRACK_FIRST ? table.first + b.tbase : nullptr;
  b.special_flag = private_tables ? &st->xspecial[tbl] : &st->special_used;
  b.special = cap_mask + 1u;
  b.cap_mask = cap_mask;
  b.budget = budget;
  __shared__ unsigned long long ckey[BUILD_CACHE];
  __shared__ uint32_t cslot[BUILD_CACHE];
  for (uint32_t j = threadIdx.x; j < BUILD_CACHE; j += blockDim.x) {
    ckey[j] = EMPTY;   // never a cached key: rows with the EMPTY key take the side slot
    cslot[j] = NULL_SLOT;
  }
  __syncthreads();
  b.ckey = ckey;
  b.cslot = cslot;
  b.nulls = 0;
  b.overflow = false;
  const uint64_t stride = uint64_t(gsize) * BUILD_ROWS;
  const int steps = int((uint64_t(n) + stride - 1u) / stride);   // the same for every thread
  if (steps > 0) {
    // Two row buffers take turns, so that no loaded register is copied (a copy would wait for it).
    BuildRows xa, xb;
    build_load<W, NV>(vals, valid, voff, n, 0, gsize, tid, xa);
    for (int t = 0; t < steps; t += 2) {
      const uint64_t base = uint64_t(t) * stride;
      build_step<TRACK_FIRST, W, KT, NV>(vals, type, valid, voff, n, base, base + stride, gsize, tid, xa, xb, slot_of, st, b);
      if (t + 1 >= steps) break;
      build_step<TRACK_FIRST, W, KT, NV>(vals, type, valid, voff, n, base + stride, base + 2 * stride, gsize, tid, xb, xa,
                                 slot_of, st, b);
    }
  }
  if (b.overflow) atomicOr(&st->overflow, 1u);
  unsigned long long nulls = b.nulls;
  for (int off = 32; off > 0; off >>= 1) nulls += __shfl_down(nulls, off, 64);
  if ((threadIdx.x & 63) == 0 && nulls) atomicAdd(&st->nulls, nulls);
}

// Folds the eight XCD-private tables into one table of cap + 1 entries: every occupied private slot
// inserts its key (a few per distinct key, not per row) and records where it went in xmap; lowest
// rows merge by minimum.
__global__ __launch_bounds__(256) void gb_merge_tables(const unsigned long long* xkeys, const uint32_t* xfirst,
                                                       uint32_t cap, BuildStatus* st, Table g, uint32_t* xmap) {
  const uint32_t per = cap + 1u;
  const uint32_t total = uint32_t(NXCD) * per;
  bool overflow = false;
  for (uint32_t xs = blockIdx.x * blockDim.x + threadIdx.x; xs < total; xs += gridDim.x * blockDim.x) {
    const uint32_t t = xs / per, sl = xs - t * per;
    uint32_t gs = NULL_SLOT;
    if (sl == cap) {
      if (st->xspecial[t]) {
        gs = cap;
        if (!st->special_used) atomicExch(&st->special_used, 1u);
      }
    } else {
      const unsigned long long k = xkeys[xs];
      if (k != EMPTY) {
        const uint32_t h = uint32_t(mix64(k)) & (cap - 1u);
        gs = probe_insert(g.key, k, h, probe_load(&g.key[h]), cap - 1u, cap);
        if (gs == NULL_SLOT) overflow = true;
      }
    }
    xmap[xs] = gs;
    if (xfirst && gs != NULL_SLOT) lower_first(&g.first[gs], xfirst[xs]);
  }
  if (overflow) atomicOr(&st->overflow, 1u);
}

// Group id of every private slot, through the merged table.
__global__ void gb_compose(const uint32_t* xmap, uint32_t total, const uint32_t* gid_of_slot, uint32_t* gid_x) {
  for (uint32_t xs = blockIdx.x * blockDim.x + threadIdx.x; xs < total; xs += gridDim.x * blockDim.x) {
    const uint32_t gs = xmap[xs];
    gid_x[xs] = gs == NULL_SLOT ? NULL_SLOT : gid_of_slot[gs];
  }
}

// Occupied slots -> (order key, slot), in slot order. One block per contiguous range of slots: a
// counting sweep, one atomic for the block's base, then the writing sweep. The special slot C is
// occupied when the EMPTY key was seen.
__global__ __launch_bounds__(256) void gb_compact(const unsigned long long* tkey, uint32_t cap, int type, BuildStatus* st,
                                                  unsigned long long* okeys, uint32_t* slots, uint32_t per_block) {
  __shared__ uint32_t wave_cnt[4];
  __shared__ uint32_t s_base;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t begin = blockIdx.x * per_block;
  const uint32_t end = min(cap + 1u, begin + per_block);
  auto occupied = [&](uint32_t x, unsigned long long& key) -> bool {
    if (x < cap) {
      key = tkey[x];
      return key != EMPTY;
    }
    key = EMPTY;
    return x == cap && st->special_used != 0;
  };
  uint32_t mine = 0;
  for (uint32_t x = begin + threadIdx.x; x - threadIdx.x < end; x += 256) {
    unsigned long long key;
    bool occ = x < end && occupied(x, key);
    mine += occ ? 1u : 0u;
  }
  for (int off = 32; off > 0; off >>= 1) mine += __shfl_down(mine, off, 64);
  if (lane == 0) wave_cnt[w] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
    s_base = tot ? atomicAdd(&st->distinct, tot) : 0u;
  }
  __syncthreads();
  uint32_t run = s_base;
  for (uint32_t x = begin + threadIdx.x; x - threadIdx.x < end; x += 256) {
    unsigned long long key = 0;
    bool occ = x < end && occupied(x, key);
    unsigned long long bal = __ballot(occ);
    // The waves of one step write in wave order: each needs the counts of the waves before it.
    __syncthreads();
    if (lane == 0) wave_cnt[w] = uint32_t(__popcll(bal));
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int v = 0; v < 4; v++) {
      before += v < w ? wave_cnt[v] : 0u;
      total += wave_cnt[v];
    }
    if (occ) {
      uint32_t at = run + before + uint32_t(__popcll(bal & ((1ull << lane) - 1ull)));
      okeys[at] = order_key(key, type);
      slots[at] = x;
    }
    run += total;
  }
}

__global__ void gb_relabel(const uint32_t* perm, const uint32_t* slots, uint32_t D, const uint32_t* tfirst,
                           uint32_t* gid_of_slot, uint32_t* rep, const BuildStatus* st, int has_null, int track) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r < D) {
    uint32_t s = slots[perm[r]];
    gid_of_slot[s] = r;
    if (track) rep[r] = tfirst[s];
  } else if (r == D && has_null && track) {
    rep[D] = st->null_first;
  }
}

#ifndef AM_IDS_CACHE_BITS
#define AM_IDS_CACHE_BITS 13
#endif
constexpr int IDS_CACHE_BITS = AM_IDS_CACHE_BITS;
constexpr uint32_t IDS_CACHE = 1u << IDS_CACHE_BITS;   // (slot * 2654435761) >> (32 - bits) indexes it

// The slots of one step's rows (BUILD_ROWS per thread, grid-strided), loaded together.
struct IdsStep {
  uint32_t sl[BUILD_ROWS];
};

__device__ __forceinline__ void ids_load(const uint32_t* slot_of, uint32_t n, uint64_t base, uint32_t gsize,
                                         uint32_t tid, IdsStep& x) {
#pragma unroll
  for (int r = 0; r < BUILD_ROWS; r++) {
    const uint64_t i = base + uint64_t(r) * gsize + tid;
    x.sl[r] = __builtin_nontemporal_load(slot_of + (i < n ? i : uint64_t(n) - 1u));
  }
}

__device__ __forceinline__ void ids_step(uint32_t* slot_of, uint32_t n, const uint32_t* gid_of_slot, uint32_t null_gid,
                                         uint64_t base, uint32_t gsize, uint32_t tid, const IdsStep& x,
                                         unsigned long long* cache) {
  auto entry = [](uint32_t sl) { return (sl * 2654435761u) >> (32 - IDS_CACHE_BITS); };   // [0, IDS_CACHE)
#pragma unroll
  for (int r = 0; r < BUILD_ROWS; r++) asm volatile("" ::"v"(x.sl[r]));   // every loaded register taken here
  uint32_t g[BUILD_ROWS];
  bool miss[BUILD_ROWS];
#pragma unroll
  for (int r = 0; r < BUILD_ROWS; r++) {
    // A row past the end holds the last row's word, which may already be an id: it is ignored.
    const uint32_t sl = base + uint64_t(r) * gsize + tid < n ? x.sl[r] : NULL_SLOT;
    miss[r] = false;
    if (sl == NULL_SLOT) {
      g[r] = null_gid;
      continue;
    }
    const unsigned long long e = cache[entry(sl)];
    miss[r] = uint32_t(e) != sl + 1u;
    g[r] = miss[r] ? load_global(gid_of_slot + sl) : uint32_t(e >> 32);
  }
#pragma unroll
  for (int r = 0; r < BUILD_ROWS; r++) {
    const uint64_t i = base + uint64_t(r) * gsize + tid;
    if (miss[r]) cache[entry(x.sl[r])] = (unsigned long long)(x.sl[r] + 1u) | ((unsigned long long)g[r] << 32);   // i < n
    if (i < n) __builtin_nontemporal_store(g[r], slot_of + i);
  }
}

// Rows' slots to group ids, in place. Each workgroup caches slot -> id pairs in LDS (one 64-bit word
// per entry, (slot + 1) | id << 32, direct mapped; any word a reader finds is a whole pair), so rows
// of hot groups do not all read the same few lines of the id table. Two steps' slots take turns: the
// next step's are loaded before this step's ids are looked up and stored. A step only reads slots that
// no earlier step rewrites (each row is its own thread's), so the prefetch never sees an id.
__global__ __launch_bounds__(BUILD_THREADS) void gb_ids(uint32_t* slot_of, uint32_t n, const uint32_t* gid_of_slot,
                                              uint32_t null_gid) {
  __shared__ unsigned long long cache[IDS_CACHE];
  for (uint32_t j = threadIdx.x; j < IDS_CACHE; j += blockDim.x) cache[j] = 0ull;
  __syncthreads();
  const uint32_t gsize = gridDim.x * blockDim.x;
  const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t stride = uint64_t(gsize) * BUILD_ROWS;
  const int steps = int((uint64_t(n) + stride - 1) / stride);   // the same for every thread
  if (steps == 0) return;
  IdsStep xa, xb;
  ids_load(slot_of, n, 0, gsize, tid, xa);
  for (int t = 0; t < steps; t += 2) {
    const uint64_t base = uint64_t(t) * stride;
    ids_load(slot_of, n, base + stride, gsize, tid, xb);
    ids_step(slot_of, n, gid_of_slot, null_gid, base, gsize, tid, xa, cache);
    if (t + 1 >= steps) break;
    ids_load(slot_of, n, base + 2 * stride, gsize, tid, xa);
    ids_step(slot_of, n, gid_of_slot, null_gid, base + stride, gsize, tid, xb, cache);
  }
}

// The group keys of an integer column, straight from the sorted distinct order keys (no
// representative rows needed: every row of a group holds the same integer). The null group, last,
// is written as 0 and marked null.
__global__ void gb_int_keys(const unsigned long long* okeys, const uint32_t* perm, uint32_t D, int type,
                            void* out, uint8_t* vb, int has_null) {
  uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > D || (r == D && !has_null)) return;
  unsigned long long v = 0;
  if (r < D) {
    unsigned long long ok = okeys[perm[r]];
    v = (type == T_I8 || type == T_I16 || type == T_I32 || type == T_I64) ? ok ^ 0x8000000000000000ull : ok;
  }
  vb[r] = r < D;
  switch (type_width(type)) {
    case 1: static_cast<uint8_t*>(out)[r] = uint8_t(v); break;
    case 2: static_cast<uint16_t*>(out)[r] = uint16_t(v); break;
    case 4: static_cast<uint32_t*>(out)[r] = uint32_t(v); break;
    default: static_cast<unsigned long long*>(out)[r] = v; break;
  }
}

__global__ void gb_combine(const int32_t* a, const int32_t* b, uint32_t n, long long kb, long long* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (; i < n; i += stride) out[i] = (long long)a[i] * kb + (long long)b[i];
}

int grid_for(int64_t n, int per_cu = 8) {
  int64_t b = (n + 255) / 256;
  return int(std::max<int64_t>(1, std::min<int64_t>(b, int64_t(device().cu_count) * per_cu)));
}

// ================================================================ aggregation framework

enum WordOp : int { W_ADD = 0, W_MIN = 1, W_MAX = 2, W_OR = 3, W_ADD128 = 4 /* lo */, W_HI = 5 /* its hi */ };

// Where a row's group id comes from: the int32 ids of a key mapping (always in [0, K)), or a dense
// key column of int32 / int64 / uint32 whose null and out-of-range keys are skipped (am_group_by).
struct IdSource {
  const void* ids;
  int type;               // T_I32, T_I64 or T_U32
  const uint8_t* valid;   // key validity, dense form only
  int64_t voff;
  long long bias = 0;     // added to every key before the range check (am_query's (sub key N))
};

__device__ __forceinline__ long long load_gid(const IdSource& s, int64_t i, long long K) {
  long long g;
  if (s.type == T_I32) g = ld_stream<int32_t>(s.ids, i);
  else if (s.type == T_U32) g = ld_stream<uint32_t>(s.ids, i);
  else g = ld_stream<long long>(s.ids, i);
  if (s.valid && !bit_get(s.valid, s.voff + i)) return -1;
  g += s.bias;
  return (g < 0 || g >= K) ? -1 : g;
}

// Integer transforms of a value column before it is aggregated (am_query's (shr c N), (bit_and c N)).
enum ValOp : int { VX_NONE = 0, VX_SHR = 1, VX_AND = 2 };

struct ValSource {
  const void* vals;
  int type;
  const uint8_t* valid;
  int64_t voff;
  int xop = VX_NONE;      // applied to the widened 64-bit value
  long long xarg = 0;
};

__device__ __forceinline__ unsigned long long raw64(const ValSource& v, int64_t i) {
  switch (v.type) {
    case T_I64: case T_U64: case T_F64: return ld_stream<unsigned long long>(v.vals, i);
    case T_I32: return (unsigned long long)(long long)ld_stream<int32_t>(v.vals, i);
    case T_I16: return (unsigned long long)(long long)ld_stream<int16_t>(v.vals, i);
    case T_I8: return (unsigned long long)(long long)ld_stream<int8_t>(v.vals, i);
    case T_U32: case T_F32: return ld_stream<uint32_t>(v.vals, i);
    case T_U16: return ld_stream<uint16_t>(v.vals, i);
    default: return ld_stream<uint8_t>(v.vals, i);
  }
}

__device__ __forceinline__ bool val_valid(const ValSource& v, int64_t i) {
  return !v.valid || bit_get(v.valid, v.voff + i);
}

// A value's 64-bit word from its raw bits (what raw64 reads): signed integers sign-extended, the rest
// zero-extended (registers only).
__device__ __forceinline__ unsigned long long val_of_raw(unsigned long long r, int type) {
  switch (type) {
    case T_I32: return (unsigned long long)(long long)int32_t(uint32_t(r));
    case T_I16: return (unsigned long long)(long long)int16_t(uint16_t(r));
    case T_I8: return (unsigned long long)(long long)int8_t(uint8_t(r));
    default: return r;
  }
}

// The value a row contributes: its widened word, transformed when the column asks for it. A shift
// right of a signed value is arithmetic (Arrow's shift_right); the transforms take integer columns.
__device__ __forceinline__ unsigned long long val_in(const ValSource& v, unsigned long long r) {
  unsigned long long x = val_of_raw(r, v.type);
  if (v.xop == VX_SHR) {
    const int sh = int(v.xarg & 63);
    x = type_is_signed(v.type) ? (unsigned long long)((long long)x >> sh) : x >> sh;
  } else if (v.xop == VX_AND) {
    x &= (unsigned long long)v.xarg;
  }
  return x;
}

// Aggregate families. Each has W 64-bit words per group with an op per word, an identity per word,
// and a per-row contribution (false when the row contributes nothing).
struct AggArgs {
  IdSource ids;
  ValSource v;
  int64_t n;
  long long K;
  const void* aux;               // family-specific (exact float sums: the groups' grid words)
  const long long* hot;          // HOT_GROUPS ids (-1: none) summed in registers, or null
};

// Up to this many of a column's most frequent groups (from a sample of its ids) are summed in each
// lane's registers by the families whose words are all additions (sums and counts), so that the rows
// of a hot group take no atomic at all: with Zipf(1.1) keys the top group alone is ~13% of the rows,
// about nine lanes of every wave, whose LDS atomics on one word would otherwise serialize.
#ifndef AM_HOT_GROUPS
#define AM_HOT_GROUPS 8
#endif
constexpr int HOT_GROUPS = AM_HOT_GROUPS;

// count_all (values ignored) and count (non-null values).
template <bool VALID_ONLY>
struct CountP {
  static constexpr int W = 1;
  static constexpr bool SKIP_NULL = VALID_ONLY;
  __host__ __device__ static constexpr int op(int) { return W_ADD; }
  static constexpr bool USES_VALUE = false;
  static constexpr int FIXED_VW = 0;
  static constexpr bool AUX = false;
  static constexpr bool PART = true;        // may run through the partitioned passes
  static constexpr bool NEEDS_ROW = false;  // its contribution depends on the row number
  __device__ static bool contrib(const AggArgs&, int64_t, long long, unsigned long long, unsigned long long (&c)[1]) {
    c[0] = 1;
    return true;
  }
};

// Integer sum (wrapping) with the count of non-null values (null result where it is zero).
struct SumCountP {
  static constexpr int W = 2;
  static constexpr bool SKIP_NULL = true;
  __host__ __device__ static constexpr int op(int) { return W_ADD; }
  static constexpr bool USES_VALUE = true;
  static constexpr int FIXED_VW = 0;
  static constexpr bool AUX = false;
  static constexpr bool PART = true;
  static constexpr bool NEEDS_ROW = false;
  __device__ static bool contrib(const AggArgs&, int64_t, long long, unsigned long long x, unsigned long long (&c)[2]) {
    c[0] = x;
    c[1] = 1;
    return true;
  }
};

// min and max of an order-preserving unsigned key; NaN skipped. A group has a value exactly when
// its min key <= its max key, which the identities (all ones, zero) never satisfy.
struct MinMaxP {
  static constexpr int W = 2;
  static constexpr bool SKIP_NULL = true;
  __host__ __device__ static constexpr int op(int w) { return w == 0 ? W_MIN : W_MAX; }
  static constexpr bool USES_VALUE = true;
  static constexpr int FIXED_VW = 0;
  static constexpr bool AUX = false;
  static constexpr bool PART = true;
  static constexpr bool NEEDS_ROW = false;
  __device__ static bool contrib(const AggArgs& a, int64_t, long long, unsigned long long x, unsigned long long (&c)[2]) {
    unsigned long long b = x, key;
    switch (a.v.type) {
      case T_F64:
        if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) return false;
        key = (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
        break;
      case T_F32: {
        uint32_t x = uint32_t(b);
        if ((x & 0x7FFFFFFFu) > 0x7F800000u) return false;
        key = (x & 0x80000000u) ? uint32_t(~x) : (x | 0x80000000u);
        break;
      }
      case T_I8: case T_I16: case T_I32: case T_I64: key = b ^ 0x8000000000000000ull; break;
      default: key = b; break;
    }
    c[0] = key;
    c[1] = key;
    return true;
  }
};

// ---- exact float sums (arithmetic in fsum.hpp)
//
// Pass 1, FsScanP: per group, word 0 the largest biased exponent of the finite non-zero values (1 for
// subnormals), word 1 the value classes seen (FS_*), word 2 the count of non-null values, word 3 the
// lowest row holding a NaN. fs_window then turns words 0-2 into the group's grid: the bit offset
// `lsb` of the fixed-point unit above 2^-1074, or FS_SKIP when the group has no finite non-zero value
// or has a NaN or an infinity (its result does not depend on the finite values). Pass 2, FsFixP: per
// group, the 128-bit two's-complement sum of floor(value / unit) (words 0-1) and the number of values
// that lost bits to the floor (word 2). VW is the width of the values: 8 for float64, 4 for float32,
// which is widened exactly on load.
constexpr unsigned long long FS_POS_INF = 1ull, FS_NEG_INF = 2ull, FS_NAN = 4ull, FS_POS_ZERO = 8ull;
constexpr uint32_t FS_SKIP = 0xFFFFFFFFu;

template <int VW>
__device__ __forceinline__ unsigned long long fs_bits(unsigned long long raw) {
  if constexpr (VW == 4) return fs::widen_f32(uint32_t(raw));
  else return raw;
}

template <int VW>
struct FsScanP {
  static constexpr int W = 4;
  static constexpr bool SKIP_NULL = true;
  __host__ __device__ static constexpr int op(int w) { return w == 0 ? W_MAX : (w == 1 ? W_OR : (w == 2 ? W_ADD : W_MIN)); }
  static constexpr bool USES_VALUE = true;
  static constexpr int FIXED_VW = VW;
  static constexpr bool AUX = false;
  static constexpr bool PART = true;
  static constexpr bool NEEDS_ROW = true;   // word 3 is the lowest NaN row
  __device__ static bool contrib(const AggArgs&, int64_t row, long long, unsigned long long raw,
                                 unsigned long long (&c)[4]) {
    const unsigned long long v = fs_bits<VW>(raw);
    const uint32_t be = uint32_t(v >> 52) & 0x7FFu;
    const bool neg = (v >> 63) != 0ull, zero_frac = (v & 0xFFFFFFFFFFFFFull) == 0ull;
    c[0] = 0ull;               // identities of MAX / OR / MIN: no effect
    c[1] = 0ull;
    c[2] = 1ull;
    c[3] = ~0ull;
    switch (be) {
      case 0x7FFu:
        if (zero_frac) c[1] = neg ? FS_NEG_INF : FS_POS_INF;
        else { c[1] = FS_NAN; c[3] = (unsigned long long)row; }
        break;
      case 0u:
        if (zero_frac) { c[1] = neg ? 0ull : FS_POS_ZERO; break; }
        c[0] = 1ull;           // a subnormal: its unit is that of exponent 1
        break;
      default:
        c[0] = be;
        break;
    }
    return true;
  }
};

template <int VW>
struct FsFixP {
  static constexpr int W = 3;
  static constexpr bool SKIP_NULL = true;
  __host__ __device__ static constexpr int op(int w) { return w == 0 ? W_ADD128 : (w == 1 ? W_HI : W_ADD); }
  static constexpr bool USES_VALUE = true;
  static constexpr int FIXED_VW = VW;
  // The group's grid word from fs_window, loaded for all rows in flight before the contributions.
  static constexpr bool AUX = true;
  static constexpr bool PART = true;
  static constexpr bool NEEDS_ROW = false;
  using AuxT = uint32_t;
  __device__ static AuxT aux_load(const AggArgs& a, long long g) {
    return reinterpret_cast<const uint32_t*>(a.aux)[g];
  }
  __device__ static bool contrib(const AggArgs&, int64_t, long long, unsigned long long raw, AuxT lsb,
                                 unsigned long long (&c)[3]) {
    if (lsb == FS_SKIP) return false;
    const unsigned long long v = fs_bits<VW>(raw);
    const uint32_t be = uint32_t(v >> 52) & 0x7FFu;
    if (be == 0x7FFu || (v << 1) == 0ull) return false;   // specials and zeros add nothing
    bool lost;
    const fs::U128 x = fs::to_grid(v, int(lsb), lost);
    c[0] = x.lo;
    c[1] = x.hi;
    c[2] = lost ? 1ull : 0ull;
    return (c[0] | c[1] | c[2]) != 0ull;
  }
};

__device__ __forceinline__ unsigned long long identity(int op) { return op == W_MIN ? ~0ull : 0ull; }


// One word update in LDS or in this XCD's private global table; workgroup-scope atomics in both
// (ds_* in LDS, L2-resident in global memory). A contribution equal to the op's identity touches
// nothing, not even a read. MIN / MAX / OR read the word first (an atomic relaxed load of the same
// address: one 8-byte access, never torn) and skip the atomic when it cannot change anything, which
// keeps hot groups read-mostly. Those words only ever move one way (MIN down, MAX up, OR gaining
// bits), so any value the read returns, however stale, lies on the far side of the current value:
// x >= read >= current makes the MIN a no-op, and likewise for MAX and OR.
// Accumulator words with their address space spelt out: an atomic through a generic pointer is a
// flat instruction, and a flat operation in flight makes every later wait a full drain of both the
// vector-memory and the LDS counters.
typedef __attribute__((address_space(1))) unsigned long long gword;   // global memory (this XCD's table)
typedef __attribute__((address_space(3))) unsigned long long lword;   // LDS

__device__ __forceinline__ gword* as_global(unsigned long long* p) { return (gword*)p; }
__device__ __forceinline__ lword* as_lds(unsigned long long* p) { return (lword*)p; }

template <int OP, typename Word>
__device__ __forceinline__ void word_update(Word* p, unsigned long long x) {
  auto seen = [p] { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); };
  if (OP == W_ADD) {
    if (x) __hip_atomic_fetch_add(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else if (OP == W_MIN) {
    if (x != ~0ull && x < seen()) __hip_atomic_fetch_min(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else if (OP == W_MAX) {
    if (x != 0ull && x > seen()) __hip_atomic_fetch_max(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else if (OP == W_OR) {
    if (x != 0ull && (x & ~seen())) __hip_atomic_fetch_or(p, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

// 128-bit two's-complement add at p[0] (lo), p[1] (hi): exact modulo 2^128 under any interleaving,
// because each carry comes from the lo word's own read-modify-write.
template <typename Word>
__device__ __forceinline__ void add128(Word* p, unsigned long long lo, unsigned long long hi) {
  unsigned long long carry = 0;
  if (lo) {
    unsigned long long old = __hip_atomic_fetch_add(&p[0], lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    carry = (old + lo < old) ? 1ull : 0ull;
  }
  unsigned long long h = hi + carry;
  if (h) __hip_atomic_fetch_add(&p[1], h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <class P, typename Word>
__device__ __forceinline__ void apply(Word* acc, const unsigned long long (&c)[P::W]) {
#pragma unroll
  for (int w = 0; w < P::W; w++) {
    const int op = P::op(w);
    if (op == W_ADD) word_update<W_ADD>(&acc[w], c[w]);
    else if (op == W_MIN) word_update<W_MIN>(&acc[w], c[w]);
    else if (op == W_MAX) word_update<W_MAX>(&acc[w], c[w]);
    else if (op == W_OR) word_update<W_OR>(&acc[w], c[w]);
    else if (op == W_ADD128) add128(&acc[w], c[w], c[w + 1 < P::W ? w + 1 : w]);
    // W_HI: added together with its W_ADD128 word
  }
}

// Merges one LDS accumulator into this XCD's table (skipping identities).
template <class P>
__device__ __forceinline__ void flush_one(gword* gacc, const lword* lacc) {
  unsigned long long c[P::W];
  bool any = false;
#pragma unroll
  for (int w = 0; w < P::W; w++) {
    c[w] = lacc[w];
    any |= c[w] != identity(P::op(w));
  }
  if (any) apply<P>(gacc, c);
}

#ifndef AM_AGG_ROWS
#define AM_AGG_ROWS 2
#endif
constexpr int AGG_ROWS = AM_AGG_ROWS;   // rows per thread per step, whose loads are issued together

// 1024 threads per workgroup and one workgroup per CU: the whole LDS budget goes to one table.
#ifndef AM_AGG_THREADS
#define AM_AGG_THREADS 1024
#endif
#ifndef AM_AGG_PROBES
#define AM_AGG_PROBES 8   // LDS cache slots a row probes before its group goes to the XCD's table
#endif
#ifndef AM_AGG_HOT
#define AM_AGG_HOT 0      // sampled hot groups summed in registers (1), or not (0)
#endif

template <class P>
__host__ __device__ constexpr bool all_add() {
  for (int w = 0; w < P::W; w++)
    if (P::op(w) != W_ADD) return false;
  return true;
}

// The rows one thread has in flight: group ids, values and validity bytes exactly as loaded. The bits
// are taken out only when the rows are used, so that nothing that is loaded is waited on before the
// whole step's loads are issued (a bool made from a loaded byte is a lane mask, and making it waits).
template <int IDW>
struct AggRows {
  typename std::conditional<IDW == 8, unsigned long long, uint32_t>::type idr[AGG_ROWS];
  unsigned long long vr[AGG_ROWS];
  uint32_t ib[AGG_ROWS], vb[AGG_ROWS];
};


// Issues the loads of the step at `base` (rows past the end read the last row; end > 0).
template <class P, int IDW, int VW>
__device__ __forceinline__ void agg_load(const AggArgs& a, int64_t base, int64_t end, AggRows<IDW>& r) {
#pragma unroll
  for (int k = 0; k < AGG_ROWS; k++) {
    const int64_t i = base + int64_t(k) * blockDim.x + threadIdx.x;
    const int64_t ii = i < end ? i : end - 1;
    r.idr[k] = ld_raw<IDW>(a.ids.ids, ii);
    r.ib[k] = ld_vbyte(a.ids.valid, a.ids.voff, ii);
    r.vr[k] = 0;
    if constexpr (VW != 0) r.vr[k] = ld_raw<VW>(a.v.vals, ii);
    r.vb[k] = P::SKIP_NULL ? ld_vbyte(a.v.valid, a.v.voff, ii) : 0xFFu;
  }
}

// What a workgroup of gb_agg holds while it runs over its rows.
template <class P>
struct AggState {
  unsigned long long* lacc;   // LDS accumulators [cap][W]
  unsigned int* lkey;         // LDS cache keys [cap] (mode 1)
  unsigned long long* gtab;   // this XCD's table [K][W]
  int mode;
  uint32_t cap;
  long long hot[HOT_GROUPS];
  unsigned long long hacc[HOT_GROUPS][P::W];
};

// The updates of one step's rows (loaded by agg_load at `base`).
template <class P, int IDW, int VW>
__device__ __forceinline__ void agg_rows(const AggArgs& a, int64_t base, int64_t end, const AggRows<IDW>& cur,
                                         AggState<P>& st) {
  constexpr bool HOTREG = AM_AGG_HOT && all_add<P>();
  // Every loaded register is taken here, on every path: a load whose register is only read on some
  // paths is still in flight, as far as the compiler knows, when that register is next written, and
  // that write then waits for all memory operations.
#pragma unroll
  for (int r = 0; r < AGG_ROWS; r++) asm volatile("" ::"v"(cur.idr[r]), "v"(cur.vr[r]), "v"(cur.ib[r]), "v"(cur.vb[r]));
  long long gs[AGG_ROWS];
  unsigned long long cs[AGG_ROWS][P::W];
#pragma unroll
  for (int r = 0; r < AGG_ROWS; r++) {
    const int64_t i = base + int64_t(r) * blockDim.x + threadIdx.x;
    long long g;
    if constexpr (IDW == 8) g = (long long)cur.idr[r];
    else g = a.ids.type == T_I32 ? (long long)int32_t(cur.idr[r]) : (long long)cur.idr[r];
    g += a.ids.bias;
    const uint32_t ok = uint32_t(i < end) & vbit(a.ids.voff, i, cur.ib[r]) &
                        (P::SKIP_NULL ? vbit(a.v.voff, i, cur.vb[r]) : 1u);
    gs[r] = (ok && g >= 0 && g < a.K) ? g : -1;
  }
  // Every row's contribution is computed (registers only) and then dropped where the row has none,
  // again so that each loaded value is used on every path.
  if constexpr (P::AUX) {
    typename P::AuxT ax[AGG_ROWS];
#pragma unroll
    for (int r = 0; r < AGG_ROWS; r++) ax[r] = P::aux_load(a, gs[r] >= 0 ? gs[r] : 0);   // all loads first
#pragma unroll
    for (int r = 0; r < AGG_ROWS; r++) {
      const int64_t i = base + int64_t(r) * blockDim.x + threadIdx.x;
      if (!P::contrib(a, i, gs[r], val_in(a.v, cur.vr[r]), ax[r], cs[r])) gs[r] = -1;
    }
  } else {
#pragma unroll
    for (int r = 0; r < AGG_ROWS; r++) {
      const int64_t i = base + int64_t(r) * blockDim.x + threadIdx.x;
      if (!P::contrib(a, i, gs[r], val_in(a.v, cur.vr[r]), cs[r])) gs[r] = -1;
    }
  }
  if (HOTREG && st.hot[0] >= 0) {   // uniform; the sampler fills hot[] most frequent first
    // A hot group's contribution goes to this lane's registers (the hot ids are distinct, and an
    // unused entry's accumulator is never flushed).
#pragma unroll
    for (int r = 0; r < AGG_ROWS; r++) {
      bool any = false;
#pragma unroll
      for (int j = 0; j < HOT_GROUPS; j++) {
        const bool m = gs[r] == st.hot[j];
#pragma unroll
        for (int w = 0; w < P::W; w++) st.hacc[j][w] += m ? cs[r][w] : 0ull;
        any = any || m;
      }
      if (any) gs[r] = -1;
    }
  }
  if (st.mode == 0) {
#pragma unroll
    for (int r = 0; r < AGG_ROWS; r++)
      if (gs[r] >= 0) apply<P>(as_lds(&st.lacc[size_t(gs[r]) * P::W]), cs[r]);
    return;
  }
  // Mode 1: every row's home slot is read at once; a row whose group holds it updates LDS at once,
  // the others probe on (claiming a free slot by CAS) and fall back to the XCD's table.
  uint32_t sl[AGG_ROWS], seen[AGG_ROWS];
#pragma unroll
  for (int r = 0; r < AGG_ROWS; r++) {
    const uint32_t h = uint32_t(gs[r]) * 2654435761u;
    sl[r] = __umulhi(h ^ (h >> 15), st.cap);   // [0, cap) for any cap
    seen[r] = gs[r] >= 0 ? st.lkey[sl[r]] : 0u;
  }
#pragma unroll
  for (int r = 0; r < AGG_ROWS; r++) {
    const long long g = gs[r];
    if (g < 0) continue;
    if (seen[r] == uint32_t(g)) {
      apply<P>(as_lds(&st.lacc[size_t(sl[r]) * P::W]), cs[r]);
      continue;
    }
    uint32_t s = sl[r];
    uint32_t c = seen[r];
    bool done = false;
    for (int probe = 0; probe < AM_AGG_PROBES; probe++) {
      if (probe > 0) c = st.lkey[s];
      if (c == 0xFFFFFFFFu) {
        const unsigned int prev = atomicCAS(&st.lkey[s], 0xFFFFFFFFu, uint32_t(g));
        c = prev == 0xFFFFFFFFu ? uint32_t(g) : prev;
      }
      if (c == uint32_t(g)) {
        apply<P>(as_lds(&st.lacc[size_t(s) * P::W]), cs[r]);
        done = true;
        break;
      }
      s = s + 1u == st.cap ? 0u : s + 1u;
    }
    if (!done) apply<P>(as_global(&st.gtab[size_t(g) * P::W]), cs[r]);
  }
}

// IDW: byte width of the group ids (4: int32 / uint32, 8: int64); VW: byte width of the values (0 when
// the family reads none).
// Mode 0: every group has a private LDS accumulator (K <= cap). Mode 1: an LDS cache of `cap`
// groups, claimed by CAS on a key word, in front of this XCD's table in global memory (L2); a row
// whose group cannot be cached within AM_AGG_PROBES probes goes to that table directly. The tables of
// the eight XCDs are merged by gb_agg_merge.
// Two row buffers take turns: the next step's loads are issued before the current step's updates, and
// the loop is unrolled over the pair so that no loaded register is copied (a copy would wait for it).
template <class P, int IDW, int VW>
__global__ __launch_bounds__(AM_AGG_THREADS) void gb_agg(AggArgs a, int mode, uint32_t cap, int64_t rpb,
                                              unsigned long long* gtab8) {
  extern __shared__ unsigned long long lds[];
  constexpr bool HOTREG = AM_AGG_HOT && all_add<P>();
  AggState<P> st;
  st.lacc = lds;
  st.lkey = reinterpret_cast<unsigned int*>(lds + size_t(cap) * P::W);
  st.gtab = gtab8 + size_t(xcd_id()) * size_t(a.K) * P::W;
  st.mode = mode;
  st.cap = cap;
#pragma unroll
  for (int j = 0; j < HOT_GROUPS; j++) {
    const long long h = HOTREG && a.hot ? a.hot[j] : -1;   // the same for every lane
    st.hot[j] = (long long)(((unsigned long long)uint32_t(__builtin_amdgcn_readfirstlane(int(h >> 32))) << 32) |
                            uint32_t(__builtin_amdgcn_readfirstlane(int(h))));
#pragma unroll
    for (int w = 0; w < P::W; w++) st.hacc[j][w] = 0ull;
  }
  for (uint32_t s = threadIdx.x; s < cap; s += blockDim.x) {
#pragma unroll
    for (int w = 0; w < P::W; w++) st.lacc[size_t(s) * P::W + w] = identity(P::op(w));
    if (mode == 1) st.lkey[s] = 0xFFFFFFFFu;
  }
  __syncthreads();
  const int64_t begin = int64_t(blockIdx.x) * rpb;
  const int64_t end = std::min<int64_t>(a.n, begin + rpb);
  const int64_t step = int64_t(blockDim.x) * AGG_ROWS;
  // The step count, read from the first lane so that the compiler keeps the loop uniform.
  const int steps = __builtin_amdgcn_readfirstlane(int(begin < end ? (end - begin + step - 1) / step : 0));
  if (steps > 0) {
    AggRows<IDW> ra, rb;
    agg_load<P, IDW, VW>(a, begin, end, ra);
    for (int t = 0; t < steps; t += 2) {
      const int64_t base = begin + int64_t(t) * step;
      agg_load<P, IDW, VW>(a, base + step, end, rb);
      agg_rows<P, IDW, VW>(a, base, end, ra, st);
      if (t + 1 >= steps) break;
      agg_load<P, IDW, VW>(a, base + 2 * step, end, ra);
      agg_rows<P, IDW, VW>(a, base + step, end, rb, st);
    }
  }
  if constexpr (HOTREG) {
    // Each wave's sum of its lanes' hot accumulators, added once into this XCD's table.
#pragma unroll
    for (int j = 0; j < HOT_GROUPS; j++) {
      if (st.hot[j] < 0) continue;   // uniform
#pragma unroll
      for (int w = 0; w < P::W; w++) {
        unsigned long long v = st.hacc[j][w];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        st.hacc[j][w] = v;
      }
      if ((threadIdx.x & 63) == 0) apply<P>(as_global(&st.gtab[size_t(st.hot[j]) * P::W]), st.hacc[j]);
    }
  }
  __syncthreads();
  for (uint32_t sl = threadIdx.x; sl < cap; sl += blockDim.x) {
    long long g;
    if (mode == 0) g = sl < uint32_t(a.K) ? (long long)sl : -1;
    else g = st.lkey[sl] == 0xFFFFFFFFu ? -1 : (long long)st.lkey[sl];
    if (g >= 0) flush_one<P>(as_global(&st.gtab[size_t(g) * P::W]), as_lds(&st.lacc[size_t(sl) * P::W]));
  }
}

// The most frequent group ids of a sample of the rows: S rows at even strides counted in an LDS
// table, then the HOT_GROUPS largest counts of at least S / 64 (about one lane per wave). One
// workgroup of 1024 threads (four rows each); hot[j] = -1 where there is no such group.
constexpr uint32_t HOT_SAMPLE = 4096;   // a group of >= 1/64 of the rows shows ~64 times
constexpr uint32_t HOT_SLOTS = 8192;    // twice the sample: probe chains stay short
__global__ __launch_bounds__(1024) void gb_hot_sample(IdSource ids, int64_t n, long long K, long long* hot) {
  __shared__ uint32_t skey[HOT_SLOTS];
  __shared__ uint32_t scnt[HOT_SLOTS];
  __shared__ unsigned long long best[32];
  for (uint32_t i = threadIdx.x; i < HOT_SLOTS; i += blockDim.x) {
    skey[i] = 0xFFFFFFFFu;
    scnt[i] = 0u;
  }
  __syncthreads();
  const uint32_t S = uint32_t(std::min<int64_t>(n, HOT_SAMPLE));
  constexpr int IN_FLIGHT = 4;
  for (uint32_t b = threadIdx.x; b < S; b += blockDim.x * IN_FLIGHT) {
    long long g[IN_FLIGHT];
#pragma unroll
    for (int k = 0; k < IN_FLIGHT; k++) {
      const uint32_t t = b + uint32_t(k) * blockDim.x;
      const int64_t row = t < S ? int64_t(uint64_t(t) * uint64_t(n) / S) : 0;   // n < 2^33, S <= 2^16
      g[k] = t < S ? load_gid(ids, row, K) : -1;
    }
#pragma unroll
    for (int k = 0; k < IN_FLIGHT; k++) {
      if (g[k] < 0) continue;
      uint32_t h = uint32_t(g[k]) * 2654435761u;
      uint32_t sl = (h ^ (h >> 15)) & (HOT_SLOTS - 1);
      for (uint32_t probe = 0; probe < 64; probe++) {   // a full table drops the row from the sample
        uint32_t c = skey[sl];
        if (c == 0xFFFFFFFFu) {
          const uint32_t prev = atomicCAS(&skey[sl], 0xFFFFFFFFu, uint32_t(g[k]));
          c = prev == 0xFFFFFFFFu ? uint32_t(g[k]) : prev;
        }
        if (c == uint32_t(g[k])) {
          atomicAdd(&scnt[sl], 1u);
          break;
        }
        sl = (sl + 1u) & (HOT_SLOTS - 1);
      }
    }
  }
  __syncthreads();
  const uint32_t floor = std::max<uint32_t>(S / 64u, 2u);
  for (int j = 0; j < HOT_GROUPS; j++) {
    // (count << 32 | ~slot): the largest count, lowest slot on ties.
    unsigned long long mine = 0ull;
    for (uint32_t i = threadIdx.x; i < HOT_SLOTS; i += blockDim.x) {
      const unsigned long long c = (unsigned long long)scnt[i] << 32 | (0xFFFFFFFFull - i);
      mine = c > mine ? c : mine;
    }
    for (int off = 32; off >= 1; off >>= 1) {
      const unsigned long long o = __shfl_xor(mine, off, 64);
      mine = o > mine ? o : mine;
    }
    if ((threadIdx.x & 63) == 0) best[threadIdx.x >> 6] = mine;
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long m = 0ull;
      for (uint32_t w = 0; w < blockDim.x / 64; w++) m = best[w] > m ? best[w] : m;
      const uint32_t slot = 0xFFFFFFFFu - uint32_t(m);
      const uint32_t cnt = uint32_t(m >> 32);
      hot[j] = cnt >= floor ? (long long)skey[slot] : -1;
      if (cnt > 0u) scnt[slot] = 0u;
    }
    __syncthreads();
  }
}

template <class P>
__global__ void gb_agg_init(unsigned long long* gtab, long long K) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < K * P::W; i += stride) gtab[i] = identity(P::op(int(i % P::W)));
}

// Folds the eight per-XCD tables [8][K][W] into one [K][W] table.
template <class P>
__global__ void gb_agg_merge(const unsigned long long* gtab8, long long K, unsigned long long* out) {
  long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= K) return;
#pragma unroll
  for (int w = 0; w < P::W; w++) {
    const int op = P::op(w);
    if (op == W_HI) continue;
    unsigned long long acc = identity(op), hi = 0;
    for (int x = 0; x < NXCD; x++) {
      const unsigned long long* t = gtab8 + (size_t(x) * size_t(K) + size_t(g)) * P::W;
      unsigned long long v = t[w];
      if (op == W_ADD) acc += v;
      else if (op == W_MIN) acc = v < acc ? v : acc;
      else if (op == W_MAX) acc = v > acc ? v : acc;
      else if (op == W_OR) acc |= v;
      else {   // W_ADD128: lo in word w, hi in word w + 1
        unsigned long long lo2 = acc + v;
        hi += t[w + 1 < P::W ? w + 1 : w] + (lo2 < acc ? 1ull : 0ull);
        acc = lo2;
      }
    }
    out[size_t(g) * P::W + w] = acc;
    if (op == W_ADD128 && w + 1 < P::W) out[size_t(g) * P::W + w + 1] = hi;
  }
}

template <class P, int IDW, int VW>
void launch_agg(int blocks, size_t shmem, hipStream_t s, const AggArgs& a, int mode, uint32_t cap, int64_t rpb,
                unsigned long long* g8) {
  if (shmem > 65536)
    AM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(gb_agg<P, IDW, VW>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, int(shmem)));
  hipLaunchKernelGGL((gb_agg<P, IDW, VW>), dim3(blocks), dim3(AM_AGG_THREADS), shmem, s, a, mode, cap, rpb, g8);
}

// The instantiation for the values' width: none for families that read no value, the fixed width of
// the float64 families, otherwise 1, 2, 4 or 8 bytes.
template <class P, int IDW>
void launch_agg_vw(int vw, int blocks, size_t shmem, hipStream_t s, const AggArgs& a, int mode, uint32_t cap,
                   int64_t rpb, unsigned long long* g8) {
  if constexpr (!P::USES_VALUE) {
    launch_agg<P, IDW, 0>(blocks, shmem, s, a, mode, cap, rpb, g8);
  } else if constexpr (P::FIXED_VW != 0) {
    launch_agg<P, IDW, P::FIXED_VW>(blocks, shmem, s, a, mode, cap, rpb, g8);
  } else {
    switch (vw) {
      case 8: launch_agg<P, IDW, 8>(blocks, shmem, s, a, mode, cap, rpb, g8); break;
      case 4: launch_agg<P, IDW, 4>(blocks, shmem, s, a, mode, cap, rpb, g8); break;
      case 2: launch_agg<P, IDW, 2>(blocks, shmem, s, a, mode, cap, rpb, g8); break;
      default: launch_agg<P, IDW, 1>(blocks, shmem, s, a, mode, cap, rpb, g8); break;
    }
  }
}

#ifndef AM_AGG_LDS
#define AM_AGG_LDS (144 * 1024)
#endif
// The groups whose accumulators all fit in one workgroup's LDS (the private-table mode of gb_agg).
template <class P>
uint32_t agg_dense_cap() {
  return uint32_t(std::min<size_t>(AM_AGG_LDS, device().lds_per_block) / (P::W * 8));
}

// Runs family P over the rows into `o` (K x W words), with `g8` (NXCD x K x W words) for the XCDs'
// partial tables; allocates nothing.
template <class P>
void run_agg_into(const AggArgs& a, hipStream_t s, const char* name, unsigned long long* g8, unsigned long long* o) {
  const long long K = std::max<long long>(a.K, 1);
  if (a.n == 0) {
    hipLaunchKernelGGL(gb_agg_init<P>, dim3(grid_for(K * P::W, 4)), dim3(256), 0, s, o, K);
    AM_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(gb_agg_init<P>, dim3(grid_for(NXCD * K * P::W, 4)), dim3(256), 0, s, g8, NXCD * K);
  AM_HIP(hipGetLastError());
  // LDS budget: a private table when the groups fit, otherwise a cache of that many slots.
  const size_t budget = std::min<size_t>(AM_AGG_LDS, device().lds_per_block);
  const uint32_t per_slot = P::W * 8;
  const uint32_t dense_cap = agg_dense_cap<P>();
  int mode;
  uint32_t cap;
  if (a.K <= (long long)dense_cap) {
    mode = 0;
    cap = uint32_t(a.K);
  } else {
    mode = 1;
    cap = uint32_t(budget / (per_slot + 4));
  }
  size_t shmem = size_t(cap) * per_slot + (mode == 1 ? size_t(cap) * 4 : 0);
  int blocks = int(std::min<int64_t>((a.n + 4095) / 4096, int64_t(device().cu_count) * 4 * 256 / AM_AGG_THREADS));
  blocks = std::max(blocks, 1);
  int64_t rpb = (a.n + blocks - 1) / blocks;
  {
    ProfScope ps(name);
    const int idw = a.ids.type == T_I64 ? 8 : 4;
    const int vw = !P::USES_VALUE ? 0 : (P::FIXED_VW ? P::FIXED_VW : type_width(a.v.type));
    if (idw == 8) launch_agg_vw<P, 8>(vw, blocks, shmem, s, a, mode, cap, rpb, g8);
    else launch_agg_vw<P, 4>(vw, blocks, shmem, s, a, mode, cap, rpb, g8);
    AM_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(gb_agg_merge<P>, dim3(uint32_t((K + 255) / 256)), dim3(256), 0, s, g8, K, o);
  AM_HIP(hipGetLastError());
}

// ---- partitioned aggregation: many groups and no global atomics
//
// A global atomic is carried out at the memory side here, not in L2, at about 23 G operations/s over the
// whole chip, so an LDS cache in front of a global table spends most of its time on the rows it cannot
// cache. Instead, with more groups than LDS holds, the rows are partitioned by group-id bucket (PART_G
// consecutive ids) and each bucket is reduced in LDS:
//   gb_part_hist     rows that contribute, per tile and bucket (bucket-major counts)
//   gb_part_scan     per bucket, those counts -> each tile's offset in the bucket, and its total
//   gb_part_bases    the buckets' bases and each bucket's first reduce workgroup
//   gb_part_scatter  every contributing row's record (group id, raw value) to its bucket's run, staged
//                    in LDS by bucket so that each run is written in consecutive addresses
//   gb_part_reduce   one chunk of one bucket's records -> a partial table of PART_G groups (plain stores)
//   gb_part_merge    each group's partials over its bucket's chunks -> the result
// Every pass streams; the only atomics are LDS ones.
constexpr int PART_LOG_G = 12;
constexpr uint32_t PART_G = 1u << PART_LOG_G;      // groups per bucket (x W words of LDS in the reduce)
constexpr int PART_ROWS = 16;                      // rows per thread of a 256-thread tile
constexpr uint32_t PART_TILE = 256u * PART_ROWS;   // 4096 rows
constexpr uint32_t PART_MAX_NB = 512;              // buckets, so at most 2^21 groups
constexpr uint32_t PART_CHUNK = 65536;             // records per reduce workgroup

// The id of row i when it contributes a record (a valid id in [0, K), and a non-null value for the
// families that skip nulls), else -1. The same test in the count and the scatter.
template <class P, int IDW>
__device__ __forceinline__ long long part_gid(const AggArgs& a, int64_t i, int64_t n, unsigned long long idr,
                                              uint32_t ib, uint32_t vb) {
  long long g;
  if constexpr (IDW == 8) g = (long long)idr;
  else g = a.ids.type == T_I32 ? (long long)int32_t(uint32_t(idr)) : (long long)uint32_t(idr);
  g += a.ids.bias;
  const uint32_t ok = uint32_t(i < n) & vbit(a.ids.voff, i, ib) & (P::SKIP_NULL ? vbit(a.v.voff, i, vb) : 1u);
  return (ok && g >= 0 && g < a.K) ? g : -1;
}

// Counts row `valid` in bucket b of s_cnt and returns its rank there. The lanes of the wave whose bucket is
// the first lane's (most of them, where one bucket is hot) take one LDS atomic together; the others
// take one each.
__device__ __forceinline__ uint32_t part_rank(uint32_t* s_cnt, bool valid, uint32_t b) {
  const int lane = threadIdx.x & 63;
  const uint32_t lead = __builtin_amdgcn_readfirstlane(valid ? b : 0xFFFFFFFFu);
  const unsigned long long m = __ballot(valid && b == lead);
  uint32_t r0 = 0;
  if (m) {   // uniform
    const int leader = __builtin_ctzll(m);
    if (lane == leader) r0 = atomicAdd(&s_cnt[lead], uint32_t(__popcll(m)));
    r0 = __shfl(r0, leader, 64);
  }
  if (valid && b == lead) return r0 + uint32_t(__popcll(m & ((1ull << lane) - 1ull)));
  return valid ? atomicAdd(&s_cnt[b], 1u) : 0u;
}

#ifndef AM_HIST_THREADS
#define AM_HIST_THREADS 256
#endif
constexpr uint32_t HIST_THREADS = AM_HIST_THREADS;          // threads per workgroup of the histogram
constexpr int HIST_ROWS = int(PART_TILE / HIST_THREADS);    // its rows per thread (one tile per workgroup)

template <class P, int IDW>
__global__ __launch_bounds__(HIST_THREADS) void gb_part_hist(AggArgs a, uint32_t nb, uint32_t ntiles, uint32_t* hist) {
  __shared__ uint32_t s_cnt[PART_MAX_NB];
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) s_cnt[b] = 0u;
  __syncthreads();
  const int64_t t0 = int64_t(blockIdx.x) * PART_TILE;
  unsigned long long idr[HIST_ROWS];
  uint32_t ib[HIST_ROWS], vb[HIST_ROWS];
#pragma unroll
  for (int k = 0; k < HIST_ROWS; k++) {   // every load first; rows past the end read the last row
    const int64_t i = t0 + int64_t(k) * blockDim.x + threadIdx.x;
    const int64_t ii = i < a.n ? i : a.n - 1;
    idr[k] = ld_raw<IDW>(a.ids.ids, ii);
    ib[k] = ld_vbyte(a.ids.valid, a.ids.voff, ii);
    vb[k] = P::SKIP_NULL ? ld_vbyte(a.v.valid, a.v.voff, ii) : 0xFFu;
  }
#pragma unroll
  for (int k = 0; k < HIST_ROWS; k++) {
    const int64_t i = t0 + int64_t(k) * blockDim.x + threadIdx.x;
    const long long g = part_gid<P, IDW>(a, i, a.n, idr[k], ib[k], vb[k]);
    (void)part_rank(s_cnt, g >= 0, uint32_t(g) >> PART_LOG_G);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) hist[size_t(b) * ntiles + blockIdx.x] = s_cnt[b];
}

// The exclusive scan of 1024 values held one per thread (s: 1024 words of LDS); returns this thread's
// exclusive prefix and sets `total`.
__device__ __forceinline__ uint32_t block_scan_1024(uint32_t x, uint32_t* s, uint32_t& total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) s[w] = incl;
  __syncthreads();
  uint32_t wb = 0, all = 0;
  for (int j = 0; j < 16; j++) {
    const uint32_t v = s[j];
    wb += j < w ? v : 0u;
    all += v;
  }
  __syncthreads();
  total = all;
  return wb + incl - x;
}

// Workgroup b: the exclusive scan of bucket b's tile counts in place (hist[b][t] becomes the first
// record of tile t within bucket b) and the bucket's total. Four consecutive tiles per thread per step.
__global__ __launch_bounds__(1024) void gb_part_scan(uint32_t* hist, uint32_t ntiles, uint32_t* total) {
  __shared__ uint32_t s[16];
  uint32_t* h = hist + size_t(blockIdx.x) * ntiles;
  uint32_t carry = 0;
  for (uint32_t t0 = 0; t0 < ntiles; t0 += 4u * blockDim.x) {
    uint32_t v[4], x = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t t = t0 + 4u * threadIdx.x + uint32_t(k);
      v[k] = t < ntiles ? h[t] : 0u;
      x += v[k];
    }
    uint32_t step_total;
    uint32_t run = carry + block_scan_1024(x, s, step_total);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t t = t0 + 4u * threadIdx.x + uint32_t(k);
      if (t < ntiles) h[t] = run;
      run += v[k];
    }
    carry += step_total;
  }
  if (threadIdx.x == 0) total[blockIdx.x] = carry;
}

// One workgroup over the buckets (nb <= 512 < 1024): base[b] = the records of the buckets before b
// (base[nb] = all records), first[b] = the reduce workgroups (chunks) of the buckets before b.
__global__ __launch_bounds__(1024) void gb_part_bases(const uint32_t* total, uint32_t nb, uint32_t* base, uint32_t* first) {
  __shared__ uint32_t s[16];
  const uint32_t b = threadIdx.x;
  const uint32_t cnt = b < nb ? total[b] : 0u;
  uint32_t all;
  const uint32_t rb = block_scan_1024(cnt, s, all);
  const uint32_t chunks = (cnt + PART_CHUNK - 1) / PART_CHUNK;
  uint32_t all_chunks;
  const uint32_t cb = block_scan_1024(chunks, s, all_chunks);
  if (b <= nb) {
    base[b] = b < nb ? rb : all;
    first[b] = b < nb ? cb : all_chunks;
  }
}

// Each contributing row's record: rec_gid[r] = its group id, rec_val[r] = its raw value bits (none for
// families that read no value). Within a tile the records are ranked per bucket in LDS (their order
// does not matter to any family here) and staged in bucket order, then written out in order: the run
// of bucket b starts at base[b] + off[b][tile].
#ifndef AM_SCAT_THREADS
#define AM_SCAT_THREADS 1024
#endif
constexpr uint32_t SCAT_THREADS = AM_SCAT_THREADS;              // threads per workgroup of the scatter
constexpr int SCAT_ROWS = int(PART_TILE / SCAT_THREADS);        // its rows per thread (one tile per workgroup)
constexpr uint32_t SCAT_PER = (PART_MAX_NB + SCAT_THREADS - 1) / SCAT_THREADS;   // buckets per thread in its scan

template <class P, int IDW, int VW, bool ROWS>
__global__ __launch_bounds__(SCAT_THREADS) void gb_part_scatter(AggArgs a, uint32_t nb, uint32_t ntiles, const uint32_t* off,
                                                       const uint32_t* base, uint32_t* rec_gid,
                                                       unsigned long long* rec_val, uint32_t* rec_row) {
  // s_cnt holds the tile's count per bucket, then (scanned in place) its offset: 52 KiB of LDS in all,
  // so that two 1024-thread workgroups (the CU's 32 waves) fit on a CU.
  __shared__ uint32_t s_cnt[PART_MAX_NB], s_dst[PART_MAX_NB];
  __shared__ uint32_t s_gid[PART_TILE];
  __shared__ unsigned long long s_val[VW ? PART_TILE : 1];
  __shared__ uint32_t s_row[ROWS ? PART_TILE : 1];
  __shared__ uint32_t s_wsum[SCAT_THREADS / 64];
  for (uint32_t b = threadIdx.x; b < nb; b += blockDim.x) {
    s_cnt[b] = 0u;
    s_dst[b] = base[b] + off[size_t(b) * ntiles + blockIdx.x];
  }
  __syncthreads();
  const int64_t t0 = int64_t(blockIdx.x) * PART_TILE;
  unsigned long long idr[SCAT_ROWS], vr[SCAT_ROWS];
  uint32_t ib[SCAT_ROWS], vb[SCAT_ROWS];
#pragma unroll
  for (int k = 0; k < SCAT_ROWS; k++) {
    const int64_t i = t0 + int64_t(k) * blockDim.x + threadIdx.x;
    const int64_t ii = i < a.n ? i : a.n - 1;
    idr[k] = ld_raw<IDW>(a.ids.ids, ii);
    ib[k] = ld_vbyte(a.ids.valid, a.ids.voff, ii);
    vr[k] = 0;
    if constexpr (VW != 0) vr[k] = ld_raw<VW>(a.v.vals, ii);
    vb[k] = P::SKIP_NULL ? ld_vbyte(a.v.valid, a.v.voff, ii) : 0xFFu;
  }
  long long gs[SCAT_ROWS];
  uint32_t rank[SCAT_ROWS];
#pragma unroll
  for (int k = 0; k < SCAT_ROWS; k++) {
    const int64_t i = t0 + int64_t(k) * blockDim.x + threadIdx.x;
    gs[k] = part_gid<P, IDW>(a, i, a.n, idr[k], ib[k], vb[k]);
    rank[k] = part_rank(s_cnt, gs[k] >= 0, uint32_t(gs[k]) >> PART_LOG_G);
  }
  __syncthreads();
  // Exclusive scan of the tile's bucket counts (nb <= 512: SCAT_PER consecutive buckets per thread,
  // then the waves).
  uint32_t c[SCAT_PER];
  uint32_t x = 0;
#pragma unroll
  for (uint32_t q = 0; q < SCAT_PER; q++) {
    const uint32_t bq = SCAT_PER * threadIdx.x + q;
    c[q] = bq < nb ? s_cnt[bq] : 0u;
    x += c[q];
  }
  uint32_t incl = x;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  if (lane == 63) s_wsum[threadIdx.x >> 6] = incl;
  __syncthreads();
  uint32_t wbase = 0, tile_count = 0;
  for (uint32_t w = 0; w < SCAT_THREADS / 64; w++) {
    wbase += w < (threadIdx.x >> 6) ? s_wsum[w] : 0u;
    tile_count += s_wsum[w];
  }
  uint32_t run = wbase + incl - x;   // this thread's own counts were read before the barrier above
#pragma unroll
  for (uint32_t q = 0; q < SCAT_PER; q++) {
    const uint32_t bq = SCAT_PER * threadIdx.x + q;
    if (bq < nb) s_cnt[bq] = run;
    run += c[q];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SCAT_ROWS; k++) {
    if (gs[k] < 0) continue;
    const uint32_t p = s_cnt[uint32_t(gs[k]) >> PART_LOG_G] + rank[k];
    s_gid[p] = uint32_t(gs[k]);
    if constexpr (VW != 0) s_val[p] = vr[k];
    if constexpr (ROWS) s_row[p] = uint32_t(t0 + int64_t(k) * blockDim.x + threadIdx.x);
  }
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < tile_count; p += blockDim.x) {
    const uint32_t g = s_gid[p];
    const uint32_t b = g >> PART_LOG_G;
    const uint32_t dst = s_dst[b] + (p - s_cnt[b]);
    rec_gid[dst] = g;
    if constexpr (VW != 0) rec_val[dst] = s_val[p];
    if constexpr (ROWS) rec_row[dst] = s_row[p];
  }
}

constexpr int PART_RED_ROWS = 4;   // records per thread per step of the reduce
#ifndef AM_RED_THREADS
#define AM_RED_THREADS 1024
#endif
constexpr uint32_t RED_THREADS = AM_RED_THREADS;   // threads per workgroup of the reduce
#ifndef AM_PART_PEEL
#define AM_PART_PEEL 16
#endif
constexpr int PART_PEEL = AM_PART_PEEL;   // lanes on one group from which a wave combines them before the LDS atomic

// One step of wave_combine64: x combined with its DPP-moved copy (lanes without a source see the
// identity), each 64-bit word moved as two halves.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ unsigned long long dpp_combine(unsigned long long x, unsigned long long id, int op) {
  const uint32_t lo = uint32_t(__builtin_amdgcn_update_dpp(int(uint32_t(id)), int(uint32_t(x)), CTRL, ROW_MASK, 0xF, false));
  const uint32_t hi = uint32_t(__builtin_amdgcn_update_dpp(int(uint32_t(id >> 32)), int(uint32_t(x >> 32)), CTRL, ROW_MASK, 0xF, false));
  const unsigned long long y = (unsigned long long)hi << 32 | lo;
  return op == W_ADD ? x + y : (op == W_MIN ? (y < x ? y : x) : (op == W_MAX ? (y > x ? y : x) : (x | y)));
}

// The combination under `op` of x over the wave, in lane 63: an inclusive scan by DPP row shifts and
// row broadcasts (VALU only; no LDS traffic).
__device__ __forceinline__ unsigned long long wave_combine64(unsigned long long x, int op) {
  const unsigned long long id = identity(op);
  x = dpp_combine<0x111, 0xF>(x, id, op);   // row_shr:1
  x = dpp_combine<0x112, 0xF>(x, id, op);   // row_shr:2
  x = dpp_combine<0x114, 0xF>(x, id, op);   // row_shr:4
  x = dpp_combine<0x118, 0xF>(x, id, op);   // row_shr:8
  x = dpp_combine<0x142, 0xA>(x, id, op);   // row_bcast:15 into rows 1 and 3
  x = dpp_combine<0x143, 0xC>(x, id, op);   // row_bcast:31 into rows 2 and 3
  return x;
}

// The records of one reduce step (PART_RED_ROWS per thread): group ids, raw values, and row numbers for
// the families that need them.
template <int VW, bool ROWS>
struct PartStep {
  uint32_t g[PART_RED_ROWS];
  unsigned long long raw[PART_RED_ROWS];
  uint32_t row[PART_RED_ROWS];
};

// The partitioned records of one set of rows.
struct PartIn {
  const uint32_t* gid;
  const unsigned long long* val;
  const uint32_t* row;
};

// Loads the records of the step at r (records past r1 read the last one; r1 > 0).
template <int VW, bool ROWS>
__device__ __forceinline__ void part_load(const PartIn& in, int64_t r, int64_t r1, PartStep<VW, ROWS>& x) {
#pragma unroll
  for (int k = 0; k < PART_RED_ROWS; k++) {
    const int64_t i = r + int64_t(k) * blockDim.x + threadIdx.x;
    const int64_t ii = i < r1 ? i : r1 - 1;
    x.g[k] = load_global(in.gid + ii);
    x.raw[k] = VW ? load_global(in.val + ii) : 0ull;
    x.row[k] = ROWS ? load_global(in.row + ii) : 0u;
  }
}

// Whether every word of family P combines by ADD / MIN / MAX / OR (a 128-bit sum's two words do not
// combine word by word across a wave).
template <class P>
__host__ __device__ constexpr bool wordwise() {
  for (int w = 0; w < P::W; w++)
    if (P::op(w) == W_ADD128 || P::op(w) == W_HI) return false;
  return true;
}

// Applies the step's records at r to the LDS table. A hot group (many lanes of the wave on the first
// lane's group) is combined across the wave first and applied once, instead of one LDS atomic per
// lane on one address.
template <class P, int VW, bool ROWS>
__device__ __forceinline__ void part_apply(const AggArgs& a, int64_t r, int64_t r1, const PartStep<VW, ROWS>& x,
                                           unsigned long long* lds) {
#pragma unroll
  for (int k = 0; k < PART_RED_ROWS; k++) asm volatile("" ::"v"(x.g[k]), "v"(x.raw[k]), "v"(x.row[k]));   // every loaded register taken here
  unsigned long long cs[PART_RED_ROWS][P::W];
  bool act[PART_RED_ROWS];
  if constexpr (P::AUX) {
    typename P::AuxT ax[PART_RED_ROWS];
#pragma unroll
    for (int k = 0; k < PART_RED_ROWS; k++) ax[k] = P::aux_load(a, (long long)x.g[k]);   // all loads first
#pragma unroll
    for (int k = 0; k < PART_RED_ROWS; k++) {
      const int64_t i = r + int64_t(k) * blockDim.x + threadIdx.x;
      act[k] = P::contrib(a, int64_t(x.row[k]), (long long)x.g[k], val_in(a.v, x.raw[k]), ax[k], cs[k]) && i < r1;
    }
  } else {
#pragma unroll
    for (int k = 0; k < PART_RED_ROWS; k++) {
      const int64_t i = r + int64_t(k) * blockDim.x + threadIdx.x;
      act[k] = P::contrib(a, int64_t(x.row[k]), (long long)x.g[k], val_in(a.v, x.raw[k]), cs[k]) && i < r1;
    }
  }
#pragma unroll
  for (int k = 0; k < PART_RED_ROWS; k++) {
    const uint32_t g = x.g[k];
    if constexpr (wordwise<P>()) {
      const uint32_t lead = __builtin_amdgcn_readfirstlane(act[k] ? g : 0xFFFFFFFFu);
      const unsigned long long m = __ballot(act[k] && g == lead);
      if (__popcll(m) >= PART_PEEL) {   // uniform
        unsigned long long v[P::W];
#pragma unroll
        for (int w = 0; w < P::W; w++) {
          const int op = P::op(w);
          v[w] = wave_combine64((act[k] && g == lead) ? cs[k][w] : identity(op), op);
        }
        if ((threadIdx.x & 63) == 63) apply<P>(as_lds(&lds[size_t(lead & (PART_G - 1)) * P::W]), v);
        act[k] = act[k] && g != lead;
      }
    }
    if (act[k]) apply<P>(as_lds(&lds[size_t(g & (PART_G - 1)) * P::W]), cs[k]);
  }
}

// Workgroup c reduces chunk c - first[b] of bucket b into part[c] (PART_G x W words, identities where
// no record fell). Workgroups past the last chunk do nothing.
template <class P, int VW, bool ROWS>
__global__ __launch_bounds__(RED_THREADS) void gb_part_reduce(AggArgs a, uint32_t nb, const uint32_t* base, const uint32_t* first,
                                                      PartIn in, unsigned long long* part) {
  extern __shared__ unsigned long long lds[];
  const uint32_t c = blockIdx.x;
  if (c >= first[nb]) return;   // uniform
  uint32_t lo = 0, hi = nb;     // the bucket b with first[b] <= c < first[b + 1]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (first[mid] <= c) lo = mid; else hi = mid;
  }
  const uint32_t b = lo;
  const int64_t r0 = int64_t(base[b]) + int64_t(c - first[b]) * PART_CHUNK;
  const int64_t r1 = std::min<int64_t>(int64_t(base[b + 1]), r0 + PART_CHUNK);
  for (uint32_t j = threadIdx.x; j < PART_G; j += blockDim.x)
#pragma unroll
    for (int w = 0; w < P::W; w++) lds[size_t(j) * P::W + w] = identity(P::op(w));
  __syncthreads();
  const int64_t step = int64_t(blockDim.x) * PART_RED_ROWS;
  // Two record buffers take turns, as in gb_agg: the next step's loads are in flight while this step's
  // records are applied (LDS only, so nothing here waits on a global write).
  const int steps = __builtin_amdgcn_readfirstlane(int((r1 - r0 + step - 1) / step));
  PartStep<VW, ROWS> ra, rb;
  part_load<VW, ROWS>(in, r0, r1, ra);
  for (int t = 0; t < steps; t += 2) {
    const int64_t r = r0 + int64_t(t) * step;
    part_load<VW, ROWS>(in, r + step, r1, rb);
    part_apply<P, VW, ROWS>(a, r, r1, ra, lds);
    if (t + 1 >= steps) break;
    part_load<VW, ROWS>(in, r + 2 * step, r1, ra);
    part_apply<P, VW, ROWS>(a, r + step, r1, rb, lds);
  }
  __syncthreads();
  unsigned long long* dst = part + size_t(c) * PART_G * P::W;
  for (uint32_t j = threadIdx.x; j < PART_G * P::W; j += blockDim.x) dst[j] = lds[j];
}

// acc <- acc combined with v, word by word under family P's ops; a 128-bit sum (lo word W_ADD128, hi
// word W_HI) adds as one number, the lo words' carry into the hi word.
template <class P>
__device__ __forceinline__ void combine_words(unsigned long long (&acc)[P::W], const unsigned long long (&v)[P::W]) {
#pragma unroll
  for (int w = 0; w < P::W; w++) {
    const int op = P::op(w);
    if (op == W_ADD) acc[w] += v[w];
    else if (op == W_MIN) acc[w] = v[w] < acc[w] ? v[w] : acc[w];
    else if (op == W_MAX) acc[w] = v[w] > acc[w] ? v[w] : acc[w];
    else if (op == W_OR) acc[w] |= v[w];
    else if (op == W_ADD128) {
      const unsigned long long lo = acc[w] + v[w];
      const unsigned long long carry = lo < acc[w] ? 1ull : 0ull;
      acc[w] = lo;
      if (w + 1 < P::W) acc[w + 1] += v[w + 1] + carry;
    }
    // W_HI: added with its W_ADD128 word
  }
}

// out[g] = the combination of part[c][g mod PART_G] over the chunks c of g's bucket. A hot bucket has
// hundreds of chunks, so a workgroup of 16 waves covers 64 groups: wave w combines the chunks
// c0 + w, c0 + w + 16, ... (each load of a wave reads 64 consecutive groups of one chunk, several
// chunks in flight), then the 16 wave results of each group are combined through LDS.
constexpr int PART_MERGE_LOADS = 4;
constexpr int PART_MERGE_WAVES = 16;
template <class P>
__global__ __launch_bounds__(1024) void gb_part_merge(const unsigned long long* part, const uint32_t* first, long long K,
                                                      unsigned long long* out) {
  __shared__ unsigned long long s_acc[PART_MERGE_WAVES][64][P::W];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long g = (long long)blockIdx.x * 64 + lane;
  const long long gg = g < K ? g : K - 1;
  const uint32_t b = uint32_t(gg >> PART_LOG_G), j = uint32_t(gg & (PART_G - 1));
  const uint32_t c0 = first[b], c1 = first[b + 1];   // the same for the 64 groups unless they straddle buckets
  unsigned long long acc[P::W];
#pragma unroll
  for (int w = 0; w < P::W; w++) acc[w] = identity(P::op(w));
  for (uint32_t c = c0 + uint32_t(wave); c < c1; c += PART_MERGE_WAVES * PART_MERGE_LOADS) {
    unsigned long long v[PART_MERGE_LOADS][P::W];
#pragma unroll
    for (int u = 0; u < PART_MERGE_LOADS; u++) {   // every load first; chunks past c1 read chunk c
      const uint32_t cu = c + uint32_t(u * PART_MERGE_WAVES);
      const uint32_t cc = cu < c1 ? cu : c;
#pragma unroll
      for (int w = 0; w < P::W; w++) v[u][w] = part[(size_t(cc) * PART_G + j) * P::W + w];
    }
#pragma unroll
    for (int u = 0; u < PART_MERGE_LOADS; u++) {
      if (c + uint32_t(u * PART_MERGE_WAVES) >= c1) break;
      combine_words<P>(acc, v[u]);
    }
  }
#pragma unroll
  for (int w = 0; w < P::W; w++) s_acc[wave][lane][w] = acc[w];
  __syncthreads();
  if (wave == 0 && g < K) {
    for (int x = 1; x < PART_MERGE_WAVES; x++) {
      unsigned long long y[P::W];
#pragma unroll
      for (int w = 0; w < P::W; w++) y[w] = s_acc[x][lane][w];
      combine_words<P>(acc, y);
    }
#pragma unroll
    for (int w = 0; w < P::W; w++) out[size_t(g) * P::W + w] = acc[w];
  }
}

// The partitioned records of one set of rows (the rows with a valid id, and a non-null value for the
// families that skip nulls): kept by a GroupBy for its last value column, so that the aggregates of
// one column (min and max, then the exact sum's two passes) share one count and scatter.
struct PartSet {
  uint64_t col = 0;          // the value column's unique id (0: none)
  bool skip_null = false;    // records of rows with a non-null value only
  bool rows = false;         // row numbers kept (rec_row)
  int vw = 0;                // 0: no values kept
  uint32_t nb = 0, ntiles = 0, chunk_bound = 0;
  DevBuf meta, gid, val, row;   // meta: base[nb + 1], first[nb + 1]
};

// Family P over many groups through the partitioned passes into `o` (K x W words); false, with
// nothing launched, where they do not apply (too many groups, too few rows). `cache` (or null) holds
// the records of an earlier family over the same column, used again when they fit; `keep_rows` keeps
// row numbers in new records even if P does not need them (a float column, whose exact sum will).
template <class P>
bool run_agg_part(const AggArgs& a, hipStream_t s, const char* name, unsigned long long* o, PartSet* cache,
                  uint64_t col, bool keep_rows) {
  if constexpr (!P::PART) {
    return false;
  } else {
    const long long K = a.K;
    const uint32_t nb = uint32_t((K + PART_G - 1) / PART_G);
    if (a.n < int64_t(PART_TILE) * 16 || nb == 0 || nb > PART_MAX_NB || a.n >= (int64_t(1) << 32)) return false;
    const int vw = !P::USES_VALUE ? 0 : type_width(a.v.type);
    const bool reuse = cache && col && cache->col == col && cache->skip_null == P::SKIP_NULL && cache->nb == nb &&
                       (cache->vw != 0 || vw == 0) && (cache->rows || !P::NEEDS_ROW) && a.v.xop == VX_NONE;
    PartSet fresh;
    PartSet& ps_ = reuse ? *cache : fresh;
    if (!reuse) {
      ps_.col = col;
      ps_.skip_null = P::SKIP_NULL;
      ps_.rows = P::NEEDS_ROW || keep_rows;
      ps_.vw = vw;
      ps_.nb = nb;
      ps_.ntiles = uint32_t((a.n + PART_TILE - 1) / PART_TILE);
      ps_.chunk_bound = uint32_t((a.n + PART_CHUNK - 1) / PART_CHUNK) + nb;
      const uint32_t ntiles = ps_.ntiles;
      const int idw = a.ids.type == T_I64 ? 8 : 4;
      DevBuf hist = device_alloc(size_t(nb) * ntiles * 4);
      ps_.meta = device_alloc(size_t(3) * (nb + 1) * 4);
      uint32_t* base = static_cast<uint32_t*>(ps_.meta->ptr);
      uint32_t* first = base + nb + 1;
      uint32_t* total = first + nb + 1;
      ps_.gid = device_alloc(size_t(a.n) * 4);
      ps_.val = device_alloc(vw ? size_t(a.n) * 8 : 8);
      ps_.row = device_alloc(ps_.rows ? size_t(a.n) * 4 : 4);
      uint32_t* h = static_cast<uint32_t*>(hist->ptr);
      uint32_t* rg = static_cast<uint32_t*>(ps_.gid->ptr);
      unsigned long long* rv = static_cast<unsigned long long*>(ps_.val->ptr);
      uint32_t* rr = static_cast<uint32_t*>(ps_.row->ptr);
      {
        ProfScope ps("gb_part_hist");
        if (idw == 8) hipLaunchKernelGGL((gb_part_hist<P, 8>), dim3(ntiles), dim3(HIST_THREADS), 0, s, a, nb, ntiles, h);
        else hipLaunchKernelGGL((gb_part_hist<P, 4>), dim3(ntiles), dim3(HIST_THREADS), 0, s, a, nb, ntiles, h);
        AM_HIP(hipGetLastError());
      }
      {
        ProfScope ps("gb_part_scan");
        hipLaunchKernelGGL(gb_part_scan, dim3(nb), dim3(1024), 0, s, h, ntiles, total);
        AM_HIP(hipGetLastError());
        hipLaunchKernelGGL(gb_part_bases, dim3(1), dim3(1024), 0, s, total, nb, base, first);
        AM_HIP(hipGetLastError());
      }
      {
        ProfScope ps("gb_part_scatter");
#define AM_PART_SCATTER(IW, VV)                                                                                  \
  do {                                                                                                           \
    if (ps_.rows)                                                                                                \
      hipLaunchKernelGGL((gb_part_scatter<P, IW, VV, true>), dim3(ntiles), dim3(SCAT_THREADS), 0, s, a, nb, ntiles, h, \
                         base, rg, rv, rr);                                                                      \
    else                                                                                                         \
      hipLaunchKernelGGL((gb_part_scatter<P, IW, VV, false>), dim3(ntiles), dim3(SCAT_THREADS), 0, s, a, nb, ntiles, h, \
                         base, rg, rv, rr);                                                                      \
  } while (0)
        if (idw == 8) {
          switch (vw) {
            case 0: AM_PART_SCATTER(8, 0); break;
            case 8: AM_PART_SCATTER(8, 8); break;
            case 4: AM_PART_SCATTER(8, 4); break;
            case 2: AM_PART_SCATTER(8, 2); break;
            default: AM_PART_SCATTER(8, 1); break;
          }
        } else {
          switch (vw) {
            case 0: AM_PART_SCATTER(4, 0); break;
            case 8: AM_PART_SCATTER(4, 8); break;
            case 4: AM_PART_SCATTER(4, 4); break;
            case 2: AM_PART_SCATTER(4, 2); break;
            default: AM_PART_SCATTER(4, 1); break;
          }
        }
#undef AM_PART_SCATTER
        AM_HIP(hipGetLastError());
      }
      if (cache && col) *cache = ps_;
    }
    uint32_t* base = static_cast<uint32_t*>(ps_.meta->ptr);
    uint32_t* first = base + nb + 1;
    const PartIn in{static_cast<const uint32_t*>(ps_.gid->ptr), static_cast<const unsigned long long*>(ps_.val->ptr),
                    static_cast<const uint32_t*>(ps_.row->ptr)};
    DevBuf part = device_alloc(size_t(ps_.chunk_bound) * PART_G * P::W * 8);
    {
      ProfScope ps(name);
      const size_t shmem = size_t(PART_G) * P::W * 8;
      auto launch = [&](auto kern) {
        if (shmem > 65536)
          AM_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                     int(shmem)));
        hipLaunchKernelGGL(kern, dim3(ps_.chunk_bound), dim3(RED_THREADS), shmem, s, a, nb, base, first, in,
                           static_cast<unsigned long long*>(part->ptr));
      };
      if (vw) {
        if constexpr (P::NEEDS_ROW) launch(gb_part_reduce<P, 8, true>);
        else launch(gb_part_reduce<P, 8, false>);
      } else {
        launch(gb_part_reduce<P, 0, false>);
      }
      AM_HIP(hipGetLastError());
    }
    {
      ProfScope ps("gb_part_merge");
      hipLaunchKernelGGL(gb_part_merge<P>, dim3(uint32_t((K + 63) / 64)), dim3(64 * PART_MERGE_WAVES), 0, s,
                         static_cast<const unsigned long long*>(part->ptr), first, K, o);
      AM_HIP(hipGetLastError());
    }
    return true;
  }
}

// Runs family P over the rows into a fresh global table of K x W words.
template <class P>
DevBuf run_agg(const AggArgs& a, hipStream_t s, const char* name, PartSet* cache = nullptr, uint64_t col = 0,
               bool keep_rows = false) {
  const long long K = std::max<long long>(a.K, 1);
  DevBuf out = device_alloc(size_t(K) * P::W * 8);
  if (a.K > (long long)agg_dense_cap<P>() &&
      run_agg_part<P>(a, s, name, static_cast<unsigned long long*>(out->ptr), cache, col, keep_rows))
    return out;
  DevBuf tabs;
  if (a.n > 0) tabs = device_alloc(size_t(NXCD) * size_t(K) * P::W * 8);
  run_agg_into<P>(a, s, name, tabs ? static_cast<unsigned long long*>(tabs->ptr) : nullptr,
                  static_cast<unsigned long long*>(out->ptr));
  return out;
}

// ---------------------------------------------------------------- finalize kernels

// Validity bytes -> Arrow bitmap (one ballot per 64 groups) and the null count.
__global__ void pack_validity(const uint8_t* vb, long long K, unsigned long long* bits, unsigned long long* nulls) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i - lane < K; i += stride) {
    bool v = i < K && vb[i];
    unsigned long long bal = __ballot(v);
    if (lane == 0) {
      long long rows = std::min<long long>(64, K - (i - lane));
      unsigned long long mask = rows >= 64 ? ~0ull : ((1ull << rows) - 1ull);
      bits[(i - lane) >> 6] = bal & mask;
      unsigned long long z = (unsigned long long)rows - __popcll(bal & mask);
      if (z) atomicAdd(nulls, z);
    }
  }
}

// kind: 0 sum (int64 / uint64 bits), 1 integer mean (signed), 2 integer mean (unsigned),
// 3 count, 4 min, 5 max (elements written at `width` bytes, from the unkeyed order key).
// Finalize kinds: 0 integer sum, 1 signed mean, 2 unsigned mean, 3 count (word 0), 4 min, 5 max, and
// the count of non-null values kept in word 1 of a sum table.
constexpr int KIND_SUM_COUNT = 6;

__global__ void gb_finalize_simple(const unsigned long long* gtab, int W, long long K, int kind, int vtype,
                                   void* out, uint8_t* vb) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const unsigned long long* t = gtab + k * W;
  if (kind == 3 || kind == KIND_SUM_COUNT) {   // a count: word 0 of a count table, word 1 of a sum table
    static_cast<unsigned long long*>(out)[k] = kind == 3 ? t[0] : t[1];
    vb[k] = 1;
    return;
  }
  if (kind <= 2) {
    unsigned long long s = t[0], c = t[1];
    vb[k] = c != 0;
    if (kind == 0) {
      static_cast<unsigned long long*>(out)[k] = c ? s : 0ull;
    } else {
      double m = 0.0;
      if (c) m = (kind == 1 ? double((long long)s) : double(s)) / double(c);
      static_cast<double*>(out)[k] = m;
    }
    return;
  }
  unsigned long long mn = t[0], mx = t[1];
  bool has = mn <= mx;
  vb[k] = has;
  unsigned long long key = kind == 4 ? mn : mx, v = 0;
  if (has) {
    switch (vtype) {
      case T_F64: v = (key & 0x8000000000000000ull) ? (key & 0x7FFFFFFFFFFFFFFFull) : ~key; break;
      case T_F32: {
        uint32_t u = uint32_t(key);
        v = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : uint32_t(~u);
        break;
      }
      case T_I8: case T_I16: case T_I32: case T_I64: v = key ^ 0x8000000000000000ull; break;
      default: v = key; break;
    }
  }
  switch (type_width(vtype)) {
    case 1: static_cast<uint8_t*>(out)[k] = uint8_t(v); break;
    case 2: static_cast<uint16_t*>(out)[k] = uint16_t(v); break;
    case 4: static_cast<uint32_t*>(out)[k] = uint32_t(v); break;
    default: static_cast<unsigned long long*>(out)[k] = v; break;
  }
}

// ---------------------------------------------------------------- exact float sums: grids and results

// One thread per group: the fixed-point grid of the group (see FsScanP / FsFixP).
__global__ void fs_window(const unsigned long long* scan, long long K, uint32_t* grid) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const unsigned long long* t = scan + k * 4;
  const uint32_t emax = uint32_t(t[0]);
  const unsigned long long cnt = t[2];
  if (emax == 0u || (t[1] & (FS_POS_INF | FS_NEG_INF | FS_NAN))) {
    grid[k] = FS_SKIP;
    return;
  }
  grid[k] = uint32_t(fs::grid_lsb(emax, cnt));
}

// A result pattern of the output format: binary64 as it is, binary32 widened to binary64 bits.
template <int P>
__device__ __forceinline__ unsigned long long fs_out(unsigned long long bits) {
  if constexpr (P == 24) return fs::widen_f32(uint32_t(bits));
  else return bits;
}

// The result of a group whose sum does not come from its finite values (FS_SKIP grid), by the
// reference's special-value rules (GroupSumExact.swift:24-60): any NaN gives the first NaN in row
// order, quieted; +inf and -inf together the default NaN; one infinity that infinity; only zeros
// -0.0 when every zero is -0.0, else +0.0. Patterns in the output format P.
template <int P, int VW>
__device__ unsigned long long fs_special(unsigned long long classes, unsigned long long nan_row, const void* vals) {
  constexpr int SH = P == 53 ? 52 : 23;
  constexpr unsigned long long EXP = P == 53 ? 0x7FFull : 0xFFull;
  constexpr unsigned long long SIGN = 1ull << (P == 53 ? 63 : 31);
  if (classes & FS_NAN) {
    unsigned long long b;
    if constexpr (VW == 8) b = static_cast<const unsigned long long*>(vals)[nan_row];
    else b = static_cast<const uint32_t*>(vals)[nan_row];
    if constexpr (P == 53 && VW == 4) b = fs::widen_f32(uint32_t(b));
    return b | (1ull << (SH - 1));            // the quiet bit
  }
  const unsigned long long inf = EXP << SH;
  switch (classes & (FS_POS_INF | FS_NEG_INF)) {
    case FS_POS_INF | FS_NEG_INF: return inf | (1ull << (SH - 1));
    case FS_POS_INF: return inf;
    case FS_NEG_INF: return SIGN | inf;
    default: return (classes & FS_POS_ZERO) ? 0ull : SIGN;
  }
}

// One thread per group: sum and mean (binary64 bits; a binary32 result is widened), their validity,
// and a slot in the list of groups to sum again at full width when the grid cannot decide.
template <int P, int EMIN, int VW>
__global__ void fs_settle(const unsigned long long* scan, const unsigned long long* fix, const uint32_t* grid,
                          long long K, const void* vals, unsigned long long* out_sum, unsigned long long* out_mean,
                          uint8_t* vb, uint32_t* redo, uint32_t* redo_count) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  const unsigned long long* t = scan + k * 4;
  const unsigned long long cnt = t[2];
  redo[k] = FS_SKIP;
  vb[k] = cnt != 0ull;
  if (cnt == 0ull) {
    out_sum[k] = 0ull;
    out_mean[k] = 0ull;
    return;
  }
  const uint32_t lsb = grid[k];
  if (lsb == FS_SKIP) {
    const unsigned long long r = fs_out<P>(fs_special<P, VW>(t[1], t[3], vals));
    out_sum[k] = r;
    out_mean[k] = r;
    return;
  }
  const unsigned long long* f = fix + k * 3;
  unsigned long long s, m;
  if (fs::settle<P, EMIN>(fs::U128{f[0], f[1]}, f[2], int(lsb), cnt, s, m)) {
    out_sum[k] = fs_out<P>(s);
    out_mean[k] = fs_out<P>(m);
  } else {
    redo[k] = atomicAdd(redo_count, 1u);
  }
}

// Full-width accumulation for the groups the grid could not decide: 72 signed 64-bit digits per group,
// digit j collecting the value bits at [32 j, 32 j + 32) above 2^-1074. Each value touches three
// digits; no digit can overflow below 2^31 rows.
constexpr int XD = 72;

template <int VW>
__global__ void fs_redo_rows(AggArgs a, const uint32_t* redo, long long* digits) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  for (; i < a.n; i += stride) {
    long long g = load_gid(a.ids, i, a.K);
    if (g < 0) continue;
    uint32_t slot = redo[g];
    if (slot == FS_SKIP || !val_valid(a.v, i)) continue;
    const unsigned long long v = fs_bits<VW>(ld_raw<VW>(a.v.vals, i));
    uint32_t e = uint32_t((v >> 52) & 0x7FF);
    if (e == 0x7FF || (v & 0x7FFFFFFFFFFFFFFFull) == 0) continue;
    unsigned long long m = (v & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
    uint32_t off = (e ? e : 1u) - 1u;
    uint32_t d0 = off >> 5, sh = off & 31u;
    unsigned long long lo64 = m << sh, hi64 = sh ? (m >> (64 - sh)) : 0ull;
    long long p0 = (long long)(lo64 & 0xFFFFFFFFull), p1 = (long long)(lo64 >> 32), p2 = (long long)hi64;
    if (v >> 63) { p0 = -p0; p1 = -p1; p2 = -p2; }
    long long* D = digits + size_t(slot) * XD + d0;
    if (p0) atomicAdd(reinterpret_cast<unsigned long long*>(&D[0]), (unsigned long long)p0);
    if (p1) atomicAdd(reinterpret_cast<unsigned long long*>(&D[1]), (unsigned long long)p1);
    if (p2) atomicAdd(reinterpret_cast<unsigned long long*>(&D[2]), (unsigned long long)p2);
  }
}

template <int P, int EMIN>
__global__ __launch_bounds__(256) void fs_redo_settle(const uint32_t* redo, long long K, const long long* digits,
                               const unsigned long long* scan, unsigned long long* out_sum,
                               unsigned long long* out_mean) {
  long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K || redo[k] == FS_SKIP) return;
  const long long* D = digits + size_t(redo[k]) * XD;
  // Normalise the carry-save digits into a two's-complement number of 36 64-bit words.
  constexpr int NW = XD / 2;
  unsigned long long w[NW];
  long long carry = 0;
  for (int j = 0; j < XD; j++) {
    // value of digit j plus the carry, split into the low 32 bits and an arithmetic carry
    long long x = D[j] + carry;            // |x| < 2^63 for fewer than 2^31 rows
    uint32_t low = uint32_t(x);
    carry = (x - (long long)low) >> 32;   // exact: x - low is a multiple of 2^32
    if (j & 1) w[j >> 1] |= (unsigned long long)low << 32;
    else w[j >> 1] = low;
  }
  const bool neg = carry < 0;
  if (neg) {
    unsigned long long c = 1;
    for (int j = 0; j < NW; j++) {
      unsigned long long x = ~w[j] + c;
      c = (c && x == 0) ? 1ull : 0ull;
      w[j] = x;
    }
  }
  int lead = -1;
  for (int j = NW - 1; j >= 0 && lead < 0; j--)
    if (w[j]) lead = j * 64 + 63 - __clzll(w[j]);
  if (lead < 0) {   // an exact cancellation: +0.0
    out_sum[k] = 0ull;
    out_mean[k] = 0ull;
    return;
  }
  int low;
  bool tail;
  const fs::U128 top = fs::top128(w, NW, lead, low, tail);
  constexpr unsigned long long SIGN = 1ull << (P == 53 ? 63 : 31);
  const unsigned long long s = neg ? SIGN : 0ull;
  out_sum[k] = fs_out<P>(s | fs::rne<P, EMIN>(top, tail, 0, low - 1074));
  out_mean[k] = fs_out<P>(s | fs::rne_div<P, EMIN>(top, tail, 0, scan[k * 4 + 2], low - 1074));
}

}  // namespace

// ================================================================ host: key mapping

struct GroupBy {
  std::vector<Array*> columns;
  int64_t n = 0;
  int64_t K = 0;
  DevBuf ids;     // int32 [n]
  DevBuf rep;     // uint32 [K], the lowest row of each group (when tracked)
  bool has_rep = false;
  // An integer single-column mapping keeps its sorted distinct keys instead of representative rows.
  DevBuf okeys, perm;
  uint32_t D = 0;
  bool has_null = false;
  std::vector<Array*> key_results;
  // A min and a max of one column share a pass; a sum and a mean of one float64 column share theirs;
  // an integer sum also yields the column's count of non-null values. Keyed by the column's unique
  // id, never its address, which a later array may reuse.
  uint64_t mm_col = 0;
  DevBuf mm_tab;
  uint64_t fx_col = 0;
  DevBuf fx_sum, fx_mean, fx_valid;
  uint64_t cnt_col = 0;
  DevBuf cnt_tab;          // K x 2 integer sum and count of non-null values of column cnt_col
  bool cnt_all_rows = false;   // column cnt_col had no nulls: its counts are the rows per group
  DevBuf count_all_tab;    // K x 1 rows per group
  PartSet part;            // the partitioned records of the last value column (many groups)
  DevBuf hot;              // HOT_GROUPS sampled frequent ids (sums and counts in registers)
  // The mapping is left running when am_group_by_keys returns: `ready` is recorded on `owner` after
  // its last kernel. Calls on another thread's stream wait for it, and the handle's release waits
  // before the key columns can be freed.
  hipEvent_t ready = nullptr;
  hipStream_t owner = nullptr;
  ~GroupBy() {
    if (ready) {
      (void)hipEventSynchronize(ready);
      (void)hipEventDestroy(ready);
    }
    for (Array* a : columns) release_array(a);
    for (Array* a : key_results) release_array(a);
  }
};

// Orders the calling thread's stream after the handle's mapping.
void gb_use(const GroupBy* g) {
  if (g->ready && g->owner != stream()) AM_HIP(hipStreamWaitEvent(stream(), g->ready, 0));
}

namespace {

struct Mapped {
  int64_t K = 0;
  uint32_t D = 0;
  bool has_null = false;
  DevBuf ids, rep, okeys, perm;
};

// ---------------------------------------------------------------- the range path

// ARROWMETAL_GROUPBY_RANGE=0 turns the range path off (the hash path then maps every key column), so
// that tests can compare the two on the same keys.
bool range_path_enabled() {
  const char* e = getenv("ARROWMETAL_GROUPBY_RANGE");
  return !(e && e[0] == '0');
}
//
// Integer keys whose values span a small range (GroupByKeys.swift:268-337: span <= 2^24 and
// <= max(2^16, 4n)) are mapped without hashing: one pass finds the smallest and largest order key, one
// marks every present key in an occupancy array over the span (and its lowest row when representative
// rows are tracked), a scan of the occupancy turns each present key into its rank, which is its id
// in ascending key order, and one pass writes every row's id (the null group, last, is D).

// occ[k - kmin] = 1 for every present key; first[k - kmin] = its lowest row when tracked. Both only
// ever move one way, so a read that already shows the value skips the write. A step's reads are
// issued together (a null row reads entry 0, which exists, and writes nothing).
template <int W, int TYPE, bool HAS_VALID, bool TRACK>
__global__ __launch_bounds__(256) void gb_range_occ(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n,
                                                    unsigned long long kmin, uint32_t* occ, uint32_t* first) {
  const uint32_t gsize = gridDim.x * blockDim.x;
  scan_steps<W, HAS_VALID>(vals, valid, voff, n, [&](uint64_t row0, const ScanRows& x) {
    scan_take<HAS_VALID>(x);
    uint32_t xs[SCAN_ROWS];
    bool ok[SCAN_ROWS];
#pragma unroll
    for (int r = 0; r < SCAN_ROWS; r++) {
      const uint64_t i = row0 + uint64_t(r) * gsize;
      const int64_t ii = int64_t(i < n ? i : uint64_t(n) - 1u);
      ok[r] = i < n && (HAS_VALID ? vbit(voff, ii, x.vb[r]) != 0u : true);
      xs[r] = ok[r] ? uint32_t(order_key(key_of_raw(x.raw[r], TYPE), TYPE) - kmin) : 0u;
    }
    uint32_t cur[SCAN_ROWS], low[SCAN_ROWS];
#pragma unroll
    for (int r = 0; r < SCAN_ROWS; r++) {
      cur[r] = occ[xs[r]];
      if (TRACK) low[r] = first[xs[r]];
    }
    // Plain accesses: a stale 0 costs one more store of the same 1; a stale row is larger than the
    // current lowest one, so it never skips a needed atomicMin.
#pragma unroll
    for (int r = 0; r < SCAN_ROWS; r++) {
      if (ok[r] && cur[r] == 0u) occ[xs[r]] = 1u;
      const uint32_t i = uint32_t(row0 + uint64_t(r) * gsize);
      if (TRACK && ok[r] && i < low[r]) atomicMin(&first[xs[r]], i);
    }
  });
}

// The occupancy of a span of at most RANGE_BITS_MAX keys without representative rows: each workgroup
// marks its rows' keys in a bitmap of the span in LDS, then ORs its words into the global bitmap (a word
// already holding the bits is only read). gb_range_unpack turns the bitmap into occ[].
constexpr uint32_t RANGE_BITS_MAX = 1u << 19;   // 64 KiB of LDS

template <int W, int TYPE, bool HAS_VALID>
__global__ __launch_bounds__(256) void gb_range_occ_bits(const void* vals, const uint8_t* valid, int64_t voff,
                                                         uint32_t n, unsigned long long kmin, uint32_t words,
                                                         uint32_t* bits) {
  extern __shared__ uint32_t s_bits[];
  for (uint32_t j = threadIdx.x; j < words; j += blockDim.x) s_bits[j] = 0u;
  __syncthreads();
  const uint32_t gsize = gridDim.x * blockDim.x;
  scan_steps<W, HAS_VALID>(vals, valid, voff, n, [&](uint64_t row0, const ScanRows& x) {
    scan_take<HAS_VALID>(x);
    // The step's words are read together; a bit already set (a frequent key) needs no atomic.
    uint32_t wd[SCAN_ROWS], bit[SCAN_ROWS], cur[SCAN_ROWS];
#pragma unroll
    for (int r = 0; r < SCAN_ROWS; r++) {
      const uint64_t i = row0 + uint64_t(r) * gsize;
      const int64_t ii = int64_t(i < n ? i : uint64_t(n) - 1u);
      const bool ok = i < n && (HAS_VALID ? vbit(voff, ii, x.vb[r]) != 0u : true);
      const uint32_t xk = ok ? uint32_t(order_key(key_of_raw(x.raw[r], TYPE), TYPE) - kmin) : 0u;
      wd[r] = xk >> 5;
      bit[r] = ok ? 1u << (xk & 31u) : 0u;
    }
#pragma unroll
    for (int r = 0; r < SCAN_ROWS; r++) cur[r] = s_bits[wd[r]];
#pragma unroll
    for (int r = 0; r < SCAN_ROWS; r++)
      if ((cur[r] & bit[r]) != bit[r]) atomicOr(&s_bits[wd[r]], bit[r]);
  });
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < words; j += blockDim.x) {
    const uint32_t v = s_bits[j];
    if (v && (load_global(&bits[j]) & v) != v) atomicOr(&bits[j], v);
  }
}

__global__ void gb_range_unpack(const uint32_t* bits, uint32_t span, uint32_t* occ) {
  const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x < span) occ[x] = (bits[x >> 5] >> (x & 31u)) & 1u;
}

constexpr uint32_t RANGE_TILE = 4096;   // occupancy entries per workgroup of the scan (16 per thread)

// Per tile of the occupancy: its count of present keys.
__global__ __launch_bounds__(256) void gb_range_count(const uint32_t* occ, uint32_t span, uint32_t* tile_sum) {
  __shared__ uint32_t s[4];
  const uint32_t t0 = blockIdx.x * RANGE_TILE;
  uint32_t c = 0;
  for (uint32_t x = t0 + threadIdx.x; x < min(span, t0 + RANGE_TILE); x += 256) c += occ[x] ? 1u : 0u;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

// One workgroup: exclusive prefix of the tile counts in place; the total is the distinct-key count.
__global__ __launch_bounds__(1024) void gb_range_scan_tiles(uint32_t* tile_sum, uint32_t tiles, uint32_t* distinct) {
  __shared__ uint32_t s[16];
  uint32_t carry = 0;
  for (uint32_t t0 = 0; t0 < tiles; t0 += 1024) {
    const uint32_t t = t0 + threadIdx.x;
    const uint32_t v = t < tiles ? tile_sum[t] : 0u;
    uint32_t x = v;   // inclusive scan in the wave, then across the 16 waves
    const int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) s[threadIdx.x >> 6] = x;
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int w = 0; w < 16; w++) {
      before += w < int(threadIdx.x >> 6) ? s[w] : 0u;
      total += s[w];
    }
    if (t < tiles) tile_sum[t] = carry + before + x - v;
    carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) *distinct = carry;
}

// Each present key's rank (its id) in place of its occupancy mark, the ascending order keys, and the
// lowest row of each group when tracked.
__global__ __launch_bounds__(256) void gb_range_rank(uint32_t* occ, uint32_t span, const uint32_t* tile_base,
                                                     unsigned long long kmin, unsigned long long* okeys,
                                                     const uint32_t* first, uint32_t* rep) {
  __shared__ uint32_t s[4];
  const uint32_t t0 = blockIdx.x * RANGE_TILE;
  uint32_t run = tile_base[blockIdx.x];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t x0 = t0; x0 < min(span, t0 + RANGE_TILE); x0 += 256) {
    const uint32_t x = x0 + threadIdx.x;
    const bool present = x < span && occ[x] != 0u;
    const unsigned long long bal = __ballot(present);
    if (lane == 0) s[w] = uint32_t(__popcll(bal));
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int k = 0; k < 4; k++) {
      before += k < w ? s[k] : 0u;
      total += s[k];
    }
    if (present) {
      const uint32_t r = run + before + uint32_t(__popcll(bal & ((1ull << lane) - 1ull)));
      occ[x] = r;
      okeys[r] = kmin + x;
      if (rep) rep[r] = first[x];
    }
    run += total;
    __syncthreads();
  }
}

// Every row's id: the rank of its key's occupancy entry, or null_gid for a null row. With DENSE every
// key of the span is present, so a key's rank is its offset in the span and nothing is looked up.
template <int W, int TYPE, bool HAS_VALID, bool DENSE>
__global__ __launch_bounds__(256) void gb_range_ids(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n,
                                                    unsigned long long kmin, const uint32_t* rank, uint32_t null_gid,
                                                    uint32_t* ids) {
  const uint32_t gsize = gridDim.x * blockDim.x;
  scan_steps<W, HAS_VALID>(vals, valid, voff, n, [&](uint64_t row0, const ScanRows& x) {
    scan_take<HAS_VALID>(x);
    uint32_t xs[SCAN_ROWS];
    bool ok[SCAN_ROWS];
#pragma unroll
    for (int r = 0; r < SCAN_ROWS; r++) {
      const uint64_t i = row0 + uint64_t(r) * gsize;
      const int64_t ii = int64_t(i < n ? i : uint64_t(n) - 1u);
      ok[r] = HAS_VALID ? vbit(voff, ii, x.vb[r]) != 0u : true;
      xs[r] = ok[r] ? uint32_t(order_key(key_of_raw(x.raw[r], TYPE), TYPE) - kmin) : 0u;
    }
    uint32_t g[SCAN_ROWS];
#pragma unroll
    for (int r = 0; r < SCAN_ROWS; r++) g[r] = DENSE ? xs[r] : rank[xs[r]];
#pragma unroll
    for (int r = 0; r < SCAN_ROWS; r++) {
      const uint64_t i = row0 + uint64_t(r) * gsize;
      if (i < n) __builtin_nontemporal_store(ok[r] ? g[r] : null_gid, ids + i);
    }
  });
}

__global__ void gb_iota(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

__global__ void gb_fill_u32(uint32_t* p, size_t n, uint32_t v) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

__global__ void gb_put_u32(uint32_t* p, uint32_t v) { *p = v; }

// The range path's mapping from the keys' statistics, or false when their span is too wide for it.
bool map_range(const void* vals, int type, const uint8_t* valid, int64_t voff, int64_t n, bool track,
               const RangeStat& host, Mapped& out) {
  hipStream_t s = stream();
  const uint32_t un = uint32_t(n);
  const unsigned long long limit = std::min<unsigned long long>(1ull << 24, std::max<unsigned long long>(1ull << 16, 4ull * uint64_t(n)));
  {
    const bool any = host.kmin <= host.kmax;
    const unsigned long long span64 = any ? host.kmax - host.kmin + 1ull : 0ull;
    if (any && (span64 == 0ull || span64 > limit)) return false;   // span64 == 0: the full 2^64 range
  }
  const bool any = host.kmin <= host.kmax;
  const unsigned long long span64 = any ? host.kmax - host.kmin + 1ull : 0ull;
  const int grid = grid_for((n + SCAN_ROWS - 1) / SCAN_ROWS, 8);
  const uint32_t span = uint32_t(std::max<unsigned long long>(span64, 1ull));
  DevBuf occ = device_alloc(size_t(span) * 4);
  DevBuf first;
  if (track) first = device_alloc(size_t(span) * 4);
  hipLaunchKernelGGL(gb_fill_u32, dim3(grid_for(span, 4)), dim3(256), 0, s, static_cast<uint32_t*>(occ->ptr), size_t(span), 0u);
  if (track)
    hipLaunchKernelGGL(gb_fill_u32, dim3(grid_for(span, 4)), dim3(256), 0, s, static_cast<uint32_t*>(first->ptr),
                       size_t(span), 0xFFFFFFFFu);
  AM_HIP(hipGetLastError());
  auto* o = static_cast<uint32_t*>(occ->ptr);
  auto* f = track ? static_cast<uint32_t*>(first->ptr) : nullptr;
  if (any) {
    ProfScope ps("gb_range_occ");
    with_key_type(type, [&](auto tc) {
      constexpr int TT = decltype(tc)::value;
      if constexpr (TT != T_F32 && TT != T_F64) {
        constexpr int WW = width_of(TT);
        if (!track && span <= RANGE_BITS_MAX) {
          const uint32_t words = (span + 31u) / 32u;
          DevBuf bits = device_alloc(size_t(words) * 4);
          auto* b = static_cast<uint32_t*>(bits->ptr);
          hipLaunchKernelGGL(gb_fill_u32, dim3(grid_for(words, 4)), dim3(256), 0, s, b, size_t(words), 0u);
          const dim3 bg(grid_for((n + SCAN_ROWS - 1) / SCAN_ROWS, AM_OCC_BLOCKS_PER_CU));
          const size_t lds = size_t(words) * 4;
          if (valid) hipLaunchKernelGGL((gb_range_occ_bits<WW, TT, true>), bg, dim3(256), lds, s, vals, valid, voff, un, host.kmin, words, b);
          else hipLaunchKernelGGL((gb_range_occ_bits<WW, TT, false>), bg, dim3(256), lds, s, vals, valid, voff, un, host.kmin, words, b);
          hipLaunchKernelGGL(gb_range_unpack, dim3((span + 255) / 256), dim3(256), 0, s, static_cast<const uint32_t*>(b), span, o);
        } else if (valid) {
          if (track) hipLaunchKernelGGL((gb_range_occ<WW, TT, true, true>), dim3(grid), dim3(256), 0, s, vals, valid, voff, un, host.kmin, o, f);
          else hipLaunchKernelGGL((gb_range_occ<WW, TT, true, false>), dim3(grid), dim3(256), 0, s, vals, valid, voff, un, host.kmin, o, f);
        } else {
          if (track) hipLaunchKernelGGL((gb_range_occ<WW, TT, false, true>), dim3(grid), dim3(256), 0, s, vals, valid, voff, un, host.kmin, o, f);
          else hipLaunchKernelGGL((gb_range_occ<WW, TT, false, false>), dim3(grid), dim3(256), 0, s, vals, valid, voff, un, host.kmin, o, f);
        }
      }
    });
    AM_HIP(hipGetLastError());
  }
  const uint32_t tiles = (span + RANGE_TILE - 1) / RANGE_TILE;
  DevBuf tsum = device_alloc(size_t(tiles) * 4 + 4);
  auto* ts = static_cast<uint32_t*>(tsum->ptr);
  out.okeys = device_alloc(size_t(span) * 8);
  out.rep = device_alloc(size_t(span + 1) * 4);
  {
    ProfScope ps("gb_range_scan");
    hipLaunchKernelGGL(gb_range_count, dim3(tiles), dim3(256), 0, s, o, span, ts);
    hipLaunchKernelGGL(gb_range_scan_tiles, dim3(1), dim3(1024), 0, s, ts, tiles, ts + tiles);
    hipLaunchKernelGGL(gb_range_rank, dim3(tiles), dim3(256), 0, s, o, span, ts, host.kmin,
                       static_cast<unsigned long long*>(out.okeys->ptr), f,
                       track ? static_cast<uint32_t*>(out.rep->ptr) : nullptr);
    AM_HIP(hipGetLastError());
  }
  const uint32_t D = read_back(ts + tiles);
  out.D = any ? D : 0u;
  out.has_null = host.nulls > 0;
  out.K = int64_t(out.D) + (out.has_null ? 1 : 0);
  if (track && out.has_null)
    hipLaunchKernelGGL(gb_put_u32, dim3(1), dim3(1), 0, s, static_cast<uint32_t*>(out.rep->ptr) + out.D, host.null_first);
  out.perm = device_alloc(size_t(std::max<uint32_t>(out.D, 1)) * 4);
  hipLaunchKernelGGL(gb_iota, dim3((out.D + 255) / 256 + 1), dim3(256), 0, s, static_cast<uint32_t*>(out.perm->ptr), out.D);
  out.ids = device_alloc(size_t(std::max<int64_t>(n, 1)) * 4);
  {
    ProfScope ps("gb_ids");
    auto* ids = static_cast<uint32_t*>(out.ids->ptr);
    with_key_type(type, [&](auto tc) {
      constexpr int TT = decltype(tc)::value;
      if constexpr (TT != T_F32 && TT != T_F64) {
        constexpr int WW = width_of(TT);
        const bool dense = out.D == span;
#define AM_RANGE_IDS(HV, DN) hipLaunchKernelGGL((gb_range_ids<WW, TT, HV, DN>), dim3(grid), dim3(256), 0, s, vals, valid, voff, un, host.kmin, o, out.D, ids)
        if (valid) {
          if (dense) AM_RANGE_IDS(true, true);
          else AM_RANGE_IDS(true, false);
        } else {
          if (dense) AM_RANGE_IDS(false, true);
          else AM_RANGE_IDS(false, false);
        }
#undef AM_RANGE_IDS
      }
    });
    AM_HIP(hipGetLastError());
  }
  return true;
}

// gb_build for the key column's element width: 64-bit integer keys without a validity bitmap (the
// common case) take an instantiation with the key mapping and the validity fixed.
template <bool TRACK>
void launch_build(int w, int grid, hipStream_t s, const void* vals, int type, const uint8_t* valid, int64_t voff,
                  uint32_t n, Table T, uint32_t mask, uint32_t budget, uint32_t* slot_of, BuildStatus* st, bool priv) {
  const int p = priv ? 1 : 0;
  if (!valid && (type == T_I64 || type == T_U64)) {
    hipLaunchKernelGGL((gb_build<TRACK, 8, T_U64, true>), dim3(grid), dim3(BUILD_THREADS), 0, s, vals, type, valid, voff, n, T, mask, budget, slot_of, st, p);
    return;
  }
  switch (w) {
    case 8: hipLaunchKernelGGL((gb_build<TRACK, 8, -1, false>), dim3(grid), dim3(BUILD_THREADS), 0, s, vals, type, valid, voff, n, T, mask, budget, slot_of, st, p); break;
    case 4: hipLaunchKernelGGL((gb_build<TRACK, 4, -1, false>), dim3(grid), dim3(BUILD_THREADS), 0, s, vals, type, valid, voff, n, T, mask, budget, slot_of, st, p); break;
    case 2: hipLaunchKernelGGL((gb_build<TRACK, 2, -1, false>), dim3(grid), dim3(BUILD_THREADS), 0, s, vals, type, valid, voff, n, T, mask, budget, slot_of, st, p); break;
    default: hipLaunchKernelGGL((gb_build<TRACK, 1, -1, false>), dim3(grid), dim3(BUILD_THREADS), 0, s, vals, type, valid, voff, n, T, mask, budget, slot_of, st, p); break;
  }
}

// Maps one key column (values, type, validity) to dense ascending ids; with `track` it also finds
// the lowest row of every group.
Mapped map_column(const void* vals, int type, const uint8_t* valid, int64_t voff, int64_t n, bool track) {
  hipStream_t s = stream();
  uint32_t un = uint32_t(n);
  Mapped out;
  // Capacity: about 5 slots per distinct key, from a HyperLogLog estimate on inputs large enough for
  // the table size to matter; otherwise two slots per row. A load of at most ~0.2 keeps the rows
  // whose home slot holds another key few (the build over 50M rows of 100k Zipf keys takes 0.36 ms
  // at 5 slots per key against 0.40 ms at 2.5). Never more than two slots per row, and
  // grown (eightfold) when a build runs out of probes. The same pass gives an integer column's span,
  // which sends a small one to the range path.
  const uint64_t want = std::max<uint64_t>(1024, 2 * uint64_t(n));
  uint64_t full = 1024;
  while (full < want) full <<= 1;
  uint64_t cap = 1024;
  if (n >= (int64_t(1) << 16)) {
    const bool range = type_is_int(type) && range_path_enabled();
    DevBuf regs = device_alloc(HLL_M * 4);
    DevBuf rsb = device_alloc(sizeof(RangeStat));
    std::shared_ptr<void> landing = pinned_slot();
    auto* ks = static_cast<KeyStat*>(landing.get());
    {
      ProfScope ps("gb_key_scan");
      auto* rg = static_cast<uint32_t*>(regs->ptr);
      auto* rs = static_cast<RangeStat*>(rsb->ptr);
      hipLaunchKernelGGL(gb_key_init, dim3(1), dim3(256), 0, s, rg, rs);
      key_scan(type, s, vals, valid, voff, un, rg, rs);
      hipLaunchKernelGGL(gb_key_estimate, dim3(1), dim3(256), 0, s, static_cast<const uint32_t*>(rg),
                         static_cast<const RangeStat*>(rs), ks);
      AM_HIP(hipGetLastError());
    }
    AM_HIP(hipStreamSynchronize(s));
    const KeyStat stat = *ks;
    if (range && map_range(vals, type, valid, voff, n, track, stat.r, out)) return out;
    const unsigned long long d = stat.est;
#ifndef AM_CAP_HALVES
#define AM_CAP_HALVES 10   // table slots per distinct key, in halves
#endif
    const uint64_t target = std::max<uint64_t>(1024, d * AM_CAP_HALVES / 2);
    while (cap < target) cap <<= 1;
  } else {
    while (cap < want && cap < (1ull << 21)) cap <<= 1;
  }
  cap = std::min(cap, full);
  out.ids = device_alloc(size_t(std::max<int64_t>(n, 1)) * 4);
  DevBuf stb = device_alloc(sizeof(BuildStatus));
  auto* st = static_cast<BuildStatus*>(stb->ptr);
  BuildStatus host{};
  DevBuf tkeys, tfirst, gkeys, gfirst, xmap, slots;
  Table T{};          // the table the rows were inserted into (eight private copies, or one)
  Table G{};          // the merged table of distinct keys
  bool priv = false;
  const int build_grid = int(std::max<int64_t>(1, std::min<int64_t>((n + BUILD_ROWS * BUILD_THREADS - 1) / (BUILD_ROWS * BUILD_THREADS),
                                                                     int64_t(device().cu_count) * AM_BUILD_BLOCKS_PER_CU)));
  for (int attempt = 0;; attempt++) {
    bool last_try = cap >= full;
    // Tables up to 2^20 slots are private to each XCD (eight copies): a key line is then only changed
    // by atomics of the XCD whose L2 holds it. Larger ones are shared.
    priv = cap <= (1u << 20);
    const size_t ntab = priv ? NXCD : 1;
    tkeys = device_alloc(ntab * size_t(cap + 1) * 8);
    if (track) tfirst = device_alloc(ntab * size_t(cap + 1) * 4);
    T.key = static_cast<unsigned long long*>(tkeys->ptr);
    T.first = track ? static_cast<uint32_t*>(tfirst->ptr) : nullptr;
    hipLaunchKernelGGL(gb_table_init, dim3(grid_for(int64_t(ntab * (cap + 1)), 4)), dim3(256), 0, s, T,
                       ntab * size_t(cap + 1), st);
    AM_HIP(hipGetLastError());
    {
      ProfScope ps("gb_build");
      uint32_t budget = last_try ? uint32_t(cap) : 128u;
      if (track)
        launch_build<true>(type_width(type), build_grid, s, vals, type, valid, voff, un, T, uint32_t(cap - 1), budget,
                           static_cast<uint32_t*>(out.ids->ptr), st, priv);
      else
        launch_build<false>(type_width(type), build_grid, s, vals, type, valid, voff, un, T, uint32_t(cap - 1), budget,
                            static_cast<uint32_t*>(out.ids->ptr), st, priv);
      AM_HIP(hipGetLastError());
    }
    G = T;
    if (priv) {
      gkeys = device_alloc(size_t(cap + 1) * 8);
      if (track) gfirst = device_alloc(size_t(cap + 1) * 4);
      xmap = device_alloc(size_t(NXCD) * size_t(cap + 1) * 4);
      G.key = static_cast<unsigned long long*>(gkeys->ptr);
      G.first = track ? static_cast<uint32_t*>(gfirst->ptr) : nullptr;
      hipLaunchKernelGGL(gb_table_init, dim3(grid_for(int64_t(cap + 1), 4)), dim3(256), 0, s, G, size_t(cap + 1),
                         static_cast<BuildStatus*>(nullptr));
      ProfScope ps("gb_merge");
      hipLaunchKernelGGL(gb_merge_tables, dim3(grid_for(int64_t(NXCD) * (cap + 1), 4)), dim3(256), 0, s,
                         static_cast<const unsigned long long*>(T.key), static_cast<const uint32_t*>(T.first),
                         uint32_t(cap), st, G, static_cast<uint32_t*>(xmap->ptr));
      AM_HIP(hipGetLastError());
    }
    // Compact the occupied slots before the status is read: a build that ran out of probes left at
    // most cap + 1 of them, which the compacted arrays hold, and is built again.
    out.okeys = device_alloc(size_t(cap + 1) * 8);
    slots = device_alloc(size_t(cap + 1) * 4);
    {
      ProfScope ps("gb_compact");
      uint32_t blocks = uint32_t(std::min<uint64_t>((cap + 1 + 4095) / 4096, uint64_t(device().cu_count) * 4));
      uint32_t per_block = uint32_t((cap + 1 + blocks - 1) / blocks);
      hipLaunchKernelGGL(gb_compact, dim3(blocks), dim3(256), 0, s, G.key, uint32_t(cap), type, st,
                         static_cast<unsigned long long*>(out.okeys->ptr), static_cast<uint32_t*>(slots->ptr), per_block);
      AM_HIP(hipGetLastError());
    }
    host = read_back(st);
    if (!host.overflow) break;
    if (attempt > 8) fail("group-by: the key table could not be sized");
    cap = std::min<uint64_t>(cap * 8, full);
  }
  out.D = host.distinct;
  out.has_null = host.nulls > 0;
  out.K = int64_t(out.D) + (out.has_null ? 1 : 0);
  out.perm = device_alloc(size_t(std::max<uint32_t>(out.D, 1)) * 4);
  if (out.D > 0)
    argsort_device(out.okeys->ptr, T_U64, nullptr, 0, out.D, false, 0, 1, static_cast<uint32_t*>(out.perm->ptr), s);
  DevBuf gid_of_slot = device_alloc(size_t(cap + 1) * 4);
  out.rep = device_alloc(size_t(std::max<int64_t>(out.K, 1)) * 4);
  {
    ProfScope ps("gb_relabel");
    hipLaunchKernelGGL(gb_relabel, dim3(uint32_t((out.D + 1 + 255) / 256)), dim3(256), 0, s,
                       static_cast<const uint32_t*>(out.perm->ptr), static_cast<const uint32_t*>(slots->ptr), out.D, G.first,
                       static_cast<uint32_t*>(gid_of_slot->ptr), static_cast<uint32_t*>(out.rep->ptr), st,
                       out.has_null ? 1 : 0, track ? 1 : 0);
    AM_HIP(hipGetLastError());
  }
  // Rows hold private-table slots: map those through the merged table first.
  DevBuf gid_x;
  const uint32_t* row_gid = static_cast<const uint32_t*>(gid_of_slot->ptr);
  if (priv) {
    const uint32_t total = uint32_t(NXCD) * uint32_t(cap + 1);
    gid_x = device_alloc(size_t(total) * 4);
    hipLaunchKernelGGL(gb_compose, dim3(grid_for(total, 4)), dim3(256), 0, s, static_cast<const uint32_t*>(xmap->ptr),
                       total, row_gid, static_cast<uint32_t*>(gid_x->ptr));
    AM_HIP(hipGetLastError());
    row_gid = static_cast<const uint32_t*>(gid_x->ptr);
  }
  if (n > 0) {
    ProfScope ps("gb_ids");
    hipLaunchKernelGGL(gb_ids, dim3(build_grid), dim3(BUILD_THREADS), 0, s, static_cast<uint32_t*>(out.ids->ptr), un, row_gid,
                       out.D);
    AM_HIP(hipGetLastError());
  }
  return out;
}

GroupBy* group_by_keys_impl(Array* const* columns, int64_t count, bool need_rep) {
  if (count < 1) fail_invalid("group-by needs at least one key column");
  int64_t n = columns[0]->length;
  for (int64_t c = 1; c < count; c++)
    if (columns[c]->length != n)
      fail_length_mismatch(n, columns[c]->length);
  check_length(n);
  for (int64_t c = 0; c < count; c++) {
    int t = columns[c]->type;
    if (!type_is_primitive(t))
      fail_unsupported("group-by over a " + columns[c]->format + " key column is not implemented");
  }
  auto* g = new GroupBy();
  try {
    for (int64_t c = 0; c < count; c++) {
      retain_array(columns[c]);
      g->columns.push_back(columns[c]);
    }
    g->n = n;
    if (n == 0) {
      g->K = 0;
      g->ids = device_alloc(4);
      g->rep = device_alloc(4);
      g->has_rep = true;
      return g;
    }
    hipStream_t s = stream();
    const Array* a0 = columns[0];
    // Representative rows are needed for float keys (-0.0 and 0.0, NaN payloads share a group but
    // not their bits) and for several columns; integer keys are answered from the sorted keys.
    bool int_single = count == 1 && type_is_int(a0->type) && !need_rep;
    Mapped m0 = map_column(a0->data(), a0->type, a0->has_nulls() ? a0->valid_bits() : nullptr, a0->offset, n,
                           count == 1 && !int_single);
    int64_t K = m0.K;
    g->ids = m0.ids;
    if (int_single) {
      g->okeys = m0.okeys;
      g->perm = m0.perm;
      g->D = m0.D;
      g->has_null = m0.has_null;
    } else if (count == 1) {
      g->rep = m0.rep;
      g->has_rep = true;
      g->has_null = m0.has_null;
    }
    // Further columns fold pairwise: ids a (K groups) and b (K2) become the key a * K2 + b, which is
    // injective and ascending in (a, b), and that key is mapped again (GroupByKeys.swift:103-112).
    for (int64_t c = 1; c < count; c++) {
      const Array* ac = columns[c];
      Mapped mc = map_column(ac->data(), ac->type, ac->has_nulls() ? ac->valid_bits() : nullptr, ac->offset, n, false);
      int64_t K2 = mc.K;
      if (K2 > 0 && K > INT64_MAX / K2)
        fail_invalid("multi-column group-by overflowed the 64-bit composite key (" + std::to_string(K) + " x " +
             std::to_string(K2) + " distinct)");
      DevBuf comp = device_alloc(size_t(n) * 8);
      hipLaunchKernelGGL(gb_combine, dim3(grid_for(n)), dim3(256), 0, s, static_cast<const int32_t*>(g->ids->ptr),
                         static_cast<const int32_t*>(mc.ids->ptr), uint32_t(n), (long long)K2,
                         static_cast<long long*>(comp->ptr));
      AM_HIP(hipGetLastError());
      bool last = c == count - 1;
      Mapped mm = map_column(comp->ptr, T_I64, nullptr, 0, n, last);
      K = mm.K;
      g->ids = mm.ids;
      if (last) {
        g->rep = mm.rep;
        g->has_rep = true;
      }
    }
    if (K > int64_t(INT32_MAX)) fail("group-by: more than 2^31 groups");
    g->K = K;
    // No wait here: the ids are only read by later calls, which are ordered after the mapping (the
    // same stream, or `ready`), so the caller's next call is prepared while the last kernels run.
    if (!in_batch()) {
      AM_HIP(hipEventCreateWithFlags(&g->ready, hipEventDisableTiming));
      AM_HIP(hipEventRecord(g->ready, s));
      g->owner = s;
    }
  } catch (...) {
    delete g;
    throw;
  }
  return g;
}

}  // namespace

GroupBy* group_by_keys(Array* const* columns, int64_t count) { return group_by_keys_impl(columns, count, false); }

int64_t group_by_group_count(const GroupBy* g) { return g->K; }

Array* group_by_ids(GroupBy* g) {
  gb_use(g);
  sync();   // the returned array is complete, like every other result
  Array* out = new_array("i", T_I32, g->n);
  out->values = g->ids;
  out->null_count = 0;
  return out;
}

Array* group_by_keys_result(GroupBy* g, int64_t i) {
  gb_use(g);
  if (i < 0 || i >= int64_t(g->columns.size()))
    fail_invalid("key column " + std::to_string(i) + " of " + std::to_string(g->columns.size()));
  if (g->key_results.empty()) g->key_results.assign(g->columns.size(), nullptr);
  if (!g->key_results[size_t(i)]) {
    const Array* col = g->columns[size_t(i)];
    Array* keys = nullptr;
    if (!g->has_rep) {
      // Integer keys: the sorted distinct keys themselves, the null group (last) null.
      hipStream_t s = stream();
      keys = new_array(col->format, col->type, g->K);
      try {
        keys->values = device_alloc(size_t(std::max<int64_t>(g->K, 1)) * col->width());
        DevBuf vb = device_alloc(size_t(std::max<int64_t>(g->K, 1)));
        hipLaunchKernelGGL(gb_int_keys, dim3(uint32_t((g->D + 1 + 255) / 256)), dim3(256), 0, s,
                           static_cast<const unsigned long long*>(g->okeys->ptr),
                           static_cast<const uint32_t*>(g->perm->ptr), g->D, col->type, keys->values->ptr,
                           static_cast<uint8_t*>(vb->ptr), g->has_null ? 1 : 0);
        AM_HIP(hipGetLastError());
        if (g->has_null) {
          keys->validity = device_alloc(size_t((g->K + 63) / 64) * 8);
          DevBuf nz = device_alloc(8);
          hipLaunchKernelGGL(gb_agg_init<CountP<false>>, dim3(1), dim3(64), 0, s,
                             static_cast<unsigned long long*>(nz->ptr), 1);
          hipLaunchKernelGGL(pack_validity, dim3(grid_for(g->K, 4)), dim3(256), 0, s,
                             static_cast<const uint8_t*>(vb->ptr), (long long)g->K,
                             static_cast<unsigned long long*>(keys->validity->ptr),
                             static_cast<unsigned long long*>(nz->ptr));
          AM_HIP(hipGetLastError());
          keys->null_count = 1;
        } else {
          keys->null_count = 0;
        }
        sync();
      } catch (...) {
        release_array(keys);
        throw;
      }
    } else {
      Array* rep = new_array("I", T_U32, g->K);
      rep->values = g->rep;
      rep->null_count = 0;
      try {
        keys = take(*col, *rep);
      } catch (...) {
        release_array(rep);
        throw;
      }
      release_array(rep);
    }
    g->key_results[size_t(i)] = keys;
  }
  Array* r = g->key_results[size_t(i)];
  retain_array(r);
  return r;
}

void group_by_release(GroupBy* g) { delete g; }

// ================================================================ host: aggregates

namespace {

Array* finalize_simple(const DevBuf& gtab, int W, long long K, int kind, int vtype, const std::string& fmt,
                       int otype) {
  hipStream_t s = stream();
  Array* out = new_array(fmt, otype, K);
  try {
    out->values = device_alloc(size_t(std::max<long long>(K, 1)) * std::max(type_width(otype), 1));
    DevBuf vb = device_alloc(size_t(std::max<long long>(K, 1)));
    if (K > 0) {
      hipLaunchKernelGGL(gb_finalize_simple, dim3(uint32_t((K + 255) / 256)), dim3(256), 0, s,
                         static_cast<const unsigned long long*>(gtab->ptr), W, K, kind, vtype, out->values->ptr,
                         static_cast<uint8_t*>(vb->ptr));
      AM_HIP(hipGetLastError());
    }
    if (kind != 3 && kind != KIND_SUM_COUNT && K > 0) {
      out->validity = device_alloc(size_t((K + 63) / 64) * 8);
      DevBuf nz = device_alloc(8);
      hipLaunchKernelGGL(gb_agg_init<CountP<false>>, dim3(1), dim3(64), 0, s, static_cast<unsigned long long*>(nz->ptr), 1);
      hipLaunchKernelGGL(pack_validity, dim3(grid_for(K, 4)), dim3(256), 0, s, static_cast<const uint8_t*>(vb->ptr), K,
                         static_cast<unsigned long long*>(out->validity->ptr), static_cast<unsigned long long*>(nz->ptr));
      AM_HIP(hipGetLastError());
      const unsigned long long nulls = read_back(static_cast<const unsigned long long*>(nz->ptr));
      out->null_count = int64_t(nulls);
      if (nulls == 0) out->validity.reset();
    } else {
      out->null_count = 0;
      sync();
    }
  } catch (...) {
    release_array(out);
    throw;
  }
  return out;
}

AggArgs make_args(const IdSource& ids, const Array* v, int64_t n, long long K) {
  AggArgs a{};
  a.ids = ids;
  if (v) {
    a.v.vals = v->data();
    a.v.type = v->type;
    a.v.valid = v->has_nulls() ? v->valid_bits() : nullptr;
    a.v.voff = v->offset;
  }
  a.n = n;
  a.K = K;
  return a;
}

// The correctly rounded sum and mean per group of a float64 (vw 8) or float32 (vw 4) column, as
// binary64 bits (prec 53), or as binary32 results widened to binary64 bits (prec 24), with validity
// bytes.
template <int P, int EMIN, int VW>
void exact_fsum_t(const AggArgs& base, DevBuf& sum, DevBuf& mean, DevBuf& vb, PartSet* recs, uint64_t col) {
  hipStream_t s = stream();
  const long long K = base.K;
  const size_t k1 = size_t(std::max<long long>(K, 1));
  AggArgs a = base;
  DevBuf scan = run_agg<FsScanP<VW>>(a, s, "fs_scan", recs, col, true);
  DevBuf grid = device_alloc(k1 * 4);
  hipLaunchKernelGGL(fs_window, dim3(uint32_t((K + 255) / 256)), dim3(256), 0, s,
                     static_cast<const unsigned long long*>(scan->ptr), K, static_cast<uint32_t*>(grid->ptr));
  AM_HIP(hipGetLastError());
  a.aux = grid->ptr;
  DevBuf fix = run_agg<FsFixP<VW>>(a, s, "fs_fixed", recs, col, true);
  sum = device_alloc(k1 * 8);
  mean = device_alloc(k1 * 8);
  vb = device_alloc(k1);
  DevBuf redo = device_alloc(k1 * 4), rc = device_alloc(8);
  hipLaunchKernelGGL(gb_agg_init<CountP<false>>, dim3(1), dim3(64), 0, s, static_cast<unsigned long long*>(rc->ptr), 1);
  {
    ProfScope ps("fs_settle");
    hipLaunchKernelGGL((fs_settle<P, EMIN, VW>), dim3(uint32_t((K + 255) / 256)), dim3(256), 0, s,
                       static_cast<const unsigned long long*>(scan->ptr), static_cast<const unsigned long long*>(fix->ptr),
                       static_cast<const uint32_t*>(grid->ptr), K, base.v.vals, static_cast<unsigned long long*>(sum->ptr),
                       static_cast<unsigned long long*>(mean->ptr), static_cast<uint8_t*>(vb->ptr),
                       static_cast<uint32_t*>(redo->ptr), static_cast<uint32_t*>(rc->ptr));
    AM_HIP(hipGetLastError());
  }
  const uint32_t nredo = read_back(static_cast<const uint32_t*>(rc->ptr));
  if (nredo == 0) return;
  if (base.n >= (int64_t(1) << 31)) fail("exact float sum: more than 2^31 rows in a group the grid cannot settle");
  DevBuf digits = device_alloc(size_t(nredo) * XD * 8);
  hipLaunchKernelGGL(gb_agg_init<CountP<false>>, dim3(grid_for(int64_t(nredo) * XD, 4)), dim3(256), 0, s,
                     static_cast<unsigned long long*>(digits->ptr), (long long)nredo * XD);
  ProfScope ps("fs_redo");
  hipLaunchKernelGGL(fs_redo_rows<VW>, dim3(grid_for(base.n)), dim3(256), 0, s, a,
                     static_cast<const uint32_t*>(redo->ptr), static_cast<long long*>(digits->ptr));
  hipLaunchKernelGGL((fs_redo_settle<P, EMIN>), dim3(uint32_t((K + 255) / 256)), dim3(256), 0, s,
                     static_cast<const uint32_t*>(redo->ptr), K, static_cast<const long long*>(digits->ptr),
                     static_cast<const unsigned long long*>(scan->ptr), static_cast<unsigned long long*>(sum->ptr),
                     static_cast<unsigned long long*>(mean->ptr));
  AM_HIP(hipGetLastError());
  AM_HIP(hipStreamSynchronize(s));
}

// Float64 values to binary64 results; float32 values to binary64 results (am_group_agg_ex) or to
// binary32 results (the dense am_group_by sum).
// recs / col (or null / 0): the group-by's shared partitioned records and the value column's id.
void exact_fsum(const AggArgs& base, bool f32_result, DevBuf& sum, DevBuf& mean, DevBuf& vb, PartSet* recs = nullptr,
                uint64_t col = 0) {
  if (base.v.type == T_F64) exact_fsum_t<53, -1022, 8>(base, sum, mean, vb, recs, col);
  else if (f32_result) exact_fsum_t<24, -126, 4>(base, sum, mean, vb, recs, col);
  else exact_fsum_t<53, -1022, 4>(base, sum, mean, vb, recs, col);
}

Array* bits_array(const DevBuf& vals, const DevBuf& vb, long long K, const std::string& fmt, int type) {
  hipStream_t s = stream();
  Array* out = new_array(fmt, type, K);
  out->values = vals;
  out->null_count = 0;
  if (K > 0) {
    try {
      out->validity = device_alloc(size_t((K + 63) / 64) * 8);
      DevBuf nz = device_alloc(8);
      hipLaunchKernelGGL(gb_agg_init<CountP<false>>, dim3(1), dim3(64), 0, s, static_cast<unsigned long long*>(nz->ptr), 1);
      hipLaunchKernelGGL(pack_validity, dim3(grid_for(K, 4)), dim3(256), 0, s, static_cast<const uint8_t*>(vb->ptr), K,
                         static_cast<unsigned long long*>(out->validity->ptr), static_cast<unsigned long long*>(nz->ptr));
      AM_HIP(hipGetLastError());
      const unsigned long long nulls = read_back(static_cast<const unsigned long long*>(nz->ptr));
      out->null_count = int64_t(nulls);
      if (nulls == 0) out->validity.reset();
    } catch (...) {
      release_array(out);
      throw;
    }
  }
  return out;
}

// The shared op table of am_group_agg_ex and am_group_by over an id source.
// The sampled hot groups of the ids (gb_hot_sample), once per mapping; none for small inputs.
const long long* hot_groups(const IdSource& ids, int64_t n, long long K, GroupBy* cache, DevBuf& holder) {
  if (!AM_AGG_HOT || n < (int64_t(1) << 20)) return nullptr;
  if (cache && cache->hot) return static_cast<const long long*>(cache->hot->ptr);
  holder = device_alloc(HOT_GROUPS * sizeof(long long));
  retain(holder);
  {
    ProfScope ps("gb_hot_sample");
    hipLaunchKernelGGL(gb_hot_sample, dim3(1), dim3(1024), 0, stream(), ids, n, K,
                       static_cast<long long*>(holder->ptr));
    AM_HIP(hipGetLastError());
  }
  if (cache) cache->hot = holder;
  return static_cast<const long long*>(holder->ptr);
}

// A transform of the value column (am_query's (shr c N) / (bit_and c N)).
struct ValX {
  int op = VX_NONE;
  long long arg = 0;
};

Array* aggregate(const IdSource& ids, int64_t n, long long K, const Array* v, int op, GroupBy* cache, ValX vx = {}) {
  hipStream_t s = stream();
  DevBuf hot_holder;
  if (op == 1) {
    DevBuf t;
    // An integer sum of a column without nulls already counted every row of each group (word 1).
    if (cache && !cache->count_all_tab && cache->cnt_tab && cache->cnt_all_rows)
      return finalize_simple(cache->cnt_tab, 2, K, KIND_SUM_COUNT, T_I64, "l", T_I64);
    if (cache && cache->count_all_tab) {
      t = cache->count_all_tab;
    } else {
      AggArgs a = make_args(ids, nullptr, n, K);
      a.hot = hot_groups(ids, n, K, cache, hot_holder);
      t = run_agg<CountP<false>>(a, s, "gb_count");
      if (cache) cache->count_all_tab = t;
    }
    return finalize_simple(t, 1, K, 3, T_I64, "l", T_I64);
  }
  if (!v) fail_invalid("am_group_agg_ex op " + std::to_string(op) + " needs a value column");
  if (v->length != n) fail_length_mismatch(n, v->length);
  // hash_count needs only the validity, so it takes a column of any type.
  if (op == 2) {
    // An integer sum of the same column already counted its non-null values (word 1).
    if (cache && cache->cnt_col == v->uid && cache->cnt_tab) return finalize_simple(cache->cnt_tab, 2, K, KIND_SUM_COUNT, T_I64, "l", T_I64);
    // A column without nulls has as many non-null values per group as rows.
    if (cache && cache->count_all_tab && !v->has_nulls()) return finalize_simple(cache->count_all_tab, 1, K, 3, T_I64, "l", T_I64);
    AggArgs a = make_args(ids, v, n, K);
    a.hot = hot_groups(ids, n, K, cache, hot_holder);
    DevBuf t = run_agg<CountP<true>>(a, s, "gb_count", cache ? &cache->part : nullptr, v->uid);
    return finalize_simple(t, 1, K, 3, T_I64, "l", T_I64);
  }
  if (!type_is_primitive(v->type))
    fail_unsupported("grouped aggregate over a " + v->format + " column is not supported by this build");
  AggArgs a = make_args(ids, v, n, K);
  a.v.xop = vx.op;
  a.v.xarg = vx.arg;
  if (vx.op != VX_NONE && (cache || !type_is_int(v->type)))
    fail("a value transform takes the dense group-by of an integer column");
  if (op == 0 || op == 3) {
    if (type_is_float(v->type)) {
      // float64: correctly rounded (GroupSumExact.swift); float32: summed as float64 values, the
      // exact sum rounded once (Segmented.swift:103-110 sums them in float64).
      DevBuf sum, mean, vb;
      if (cache && cache->fx_col == v->uid) {
        sum = cache->fx_sum; mean = cache->fx_mean; vb = cache->fx_valid;
      } else {
        exact_fsum(a, false, sum, mean, vb, cache ? &cache->part : nullptr, v->uid);
        if (cache) { cache->fx_col = v->uid; cache->fx_sum = sum; cache->fx_mean = mean; cache->fx_valid = vb; }
      }
      return bits_array(op == 0 ? sum : mean, vb, K, "g", T_F64);
    }
    DevBuf t;
    if (cache && cache->cnt_col == v->uid && cache->cnt_tab) {
      t = cache->cnt_tab;
    } else {
      a.hot = hot_groups(ids, n, K, cache, hot_holder);
      t = run_agg<SumCountP>(a, s, "gb_sum", cache ? &cache->part : nullptr, v->uid);
      if (cache) { cache->cnt_col = v->uid; cache->cnt_tab = t; cache->cnt_all_rows = !v->has_nulls(); }
    }
    if (op == 0) {
      bool u64 = v->type == T_U64;
      return finalize_simple(t, 2, K, 0, v->type, u64 ? "L" : "l", u64 ? T_U64 : T_I64);
    }
    return finalize_simple(t, 2, K, type_is_signed(v->type) ? 1 : 2, v->type, "g", T_F64);
  }
  if (op == 4 || op == 5) {
    DevBuf t;
    if (cache && cache->mm_col == v->uid) {
      t = cache->mm_tab;
    } else {
      t = run_agg<MinMaxP>(a, s, "gb_minmax", cache ? &cache->part : nullptr, v->uid, type_is_float(v->type));
      if (cache) { cache->mm_col = v->uid; cache->mm_tab = t; }
    }
    return finalize_simple(t, 2, K, op, v->type, v->format, v->type);
  }
  if (op == 6) {   // hash_min_max: struct<min, max> over one min/max pass (AggregatesExtra.swift:90-96)
    Array* lo = aggregate(ids, n, K, v, 4, cache);
    Array* hi = nullptr;
    try {
      hi = aggregate(ids, n, K, v, 5, cache);
    } catch (...) {
      release_array(lo);
      throw;
    }
    Array* st = new_array("+s", T_STRUCT, K);
    st->null_count = 0;
    st->children = {lo, hi};
    st->child_names = {"min", "max"};
    return st;
  }
  if (op < 0 || op > 25) fail_invalid("unknown am_group_agg_ex op " + std::to_string(op));
  fail("am_group_agg_ex op " + std::to_string(op) + " is not supported by this build");
}

}  // namespace

Array* group_agg(GroupBy* g, const Array* values, int op, double /*p1*/) {
  gb_use(g);
  IdSource ids{g->ids->ptr, T_I32, nullptr, 0};
  if (op != 1 && values && values->length != g->n)
    fail_length_mismatch(g->n, values->length);
  // GroupBy refuses a zero key count, so an empty mapping aggregates over one empty group and the
  // result is trimmed to zero rows (GroupByKeys.swift:59-76, 159-165).
  long long K = std::max<int64_t>(g->K, 1);
  Array* r = aggregate(ids, g->n, K, values, op, g);
  if (g->K == 0) {
    r->length = 0;
    r->null_count = 0;
    r->validity.reset();
    for (Array* c : r->children) { c->length = 0; c->null_count = 0; c->validity.reset(); }
  }
  return r;
}

// The dense form's float32 sum (ArrowMetalC.swift:426, GroupBy.swift:94-106): a float32 total per
// key, returned as float64. The reference adds in float32 in whatever order its atomics run; here the
// exact sum is rounded once to float32, which is what any order gives whenever float32 addition is exact.
Array* sum_f32_as_f64(const IdSource& ids, int64_t n, long long K, const Array& v) {
  AggArgs a = make_args(ids, &v, n, K);
  DevBuf sum, mean, vb;
  exact_fsum(a, true, sum, mean, vb);
  return bits_array(sum, vb, K, "g", T_F64);
}

Array* group_by_dense(const Array& keys, int64_t key_count, int agg, const Array* values) {
  return group_by_dense_ex(keys, key_count, 0, agg, values, VX_NONE, 0);
}

// The dense group-by with the key offset by `bias` before the range check and the values transformed
// by `xop` / `xarg` (am_query's (sub key N), (shr c N), (bit_and c N)).
Array* group_by_dense_ex(const Array& keys, int64_t key_count, long long bias, int agg, const Array* values, int xop,
                         long long xarg) {
  if (key_count <= 0 || key_count > int64_t(UINT32_MAX) / 4) fail_invalid("keyCount out of range");
  if (keys.type != T_I32 && keys.type != T_I64 && keys.type != T_U32)
    fail("am_group_by: keys must be int32, int64 or uint32, got " + keys.format);
  check_length(keys.length);
  IdSource ids{keys.data(), keys.type, keys.has_nulls() ? keys.valid_bits() : nullptr, keys.offset, bias};
  ValX vx;
  vx.op = xop;
  vx.arg = xarg;
  // agg: 0 sum, 1 count rows, 2 min, 3 max, 4 mean, 5 count values (ArrowMetalC.swift:416-469).
  static const int op_of[6] = {0, 1, 4, 5, 3, 2};
  if (agg == 1) return aggregate(ids, keys.length, key_count, nullptr, 1, nullptr);
  if (!values) fail_invalid("values required for this aggregation");
  const int vt = values->type;
  const bool int_sum = type_is_int(vt);
  // The value types each agg takes there: sums and means of integers, the float32 sum (a float64
  // result), min / max of 32-bit or narrower types, counts of anything. Every other pairing is the
  // error "group-by agg N on <format>".
  bool taken = agg == 5 || ((agg == 0 || agg == 4) && int_sum) || (agg == 0 && vt == T_F32) ||
               ((agg == 2 || agg == 3) && type_is_primitive(vt) && values->width() <= 4);
  if (agg < 0 || agg > 5 || !taken) {
    if ((agg == 2 || agg == 3) && type_is_primitive(vt))
      fail_unsupported("group-by min/max needs a 32-bit or narrower type");
    fail_unsupported("group-by agg " + std::to_string(agg) + " on " + values->format);
  }
  if (values->length != keys.length) fail_length_mismatch(keys.length, values->length);
  if (agg == 0 && vt == T_F32 && xop == VX_NONE) return sum_f32_as_f64(ids, keys.length, key_count, *values);
  return aggregate(ids, keys.length, key_count, values, op_of[agg], nullptr, vx);
}

// ---------------------------------------------------------------- dense group-by without allocation
//
// am_group_by_noalloc: the integer aggregates of am_group_by written into caller buffers, for callers
// that cannot allocate device memory. Its only memory besides the inputs and outputs is `scratch`, a
// fixed array in every GPU's global memory that holds the XCDs' partial tables [8][K][2], the merged
// table [K][2] and K validity bytes for at most NOALLOC_MAX_KEYS groups (about 9 MiB, held by each GPU
// the library runs on until the process ends). Every call on a GPU shares that GPU's array: calls that
// overlap on one GPU race; each call initializes what it reads (fill kernels, no memset).
constexpr long long NOALLOC_MAX_KEYS = 65536;
constexpr size_t NOALLOC_G8_WORDS = size_t(NXCD) * NOALLOC_MAX_KEYS * 2;
constexpr size_t NOALLOC_TAB_WORDS = size_t(NOALLOC_MAX_KEYS) * 2;
constexpr size_t NOALLOC_VB_WORDS = size_t(NOALLOC_MAX_KEYS) / 8;
static __device__ __align__(256) unsigned long long scratch[NOALLOC_G8_WORDS + NOALLOC_TAB_WORDS + NOALLOC_VB_WORDS];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64] = {};

// The device address of this GPU's `scratch` (the host-side name is only a stand-in for it), looked up
// once per device with that device current.
unsigned long long* scratch_words() {
  int dev = 0;
  AM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) fail("am_group_by_noalloc: device " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    AM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail("am_group_by_noalloc: no device address for the workspace");
    g_scratch_ptr[dev] = p;
  }
  return static_cast<unsigned long long*>(g_scratch_ptr[dev]);
}

template <class P>
void noalloc_run(const AggArgs& a, int kind, int vtype, void* out_values, unsigned long long* out_valid,
                 unsigned long long* out_nulls, unsigned long long* ws, hipStream_t s) {
  unsigned long long* g8 = ws;
  unsigned long long* tab = ws + NOALLOC_G8_WORDS;
  uint8_t* vb = reinterpret_cast<uint8_t*>(ws + NOALLOC_G8_WORDS + NOALLOC_TAB_WORDS);
  run_agg_into<P>(a, s, "gb_noalloc", g8, tab);
  const long long K = a.K;
  hipLaunchKernelGGL(gb_finalize_simple, dim3(uint32_t((K + 255) / 256)), dim3(256), 0, s, tab, P::W, K, kind, vtype,
                     out_values, vb);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(gb_agg_init<CountP<false>>, dim3(1), dim3(64), 0, s, out_nulls, 1);
  AM_HIP(hipGetLastError());
  hipLaunchKernelGGL(pack_validity, dim3(grid_for(K, 4)), dim3(256), 0, s, vb, K, out_valid, out_nulls);
  AM_HIP(hipGetLastError());
}
}  // namespace

void group_by_noalloc(const Array& keys, int64_t key_count, int agg, const Array* values, void* out_values,
                      uint64_t* out_validity, int64_t* out_null_count) {
  if (key_count <= 0 || key_count > int64_t(UINT32_MAX) / 4) fail_invalid("keyCount out of range");
  if (key_count > NOALLOC_MAX_KEYS)
    fail_invalid("am_group_by_noalloc: key_count " + std::to_string(key_count) + " exceeds the " +
                 std::to_string(NOALLOC_MAX_KEYS) + " groups of its fixed workspace");
  if (keys.type != T_I32 && keys.type != T_I64 && keys.type != T_U32)
    fail("am_group_by: keys must be int32, int64 or uint32, got " + keys.format);
  check_length(keys.length);
  if (agg < 0 || agg > 5) fail_unsupported("group-by agg " + std::to_string(agg));
  if (agg != 1) {
    if (!values) fail_invalid("values required for this aggregation");
    const int vt = values->type;
    if (agg == 0 && vt == T_F32)
      fail_unsupported("am_group_by_noalloc does not take the float32 sum (its exact sum needs workspace "
                       "that grows with the groups); use am_group_by");
    const bool taken = agg == 5 || ((agg == 0 || agg == 4) && type_is_int(vt)) ||
                       ((agg == 2 || agg == 3) && type_is_primitive(vt) && values->width() <= 4);
    if (!taken) {
      if ((agg == 2 || agg == 3) && type_is_primitive(vt))
        fail_unsupported("group-by min/max needs a 32-bit or narrower type");
      fail_unsupported("group-by agg " + std::to_string(agg) + " on " + values->format);
    }
    if (values->length != keys.length) fail_length_mismatch(keys.length, values->length);
  }
  if (!out_values || !out_validity || !out_null_count) fail_invalid("am_group_by_noalloc: NULL output buffer");
  unsigned long long* ws = scratch_words();
  hipStream_t s = stream();
  IdSource ids{keys.data(), keys.type, keys.has_nulls() ? keys.valid_bits() : nullptr, keys.offset};
  AggArgs a = make_args(ids, agg == 1 ? nullptr : values, keys.length, key_count);
  auto* ov = reinterpret_cast<unsigned long long*>(out_validity);
  auto* on = reinterpret_cast<unsigned long long*>(out_null_count);
  switch (agg) {
    case 1: noalloc_run<CountP<false>>(a, 3, T_I64, out_values, ov, on, ws, s); break;
    case 5: noalloc_run<CountP<true>>(a, 3, T_I64, out_values, ov, on, ws, s); break;
    case 0: noalloc_run<SumCountP>(a, 0, values->type, out_values, ov, on, ws, s); break;
    case 4: noalloc_run<SumCountP>(a, type_is_signed(values->type) ? 1 : 2, values->type, out_values, ov, on, ws, s); break;
    default: noalloc_run<MinMaxP>(a, agg == 2 ? 4 : 5, values->type, out_values, ov, on, ws, s); break;
  }
  AM_HIP(hipStreamSynchronize(s));
}

Array* unique_rows(Array* const* columns, int64_t count) {
  if (count >= 1) {
    check_length(columns[0]->length);
    // The representative rows come from the row numbers 0 .. n-1 (GroupByKeys.swift:126-149, 436-438).
    check_index_rows(columns[0]->length, "row indices");
  }
  GroupBy* g = group_by_keys_impl(columns, count, true);
  Array* out = nullptr;
  try {
    // The representative rows sorted ascending: keep-first order (Executor.swift:134-148).
    out = new_array("I", T_U32, g->K);
    out->values = device_alloc(size_t(std::max<int64_t>(g->K, 1)) * 4);
    out->null_count = 0;
    if (g->K > 0)
      argsort_device(g->rep->ptr, T_U32, nullptr, 0, g->K, false, 0, 0, static_cast<uint32_t*>(out->values->ptr),
                     stream());
    // argsort gives the permutation; the rows themselves are rep[perm]: gather them.
    if (g->K > 0) {
      Array* perm = out;
      Array* rep = new_array("I", T_U32, g->K);
      rep->values = g->rep;
      out = take(*rep, *perm);
      release_array(rep);
      release_array(perm);
    } else {
      sync();
    }
  } catch (...) {
    if (out) release_array(out);
    group_by_release(g);
    throw;
  }
  group_by_release(g);
  return out;
}

}  // namespace am

namespace am {

// Arrow `unique` with the value order spelled out (ArrowMetalC_Options.swift:286-297,
// Kernels/UniqueOrder.swift:117-125, Kernels/Unique.swift:46-51). order 0: first appearance, the null
// kept once at the position of the first null row. order 1: ascending, nulls dropped.
Array* unique_values(const Array& a, int order) {
  if (order == 0) {
    check_length(a.length);
    if (a.length > 0) check_index_rows(a.length, "unique");   // Kernels/UniqueOrder.swift:97
    Array* col = const_cast<Array*>(&a);
    Array* rows = unique_rows(&col, 1);
    Array* out = nullptr;
    try {
      out = take(a, *rows);
    } catch (...) {
      release_array(rows);
      throw;
    }
    release_array(rows);
    return out;
  }
  Array* col = const_cast<Array*>(&a);
  GroupBy* g = group_by_keys_impl(&col, 1, true);
  Array* out = nullptr;
  try {
    const int64_t k = g->K - (g->has_null ? 1 : 0);   // the null group is last
    Array* rep = new_array("I", T_U32, k);
    rep->values = g->rep;
    rep->null_count = 0;
    try {
      out = take(a, *rep);
    } catch (...) {
      release_array(rep);
      throw;
    }
    release_array(rep);
  } catch (...) {
    group_by_release(g);
    throw;
  }
  group_by_release(g);
  return out;
}

}  // namespace am
