// The levels of the adaptive MSD radix argsort of 64-bit keys, shared by the plain argsort (sort.hip)
// and the batched plan (batched.hip).
//
// Rows are in buckets (Bucket: a range of the value block, in one of two key / row buffers, or in the
// first key buffer with implicit rows, whose keys share their digits from the top down to `level`).
// Level by level down the digits, every bucket larger than `local_max` rows is partitioned stably by
// the next digit (onesweep tiles whose look-back stops at their bucket's first tile) into the other
// buffer at the same positions; a bucket whose rows all share the digit is carried without moving; a
// run of sub-buckets whose keys differ in no lower bit leaves as one copy. Sub-buckets of at most
// `local_max` rows leave for the bucket sorts, which sort them in LDS by their remaining digits and
// write their rows to their final slots: one workgroup per bucket (hyb_local), or, for small sibling
// buckets packed into groups, one workgroup per group (hyb_group); buckets with long runs of equal
// keys are queued for LSD passes in LDS (hyb_local_lsd).
//
// The kernels take their counts from device memory (the `lc` counters and an MsdCtl), so a level can
// run from host read-backs (the plain argsort sizes each launch from them) or blind, inside a graph
// capture (the batched plan launches every digit's level with grids sized for the worst case; work-
// groups past the work exit, and a level of an inactive digit carries every bucket).
#pragma once

#include "sort_common.hpp"

namespace am {
namespace {

constexpr uint32_t LVL_CHUNK = 16384;   // rows per histogram chunk of a level
constexpr uint32_t LOCAL_MAX = 4096;    // largest bucket the LDS sort takes (local_max is at most this)
static_assert(LOCAL_MAX <= 4096u, "positions in a bucket fit the 12 low bits of the bucket sort's words");
constexpr uint32_t SUB_MAX = 64;       // longest digit run the bucket sorts rank by counting
constexpr uint32_t GROUP_SPAN = 16;    // a group's buckets have their digit in one window of this many values
constexpr uint32_t GROUP_BINS = GROUP_SPAN * RADIX;
constexpr uint32_t GROUP_ROWS = 2048;  // rows of a group at most (the group sort at 8 keys per thread)

// A bucket: rows [start, start + size) of the value block, in buffer `buf`, all sharing the digits from
// the top down to `level`.
struct Bucket {
  uint32_t start, size, buf, level;
};
// Bucket::buf: 0 = keys0 / rows0, 1 = keys1 / rows1, 2 = keys0 with implicit rows: a whole batch in
// input order, whose row at index i is i - start (the batched plan's first level reads no row buffer).
constexpr uint32_t BUF_IMPLICIT = 2;
template <typename K>
__device__ __forceinline__ const K* bucket_keys(const Bucket& b, const K* keys0, const K* keys1) {
  return b.buf == 1u ? keys1 : keys0;
}
// Null for implicit rows.
__device__ __forceinline__ const uint32_t* bucket_rows(const Bucket& b, const uint32_t* rows0, const uint32_t* rows1) {
  return b.buf == BUF_IMPLICIT ? nullptr : b.buf == 1u ? rows1 : rows0;
}
// The rows at indices src[j] of a bucket starting at `start`: from `rows`, or implicit (rows null).
template <int LI>
__device__ __forceinline__ void load_rows(const uint32_t* rows, uint32_t start, const uint32_t (&src)[LI],
                                          uint32_t (&r)[LI]) {
  if (rows) {
#pragma unroll
    for (int j = 0; j < LI; j++) r[j] = rows[src[j]];
  } else {
#pragma unroll
    for (int j = 0; j < LI; j++) r[j] = src[j] - start;
  }
}
constexpr uint32_t LEVEL_BITS = 0xFFu;    // Bucket::level: the level ...
constexpr uint32_t GROUP_FLAG = 0x100u;   // ... and, in the bucket sort's list, a group of small buckets
constexpr int SPAN_SHIFT = 12;            // ... whose digits share an aligned window of 1 << (level >> 12)
struct LvlChunk {
  uint32_t bucket, start, end, pad;
};
struct LvlTile {
  uint32_t bucket, start, end, head;   // head: the bucket's first tile
};
// What the level kernels read besides their lists: the active digits (bit d: digit d varies over
// the value keys), where the value block starts in the output, and the run's look-back epoch.
struct MsdCtl {
  uint32_t active_mask, value_base, epoch, pad;
  // The ORDER BY of whole tables by a 64-bit integer key also takes the sorted key column from the
  // bucket sorts: key_out[value_base + i] = the mapped key of output i XOR key_xor (its original
  // bits). Null when not asked for.
  unsigned long long* key_out;
  unsigned long long key_xor;
};

// The sorted key column's output as one bucket sort sees it (already offset to the value block).
struct KeyOut {
  unsigned long long* p;
  unsigned long long x;
};

__device__ __forceinline__ KeyOut key_out_of(const MsdCtl* mc) {
  return KeyOut{mc->key_out ? mc->key_out + mc->value_base : nullptr, mc->key_xor};
}

// Beside the rows a bucket sort writes (dest[j] != ~0u), their keys, when the key column is asked for:
// load_keys reads keys_in[src[j]] (src[j] is always a position of the bucket, so every load stays in
// bounds; issued with the row loads, before any store), write_keys stores them.
template <typename K, int LI, bool KEYS>
__device__ __forceinline__ void load_keys(const KeyOut& ko, const K* keys_in, const uint32_t (&src)[LI],
                                          unsigned long long (&v)[LI]) {
  if constexpr (KEYS && sizeof(K) == 8) {
    if (!ko.p) return;   // uniform
#pragma unroll
    for (int j = 0; j < LI; j++) v[j] = keys_in[src[j]];
  }
}

template <typename K, int LI, bool KEYS>
__device__ __forceinline__ void write_keys(const KeyOut& ko, const uint32_t (&dest)[LI], const unsigned long long (&v)[LI]) {
  if constexpr (KEYS && sizeof(K) == 8) {
    if (!ko.p) return;   // uniform
#pragma unroll
    for (int j = 0; j < LI; j++)
      if (dest[j] != ~0u) ko.p[dest[j]] = v[j] ^ ko.x;
  }
}

// Device counters of the levels (one small array). The host reads back the first LC_READ words. The
// two counters that every workgroup of a kernel bumps each sit on a cache line of their own: sharing a
// line with words the same workgroups read (the tile count) made the level pass 1.8x slower.
constexpr int LC_LIST = 0;        // [0], [1]: entries of the two large-bucket lists
constexpr int LC_LOCAL = 2;       // entries of the bucket sort's list
constexpr int LC_CHUNKS = 3;      // chunks of the current level
constexpr int LC_TILES = 4;       // tiles of the current level
constexpr int LC_MAXLOCAL = 5;    // largest bucket the bucket sort sorts (not counting copies)
constexpr int LC_OVERFLOW = 6;    // the bucket sort's list ran out of room
constexpr int LC_QLEN = 7;        // buckets queued for the LSD bucket sort
constexpr int LC_ROWS = 8;        // rows in the next level's large buckets (bounds its grids)
constexpr int LC_GROUPS = 9;      // groups of small buckets, at the far end of the bucket sort's list
constexpr int LC_READ = 16;
constexpr int LC_NEXT_TILE = 32;  // the next tile of the current level to take (own 128-byte line)
constexpr int LC_QNEXT = 64;      // the next queued bucket to take (own 128-byte line)
constexpr int LC_WORDS = 96;

// Appends one bucket to the bucket sort's list. `sorted` = its rows need no sorting (equal keys).
__device__ __forceinline__ void push_local(Bucket* local, uint32_t local_cap, uint32_t* lc, Bucket b, bool sorted) {
  const uint32_t i = atomicAdd(&lc[LC_LOCAL], 1u);
  if (i < local_cap) local[i] = b;
  else atomicOr(&lc[LC_OVERFLOW], 1u);
  if (!sorted) atomicMax(&lc[LC_MAXLOCAL], b.size);
}
// Groups of small buckets fill the list from its far end (entry local_cap - 1 down, counted by
// LC_GROUPS). The list ran out of room when LC_LOCAL + LC_GROUPS > local_cap.
__device__ __forceinline__ bool list_overflow(const uint32_t* lc, uint32_t local_cap) {
  return lc[LC_OVERFLOW] != 0u || uint64_t(lc[LC_LOCAL]) + lc[LC_GROUPS] > local_cap;
}

// One workgroup: where each large bucket's chunks and tiles start in a level's tables (list `cur`
// order; bofs[2b], bofs[2b + 1]); the other list's counter, the tile counter and the row count reset.
// lvl_fill writes the tables.
template <int ITEMS>
__global__ __launch_bounds__(BLOCK) void lvl_layout(const Bucket* large, int cur, uint32_t* bofs, uint32_t* lc) {
  __shared__ uint32_t tmp[WAVES];
  constexpr uint32_t TILE = BLOCK * ITEMS;
  const uint32_t nb = lc[LC_LIST + cur];
  uint32_t coff = 0, toff = 0;
  for (uint32_t b0 = 0; b0 < nb; b0 += BLOCK) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t size = b < nb ? large[b].size : 0u;
    const uint32_t nc = (size + LVL_CHUNK - 1) / LVL_CHUNK, nt = (size + TILE - 1) / TILE;
    uint32_t tc, tt;
    const uint32_t c0 = coff + block_excl_scan(nc, tmp, tc);
    const uint32_t t0 = toff + block_excl_scan(nt, tmp, tt);
    if (b < nb) {
      bofs[2 * b] = c0;
      bofs[2 * b + 1] = t0;
    }
    coff += tc;
    toff += tt;
  }
  if (threadIdx.x == 0) {
    lc[LC_CHUNKS] = coff;
    lc[LC_TILES] = toff;
    lc[LC_NEXT_TILE] = 0u;
    lc[LC_LIST + (cur ^ 1)] = 0u;
    lc[LC_ROWS] = 0u;
  }
}

// The chunks and tiles of each large bucket, a bucket per workgroup in turn, written by all threads;
// its histogram and OR / AND cleared.
template <int ITEMS>
__global__ __launch_bounds__(BLOCK) void lvl_fill(const Bucket* large, int cur, const uint32_t* lc,
                                                  const uint32_t* bofs, LvlChunk* chunks, LvlTile* tiles,
                                                  uint32_t* bucket_hist, unsigned long long* bucket_or,
                                                  unsigned long long* bucket_and) {
  constexpr uint32_t TILE = BLOCK * ITEMS;
  static_assert(RADIX == BLOCK, "a thread per digit value");
  const uint32_t nb = lc[LC_LIST + cur];
  for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const Bucket bk = large[b];
    bucket_hist[size_t(b) * RADIX + threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
      bucket_or[b] = 0ull;
      bucket_and[b] = ~0ull;
    }
    const uint32_t c0 = bofs[2 * b], t0 = bofs[2 * b + 1];
    const uint32_t nc = (bk.size + LVL_CHUNK - 1) / LVL_CHUNK, nt = (bk.size + TILE - 1) / TILE;
    for (uint32_t c = threadIdx.x; c < nc; c += BLOCK)
      chunks[c0 + c] = {b, bk.start + c * LVL_CHUNK, bk.start + min(bk.size, (c + 1u) * LVL_CHUNK), 0u};
    for (uint32_t t = threadIdx.x; t < nt; t += BLOCK)
      tiles[t0 + t] = {b, bk.start + t * TILE, bk.start + min(bk.size, (t + 1u) * TILE), t0};
  }
}

// Digit d's histogram, and the OR / AND of the keys, of every large bucket, a chunk per workgroup in
// turn (nothing when digit d is inactive).
template <typename K>
__global__ __launch_bounds__(BLOCK) void lvl_hist(const K* keys0, const K* keys1, const Bucket* large,
                                                  const LvlChunk* chunks, const uint32_t* lc, const MsdCtl* mc,
                                                  int d, uint32_t* bucket_hist, unsigned long long* bucket_or,
                                                  unsigned long long* bucket_and) {
  __shared__ uint32_t h[RADIX];
  __shared__ unsigned long long s_or, s_and;
  if (!((mc->active_mask >> d) & 1u)) return;   // uniform over the grid
  const uint32_t nch = lc[LC_CHUNKS];
  for (uint32_t ci = blockIdx.x; ci < nch; ci += gridDim.x) {
    const LvlChunk c = chunks[ci];
    const K* keys = bucket_keys(large[c.bucket], keys0, keys1);
    h[threadIdx.x] = 0u;
    if (threadIdx.x == 0) {
      s_or = 0ull;
      s_and = ~0ull;
    }
    __syncthreads();
    unsigned long long o = 0ull, a = ~0ull;
    constexpr int U = 8;
    for (uint32_t base = c.start + threadIdx.x; base < c.end; base += BLOCK * U) {
      K k[U];
#pragma unroll
      for (int u = 0; u < U; u++) k[u] = keys[min(base + uint32_t(u) * BLOCK, c.end - 1u)];
#pragma unroll
      for (int u = 0; u < U; u++) {
        // Rows past the end read the chunk's last row, which does not change the OR / AND.
        o |= (unsigned long long)k[u];
        a &= (unsigned long long)k[u];
        if (base + uint32_t(u) * BLOCK < c.end) atomicAdd(&h[uint32_t(k[u] >> (8 * d)) & 0xFFu], 1u);
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      o |= __shfl_xor(o, off, 64);
      a &= __shfl_xor(a, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {   // waves that read no row hold the identities
      atomicOr(&s_or, o);
      atomicAnd(&s_and, a);
    }
    __syncthreads();
    const uint32_t v = h[threadIdx.x];
    if (v) atomicAdd(&bucket_hist[size_t(c.bucket) * RADIX + threadIdx.x], v);
    if (threadIdx.x == 0) {
      atomicOr(&bucket_or[c.bucket], s_or);
      atomicAnd(&bucket_and[c.bucket], s_and);
    }
    __syncthreads();
  }
}

// Inclusive running maximum over the threads of a workgroup (tmp: WAVES words of LDS).
__device__ __forceinline__ uint32_t block_incl_max(uint32_t v, uint32_t* tmp) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(v, o, 64);
    if (lane >= o) v = max(v, u);
  }
  if (lane == 63) tmp[w] = v;
  __syncthreads();
  uint32_t pre = 0;
#pragma unroll
  for (int k = 0; k < WAVES; k++) pre = k < w ? max(pre, tmp[k]) : pre;
  __syncthreads();
  return max(v, pre);
}

// Each large bucket, a workgroup per bucket in turn: the start of each of its digit-d sub-buckets
// (bucket_base), and where each goes: to the bucket sort's list (at most local_max rows; or, of any
// size, equal keys: the bucket's keys differ in no bit below digit d, or no active digit is left
// below d), or to the next level's list. A bucket whose rows all share digit d (every bucket, when
// digit d is inactive) is carried (carried[b] = 1): it keeps its buffer and the pass skips it.
// Neighbouring sub-buckets are adjacent in memory and in key order, so several leave as one entry:
// - a run of sub-buckets of equal keys, as one copy;
// - sub-buckets of at most small_max rows, as a group: the sub-buckets of the widest aligned window of
//   digits (1, 2, 4, 8 or GROUP_SPAN values) around each that holds at most group_max rows and only
//   such sub-buckets, sorted together by the digits from d down (level d + 1).
__global__ __launch_bounds__(BLOCK) void lvl_plan(const Bucket* large, int cur, const uint32_t* bucket_hist,
                                                  const unsigned long long* bucket_or,
                                                  const unsigned long long* bucket_and, const MsdCtl* mc, int d,
                                                  uint32_t local_max, uint32_t small_max, uint32_t group_max,
                                                  uint32_t* bucket_base, uint32_t* carried, Bucket* next,
                                                  Bucket* local, uint32_t local_cap, uint32_t* lc) {
  enum : uint32_t { SUB_EMPTY = 0, SUB_COPY = 1, SUB_SMALL = 2, SUB_SINGLE = 3, SUB_LARGE = 4 };
  __shared__ uint32_t tmp[WAVES];
  __shared__ uint2 s_sub[RADIX];   // (rows, start) of each non-empty sub-bucket of equal keys
  __shared__ uint32_t s_base[3];   // where this bucket's entries go in the three lists
  const uint32_t nb = lc[LC_LIST + cur];
  const uint32_t mask = mc->active_mask;
  const bool active = (mask >> d) & 1u;
  const bool last = (mask & ((1u << d) - 1u)) == 0u;
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
    const Bucket bk = large[b];
    uint32_t v, base;
    bool carry;
    if (active) {
      v = bucket_hist[size_t(b) * RADIX + threadIdx.x];
      uint32_t total;
      base = bk.start + block_excl_scan(v, tmp, total);
      bucket_base[size_t(b) * RADIX + threadIdx.x] = base;
      carry = __syncthreads_count(v != 0u ? 1 : 0) == 1;
    } else {   // one sub-bucket: the whole bucket, kept by thread 0
      v = threadIdx.x == 0 ? bk.size : 0u;
      base = bk.start;
      carry = true;
    }
    if (threadIdx.x == 0) carried[b] = carry ? 1u : 0u;
    const uint32_t buf = carry ? bk.buf : (bk.buf == 1u ? 0u : 1u);   // a pass writes the other buffer
    uint32_t kind = SUB_EMPTY;
    if (v != 0u) {
      const unsigned long long diff = active ? (bucket_or[b] ^ bucket_and[b]) : ~0ull;
      const bool equal_below = (diff & ((1ull << (8 * d)) - 1ull)) == 0ull;
      kind = (last || equal_below) ? SUB_COPY : v <= small_max ? SUB_SMALL : v <= local_max ? SUB_SINGLE : SUB_LARGE;
    }
    // Groups: the rows of each aligned window (any other sub-bucket weighs more than a group holds);
    // the first non-empty sub-bucket of the chosen window stands for it.
    bool group_head = false;
    uint32_t group_rows = 0, group_span = 1;
    {
      const uint32_t w1 = kind == SUB_SMALL ? v : (v != 0u ? group_max + 1u : 0u);
      const uint32_t w2 = w1 + __shfl_xor(w1, 1, 64);
      const uint32_t w4 = w2 + __shfl_xor(w2, 2, 64);
      const uint32_t w8 = w4 + __shfl_xor(w4, 4, 64);
      const uint32_t w16 = w8 + __shfl_xor(w8, 8, 64);
      static_assert(GROUP_SPAN == 16 && GROUP_BINS == RADIX << 4, "window sums up to GROUP_SPAN digits");
      const unsigned long long nonempty = __ballot(v != 0u);
      if (kind == SUB_SMALL) {
        group_span = 16;
        group_rows = w16;
        if (group_rows > group_max) { group_span = 8; group_rows = w8; }
        if (group_rows > group_max) { group_span = 4; group_rows = w4; }
        if (group_rows > group_max) { group_span = 2; group_rows = w2; }
        if (group_rows > group_max) { group_span = 1; group_rows = w1; }
        const uint32_t first = lane & ~(group_span - 1u);
        group_head = (nonempty & ((1ull << lane) - 1ull) & ~((1ull << first) - 1ull)) == 0ull;
      }
    }
    // Copies: the non-empty sub-buckets compacted in digit order; a run of equal-key ones leaves as one
    // (thread i stands for the run that ends at compacted sub-bucket i).
    bool copy_end = false;
    uint32_t copy_start = 0, copy_rows = 0;
    uint32_t nne;
    const uint32_t ci = block_excl_scan(v != 0u ? 1u : 0u, tmp, nne);
    if (v != 0u) s_sub[ci] = make_uint2(kind == SUB_COPY ? v : 0u, base);
    if (__syncthreads_or(kind == SUB_COPY ? 1 : 0)) {   // (orders the writes above)
      const uint32_t i = threadIdx.x;
      uint2 me = make_uint2(0u, 0u), prev = me, next_sub = me;
      if (i < nne) {
        me = s_sub[i];
        if (i > 0) prev = s_sub[i - 1];
        if (i + 1 < nne) next_sub = s_sub[i + 1];
      }
      const bool copy = me.x != 0u;
      const bool run_start = copy && prev.x == 0u;
      copy_end = copy && next_sub.x == 0u;
      copy_start = block_incl_max(run_start ? me.y + 1u : 0u, tmp) - 1u;
      copy_rows = me.y + me.x - copy_start;
    }
    // Every entry of this bucket into its list with one counter bump per list.
    const bool single = kind == SUB_SINGLE, to_next = kind == SUB_LARGE;
    uint32_t n_local, n_group, n_next, n_rows;
    const uint32_t r_local = block_excl_scan((single ? 1u : 0u) + (copy_end ? 1u : 0u), tmp, n_local);
    const uint32_t r_group = block_excl_scan(group_head ? 1u : 0u, tmp, n_group);
    const uint32_t r_next = block_excl_scan(to_next ? 1u : 0u, tmp, n_next);
    block_excl_scan(to_next ? v : 0u, tmp, n_rows);
    const uint32_t biggest = block_incl_max(single ? v : 0u, tmp);   // thread BLOCK - 1 holds the maximum
    if (threadIdx.x == 0) {
      s_base[0] = n_local ? atomicAdd(&lc[LC_LOCAL], n_local) : 0u;
      s_base[1] = n_group ? atomicAdd(&lc[LC_GROUPS], n_group) : 0u;
      s_base[2] = n_next ? atomicAdd(&lc[LC_LIST + (cur ^ 1)], n_next) : 0u;
      if (n_rows) atomicAdd(&lc[LC_ROWS], n_rows);
    }
    if (threadIdx.x == BLOCK - 1 && biggest) atomicMax(&lc[LC_MAXLOCAL], biggest);
    __syncthreads();
    uint32_t at = s_base[0] + r_local;
    if (single) {
      if (at < local_cap) local[at] = Bucket{base, v, buf, uint32_t(d)};
      at++;
    }
    if (copy_end && at < local_cap) local[at] = Bucket{copy_start, copy_rows, buf, 0u};   // level 0: no digit left
    if (group_head) {
      const uint32_t g = s_base[1] + r_group;
      if (g < local_cap)
        local[local_cap - 1u - g] =
            Bucket{base, group_rows, buf, (uint32_t(d) + 1u) | GROUP_FLAG | (uint32_t(__builtin_ctz(group_span)) << SPAN_SHIFT)};
    }
    if (to_next) next[s_base[2] + r_next] = Bucket{base, v, buf, uint32_t(d)};
    __syncthreads();
  }
}

// A level's pass: every large bucket that is not carried partitioned stably by digit d from its buffer
// into the other one, at the same positions (nothing when digit d is inactive). Tile ids come from a
// counter, so a tile's look-back only ever waits on tiles already taken by running workgroups; tiles
// of carried buckets (whose whole look-back chain is skipped) are dropped.
template <typename K, int ITEMS>
struct LvlLds {
  uint32_t wc[WAVES][WC_STRIDE];
  uint32_t tile_total[WC_STRIDE];
  uint32_t bin_start[RADIX];
  uint32_t bin_base[WC_STRIDE];
  uint32_t tmp[WAVES];
  uint32_t tile;
  uint8_t digit[BLOCK * ITEMS];
  K stage[BLOCK * ITEMS];
};

template <typename K, int ITEMS>
__device__ __forceinline__ void lvl_tile(LvlLds<K, ITEMS>& L, uint32_t tile, const Bucket* large,
                                         const LvlTile* tiles, const uint32_t* carried, const uint32_t* bucket_base,
                                         unsigned long long* status, uint32_t epoch, int d, K* keys0,
                                         uint32_t* rows0, K* keys1, uint32_t* rows1) {
  constexpr int TILE = BLOCK * ITEMS;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const LvlTile t = tiles[tile];
  if (carried[t.bucket]) return;   // uniform over the workgroup
  const Bucket bk = large[t.bucket];
  const bool from1 = bk.buf == 1u;
  const K* in_keys = from1 ? keys1 : keys0;
  const uint32_t* in_rows = from1 ? rows1 : rows0;
  K* out_keys = from1 ? keys0 : keys1;
  uint32_t* out_rows = from1 ? rows0 : rows1;
  const int shift = 8 * d;
  K keys[ITEMS];
  uint32_t rows[ITEMS], bins[ITEMS];
  const uint32_t base_e = t.start + uint32_t(w) * 64u * ITEMS + uint32_t(lane);
  if (bk.buf == BUF_IMPLICIT) {
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {   // every load first (lanes past the tile read its last row)
      const uint32_t e = min(base_e + uint32_t(j) * 64u, t.end - 1u);
      keys[j] = in_keys[e];
      rows[j] = e - bk.start;
    }
  } else {
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      const uint32_t e = min(base_e + uint32_t(j) * 64u, t.end - 1u);
      keys[j] = in_keys[e];
      rows[j] = in_rows[e];
    }
  }
#pragma unroll
  for (int j = 0; j < ITEMS; j++)
    bins[j] = base_e + uint32_t(j) * 64u < t.end ? uint32_t(keys[j] >> shift) & 0xFFu : uint32_t(BIN_NONE);
  if (t.end - t.start == uint32_t(TILE)) rank_tile<ITEMS, 8>(bins, L.wc[w], lane);
  else rank_tile<ITEMS, 9>(bins, L.wc[w], lane);
  __syncthreads();
  const uint32_t vcount = tile_bin_layout<RADIX>(L.wc, L.tile_total, L.bin_start, L.tmp);
  look_back<RADIX>(status, tile, epoch, L.tile_total, L.bin_base, t.head);
  __syncthreads();
  L.bin_base[threadIdx.x] += bucket_base[size_t(t.bucket) * RADIX + threadIdx.x];
  __syncthreads();
  scatter_values<K, ITEMS, true>(keys, rows, bins, L.wc[w], L.bin_start, L.bin_base, L.digit, L.stage, vcount,
                                 out_keys, out_rows);
}

// One tile per workgroup: exactly the first LC_TILES workgroups take a tile each; the rest exit.
template <typename K, int ITEMS>
__global__ __launch_bounds__(BLOCK, 3) void lvl_pass(const Bucket* large, const LvlTile* tiles, uint32_t* lc,
                                                  const uint32_t* carried, const uint32_t* bucket_base,
                                                  unsigned long long* status, const MsdCtl* mc, uint32_t eoff,
                                                  int d, K* keys0, uint32_t* rows0, K* keys1, uint32_t* rows1) {
  __shared__ LvlLds<K, ITEMS> L;
  if (blockIdx.x >= lc[LC_TILES] || !((mc->active_mask >> d) & 1u)) return;
  for (int i = threadIdx.x; i < WAVES * WC_STRIDE; i += BLOCK) (&L.wc[0][0])[i] = 0u;
  if (threadIdx.x == 0) L.tile = atomicAdd(&lc[LC_NEXT_TILE], 1u);
  __syncthreads();
  lvl_tile<K, ITEMS>(L, L.tile, large, tiles, carried, bucket_base, status, mc->epoch + eoff, d, keys0, rows0, keys1,
                     rows1);
}

// Persistent form (blind levels, where the tile count is not known on the host): each workgroup takes
// tiles until none is left.
template <typename K, int ITEMS>
__global__ __launch_bounds__(BLOCK, 3) void lvl_pass_loop(const Bucket* large, const LvlTile* tiles, uint32_t* lc,
                                                       const uint32_t* carried, const uint32_t* bucket_base,
                                                       unsigned long long* status, const MsdCtl* mc, uint32_t eoff,
                                                       int d, K* keys0, uint32_t* rows0, K* keys1, uint32_t* rows1) {
  __shared__ LvlLds<K, ITEMS> L;
  if (!((mc->active_mask >> d) & 1u)) return;   // uniform over the grid
  const uint32_t ntiles = lc[LC_TILES];
  const uint32_t epoch = mc->epoch + eoff;
  while (true) {
    for (int i = threadIdx.x; i < WAVES * WC_STRIDE; i += BLOCK) (&L.wc[0][0])[i] = 0u;
    if (threadIdx.x == 0) L.tile = atomicAdd(&lc[LC_NEXT_TILE], 1u);
    __syncthreads();
    const uint32_t tile = L.tile;
    if (tile >= ntiles) return;   // uniform: every workgroup ends here once the tiles run out
    lvl_tile<K, ITEMS>(L, tile, large, tiles, carried, bucket_base, status, epoch, d, keys0, rows0, keys1, rows1);
    __syncthreads();
  }
}


// The bucket sort's LDS besides the bucket: the counters of its split.
struct LocalLds {
  uint32_t cnt[RADIX], start[RADIX], cursor[RADIX], tmp[WAVES];
};

// One stable LDS pass of a bucket by digit d: keys and their positions in the bucket (pos[j]) land
// in s_keys / s_pos in (digit, input) order; on return bin_start / tile_total give each digit's run.
template <typename K, int LI>
__device__ __forceinline__ void local_split(const K (&keys)[LI], const uint32_t (&pos)[LI], uint32_t base_e,
                                            uint32_t size, int d, uint32_t (*wc)[WC_STRIDE], uint32_t* tile_total,
                                            uint32_t* bin_start, uint32_t* tmp, K* s_keys, uint16_t* s_pos) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < WAVES * WC_STRIDE; i += BLOCK) (&wc[0][0])[i] = 0u;
  __syncthreads();
  uint32_t bins[LI];
#pragma unroll
  for (int j = 0; j < LI; j++)
    bins[j] = base_e + uint32_t(j) * 64u < size ? uint32_t(keys[j] >> (8 * d)) & 0xFFu : uint32_t(BIN_NONE);
  rank_tile<LI, 9>(bins, wc[w], lane);
  __syncthreads();
  tile_bin_layout<RADIX>(wc, tile_total, bin_start, tmp);
  __syncthreads();
#pragma unroll
  for (int j = 0; j < LI; j++) {
    const uint32_t b = bin_of(bins[j]);
    if (b < RADIX) {
      const uint32_t at = bin_start[b] + wc[w][b] + rank_of(bins[j]);
      s_keys[at] = keys[j];
      s_pos[at] = uint16_t(pos[j]);
    }
  }
}

// One bucket of at most LI * BLOCK rows (rows in input order inside each digit run of the levels
// before) in LDS, its rows written to their final slots. A bucket whose top varying digit is 7 is
// first split stably by it; when no digit run is longer than run_max rows, a row's place inside its
// run is the number of rows there with a smaller key, or an equal key earlier (LDS order is input
// order inside a run). Otherwise stable LSD passes run over every varying digit.
template <typename K, int LI, bool KEYS>
__device__ __forceinline__ void local_lsd(const K* keys_in, const uint32_t* rows_in, uint32_t off, uint32_t size,
                                          uint32_t low_mask, uint32_t run_max, uint32_t* out_idx, const KeyOut& ko,
                                          uint32_t (*wc)[WC_STRIDE], uint32_t* tile_total, uint32_t* bin_start,
                                          uint32_t* tmp, K* s_keys, uint16_t* s_pos) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t base_e = uint32_t(w) * 64u * LI + uint32_t(lane);
  K keys[LI];
  uint32_t pos[LI];
#pragma unroll
  for (int j = 0; j < LI; j++) {
    pos[j] = base_e + uint32_t(j) * 64u;
    keys[j] = keys_in[off + min(pos[j], size - 1u)];
  }
  if ((low_mask >> 7) & 1u) {
    local_split<K, LI>(keys, pos, base_e, size, 7, wc, tile_total, bin_start, tmp, s_keys, s_pos);
    __syncthreads();
    if (!__syncthreads_or(tile_total[threadIdx.x] > run_max ? 1 : 0)) {
      uint32_t dest[LI], src[LI];
#pragma unroll
      for (int j = 0; j < LI; j++) {
        const uint32_t e = base_e + uint32_t(j) * 64u;
        dest[j] = ~0u;
        src[j] = off;
        if (e < size) {
          const K key = s_keys[e];
          const uint32_t dd = uint32_t(key >> 56);
          const uint32_t lo = bin_start[dd], hi = lo + tile_total[dd];
          uint32_t below = 0;
#pragma unroll 4
          for (uint32_t q = lo; q < e; q++) below += s_keys[q] <= key ? 1u : 0u;
#pragma unroll 4
          for (uint32_t q = e + 1u; q < hi; q++) below += s_keys[q] < key ? 1u : 0u;
          dest[j] = off + lo + below;
          src[j] = off + s_pos[e];
        }
      }
      uint32_t r[LI];
      unsigned long long kv[LI];
      load_rows<LI>(rows_in, off, src, r);
      load_keys<K, LI, KEYS>(ko, keys_in, src, kv);
#pragma unroll
      for (int j = 0; j < LI; j++)
        if (dest[j] != ~0u) out_idx[dest[j]] = r[j];
      write_keys<K, LI, KEYS>(ko, dest, kv);
      return;
    }
  }
  for (int d = 0; d < 8; d++) {
    if (!((low_mask >> d) & 1u)) continue;   // uniform
    local_split<K, LI>(keys, pos, base_e, size, d, wc, tile_total, bin_start, tmp, s_keys, s_pos);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < LI; j++) {
      const uint32_t e = min(base_e + uint32_t(j) * 64u, size - 1u);
      keys[j] = s_keys[e];
      pos[j] = s_pos[e];
    }
    __syncthreads();
  }
  uint32_t src[LI], r[LI], dest[LI];
  unsigned long long kv[LI];
#pragma unroll
  for (int j = 0; j < LI; j++) src[j] = off + min(pos[j], size - 1u);
  load_rows<LI>(rows_in, off, src, r);
  load_keys<K, LI, KEYS>(ko, keys_in, src, kv);
#pragma unroll
  for (int j = 0; j < LI; j++) {
    const uint32_t e = base_e + uint32_t(j) * 64u;
    dest[j] = e < size ? off + e : ~0u;
    if (e < size) out_idx[off + e] = r[j];
  }
  write_keys<K, LI, KEYS>(ko, dest, kv);
}

// A group of small sibling buckets (at most GROUP_ROWS rows in all; their digits `top` in one aligned
// window of 1 << span_log values; rows in input order inside each): split with LDS counters by (digit
// top mod the window, digit `sub`, the next varying one), in that order, each row as one word, (key bits
// below digit sub) << 12 | position, its bin beside it; a row's place inside its bin is the number of
// words there below its own. Returns false, having written nothing, when a bin holds more than SUB_MAX
// rows. g: RADIX << span_log words of LDS (counts, then starts, then, after the scatter, each bin's
// end).
template <typename K, int LI, uint32_t SPAN_LOG, bool KEYS>
__device__ __forceinline__ bool local_group(const K* keys_in, const uint32_t* rows_in, uint32_t off, uint32_t size,
                                            int top, int sub, uint32_t* out, const KeyOut& ko, uint32_t* tmp, uint32_t* g,
                                            K* s_words, uint16_t* s_bin) {
  static_assert(RADIX == BLOCK, "a thread per digit value");
  constexpr uint32_t per = 1u << SPAN_LOG;   // bins per thread
  const int ts = 8 * top, ss = 8 * sub;
  uint32_t* mine = g + threadIdx.x * per;
#pragma unroll
  for (uint32_t k = 0; k < per; k++) mine[k] = 0u;
  K keys[LI];
  uint32_t bins[LI];
#pragma unroll
  for (int j = 0; j < LI; j++) {
    keys[j] = keys_in[off + min(threadIdx.x + uint32_t(j) * BLOCK, size - 1u)];
    bins[j] = ((uint32_t(keys[j] >> ts) & (per - 1u)) << 8) | (uint32_t(keys[j] >> ss) & 0xFFu);
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < LI; j++)
    if (threadIdx.x + uint32_t(j) * BLOCK < size) atomicAdd(&g[bins[j]], 1u);
  __syncthreads();
  uint32_t c[per], sum = 0, big = 0;
#pragma unroll
  for (uint32_t k = 0; k < per; k++) {
    c[k] = mine[k];
    sum += c[k];
    big |= c[k] > SUB_MAX ? 1u : 0u;
  }
  uint32_t total;
  uint32_t run = block_excl_scan(sum, tmp, total);
#pragma unroll
  for (uint32_t k = 0; k < per; k++) {
    mine[k] = run;
    run += c[k];
  }
  if (__syncthreads_or(int(big))) return false;
  const unsigned long long low_bits = (1ull << ss) - 1ull;
#pragma unroll
  for (int j = 0; j < LI; j++) {
    const uint32_t p = threadIdx.x + uint32_t(j) * BLOCK;
    if (p < size) {
      const uint32_t slot = atomicAdd(&g[bins[j]], 1u);
      s_words[slot] = K(((static_cast<unsigned long long>(keys[j]) & low_bits) << 12) | p);
      s_bin[slot] = uint16_t(bins[j]);
    }
  }
  __syncthreads();
  uint32_t dest[LI], src[LI];
#pragma unroll
  for (int j = 0; j < LI; j++) {
    const uint32_t p = threadIdx.x + uint32_t(j) * BLOCK;
    dest[j] = ~0u;
    src[j] = off;
    if (p < size) {
      const K word = s_words[p];
      const uint32_t bin = s_bin[p];
      const uint32_t lo = bin ? g[bin - 1u] : 0u, hi = g[bin];
      uint32_t below = 0;
#pragma unroll 4
      for (uint32_t q = lo; q < hi; q++) below += s_words[q] < word ? 1u : 0u;
      dest[j] = off + lo + below;
      src[j] = off + uint32_t(word & 0xFFFu);
    }
  }
  uint32_t r[LI];
  unsigned long long kv[LI];
  load_rows<LI>(rows_in, off, src, r);   // every row and key load before the first store
  load_keys<K, LI, KEYS>(ko, keys_in, src, kv);
#pragma unroll
  for (int j = 0; j < LI; j++)
    if (dest[j] != ~0u) out[dest[j]] = r[j];
  write_keys<K, LI, KEYS>(ko, dest, kv);
  return true;
}

// The bucket sort: one workgroup per entry of its list (a bucket of at most LI * BLOCK rows; the host
// picks LI from the largest). The rows of a bucket arrive in input order inside each digit run of every
// level (every partition before was stable), so the stable order is the order of (key, position)
// pairs. The bucket is split by its most significant varying digit `top` with LDS counters (any order
// inside a digit run), each row as one 64-bit word, (key bits below the digit) << 12 | position, its
// digit in a byte beside it; inside a run, words order as the pairs do, and a row's place there is the
// number of words below its own. A bucket whose words would not fit 64 bits (top = 7), or with a run
// longer than SUB_MAX rows, is queued for hyb_local_lsd instead. A bucket of equal keys (no varying
// digit left), of any size, is copied.
template <typename K, int LI, bool KEYS>
__device__ __forceinline__ void local_one(const MsdCtl* mc, const Bucket* list, uint32_t idx, const K* keys0,
                                          const uint32_t* rows0, const K* keys1, const uint32_t* rows1,
                                          uint32_t* out_idx, uint32_t* queue, uint32_t* lc, LocalLds& L,
                                          K* s_words, uint8_t* s_dig) {
  const Bucket bk = list[idx];
  const K* keys_in = bucket_keys(bk, keys0, keys1);
  const uint32_t* rows_in = bucket_rows(bk, rows0, rows1);
  const uint32_t size = bk.size, off = bk.start;
  const uint32_t low_mask = mc->active_mask & ((1u << (bk.level & LEVEL_BITS)) - 1u);
  uint32_t* out = out_idx + mc->value_base;
  const KeyOut ko = key_out_of(mc);
  if (low_mask == 0u) {   // equal keys: the bucket is in order already
    for (uint32_t p = threadIdx.x; p < size; p += BLOCK) out[off + p] = rows_in ? rows_in[off + p] : p;
    if constexpr (KEYS && sizeof(K) == 8) {
      if (ko.p)
        for (uint32_t p = threadIdx.x; p < size; p += BLOCK) ko.p[off + p] = keys_in[off + p] ^ ko.x;
    }
    return;
  }
  if (size > uint32_t(LI * BLOCK)) return;   // never: LI covers every bucket to sort
  const int top = 31 - __builtin_clz(low_mask);
  if (top > 6) {   // the key bits below the digit and the position would not fit one word
    if (threadIdx.x == 0) queue[atomicAdd(&lc[LC_QLEN], 1u)] = idx;
    return;
  }
  uint32_t dest[LI], src[LI];
  const int shift = 8 * top;   // at most 48: 48 + 12 bits per word
  L.cnt[threadIdx.x] = 0u;
  K keys[LI];
#pragma unroll
  for (int j = 0; j < LI; j++) keys[j] = keys_in[off + min(threadIdx.x + uint32_t(j) * BLOCK, size - 1u)];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < LI; j++)
    if (threadIdx.x + uint32_t(j) * BLOCK < size) atomicAdd(&L.cnt[uint32_t(keys[j] >> shift) & 0xFFu], 1u);
  __syncthreads();
  const uint32_t c = L.cnt[threadIdx.x];
  uint32_t total;
  const uint32_t st = block_excl_scan(c, L.tmp, total);
  L.start[threadIdx.x] = st;
  L.cursor[threadIdx.x] = st;
  if (__syncthreads_or(c > SUB_MAX ? 1 : 0)) {
    if (threadIdx.x == 0) queue[atomicAdd(&lc[LC_QLEN], 1u)] = idx;
    return;
  }
  const unsigned long long low_bits = (1ull << shift) - 1ull;
#pragma unroll
  for (int j = 0; j < LI; j++) {
    const uint32_t p = threadIdx.x + uint32_t(j) * BLOCK;
    if (p < size) {
      const uint32_t dd = uint32_t(keys[j] >> shift) & 0xFFu;
      const uint32_t slot = atomicAdd(&L.cursor[dd], 1u);
      s_words[slot] = K(((static_cast<unsigned long long>(keys[j]) & low_bits) << 12) | p);
      s_dig[slot] = uint8_t(dd);
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < LI; j++) {
    const uint32_t p = threadIdx.x + uint32_t(j) * BLOCK;
    dest[j] = ~0u;
    src[j] = off;
    if (p < size) {
      const K word = s_words[p];
      const uint32_t dd = s_dig[p];
      const uint32_t lo = L.start[dd], hi = lo + L.cnt[dd];
      uint32_t below = 0;
#pragma unroll 4
      for (uint32_t q = lo; q < hi; q++) below += s_words[q] < word ? 1u : 0u;
      dest[j] = off + lo + below;
      src[j] = off + uint32_t(word & 0xFFFu);
    }
  }
  uint32_t r[LI];
  unsigned long long kv[LI];
  load_rows<LI>(rows_in, off, src, r);   // every row and key load before the first store
  load_keys<K, LI, KEYS>(ko, keys_in, src, kv);
#pragma unroll
  for (int j = 0; j < LI; j++)
    if (dest[j] != ~0u) out[dest[j]] = r[j];
  write_keys<K, LI, KEYS>(ko, dest, kv);
}

// The bucket sort over its list, an entry per workgroup in turn (nothing when the list ran out of room).
// LDS for the bucket is dynamic: LI * BLOCK * (sizeof(K) + 1) bytes.
template <typename K, int LI, bool KEYS>
__global__ __launch_bounds__(BLOCK) void hyb_local(const MsdCtl* mc, const Bucket* list, uint32_t local_cap,
                                                   const K* keys0, const uint32_t* rows0, const K* keys1,
                                                   const uint32_t* rows1, uint32_t* out_idx, uint32_t* queue,
                                                   uint32_t* lc) {
  __shared__ LocalLds L;
  extern __shared__ __align__(16) unsigned char local_lds[];
  K* s_words = reinterpret_cast<K*>(local_lds);
  uint8_t* s_dig = reinterpret_cast<uint8_t*>(s_words + LI * BLOCK);
  if (list_overflow(lc, local_cap)) return;   // uniform
  const uint32_t nlocal = lc[LC_LOCAL];
  for (uint32_t i = blockIdx.x; i < nlocal; i += gridDim.x) {
    local_one<K, LI, KEYS>(mc, list, i, keys0, rows0, keys1, rows1, out_idx, queue, lc, L, s_words, s_dig);
    __syncthreads();
  }
}

// The groups of small buckets (the far end of the list), a group per workgroup in turn; a group with a
// bin longer than SUB_MAX rows is queued for hyb_local_lsd. LDS for the group is dynamic:
// GROUP_ROWS * (sizeof(K) + 2) + GROUP_BINS * 4 bytes.
template <typename K, bool KEYS>
__global__ __launch_bounds__(BLOCK) void hyb_group(const MsdCtl* mc, const Bucket* list, uint32_t local_cap,
                                                   const K* keys0, const uint32_t* rows0, const K* keys1,
                                                   const uint32_t* rows1, uint32_t* out_idx, uint32_t* queue,
                                                   uint32_t* lc) {
  constexpr int LI = int(GROUP_ROWS / BLOCK);
  __shared__ uint32_t tmp[WAVES];
  extern __shared__ __align__(16) unsigned char local_lds[];
  K* s_words = reinterpret_cast<K*>(local_lds);
  uint16_t* s_bin = reinterpret_cast<uint16_t*>(s_words + GROUP_ROWS);
  uint32_t* g_bins = reinterpret_cast<uint32_t*>(s_bin + GROUP_ROWS);
  if (list_overflow(lc, local_cap)) return;   // uniform
  const uint32_t ngroups = lc[LC_GROUPS];
  uint32_t* out = out_idx + mc->value_base;
  const KeyOut ko = key_out_of(mc);
  for (uint32_t g = blockIdx.x; g < ngroups; g += gridDim.x) {
    const uint32_t idx = local_cap - 1u - g;
    const Bucket bk = list[idx];
    const uint32_t low_mask = mc->active_mask & ((1u << (bk.level & LEVEL_BITS)) - 1u);
    const int top = 31 - __builtin_clz(low_mask);   // a group's buckets differ in digit top (its level - 1)
    const uint32_t below_top = low_mask & ((1u << top) - 1u);   // never 0: they are not of equal keys
    const uint32_t span_log = (bk.level >> SPAN_SHIFT) & 0xFu;
    const K* keys_in = bucket_keys(bk, keys0, keys1);
    const uint32_t* rows_in = bucket_rows(bk, rows0, rows1);
    const int sub = 31 - __builtin_clz(below_top);
    bool done = bk.size > GROUP_ROWS;   // never
    if (!done) {
      switch (span_log) {
        case 0: done = local_group<K, LI, 0, KEYS>(keys_in, rows_in, bk.start, bk.size, top, sub, out, ko, tmp, g_bins, s_words, s_bin); break;
        case 1: done = local_group<K, LI, 1, KEYS>(keys_in, rows_in, bk.start, bk.size, top, sub, out, ko, tmp, g_bins, s_words, s_bin); break;
        case 2: done = local_group<K, LI, 2, KEYS>(keys_in, rows_in, bk.start, bk.size, top, sub, out, ko, tmp, g_bins, s_words, s_bin); break;
        case 3: done = local_group<K, LI, 3, KEYS>(keys_in, rows_in, bk.start, bk.size, top, sub, out, ko, tmp, g_bins, s_words, s_bin); break;
        default: done = local_group<K, LI, 4, KEYS>(keys_in, rows_in, bk.start, bk.size, top, sub, out, ko, tmp, g_bins, s_words, s_bin); break;
      }
    }
    if (!done && threadIdx.x == 0) queue[atomicAdd(&lc[LC_QLEN], 1u)] = idx;
    __syncthreads();
  }
}

// The queued buckets, each by local_lsd: persistent workgroups take them from a counter until none is
// left. LDS for the bucket is dynamic: LI * BLOCK * (sizeof(K) + 2) bytes.
template <typename K, int LI, bool KEYS>
__global__ __launch_bounds__(BLOCK) void hyb_local_lsd(const MsdCtl* mc, const Bucket* list, const K* keys0,
                                                       const uint32_t* rows0, const K* keys1, const uint32_t* rows1,
                                                       uint32_t* out_idx, const uint32_t* queue, uint32_t* lc) {
  __shared__ uint32_t wc[WAVES][WC_STRIDE];
  __shared__ uint32_t tile_total[WC_STRIDE];
  __shared__ uint32_t bin_start[RADIX];
  __shared__ uint32_t tmp[WAVES];
  __shared__ uint32_t s_q;
  extern __shared__ __align__(16) unsigned char local_lds[];
  K* s_keys = reinterpret_cast<K*>(local_lds);
  uint16_t* s_pos = reinterpret_cast<uint16_t*>(s_keys + LI * BLOCK);
  const uint32_t queued = lc[LC_QLEN];
  if (blockIdx.x >= queued) return;   // the first `queued` workgroups take every entry between them
  while (true) {
    if (threadIdx.x == 0) s_q = atomicAdd(&lc[LC_QNEXT], 1u);
    __syncthreads();
    const uint32_t qi = s_q;
    if (qi >= queued) return;   // uniform: every workgroup ends here once the queue is empty
    const Bucket bk = list[queue[qi]];
    const uint32_t low_mask = mc->active_mask & ((1u << (bk.level & LEVEL_BITS)) - 1u);
    if (bk.size <= uint32_t(LI * BLOCK))   // always: the host sizes LI for every queued bucket
      local_lsd<K, LI, KEYS>(bucket_keys(bk, keys0, keys1), bucket_rows(bk, rows0, rows1), bk.start, bk.size, low_mask, SUB_MAX,
                       out_idx + mc->value_base, key_out_of(mc), wc, tile_total, bin_start, tmp, s_keys, s_pos);
    __syncthreads();
  }
}


// ---------------------------------------------------------------- host side

// The device buffers of the levels for up to m value rows, with `items` keys per thread in a pass
// (the tile tables are sized for that many or more), and extra_local more entries in the bucket
// sort's list (a batched plan's batches).
// Capacities: buckets above LOCAL_MAX / 2 rows at one level (the least local_max used); entries of the
// bucket sort's list (every entry has a row, and runs of many tiny buckets are what fills it); chunks
// and tiles of a level.
struct MsdWork {
  uint32_t large_cap = 0, local_cap = 0;
  DevBuf lists, tabs, words, wide;
  Bucket* large[2] = {nullptr, nullptr};
  Bucket* local = nullptr;
  LvlChunk* chunks = nullptr;
  LvlTile* tiles = nullptr;
  uint32_t *lc = nullptr, *bucket_hist = nullptr, *bucket_base = nullptr, *carried = nullptr, *queue = nullptr,
           *bofs = nullptr;
  unsigned long long *bucket_or = nullptr, *bucket_and = nullptr;
  MsdCtl* mc = nullptr;
};

inline MsdWork msd_work(uint32_t m, int items, DevBuf (*alloc)(size_t), uint32_t extra_local = 0) {
  MsdWork w;
  const uint32_t tile = uint32_t(BLOCK * items);
  w.large_cap = m / (LOCAL_MAX / 2) + 4096u;
  w.local_cap = m / 32u + (1u << 17) + extra_local;
  // ARROWMETAL_MSD_LIST_CAP (tests only) caps the bucket-sort list, so that a test can send a sort down
  // the overflow path: the plain argsort falls back to the LSD passes, a batched plan to its segmented
  // LSD passes.
  if (const char* e = getenv("ARROWMETAL_MSD_LIST_CAP"))
    if (atoi(e) > 0) w.local_cap = std::min(w.local_cap, uint32_t(atoi(e)));
  const uint32_t max_chunks = m / LVL_CHUNK + w.large_cap + 1u;
  const uint32_t max_tiles = m / tile + w.large_cap + 1u;
  w.lists = alloc((size_t(w.large_cap) * 2 + w.local_cap) * sizeof(Bucket));
  w.tabs = alloc(size_t(max_chunks) * sizeof(LvlChunk) + size_t(max_tiles) * sizeof(LvlTile));
  w.words = alloc((size_t(w.large_cap) * (2 * RADIX + 3) + w.local_cap + LC_WORDS) * 4);
  w.wide = alloc(size_t(w.large_cap) * 16 + sizeof(MsdCtl));
  w.large[0] = static_cast<Bucket*>(w.lists->ptr);
  w.large[1] = w.large[0] + w.large_cap;
  w.local = w.large[1] + w.large_cap;
  w.chunks = static_cast<LvlChunk*>(w.tabs->ptr);
  w.tiles = reinterpret_cast<LvlTile*>(w.chunks + max_chunks);
  w.lc = static_cast<uint32_t*>(w.words->ptr);   // first, so that its lines are its own
  w.bucket_hist = w.lc + LC_WORDS;
  w.bucket_base = w.bucket_hist + size_t(w.large_cap) * RADIX;
  w.carried = w.bucket_base + size_t(w.large_cap) * RADIX;
  w.bofs = w.carried + w.large_cap;   // 2 words per bucket
  w.queue = w.bofs + 2 * size_t(w.large_cap);
  w.bucket_or = static_cast<unsigned long long*>(w.wide->ptr);
  w.bucket_and = w.bucket_or + w.large_cap;
  w.mc = reinterpret_cast<MsdCtl*>(w.bucket_and + w.large_cap);
  return w;
}

// The levels below digit `top` and the bucket sort, on stream s, after the first list is built (in
// w.large[0] with its count, the bucket sort's list started, the counters and w.mc set).
// - Read-back form (blind = false; `mask` = the active digits, known on the host): only active
//   digits get a level, each launch is sized from 64 bytes of counters read back after each level,
//   and the bucket sort's keys per thread come from the largest bucket. Returns false, with no row
//   written that a rerun does not write again, when the bucket sort's list ran out of room.
// - Blind form (blind = true, e.g. inside a graph capture): every digit below `top` gets a level,
//   sized for the worst case, the kernels read what to do from the device, and the bucket sort takes
//   buckets of at most local_max = LOCAL_MAX / 2 rows. The caller must check LC_OVERFLOW itself.
// Look-back epochs: w.mc->epoch + 9 .. + 15.
template <typename K, int ITEMS>
bool msd_levels(const MsdWork& w, uint32_t m, int top, uint32_t mask, bool blind, uint32_t local_max,
                unsigned long long* status, K* k0, uint32_t* r0, K* k1, uint32_t* r1, uint32_t* out_idx,
                hipStream_t s, bool keys = false) {
  constexpr int P = int(sizeof(K));
  constexpr uint32_t TILE = BLOCK * ITEMS;
  const uint32_t cus = uint32_t(device().cu_count);
  uint32_t* hp = blind ? nullptr : host_plan_words();   // nothing pinned is touched in a capture
  // Sub-buckets packed into groups: up to 512 rows from host read-backs (larger ones sort faster one
  // per workgroup, where the workgroup count follows the list), any bucket sort size when blind (the
  // bucket sort's workgroups are persistent and take 8 keys per thread).
  const uint32_t small_max = blind ? GROUP_ROWS : 512u;
  auto next_digit = [&](int d) {
    while (--d >= 0 && !blind && !((mask >> d) & 1u)) {}
    return d;
  };
  // Large buckets and their rows (bounds of the grids): at most RADIX top-digit buckets first.
  uint32_t bound = blind ? w.large_cap : uint32_t(RADIX), rows = m;
  int cur = 0;
  bool read = false;
  for (int d = next_digit(top); d >= 0 && bound > 0; d = next_digit(d)) {
    {
      ProfScope ps("sort_msd_layout");
      hipLaunchKernelGGL(lvl_layout<ITEMS>, dim3(1), dim3(BLOCK), 0, s, w.large[cur], cur, w.bofs, w.lc);
      hipLaunchKernelGGL(lvl_fill<ITEMS>, dim3(std::min(bound, cus * 4u)), dim3(BLOCK), 0, s, w.large[cur], cur,
                         w.lc, w.bofs, w.chunks, w.tiles, w.bucket_hist, w.bucket_or, w.bucket_and);
      AM_HIP(hipGetLastError());
    }
    {
      ProfScope ps("sort_msd_hist");
      hipLaunchKernelGGL(lvl_hist<K>, dim3(std::min(rows / LVL_CHUNK + bound + 1u, cus * 8u)), dim3(BLOCK), 0, s,
                         k0, k1, w.large[cur], w.chunks, w.lc, w.mc, d, w.bucket_hist, w.bucket_or, w.bucket_and);
      AM_HIP(hipGetLastError());
    }
    {
      ProfScope ps("sort_msd_plan");
      hipLaunchKernelGGL(lvl_plan, dim3(std::min(bound, cus * 4u)), dim3(BLOCK), 0, s, w.large[cur], cur,
                         w.bucket_hist, w.bucket_or, w.bucket_and, w.mc, d, local_max, small_max, GROUP_ROWS,
                         w.bucket_base, w.carried, w.large[cur ^ 1], w.local, w.local_cap, w.lc);
      AM_HIP(hipGetLastError());
    }
    {
      ProfScope ps("sort_msd_scatter");
      if (blind)
        hipLaunchKernelGGL((lvl_pass_loop<K, ITEMS>), dim3(cus * 3u), dim3(BLOCK), 0, s, w.large[cur], w.tiles, w.lc,
                           w.carried, w.bucket_base, status, w.mc, 8u + uint32_t(P - 1 - d), d, k0, r0, k1, r1);
      else
        hipLaunchKernelGGL((lvl_pass<K, ITEMS>), dim3(rows / TILE + bound + 1u), dim3(BLOCK), 0, s, w.large[cur],
                           w.tiles, w.lc, w.carried, w.bucket_base, status, w.mc, 8u + uint32_t(P - 1 - d), d, k0,
                           r0, k1, r1);
      AM_HIP(hipGetLastError());
    }
    cur ^= 1;
    if (!blind) {
      AM_HIP(hipMemcpyAsync(hp, w.lc, LC_READ * 4, hipMemcpyDeviceToHost, s));
      AM_HIP(hipStreamSynchronize(s));
      read = true;
      bound = hp[LC_LIST + cur];
      rows = hp[LC_ROWS];
      if (bound > w.large_cap) fail("argsort: bucket list overflow");   // cannot happen: each holds > local_max rows
    }
  }
  int li = 8, li_lsd = 8;
  uint32_t grid = cus * 8u, group_grid = cus * 4u;
  if (!blind) {
    if (!read) {   // no level ran (no active digit below the top one)
      AM_HIP(hipMemcpyAsync(hp, w.lc, LC_READ * 4, hipMemcpyDeviceToHost, s));
      AM_HIP(hipStreamSynchronize(s));
    }
    if (hp[LC_OVERFLOW] || uint64_t(hp[LC_LOCAL]) + hp[LC_GROUPS] > w.local_cap) return false;
    const uint32_t biggest = hp[LC_MAXLOCAL];
    li = biggest <= 4u * BLOCK ? 4 : biggest <= 8u * BLOCK ? 8 : 16;
    li_lsd = hp[LC_GROUPS] != 0u ? std::max(li, int(GROUP_ROWS / BLOCK)) : li;   // a queued group fits
    grid = hp[LC_LOCAL];
    group_grid = std::min(hp[LC_GROUPS], cus * 4u);
  }
  // The LSD kernel takes the buckets the counting sorts queued (its workgroups exit at once when none
  // was).
  const dim3 lsd_grid(cus * 4u);
  const size_t lds = size_t(li) * BLOCK * (sizeof(K) + 1), lsd_lds = size_t(li_lsd) * BLOCK * (sizeof(K) + 2);
  // The bucket sorts in the form that also writes the sorted key column (keys), or not: a separate
  // instantiation, so that the plain sort keeps its register budget.
  auto bucket_sorts = [&](auto keys_tag) {
    constexpr bool KEYS = decltype(keys_tag)::value;
    if (grid != 0u) {
      ProfScope ps("sort_msd_local");
      if (li == 4)
        hipLaunchKernelGGL((hyb_local<K, 4, KEYS>), dim3(grid), dim3(BLOCK), lds, s, w.mc, w.local, w.local_cap, k0, r0,
                           k1, r1, out_idx, w.queue, w.lc);
      else if (li == 8)
        hipLaunchKernelGGL((hyb_local<K, 8, KEYS>), dim3(grid), dim3(BLOCK), lds, s, w.mc, w.local, w.local_cap, k0, r0,
                           k1, r1, out_idx, w.queue, w.lc);
      else
        hipLaunchKernelGGL((hyb_local<K, 16, KEYS>), dim3(grid), dim3(BLOCK), lds, s, w.mc, w.local, w.local_cap, k0,
                           r0, k1, r1, out_idx, w.queue, w.lc);
      AM_HIP(hipGetLastError());
    }
    if (group_grid != 0u) {
      ProfScope ps("sort_msd_group");
      hipLaunchKernelGGL((hyb_group<K, KEYS>), dim3(group_grid), dim3(BLOCK),
                         GROUP_ROWS * (sizeof(K) + 2) + GROUP_BINS * 4, s, w.mc, w.local, w.local_cap, k0, r0, k1, r1,
                         out_idx, w.queue, w.lc);
      AM_HIP(hipGetLastError());
    }
    {
      ProfScope ps("sort_msd_local_lsd");
      if (li_lsd == 4)
        hipLaunchKernelGGL((hyb_local_lsd<K, 4, KEYS>), lsd_grid, dim3(BLOCK), lsd_lds, s, w.mc, w.local, k0, r0, k1,
                           r1, out_idx, w.queue, w.lc);
      else if (li_lsd == 8)
        hipLaunchKernelGGL((hyb_local_lsd<K, 8, KEYS>), lsd_grid, dim3(BLOCK), lsd_lds, s, w.mc, w.local, k0, r0, k1,
                           r1, out_idx, w.queue, w.lc);
      else
        hipLaunchKernelGGL((hyb_local_lsd<K, 16, KEYS>), lsd_grid, dim3(BLOCK), lsd_lds, s, w.mc, w.local, k0, r0, k1,
                           r1, out_idx, w.queue, w.lc);
      AM_HIP(hipGetLastError());
    }
  };
  if constexpr (sizeof(K) == 8) {
    if (keys) bucket_sorts(std::true_type{});
    else bucket_sorts(std::false_type{});
  } else {
    bucket_sorts(std::false_type{});
  }
  return true;
}

}  // namespace
}  // namespace am
