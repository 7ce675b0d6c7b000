// Batched argsort and group-by over thousands of record batches concatenated row-wise, in a fixed
// sequence of kernels with no host readback, captured once in a hipGraph and replayed.
//
// The reference runs such workloads one batch at a time (its per-thread command-buffer batching,
// Sources/ArrowMetal/MetalContext.swift:245-312, amortises submission but every batch still runs its
// own sort; a persistent GPU worker was measured not viable on Metal,
// Sources/ArrowMetal/Resident/ResidentMode.swift:4-31). Here:
//
//   * the argsort of every batch is one segmented onesweep sort: tiles never cross a batch, each
//     batch has its own digit histograms and bases, and the decoupled look-back chains stop at the
//     batch's first tile. Row numbers are local to their batch. Every pass runs over all batches in
//     one launch (eight launches for 64-bit keys), so 2,048 batches cost what one big sort costs.
//   * the group-by sum / count of every batch runs in a persistent kernel: workgroups pull chunks of
//     up to CHUNK_ROWS rows of one batch from a queue, aggregate them in an LDS hash table, and
//     flush it into the batch's table in HBM; one workgroup per batch then sorts the batch's groups
//     by key (a bitonic sort in LDS) and they are compacted into one output.
//
// Semantics per batch are those of am_argsort_ex2 (nulls last; `ieee` NaN after every value in both
// directions) and of am_group_by_keys + am_group_agg_ex sum / count over an int64-like key column
// with int64 values (groups ascending by key; the sum wraps in 64 bits). Nulls are not taken here.
#include "msd_levels.hpp"

#include <vector>

namespace am {
namespace {

struct SegCtl {
  unsigned long long key_or, key_and;
  uint32_t active_mask, first_active, last_active, epoch;
  // Bumped by every workgroup of a kernel, so on cache lines of their own: sharing a line with the
  // plan words those workgroups read slows the pass down.
  alignas(128) uint32_t tile_counter[16];
  alignas(128) uint32_t chunk_counter;
  uint32_t pad[31];
};

struct SegTile {
  uint32_t seg, start, end, head;   // head: id of the batch's first tile
};

struct SegChunk {
  uint32_t seg, start, end;
  uint32_t whole;   // a group-by chunk that holds its whole batch (1), else 0
};

// A value key of a batched sort: NaN (ieee order) becomes the largest key after the descending
// flip, which keeps every NaN after the values, in input order, in both directions.
template <typename K>
__device__ __forceinline__ K seg_key(uint64_t raw, int type, int forder, int desc) {
  K k;
  if (classify_raw<K>(raw, true, type, forder, desc, k) != 0) k = K(~K(0));
  return k;
}

__global__ void seg_init(SegCtl* ctl, uint32_t* hist, size_t hist_words, uint32_t* epoch_word) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < hist_words; i += stride) hist[i] = 0u;
  if (blockIdx.x == 0 && threadIdx.x < 16) ctl->tile_counter[threadIdx.x] = 0u;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->key_or = 0ull;
    ctl->key_and = ~0ull;
    ctl->chunk_counter = 0u;
    // Every run (and every graph replay) takes fresh epochs, so the status table is never cleared.
    uint32_t e = *epoch_word + 16u;
    *epoch_word = e;
    ctl->epoch = e;
  }
}

template <typename K, int W>
__global__ __launch_bounds__(BLOCK) void seg_upsweep(const void* vals, int type, int forder, int desc,
                                                     const SegChunk* chunks, SegCtl* ctl, uint32_t* hist) {
  constexpr int P = int(sizeof(K));
  __shared__ uint32_t h[P][RADIX];
  __shared__ unsigned long long s_or, s_and;
  for (int i = threadIdx.x; i < P * RADIX; i += BLOCK) (&h[0][0])[i] = 0u;
  if (threadIdx.x == 0) {
    s_or = 0ull;
    s_and = ~0ull;
  }
  __syncthreads();
  const SegChunk c = chunks[blockIdx.x];
  K o = 0, a = K(~K(0));
  constexpr int UP_ITEMS = 8;
  for (uint32_t base = c.start; base < c.end; base += BLOCK * UP_ITEMS) {
    K kk[UP_ITEMS];
    uint64_t raw[UP_ITEMS];
#pragma unroll
    for (int u = 0; u < UP_ITEMS; u++)   // every load first (rows past the chunk read its last row)
      raw[u] = load_raw<W>(vals, min(base + uint32_t(u) * BLOCK + threadIdx.x, c.end - 1u));
#pragma unroll
    for (int u = 0; u < UP_ITEMS; u++) {
      uint32_t i = base + uint32_t(u) * BLOCK + threadIdx.x;
      kk[u] = i < c.end ? seg_key<K>(raw[u], type, forder, desc) : K(0);
    }
#pragma unroll
    for (int u = 0; u < UP_ITEMS; u++) {
      uint32_t i = base + uint32_t(u) * BLOCK + threadIdx.x;
      if (i >= c.end) continue;
      o |= kk[u];
      a &= kk[u];
#pragma unroll
      for (int p = 0; p < P; p++) atomicAdd(&h[p][uint32_t(kk[u] >> (8 * p)) & 0xFFu], 1u);
    }
  }
  unsigned long long o64 = o, a64 = a;
  if (sizeof(K) == 4) a64 |= 0xFFFFFFFF00000000ull;
  for (int off = 32; off > 0; off >>= 1) {
    o64 |= __shfl_down(o64, off, 64);
    a64 &= __shfl_down(a64, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicOr(&s_or, o64);
    atomicAnd(&s_and, a64);
  }
  __syncthreads();
  uint32_t* hs = hist + size_t(c.seg) * P * RADIX;
  for (int i = threadIdx.x; i < P * RADIX; i += BLOCK) {
    uint32_t v = (&h[0][0])[i];
    if (v) atomicAdd(&hs[i], v);
  }
  if (threadIdx.x == 0 && c.end > c.start) {
    atomicOr(&ctl->key_or, s_or);
    atomicAnd(&ctl->key_and, s_and);
  }
}

// One block per batch: the exclusive digit prefix of every pass, offset by the batch's first row.
// Block 0 also decides which passes run (a digit constant over every batch is skipped).
template <int P>
__global__ __launch_bounds__(BLOCK) void seg_plan(const uint32_t* hist, const uint32_t* seg_off, uint32_t* base,
                                                  SegCtl* ctl) {
  __shared__ uint32_t tmp[WAVES];
  const uint32_t seg = blockIdx.x;
  const uint32_t start = seg_off[seg];
  for (int p = 0; p < P; p++) {
    uint32_t total;
    size_t at = (size_t(seg) * P + p) * RADIX + threadIdx.x;
    base[at] = start + block_excl_scan(hist[at], tmp, total);
  }
  if (seg == 0 && threadIdx.x == 0) {
    unsigned long long diff = ctl->key_or ^ ctl->key_and;
    uint32_t mask = 0;
    for (int p = 0; p < P; p++)
      if ((diff >> (8 * p)) & 0xFFull) mask |= 1u << p;
    if (mask == 0) mask = 1;
    ctl->active_mask = mask;
    ctl->first_active = uint32_t(__builtin_ctz(mask));
    ctl->last_active = 31u - uint32_t(__builtin_clz(mask));
  }
}

// One segmented scatter pass. FIRST: the first active pass, reading the raw column (row numbers start
// local to their batch). Otherwise pass `pass` over the ping-pong buffers; every digit is launched
// and an inactive one exits at once, so the sequence needs no host decision.
template <typename K, int ITEMS, bool FIRST, int W>
__global__ __launch_bounds__(BLOCK) void seg_pass(int pass, const void* vals, int type, int forder, int desc,
                                                  const SegTile* tiles, const uint32_t* seg_off,
                                                  const uint32_t* base, SegCtl* ctl, unsigned long long* status,
                                                  K* keys0, uint32_t* rows0, K* keys1, uint32_t* rows1,
                                                  uint32_t* out_idx) {
  constexpr int TILE = BLOCK * ITEMS;
  constexpr int P = int(sizeof(K));
  __shared__ uint32_t wc[WAVES][WC_STRIDE];
  __shared__ uint32_t tile_total[WC_STRIDE];
  __shared__ uint32_t bin_start[RADIX];
  __shared__ uint32_t bin_base[WC_STRIDE];
  __shared__ uint32_t tmp[WAVES];
  __shared__ uint32_t s_tile;
  __shared__ uint8_t s_digit[TILE];
  __shared__ K s_stage[TILE];

  const uint32_t mask = ctl->active_mask;
  const uint32_t first = ctl->first_active;
  const int dpass = FIRST ? int(first) : pass;
  if (!FIRST && (!((mask >> pass) & 1u) || uint32_t(pass) == first)) return;
  const bool last = uint32_t(dpass) == ctl->last_active;
  const uint32_t rank = uint32_t(__builtin_popcount(mask & ((1u << dpass) - 1u)));
  const uint32_t epoch = ctl->epoch + 1u + uint32_t(dpass);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < WAVES * WC_STRIDE; i += BLOCK) (&wc[0][0])[i] = 0u;
  if (threadIdx.x == 0) s_tile = atomicAdd(&ctl->tile_counter[dpass], 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const SegTile t = tiles[tile];
  const uint32_t seg_start = seg_off[t.seg];
  const K* in_keys = (rank & 1u) ? keys0 : keys1;
  const uint32_t* in_rows = (rank & 1u) ? rows0 : rows1;
  K* out_keys = (rank & 1u) ? keys1 : keys0;
  uint32_t* out_rows = (rank & 1u) ? rows1 : rows0;
  if (FIRST) {
    out_keys = keys0;
    out_rows = rows0;
  }
  const int shift = 8 * dpass;

  K keys[ITEMS];
  uint32_t rows[ITEMS], bins[ITEMS];
  const uint32_t base_e = t.start + uint32_t(w) * 64u * ITEMS + uint32_t(lane);
  if constexpr (FIRST) {
    uint64_t raw[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++)   // every load first (lanes past the tile read its last row)
      raw[j] = t.end > t.start ? load_raw<W>(vals, min(base_e + uint32_t(j) * 64u, t.end - 1u)) : 0ull;
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      uint32_t e = base_e + uint32_t(j) * 64u;
      keys[j] = 0;
      rows[j] = 0;
      bins[j] = BIN_NONE;
      if (e < t.end) {
        keys[j] = seg_key<K>(raw[j], type, forder, desc);
        rows[j] = e - seg_start;
        bins[j] = uint32_t(keys[j] >> shift) & 0xFFu;
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      uint32_t e = base_e + uint32_t(j) * 64u;
      keys[j] = 0;
      rows[j] = 0;
      bins[j] = BIN_NONE;
      if (e < t.end) {
        keys[j] = in_keys[e];
        rows[j] = in_rows[e];
        bins[j] = uint32_t(keys[j] >> shift) & 0xFFu;
      }
    }
  }
  if (t.end - t.start == uint32_t(TILE)) rank_tile<ITEMS, 8>(bins, wc[w], lane);
  else rank_tile<ITEMS, 9>(bins, wc[w], lane);
  __syncthreads();
  uint32_t vcount = tile_bin_layout<RADIX>(wc, tile_total, bin_start, tmp);
  look_back<RADIX>(status, tile, epoch, tile_total, bin_base, t.head);
  __syncthreads();
  bin_base[threadIdx.x] += base[(size_t(t.seg) * P + dpass) * RADIX + threadIdx.x];
  __syncthreads();
  if (last) {
    scatter_values<K, ITEMS, false>(keys, rows, bins, wc[w], bin_start, bin_base, s_digit, s_stage, vcount,
                                    nullptr, out_idx);
  } else {
    scatter_values<K, ITEMS, true>(keys, rows, bins, wc[w], bin_start, bin_base, s_digit, s_stage, vcount,
                                   out_keys, out_rows);
  }
}

// ---------------------------------------------------------------- batched MSD argsort (64-bit keys)
//
// Every batch is one bucket of the adaptive MSD radix sort of msd_levels.hpp, in its blind form: the
// whole sequence is fixed (a map pass, then one level per digit, then the bucket sort), so it can be
// captured in the plan's graph, and every kernel reads from device memory what it has to do.

// Maps the sort keys of every batch (the same map as the LSD pass) into k0, and the OR / AND of the
// mapped keys into the control block; a 64K-row chunk of one batch per workgroup. No row buffer: the
// batches enter the levels as buckets with implicit rows (a row is its index inside its batch).
template <typename K, int W>
__global__ __launch_bounds__(BLOCK) void seg_msd_map(const void* vals, int type, int forder, int desc,
                                                     const SegChunk* chunks, SegCtl* ctl, K* k0) {
  __shared__ unsigned long long s_or, s_and;
  if (threadIdx.x == 0) {
    s_or = 0ull;
    s_and = ~0ull;
  }
  __syncthreads();
  const SegChunk c = chunks[blockIdx.x];
  unsigned long long o = 0ull, a = ~0ull;
  constexpr int U = 8;
  for (uint32_t base = c.start + threadIdx.x; base < c.end; base += BLOCK * U) {
    uint64_t raw[U];
#pragma unroll
    for (int u = 0; u < U; u++) raw[u] = load_raw<W>(vals, min(base + uint32_t(u) * BLOCK, c.end - 1u));
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint32_t i = base + uint32_t(u) * BLOCK;
      const K k = seg_key<K>(raw[u], type, forder, desc);
      o |= (unsigned long long)k;   // rows past the end repeat the chunk's last key
      a &= (unsigned long long)k;
      if (i < c.end) k0[i] = k;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    o |= __shfl_xor(o, off, 64);
    a &= __shfl_xor(a, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicOr(&s_or, o);
    atomicAnd(&s_and, a);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicOr(&ctl->key_or, s_or);
    atomicAnd(&ctl->key_and, s_and);
  }
}

// One workgroup: the level kernels' control words (the active digits from the OR / AND, the run's
// epoch) and their counters cleared.
template <int P>
__global__ void seg_msd_ctl(const SegCtl* ctl, MsdCtl* mc, uint32_t* lc) {
  if (threadIdx.x < LC_WORDS) lc[threadIdx.x] = 0u;
  if (threadIdx.x == 0) {
    const unsigned long long diff = ctl->key_or ^ ctl->key_and;
    uint32_t mask = 0;
    for (int p = 0; p < P; p++)
      if ((diff >> (8 * p)) & 0xFFull) mask |= 1u << p;
    *mc = {mask, 0u, ctl->epoch, 0u};
  }
}

// Every batch as a bucket of level P (no digit applied yet): into the first large-bucket list, or the
// bucket sort's list when it has at most local_max rows or all its keys are equal.
template <int P>
__global__ void seg_msd_lists(const uint32_t* seg_off, uint32_t nseg, const MsdCtl* mc, uint32_t local_max,
                              Bucket* large, Bucket* local, uint32_t local_cap, uint32_t* lc) {
  const uint32_t mask = mc->active_mask;
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nseg; b += gridDim.x * blockDim.x) {
    const uint32_t start = seg_off[b], size = seg_off[b + 1] - start;
    if (size == 0u) continue;
    const Bucket bk{start, size, BUF_IMPLICIT, uint32_t(P)};   // rows: index - start
    if (size > local_max && mask != 0u) large[atomicAdd(&lc[LC_LIST], 1u)] = bk;
    else push_local(local, local_cap, lc, bk, mask == 0u);
  }
}

// ---------------------------------------------------------------- batched group-by

constexpr unsigned long long BG_EMPTY = 0x8000000000000000ull;
constexpr uint32_t CHUNK_ROWS = 32768;

__device__ __forceinline__ unsigned long long bg_key(const void* keys, uint32_t i, int type) {
  switch (type) {
    case T_I64: case T_U64: return static_cast<const unsigned long long*>(keys)[i];
    case T_I32: return (unsigned long long)(long long)static_cast<const int32_t*>(keys)[i];
    case T_U32: return static_cast<const uint32_t*>(keys)[i];
    case T_I16: return (unsigned long long)(long long)static_cast<const int16_t*>(keys)[i];
    case T_U16: return static_cast<const uint16_t*>(keys)[i];
    case T_I8: return (unsigned long long)(long long)static_cast<const int8_t*>(keys)[i];
    default: return static_cast<const uint8_t*>(keys)[i];
  }
}

__device__ __forceinline__ long long bg_value(const void* vals, uint32_t i, int type) {
  switch (type) {
    case T_I64: case T_U64: return static_cast<const long long*>(vals)[i];
    case T_I32: return static_cast<const int32_t*>(vals)[i];
    case T_U32: return static_cast<const uint32_t*>(vals)[i];
    case T_I16: return static_cast<const int16_t*>(vals)[i];
    case T_U16: return static_cast<const uint16_t*>(vals)[i];
    case T_I8: return static_cast<const int8_t*>(vals)[i];
    default: return static_cast<const uint8_t*>(vals)[i];
  }
}

__device__ __forceinline__ uint32_t bg_hash(unsigned long long k, uint32_t mask) {
  k ^= 0x9E3779B97F4A7C15ull;
  k ^= k >> 33;
  k *= 0xFF51AFD7ED558CCDull;
  k ^= k >> 33;
  return uint32_t(k) & mask;
}

// Clears the per-batch tables (keys to EMPTY, sums and counts to zero), the overflow flags and the
// chunk queue.
__global__ void bg_init(unsigned long long* tkey, unsigned long long* tsum, unsigned long long* tcnt, size_t slots,
                        uint32_t* overflow, uint32_t nseg, SegCtl* ctl) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  if (i == 0) ctl->chunk_counter = 0u;
  for (size_t j = i; j < slots; j += stride) {
    tkey[j] = BG_EMPTY;
    tsum[j] = 0ull;
    tcnt[j] = 0ull;
  }
  for (size_t j = i; j < nseg; j += stride) overflow[j] = 0u;
}

// Finds (or claims) the slot of key k in a table of cap slots (+1 for the EMPTY key at slot cap);
// cap when the table is full.
__device__ __forceinline__ uint32_t bg_slot(unsigned long long* tk, unsigned long long k, uint32_t cap, bool& full) {
  if (k == BG_EMPTY) return cap;
  const uint32_t mask = cap - 1u;
  uint32_t s = bg_hash(k, mask);
  for (uint32_t probes = 0; probes < cap; probes++) {
    unsigned long long cur = tk[s];
    if (cur == k) return s;
    if (cur == BG_EMPTY) {
      unsigned long long prev = atomicCAS(&tk[s], BG_EMPTY, k);
      if (prev == BG_EMPTY || prev == k) return s;
    }
    s = (s + 1u) & mask;
  }
  full = true;
  return cap;
}

// Group-by strategy 1 (global): every row updates its batch's table in HBM (L2) directly. Rows are
// taken in grid-stride order over the chunk queue; a batch whose table fills is flagged.
__global__ __launch_bounds__(256) void bg_agg_global(const void* keys, int ktype, const void* vals, int vtype,
                                                     const SegChunk* chunks, uint32_t nchunks, SegCtl* ctl,
                                                     unsigned long long* tkey, unsigned long long* tsum,
                                                     unsigned long long* tcnt, uint32_t cap, uint32_t* overflow) {
  __shared__ uint32_t s_chunk;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) s_chunk = atomicAdd(&ctl->chunk_counter, 1u);
    __syncthreads();
    const uint32_t ci = s_chunk;
    if (ci >= nchunks) break;
    const SegChunk c = chunks[ci];
    unsigned long long* gk = tkey + size_t(c.seg) * (cap + 1);
    unsigned long long* gs = tsum + size_t(c.seg) * (cap + 1);
    unsigned long long* gc = tcnt + size_t(c.seg) * (cap + 1);
    bool full = false;
    for (uint32_t i = c.start + threadIdx.x; i < c.end && !full; i += blockDim.x) {
      unsigned long long k = bg_key(keys, i, ktype);
      unsigned long long v = (unsigned long long)bg_value(vals, i, vtype);
      uint32_t s = bg_slot(gk, k, cap, full);
      if (full) break;
      atomicAdd(&gs[s], v);
      atomicAdd(&gc[s], 1ull);
    }
    if (full) atomicOr(&overflow[c.seg], 1u);
  }
}

// The rows of one step of bg_agg: ROWS per thread, blockDim.x apart.
template <int ROWS>
struct BgRows {
  unsigned long long k[ROWS], v[ROWS];
};

// Loads the 64-bit keys and values of the step's rows from i0; rows past `end` read its last row.
template <int ROWS>
__device__ __forceinline__ void bg_load_wide(const unsigned long long* keys, const unsigned long long* vals, uint32_t i0,
                                             uint32_t end, BgRows<ROWS>& x) {
#pragma unroll
  for (int r = 0; r < ROWS; r++) {
    const uint32_t i = min(i0 + uint32_t(r) * blockDim.x, end - 1u);
    x.k[r] = keys[i];
    x.v[r] = vals[i];
  }
}

// Adds the step's rows before `end` to the LDS table (cap slots, linear probing; slot cap holds the
// EMPTY key). The rows probe in rounds: each round reads the current slot of every row not yet placed
// (those reads in flight together), then claims empty slots and moves the others one slot on, so a
// step waits on about one LDS read per round rather than one per row. Sets `over` and applies nothing
// when a key finds no slot.
template <int ROWS>
__device__ __forceinline__ void bg_apply(const BgRows<ROWS>& x, uint32_t i0, uint32_t end, unsigned long long* lkey,
                                         unsigned long long* lsum, uint32_t* lcnt, uint32_t cap, bool& over) {
#pragma unroll
  for (int r = 0; r < ROWS; r++) asm volatile("" ::"v"(x.k[r]), "v"(x.v[r]));   // every loaded register taken here
  if (over) return;
  const uint32_t mask = cap - 1u;
  uint32_t s[ROWS];
  uint32_t live = 0, todo = 0;   // rows of the chunk; rows whose slot is not known yet
#pragma unroll
  for (int r = 0; r < ROWS; r++) {
    const bool in = i0 + uint32_t(r) * blockDim.x < end;
    live |= uint32_t(in) << r;
    const bool key = in && x.k[r] != BG_EMPTY;
    todo |= uint32_t(key) << r;
    s[r] = key ? bg_hash(x.k[r], mask) : cap;
  }
  for (uint32_t round = 0; todo; round++) {
    if (round >= cap) {
      over = true;
      return;
    }
    unsigned long long cur[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; r++) cur[r] = (todo >> r) & 1u ? lkey[s[r]] : 0ull;
#pragma unroll
    for (int r = 0; r < ROWS; r++) {
      if (!((todo >> r) & 1u)) continue;
      const unsigned long long k = x.k[r];
      unsigned long long c = cur[r];
      if (c == BG_EMPTY) {
        c = atomicCAS(&lkey[s[r]], BG_EMPTY, k);
        if (c == BG_EMPTY) c = k;
      }
      if (c == k) todo &= ~(1u << r);
      else s[r] = (s[r] + 1u) & mask;
    }
  }
#pragma unroll
  for (int r = 0; r < ROWS; r++) {
    if (!((live >> r) & 1u)) continue;
    atomicAdd(&lsum[s[r]], x.v[r]);
    atomicAdd(&lcnt[s[r]], 1u);
  }
}

// Group-by strategy 0 (LDS). Persistent: workgroups take chunks from the queue until it is empty (an
// exit every wave reaches). Each chunk is aggregated in an LDS table of `cap` slots (+1 for the
// EMPTY key; counts in 32 bits, a chunk holding at most CHUNK_ROWS rows) and flushed into its batch's
// table of the same shape. A batch with more distinct keys than `cap` is flagged. WIDE: 64-bit keys
// and values, loaded directly with ROWS rows in flight per thread; otherwise one row at a time, of any
// integer types.
// Threads per workgroup of bg_agg and 64-bit rows per thread per step. The LDS table (20 B per slot)
// bounds the workgroups per CU (three at 2,048 slots), so 512 threads give twice the waves of 256 in the
// same LDS; 4 rows per step keep a wave within the ~85 VGPRs that six waves per SIMD leave (191.4M
// rows in 2,048 batches: 1.74 ms at 256 x 8, 1.20 ms at 512 x 4, 1.21 ms at 512 x 6).
#ifndef AM_BG_THREADS
#define AM_BG_THREADS 512
#endif
#ifndef AM_BG_ROWS
#define AM_BG_ROWS 4
#endif
constexpr uint32_t BG_THREADS = AM_BG_THREADS;

template <bool WIDE>
__global__ __launch_bounds__(BG_THREADS) void bg_agg(const void* keys, int ktype, const void* vals, int vtype,
                                              const SegChunk* chunks, uint32_t nchunks, SegCtl* ctl,
                                              unsigned long long* tkey, unsigned long long* tsum,
                                              unsigned long long* tcnt, uint32_t cap, uint32_t* overflow) {
  constexpr int ROWS = WIDE ? AM_BG_ROWS : 1;
  static_assert(CHUNK_ROWS <= 0xFFFFFFFFu, "chunk counts fit 32 bits");
  extern __shared__ unsigned long long lds[];
  unsigned long long* lkey = lds;                                    // [cap + 1]
  unsigned long long* lsum = lds + (cap + 1);                        // [cap + 1]
  uint32_t* lcnt = reinterpret_cast<uint32_t*>(lds + 2 * (cap + 1));   // [cap + 1]
  __shared__ uint32_t s_chunk, s_over;
  const uint32_t mask = cap - 1u;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) {
      s_chunk = atomicAdd(&ctl->chunk_counter, 1u);
      s_over = 0u;
    }
    for (uint32_t j = threadIdx.x; j <= cap; j += blockDim.x) {
      lkey[j] = BG_EMPTY;
      lsum[j] = 0ull;
      lcnt[j] = 0u;
    }
    __syncthreads();
    const uint32_t ci = s_chunk;
    if (ci >= nchunks) break;
    const SegChunk c = chunks[ci];
    bool over = false;
    if constexpr (WIDE) {
      // Two row buffers take turns: the next step's rows are loaded before this step's are applied,
      // so that the loads are in flight while this step works in LDS. The step count is the same for
      // every thread; rows past the chunk's end read its last row and are not applied.
      const auto* k64 = static_cast<const unsigned long long*>(keys);
      const auto* v64 = static_cast<const unsigned long long*>(vals);
      const uint32_t stride = blockDim.x * ROWS;
      const int steps = int((c.end - c.start + stride - 1u) / stride);
      BgRows<ROWS> xa, xb;
      bg_load_wide(k64, v64, c.start + threadIdx.x, c.end, xa);
      for (int t = 0; t < steps; t += 2) {
        const uint32_t i0 = c.start + threadIdx.x + uint32_t(t) * stride;
        if (t + 1 < steps) bg_load_wide(k64, v64, i0 + stride, c.end, xb);
        bg_apply(xa, i0, c.end, lkey, lsum, lcnt, cap, over);
        if (t + 1 >= steps) break;
        if (t + 2 < steps) bg_load_wide(k64, v64, i0 + 2u * stride, c.end, xa);
        bg_apply(xb, i0 + stride, c.end, lkey, lsum, lcnt, cap, over);
      }
    } else {
      for (uint32_t i0 = c.start + threadIdx.x; i0 < c.end && !over; i0 += blockDim.x * ROWS) {
        BgRows<ROWS> x;
#pragma unroll
        for (int r = 0; r < ROWS; r++) {
          const uint32_t i = min(i0 + uint32_t(r) * blockDim.x, c.end - 1u);
          x.k[r] = bg_key(keys, i, ktype);
          x.v[r] = (unsigned long long)bg_value(vals, i, vtype);
        }
        bg_apply(x, i0, c.end, lkey, lsum, lcnt, cap, over);
      }
    }
    if (over) atomicOr(&s_over, 1u);
    __syncthreads();
    if (s_over) {
      if (threadIdx.x == 0) atomicOr(&overflow[c.seg], 1u);
      continue;
    }
    // Flush into the batch's table.
    unsigned long long* gk = tkey + size_t(c.seg) * (cap + 1);
    unsigned long long* gs = tsum + size_t(c.seg) * (cap + 1);
    unsigned long long* gc = tcnt + size_t(c.seg) * (cap + 1);
    if (c.whole) {
      // The chunk is its batch's only one: its table, of the batch table's shape and hash, is the
      // batch's table as it stands (slot for slot, with plain stores; the rest stay empty).
      for (uint32_t j = threadIdx.x; j <= cap; j += blockDim.x) {
        const uint32_t cnt = lcnt[j];
        if (!cnt) continue;
        gk[j] = lkey[j];
        gs[j] = lsum[j];
        gc[j] = cnt;
      }
      continue;
    }
    bool gover = false;
    for (uint32_t j = threadIdx.x; j <= cap; j += blockDim.x) {
      const uint32_t cnt = lcnt[j];
      if (!cnt) continue;
      unsigned long long k = lkey[j];
      uint32_t s;
      if (j == cap) {
        s = cap;
      } else {
        s = bg_hash(k, mask);
        uint32_t probes = 0;
        while (true) {
          unsigned long long cur = gk[s];
          if (cur == k) break;
          if (cur == BG_EMPTY) {
            unsigned long long prev = atomicCAS(&gk[s], BG_EMPTY, k);
            if (prev == BG_EMPTY || prev == k) break;
          }
          if (++probes >= cap) { gover = true; break; }
          s = (s + 1u) & mask;
        }
        if (gover) break;
      }
      atomicAdd(&gs[s], lsum[j]);
      atomicAdd(&gc[s], (unsigned long long)cnt);
    }
    if (gover) atomicOr(&overflow[c.seg], 1u);
  }
}

// One block per batch: the batch's groups sorted ascending by key (signed or unsigned order) with a
// bitonic sort in LDS over `pow2` >= cap + 1 entries; writes the group count of the batch and its
// groups, in order, to the batch's staging area.
__global__ __launch_bounds__(256) void bg_sort(const unsigned long long* tkey, const unsigned long long* tsum,
                                               const unsigned long long* tcnt, uint32_t cap, uint32_t pow2,
                                               int signed_keys, unsigned long long* skey, unsigned long long* ssum,
                                               unsigned long long* scnt, uint32_t* gcount) {
  extern __shared__ unsigned long long lds[];
  unsigned long long* ok = lds;                                      // [pow2] order keys
  uint32_t* sl = reinterpret_cast<uint32_t*>(lds + pow2);           // [pow2] slots
  const uint32_t seg = blockIdx.x;
  const size_t off = size_t(seg) * (cap + 1);
  __shared__ uint32_t s_count;
  if (threadIdx.x == 0) s_count = 0u;
  __syncthreads();
  // The occupied slots, compacted in any order, then padded with all-ones keys (which sort last) up to
  // the next power of two.
  for (uint32_t j = threadIdx.x; j <= cap; j += blockDim.x) {
    if (tcnt[off + j] == 0) continue;
    const unsigned long long k = j == cap ? BG_EMPTY : tkey[off + j];
    const uint32_t at = atomicAdd(&s_count, 1u);
    ok[at] = signed_keys ? k ^ 0x8000000000000000ull : k;
    sl[at] = j;
  }
  __syncthreads();
  const uint32_t D = s_count;
  uint32_t p2 = 1;
  while (p2 < D) p2 <<= 1;   // at most pow2: D <= cap + 1
  for (uint32_t j = D + threadIdx.x; j < p2; j += blockDim.x) {
    ok[j] = ~0ull;
    sl[j] = 0xFFFFFFFFu;
  }
  __syncthreads();
  // Bitonic sort by (order key, slot) over the p2 entries.
  for (uint32_t k = 2; k <= p2; k <<= 1) {
    for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
      for (uint32_t i = threadIdx.x; i < p2; i += blockDim.x) {
        uint32_t l = i ^ jj;
        if (l > i) {
          bool up = (i & k) == 0;
          unsigned long long a = ok[i], b = ok[l];
          uint32_t sa = sl[i], sb = sl[l];
          bool gt = a > b || (a == b && sa > sb);
          if (gt == up) {
            ok[i] = b;
            ok[l] = a;
            sl[i] = sb;
            sl[l] = sa;
          }
        }
      }
      __syncthreads();
    }
  }
  for (uint32_t r = threadIdx.x; r < D; r += blockDim.x) {
    uint32_t j = sl[r];
    skey[off + r] = j == cap ? BG_EMPTY : tkey[off + j];
    ssum[off + r] = tsum[off + j];
    scnt[off + r] = tcnt[off + j];
  }
  if (threadIdx.x == 0) gcount[seg] = D;
}

// Exclusive scan of the per-batch group counts (one block, any number of batches).
__global__ __launch_bounds__(256) void bg_scan(const uint32_t* gcount, uint32_t nseg, unsigned long long* goff) {
  __shared__ uint32_t tmp[WAVES];
  __shared__ unsigned long long carry;
  if (threadIdx.x == 0) carry = 0ull;
  __syncthreads();
  for (uint32_t base = 0; base < nseg; base += BLOCK) {
    uint32_t i = base + threadIdx.x;
    uint32_t v = i < nseg ? gcount[i] : 0u;
    uint32_t total;
    uint32_t ex = block_excl_scan(v, tmp, total);
    if (i < nseg) goff[i] = carry + ex;
    __syncthreads();
    if (threadIdx.x == 0) carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) goff[nseg] = carry;
}

__global__ void bg_emit(const unsigned long long* skey, const unsigned long long* ssum,
                        const unsigned long long* scnt, const uint32_t* gcount, const unsigned long long* goff,
                        uint32_t cap, unsigned long long* okey, unsigned long long* osum,
                        unsigned long long* ocnt) {
  const uint32_t seg = blockIdx.x;
  const size_t off = size_t(seg) * (cap + 1);
  const unsigned long long at = goff[seg];
  for (uint32_t r = threadIdx.x; r < gcount[seg]; r += blockDim.x) {
    okey[at + r] = skey[off + r];
    osum[at + r] = ssum[off + r];
    ocnt[at + r] = scnt[off + r];
  }
}

}  // namespace

// ================================================================ host

// Widens group keys of any integer type to the 64-bit staging form (sign- or zero-extended).
__global__ void bg_widen(const void* keys, int ktype, uint32_t n, unsigned long long* out) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = bg_key(keys, i, ktype);
}

struct Batched {
  std::vector<uint32_t> offsets;     // host copy, nseg + 1
  uint32_t nseg = 0, n = 0;
  // Tuned launch parameters.
  int items = 16;                    // sort tile = 256 * items keys
  uint32_t cap = 2048;               // group-by hash capacity per batch (a power of two)
  int strategy = 0;                  // group-by: 0 LDS table per chunk, 1 global table per batch
  DevBuf seg_off, tiles, chunks, gb_chunks, hist, base, ctl, epoch, status;
  uint32_t ntiles = 0, nchunks = 0, ngb_chunks = 0;
  DevBuf k0, k1, r0, r1, out_idx;
  MsdWork msd;                       // the levels of the MSD argsort of 64-bit keys
  bool msd_used = false;             // the last recorded sort was the MSD one
  bool use_msd = true;               // 64-bit sort keys take the MSD sort (the tuner may pick the LSD one)
  bool msd_overflowed = false;       // a run overflowed the MSD bucket list: the LSD passes from then on
  float tuned_sort_ms = 0.f;         // the tuner's best sort time (diagnostics)
  DevBuf tkey, tsum, tcnt, overflow, skey, ssum, scnt, gcount, goff, okey, osum, ocnt;
  size_t out_slots = 0;              // capacity of okey / osum / ocnt
  // What the last run produced.
  bool ran_sort = false, ran_group = false;
  std::string key_format = "l";
  int key_type = T_I64;
  uint64_t groups = 0;
  uint32_t fallback_batches = 0;
  // Graph of the whole pipeline for the inputs it was captured with.
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* g_sort = nullptr;
  const void* g_keys = nullptr;
  const void* g_vals = nullptr;
  int g_sig = -1;
  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~Batched() { drop_graph(); }
};

namespace {

template <typename K, int W, int ITEMS>
void seg_sort_launch(Batched& b, const void* vals, int type, int desc, int forder, hipStream_t s) {
  constexpr int P = int(sizeof(K));
  auto* ctl = static_cast<SegCtl*>(b.ctl->ptr);
  size_t hist_words = size_t(b.nseg) * P * RADIX;
  hipLaunchKernelGGL(seg_init, dim3(std::max<int>(1, int(std::min<size_t>((hist_words + 255) / 256, 2048)))),
                     dim3(256), 0, s, ctl, static_cast<uint32_t*>(b.hist->ptr), hist_words,
                     static_cast<uint32_t*>(b.epoch->ptr));
  if (b.nchunks > 0)
    hipLaunchKernelGGL((seg_upsweep<K, W>), dim3(b.nchunks), dim3(BLOCK), 0, s, vals, type, forder, desc,
                       static_cast<const SegChunk*>(b.chunks->ptr), ctl, static_cast<uint32_t*>(b.hist->ptr));
  hipLaunchKernelGGL(seg_plan<P>, dim3(b.nseg), dim3(BLOCK), 0, s, static_cast<const uint32_t*>(b.hist->ptr),
                     static_cast<const uint32_t*>(b.seg_off->ptr), static_cast<uint32_t*>(b.base->ptr), ctl);
  auto* st = static_cast<unsigned long long*>(b.status->ptr);
  K* kp0 = static_cast<K*>(b.k0->ptr);
  K* kp1 = static_cast<K*>(b.k1->ptr);
  uint32_t* rp0 = static_cast<uint32_t*>(b.r0->ptr);
  uint32_t* rp1 = static_cast<uint32_t*>(b.r1->ptr);
  uint32_t* out = static_cast<uint32_t*>(b.out_idx->ptr);
  const SegTile* tiles = static_cast<const SegTile*>(b.tiles->ptr);
  const uint32_t* so = static_cast<const uint32_t*>(b.seg_off->ptr);
  const uint32_t* base = static_cast<const uint32_t*>(b.base->ptr);
  {
    ProfScope ps("seg_first");
    hipLaunchKernelGGL((seg_pass<K, ITEMS, true, W>), dim3(b.ntiles), dim3(BLOCK), 0, s, 0, vals, type, forder, desc,
                       tiles, so, base, ctl, st, kp0, rp0, kp1, rp1, out);
  }
  for (int p = 0; p < P; p++) {
    ProfScope ps("seg_pass");
    hipLaunchKernelGGL((seg_pass<K, ITEMS, false, 0>), dim3(b.ntiles), dim3(BLOCK), 0, s, p, vals, type, forder, desc,
                       tiles, so, base, ctl, st, kp0, rp0, kp1, rp1, out);
  }
  AM_HIP(hipGetLastError());
}

// The batched argsort of 64-bit keys as an adaptive MSD radix sort (blind levels, graph-capturable).
// Keys per thread in a level's pass: 16 (its persistent form stays at three workgroups per CU).
void seg_sort_msd(Batched& b, const void* vals, int type, int desc, int forder, hipStream_t s) {
  using K = unsigned long long;
  constexpr int P = 8;
  auto* ctl = static_cast<SegCtl*>(b.ctl->ptr);
  hipLaunchKernelGGL(seg_init, dim3(1), dim3(256), 0, s, ctl, static_cast<uint32_t*>(b.hist->ptr), size_t(0),
                     static_cast<uint32_t*>(b.epoch->ptr));
  K* kp0 = static_cast<K*>(b.k0->ptr);
  K* kp1 = static_cast<K*>(b.k1->ptr);
  uint32_t* rp0 = static_cast<uint32_t*>(b.r0->ptr);
  uint32_t* rp1 = static_cast<uint32_t*>(b.r1->ptr);
  if (b.nchunks > 0) {
    ProfScope ps("seg_msd_map");
    hipLaunchKernelGGL((seg_msd_map<K, 8>), dim3(b.nchunks), dim3(BLOCK), 0, s, vals, type, forder, desc,
                       static_cast<const SegChunk*>(b.chunks->ptr), ctl, kp0);
  }
  hipLaunchKernelGGL(seg_msd_ctl<P>, dim3(1), dim3(BLOCK), 0, s, ctl, b.msd.mc, b.msd.lc);
  hipLaunchKernelGGL(seg_msd_lists<P>, dim3(std::min<uint32_t>((b.nseg + 255) / 256, 1024)), dim3(256), 0, s,
                     static_cast<const uint32_t*>(b.seg_off->ptr), b.nseg, b.msd.mc, LOCAL_MAX / 2, b.msd.large[0],
                     b.msd.local, b.msd.local_cap, b.msd.lc);
  AM_HIP(hipGetLastError());
  msd_levels<K, 16>(b.msd, b.n, P, 0xFFu, true, LOCAL_MAX / 2, static_cast<unsigned long long*>(b.status->ptr), kp0,
                    rp0, kp1, rp1, static_cast<uint32_t*>(b.out_idx->ptr), s);
}

template <typename K, int W>
void seg_sort_items(Batched& b, const void* vals, int type, int desc, int forder, hipStream_t s) {
  if (b.items == 8) seg_sort_launch<K, W, 8>(b, vals, type, desc, forder, s);
  else if (b.items == 20) seg_sort_launch<K, W, 20>(b, vals, type, desc, forder, s);
  else seg_sort_launch<K, W, 16>(b, vals, type, desc, forder, s);
}

void seg_sort(Batched& b, const void* vals, int type, int desc, int forder, hipStream_t s) {
  b.msd_used = type_width(type) == 8 && b.use_msd;
  if (b.msd_used) {
    seg_sort_msd(b, vals, type, desc, forder, s);
    return;
  }
  switch (type_width(type)) {
    case 8: seg_sort_items<unsigned long long, 8>(b, vals, type, desc, forder, s); break;
    case 4: seg_sort_items<uint32_t, 4>(b, vals, type, desc, forder, s); break;
    case 2: seg_sort_items<uint32_t, 2>(b, vals, type, desc, forder, s); break;
    default: seg_sort_items<uint32_t, 1>(b, vals, type, desc, forder, s); break;
  }
}

uint32_t pow2_at_least(uint32_t v) {
  uint32_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// LDS bytes of the group-by kernels at capacity `cap`.
size_t agg_lds(uint32_t cap) { return size_t(cap + 1) * 20; }
size_t sort_lds(uint32_t cap) { return size_t(pow2_at_least(cap + 1)) * 12; }

void group_launch(Batched& b, const void* keys, int ktype, const void* vals, int vtype, hipStream_t s) {
  const size_t slots = size_t(b.nseg) * (b.cap + 1);
  auto* tk = static_cast<unsigned long long*>(b.tkey->ptr);
  auto* ts = static_cast<unsigned long long*>(b.tsum->ptr);
  auto* tc = static_cast<unsigned long long*>(b.tcnt->ptr);
  auto* ctl = static_cast<SegCtl*>(b.ctl->ptr);
  hipLaunchKernelGGL(bg_init, dim3(int(std::min<size_t>((slots + 255) / 256, 4096))), dim3(256), 0, s, tk, ts, tc,
                     slots, static_cast<uint32_t*>(b.overflow->ptr), b.nseg, ctl);
  if (b.ngb_chunks > 0) {
    ProfScope ps("bg_agg");
    // As many persistent workgroups as the LDS tables let each CU hold (at most 4).
    const int per_cu = int(std::clamp<size_t>(device().lds_per_cu / agg_lds(b.cap), 1, 4));
    int grid = std::max(1, std::min<int>(int(b.ngb_chunks), device().cu_count * per_cu));
    if (b.strategy == 1) {
      hipLaunchKernelGGL(bg_agg_global, dim3(std::min<int>(int(b.ngb_chunks), device().cu_count * 8)), dim3(256), 0, s,
                         keys, ktype, vals, vtype, static_cast<const SegChunk*>(b.gb_chunks->ptr), b.ngb_chunks, ctl,
                         tk, ts, tc, b.cap, static_cast<uint32_t*>(b.overflow->ptr));
    } else {
      const bool wide = type_width(ktype) == 8 && type_width(vtype) == 8;
      hipLaunchKernelGGL(wide ? bg_agg<true> : bg_agg<false>, dim3(grid), dim3(BG_THREADS), agg_lds(b.cap), s, keys, ktype,
                         vals, vtype, static_cast<const SegChunk*>(b.gb_chunks->ptr), b.ngb_chunks, ctl, tk, ts, tc,
                         b.cap, static_cast<uint32_t*>(b.overflow->ptr));
    }
  }
  const uint32_t pow2 = pow2_at_least(b.cap + 1);
  {
    ProfScope ps("bg_sort");
    hipLaunchKernelGGL(bg_sort, dim3(b.nseg), dim3(256), sort_lds(b.cap), s, tk, ts, tc, b.cap, pow2,
                       type_is_signed(ktype) ? 1 : 0, static_cast<unsigned long long*>(b.skey->ptr),
                       static_cast<unsigned long long*>(b.ssum->ptr), static_cast<unsigned long long*>(b.scnt->ptr),
                       static_cast<uint32_t*>(b.gcount->ptr));
  }
  hipLaunchKernelGGL(bg_scan, dim3(1), dim3(BLOCK), 0, s, static_cast<const uint32_t*>(b.gcount->ptr), b.nseg,
                     static_cast<unsigned long long*>(b.goff->ptr));
  hipLaunchKernelGGL(bg_emit, dim3(b.nseg), dim3(256), 0, s, static_cast<const unsigned long long*>(b.skey->ptr),
                     static_cast<const unsigned long long*>(b.ssum->ptr),
                     static_cast<const unsigned long long*>(b.scnt->ptr), static_cast<const uint32_t*>(b.gcount->ptr),
                     static_cast<const unsigned long long*>(b.goff->ptr), b.cap,
                     static_cast<unsigned long long*>(b.okey->ptr), static_cast<unsigned long long*>(b.osum->ptr),
                     static_cast<unsigned long long*>(b.ocnt->ptr));
  AM_HIP(hipGetLastError());
}

DevBuf persistent(size_t bytes) {
  // Plan buffers live as long as the plan, outside any batch's retention.
  void* p = nullptr;
  AM_HIP(hipMalloc(&p, std::max<size_t>(bytes, 8)));
  return device_wrap(p, bytes, [](void* q) { (void)hipFree(q); });
}

// Sort tiles (never crossing a batch) and the look-back status table for tiles of 256 * items keys.
void build_tiles(Batched& b, int items) {
  const uint32_t TILE = 256u * uint32_t(items);
  std::vector<SegTile> tiles;
  for (uint32_t sgi = 0; sgi < b.nseg; sgi++) {
    uint32_t a = b.offsets[sgi], e = b.offsets[sgi + 1];
    uint32_t head = uint32_t(tiles.size());
    for (uint32_t x = a; x < e; x += TILE) tiles.push_back({sgi, x, std::min(e, x + TILE), head});
  }
  if (tiles.empty()) tiles.push_back({0, 0, 0, 0});
  b.drop_graph();
  b.items = items;
  b.ntiles = uint32_t(tiles.size());
  b.tiles = persistent(tiles.size() * sizeof(SegTile));
  AM_HIP(hipMemcpy(b.tiles->ptr, tiles.data(), tiles.size() * sizeof(SegTile), hipMemcpyHostToDevice));
  // Status words carry the epoch of the pass that wrote them; zero is older than every epoch. Rows
  // for the LSD passes' tiles, or for a level of the MSD sort (16 keys per thread, up to one more
  // tile per large bucket), whichever is more.
  const size_t rows = std::max<size_t>(b.ntiles, size_t(b.n) / 4096u + b.msd.large_cap + 1u);
  b.status = persistent(rows * STATUS_STRIDE * 8);
  AM_HIP(hipMemset(b.status->ptr, 0, rows * STATUS_STRIDE * 8));
  AM_HIP(hipMemset(b.epoch->ptr, 0, 4));
}

// The per-batch group tables for capacity `cap`.
void build_tables(Batched& b, uint32_t cap) {
  b.drop_graph();
  b.cap = cap;
  size_t slots = size_t(b.nseg) * (cap + 1);
  b.tkey = persistent(slots * 8);
  b.tsum = persistent(slots * 8);
  b.tcnt = persistent(slots * 8);
  b.skey = persistent(slots * 8);
  b.ssum = persistent(slots * 8);
  b.scnt = persistent(slots * 8);
  b.okey = persistent(slots * 8);
  b.osum = persistent(slots * 8);
  b.ocnt = persistent(slots * 8);
  b.out_slots = slots;
}

uint32_t max_cap() {
  // Both group-by kernels keep their table in LDS: the capacity is bounded by one workgroup's LDS.
  uint32_t c = 256;
  while (agg_lds(c * 2) <= device().lds_per_block && sort_lds(c * 2) <= device().lds_per_block) c *= 2;
  return c;
}

// Batches whose group table overflowed are grouped again, one at a time, by the general hash group-by
// (am_group_by_keys + sum + count_all on a slice of the columns), and the outputs are reassembled in
// batch order.
void regroup_overflowed(Batched& b, const Array& keys, const Array& vals, const std::vector<uint32_t>& over) {
  hipStream_t s = stream();
  std::vector<uint32_t> gcount(b.nseg);
  AM_HIP(hipMemcpyAsync(gcount.data(), b.gcount->ptr, size_t(b.nseg) * 4, hipMemcpyDeviceToHost, s));
  AM_HIP(hipStreamSynchronize(s));
  struct Redo {
    uint32_t seg;
    Array *k, *sum, *cnt;
  };
  std::vector<Redo> redo;
  auto release_all = [&] {
    for (Redo& r : redo) {
      release_array(r.k);
      release_array(r.sum);
      release_array(r.cnt);
    }
  };
  try {
    for (uint32_t sg = 0; sg < b.nseg; sg++) {
      if (!over[sg]) continue;
      const int64_t at = b.offsets[sg], len = int64_t(b.offsets[sg + 1]) - at;
      Array* ks = new_array(keys.format, keys.type, len);
      ks->values = keys.values;
      ks->offset = keys.offset + at;
      Array* vs = new_array(vals.format, vals.type, len);
      vs->values = vals.values;
      vs->offset = vals.offset + at;
      GroupBy* g = nullptr;
      Redo r{sg, nullptr, nullptr, nullptr};
      try {
        g = group_by_keys(&ks, 1);
        r.k = group_by_keys_result(g, 0);
        r.sum = group_agg(g, vs, 0, 0.0);
        r.cnt = group_agg(g, nullptr, 1, 0.0);
      } catch (...) {
        if (g) group_by_release(g);
        release_array(ks);
        release_array(vs);
        if (r.k) release_array(r.k);
        if (r.sum) release_array(r.sum);
        throw;
      }
      group_by_release(g);
      release_array(ks);
      release_array(vs);
      gcount[sg] = uint32_t(r.k->length);
      redo.push_back(r);
    }
    std::vector<unsigned long long> goff(b.nseg + 1, 0ull);
    for (uint32_t sg = 0; sg < b.nseg; sg++) goff[sg + 1] = goff[sg] + gcount[sg];
    const size_t total = size_t(goff[b.nseg]);
    if (total > b.out_slots) {
      b.okey = persistent(total * 8);
      b.osum = persistent(total * 8);
      b.ocnt = persistent(total * 8);
      b.out_slots = total;
      b.drop_graph();   // the captured emit wrote to the old buffers
    }
    std::vector<uint32_t> kept(gcount);
    for (const Redo& r : redo) kept[r.seg] = 0u;
    DevBuf kept_d = device_alloc(size_t(b.nseg) * 4);
    AM_HIP(hipMemcpyAsync(kept_d->ptr, kept.data(), size_t(b.nseg) * 4, hipMemcpyHostToDevice, s));
    AM_HIP(hipMemcpyAsync(b.goff->ptr, goff.data(), goff.size() * 8, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(bg_emit, dim3(b.nseg), dim3(256), 0, s, static_cast<const unsigned long long*>(b.skey->ptr),
                       static_cast<const unsigned long long*>(b.ssum->ptr),
                       static_cast<const unsigned long long*>(b.scnt->ptr), static_cast<const uint32_t*>(kept_d->ptr),
                       static_cast<const unsigned long long*>(b.goff->ptr), b.cap,
                       static_cast<unsigned long long*>(b.okey->ptr), static_cast<unsigned long long*>(b.osum->ptr),
                       static_cast<unsigned long long*>(b.ocnt->ptr));
    AM_HIP(hipGetLastError());
    for (const Redo& r : redo) {
      const uint32_t m = gcount[r.seg];
      if (!m) continue;
      const size_t at = size_t(goff[r.seg]);
      hipLaunchKernelGGL(bg_widen, dim3((m + 255) / 256), dim3(256), 0, s, r.k->data(), r.k->type, m,
                         static_cast<unsigned long long*>(b.okey->ptr) + at);
      AM_HIP(hipGetLastError());
      AM_HIP(hipMemcpyAsync(static_cast<unsigned long long*>(b.osum->ptr) + at, r.sum->data(), size_t(m) * 8,
                            hipMemcpyDeviceToDevice, s));
      AM_HIP(hipMemcpyAsync(static_cast<unsigned long long*>(b.ocnt->ptr) + at, r.cnt->data(), size_t(m) * 8,
                            hipMemcpyDeviceToDevice, s));
    }
    AM_HIP(hipStreamSynchronize(s));
    b.groups = total;
  } catch (...) {
    release_all();
    throw;
  }
  release_all();
}

}  // namespace

Batched* batched_create(const int64_t* offsets, int64_t nbatches, int hash_capacity, int items) {
  if (nbatches < 1) fail("batched: at least one batch");
  if (offsets[0] != 0) fail("batched: offsets must start at 0");
  for (int64_t i = 0; i < nbatches; i++)
    if (offsets[i + 1] < offsets[i]) fail("batched: offsets must not decrease");
  int64_t n = offsets[nbatches];
  check_length(n);
  check_index_rows(n, "batched argsort");
  auto* b = new Batched();
  try {
    b->nseg = uint32_t(nbatches);
    b->n = uint32_t(n);
    b->offsets.assign(offsets, offsets + nbatches + 1);
    std::vector<SegChunk> chunks, gchunks;
    for (uint32_t sgi = 0; sgi < b->nseg; sgi++) {
      uint32_t a = b->offsets[sgi], e = b->offsets[sgi + 1];
      for (uint32_t x = a; x < e; x += 65536u) chunks.push_back({sgi, x, std::min(e, x + 65536u), 0});
      for (uint32_t x = a; x < e; x += CHUNK_ROWS)
        gchunks.push_back({sgi, x, std::min(e, x + CHUNK_ROWS), uint32_t(x == a && e - a <= CHUNK_ROWS)});
    }
    b->nchunks = uint32_t(chunks.size());
    b->ngb_chunks = uint32_t(gchunks.size());
    std::vector<uint32_t> so(b->offsets.begin(), b->offsets.end());
    b->seg_off = persistent(so.size() * 4);
    b->chunks = persistent(std::max<size_t>(chunks.size(), 1) * sizeof(SegChunk));
    b->gb_chunks = persistent(std::max<size_t>(gchunks.size(), 1) * sizeof(SegChunk));
    AM_HIP(hipMemcpy(b->seg_off->ptr, so.data(), so.size() * 4, hipMemcpyHostToDevice));
    if (!chunks.empty())
      AM_HIP(hipMemcpy(b->chunks->ptr, chunks.data(), chunks.size() * sizeof(SegChunk), hipMemcpyHostToDevice));
    if (!gchunks.empty())
      AM_HIP(hipMemcpy(b->gb_chunks->ptr, gchunks.data(), gchunks.size() * sizeof(SegChunk), hipMemcpyHostToDevice));
    b->hist = persistent(size_t(b->nseg) * 8 * RADIX * 4);
    b->base = persistent(size_t(b->nseg) * 8 * RADIX * 4);
    b->ctl = persistent(sizeof(SegCtl));
    b->epoch = persistent(4);
    b->msd = msd_work(b->n, 16, persistent, b->nseg);
    build_tiles(*b, (items == 8 || items == 20) ? items : 16);
    size_t nn = std::max<size_t>(n, 1);
    b->k0 = persistent(nn * 8);
    b->k1 = persistent(nn * 8);
    b->r0 = persistent(nn * 4);
    b->r1 = persistent(nn * 4);
    b->out_idx = persistent(nn * 4);
    uint32_t cap = pow2_at_least(uint32_t(std::max(hash_capacity, 256)));
    build_tables(*b, std::min(cap, max_cap()));
    b->overflow = persistent(size_t(b->nseg) * 4);
    b->gcount = persistent(size_t(b->nseg) * 4);
    b->goff = persistent(size_t(b->nseg + 1) * 8);
    AM_HIP(hipDeviceSynchronize());
  } catch (...) {
    delete b;
    throw;
  }
  return b;
}

void batched_release(Batched* b) {
  if (b) {
    (void)hipStreamSynchronize(stream());
    delete b;
  }
}

namespace {

void check_inputs(const Batched* b, const Array* sort_keys, const Array* group_keys, const Array* values,
                  int forder) {
  for (const Array* a : {sort_keys, group_keys, values}) {
    if (!a) continue;
    if (a->length != int64_t(b->n))
      fail("batched: column length " + std::to_string(a->length) + " is not the batches' total " +
           std::to_string(b->n));
    if (a->has_nulls()) fail("batched: columns with nulls are not taken by the batched path");
  }
  if (sort_keys && !type_is_primitive(sort_keys->type))
    fail("batched: unsupported sort key type " + sort_keys->format);
  if ((group_keys != nullptr) != (values != nullptr)) fail("batched: group keys and values go together");
  if (group_keys && (!type_is_int(group_keys->type) || !type_is_int(values->type)))
    fail("batched: group-by takes integer keys and integer values");
  if (forder < 0 || forder > 2) fail_invalid("unknown float_order " + std::to_string(forder));
}

// The kernel sequence of one run, on stream s.
void record(Batched* b, const Array* sort_keys, const Array* group_keys, const Array* values, int desc, int forder,
            hipStream_t s) {
  if (sort_keys) seg_sort(*b, sort_keys->data(), sort_keys->type, desc, forder, s);
  if (group_keys) group_launch(*b, group_keys->data(), group_keys->type, values->data(), values->type, s);
}

}  // namespace

// Runs the argsort of `sort_keys` (desc / float order per batch) and the group-by sum / count of
// `values` by `group_keys` per batch. With use_graph, the sequence is captured once per set of input
// buffers and replayed afterwards. The only readback is the per-batch overflow flags of the group-by.
void batched_run(Batched* b, const Array* sort_keys, const Array* group_keys, const Array* values, int desc,
                 int forder, bool use_graph) {
  check_inputs(b, sort_keys, group_keys, values, forder);
  hipStream_t s = stream();
  if (!use_graph) {
    record(b, sort_keys, group_keys, values, desc, forder, s);
  } else {
    int sig = (desc ? 1 : 0) | (forder << 1) | ((sort_keys ? sort_keys->type : 15) << 3) |
              ((group_keys ? group_keys->type : 15) << 8) | ((values ? values->type : 15) << 13);
    const void* ps = sort_keys ? sort_keys->data() : nullptr;
    const void* pk = group_keys ? group_keys->data() : nullptr;
    const void* pv = values ? values->data() : nullptr;
    if (!b->exec || b->g_sort != ps || b->g_keys != pk || b->g_vals != pv || b->g_sig != sig) {
      b->drop_graph();
      AM_HIP(hipStreamSynchronize(s));
      AM_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        record(b, sort_keys, group_keys, values, desc, forder, s);
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      AM_HIP(hipStreamEndCapture(s, &b->graph));
      AM_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      b->g_sort = ps;
      b->g_keys = pk;
      b->g_vals = pv;
      b->g_sig = sig;
    }
    ProfScope pscope("batched_graph");
    AM_HIP(hipGraphLaunch(b->exec, s));
  }
  if (sort_keys && b->msd_used) {
    // The MSD sort's bucket list has room for every batch plus m / 32 more buckets; if a run needed
    // more, this plan sorts with the LSD passes from now on, and this run's sort is done again.
    uint32_t words[LC_READ];
    AM_HIP(hipMemcpyAsync(words, b->msd.lc, sizeof(words), hipMemcpyDeviceToHost, s));
    AM_HIP(hipStreamSynchronize(s));
    if (words[LC_OVERFLOW] || uint64_t(words[LC_LOCAL]) + words[LC_GROUPS] > b->msd.local_cap) {
      b->use_msd = false;
      b->msd_overflowed = true;
      b->drop_graph();
      seg_sort(*b, sort_keys->data(), sort_keys->type, desc, forder, s);
    }
  }
  b->ran_sort = sort_keys != nullptr;
  b->ran_group = group_keys != nullptr;
  b->fallback_batches = 0;
  if (group_keys) {
    b->key_format = group_keys->format;
    b->key_type = group_keys->type;
    std::vector<uint32_t> over(b->nseg);
    unsigned long long total = 0;
    AM_HIP(hipMemcpyAsync(over.data(), b->overflow->ptr, size_t(b->nseg) * 4, hipMemcpyDeviceToHost, s));
    AM_HIP(hipMemcpyAsync(&total, static_cast<unsigned long long*>(b->goff->ptr) + b->nseg, 8,
                          hipMemcpyDeviceToHost, s));
    AM_HIP(hipStreamSynchronize(s));
    b->groups = total;
    for (uint32_t x : over) b->fallback_batches += x ? 1u : 0u;
    if (b->fallback_batches) regroup_overflowed(*b, *group_keys, *values, over);
  } else {
    AM_HIP(hipStreamSynchronize(s));
  }
}

// Picks the sort tile (items per thread), the group-by hash capacity and the group-by strategy (LDS
// table per chunk, or one global table per batch) by timing each candidate on the given columns.
// The capacity starts from the smallest that holds every batch's distinct keys.
void batched_tune(Batched* b, const Array* sort_keys, const Array* group_keys, const Array* values, int desc,
                  int forder) {
  check_inputs(b, sort_keys, group_keys, values, forder);
  hipStream_t s = stream();
  hipEvent_t e0, e1;
  AM_HIP(hipEventCreate(&e0));
  AM_HIP(hipEventCreate(&e1));
  auto time_ms = [&](const Array* sk, const Array* gk, const Array* v) {
    float best = 1e30f;
    for (int rep = 0; rep < 3; rep++) {
      AM_HIP(hipEventRecord(e0, s));
      record(b, sk, gk, v, desc, forder, s);
      AM_HIP(hipEventRecord(e1, s));
      AM_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      AM_HIP(hipEventElapsedTime(&ms, e0, e1));
      best = std::min(best, ms);
    }
    return best;
  };
  try {
    if (sort_keys) {
      // 64-bit keys: the MSD radix sort (no tile size) against the LSD passes at each tile size;
      // narrower keys: the LSD tile sizes.
      int best_items = b->items;
      float best = 1e30f;
      bool best_msd = false;
      if (type_width(sort_keys->type) == 8 && !b->msd_overflowed) {
        b->use_msd = true;
        best = time_ms(sort_keys, nullptr, nullptr);
        best_msd = true;
      }
      b->use_msd = false;
      for (int it : {8, 16, 20}) {
        build_tiles(*b, it);
        float t = time_ms(sort_keys, nullptr, nullptr);
        if (t < best) { best = t; best_items = it; best_msd = false; }
      }
      build_tiles(*b, best_items);
      b->use_msd = best_msd;
      b->msd_used = best_msd;
      b->tuned_sort_ms = best;
    }
    if (group_keys) {
      // Smallest capacity without overflow (the flags read back after each trial).
      const uint32_t top = max_cap();
      uint32_t cap = 256;
      std::vector<uint32_t> over(b->nseg);
      while (true) {
        build_tables(*b, cap);
        b->strategy = 0;
        record(b, nullptr, group_keys, values, desc, forder, s);
        AM_HIP(hipMemcpyAsync(over.data(), b->overflow->ptr, size_t(b->nseg) * 4, hipMemcpyDeviceToHost, s));
        AM_HIP(hipStreamSynchronize(s));
        bool any = false;
        for (uint32_t x : over) any |= x != 0;
        if (!any || cap >= top) break;
        cap *= 2;
      }
      uint32_t best_cap = cap;
      int best_strategy = 0;
      float best = 1e30f;
      for (uint32_t c : {cap, cap * 2}) {
        if (c > top) continue;
        build_tables(*b, c);
        for (int st : {0, 1}) {
          b->strategy = st;
          float t = time_ms(nullptr, group_keys, values);
          if (t < best) { best = t; best_cap = c; best_strategy = st; }
        }
      }
      build_tables(*b, best_cap);
      b->strategy = best_strategy;
    }
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->drop_graph();   // a graph recorded before the choices above is re-recorded with them
  b->ran_sort = b->ran_group = false;
}

void batched_config(const Batched* b, int64_t* out, int count) {
  // The sort path is the one the last recorded sort (a run, or the tuner's pick) used.
  const int64_t v[7] = {b->items, int64_t(b->cap), b->strategy, 8 /* digit bits */, int64_t(b->fallback_batches),
                        int64_t(b->ntiles), (b->msd_used && b->use_msd) ? 1 : 0};
  for (int i = 0; i < count && i < 7; i++) out[i] = v[i];
}

// Results of the last run: 0 argsort indices (uint32, local rows, concatenated by batch), 1 group keys
// (the key column's type), 2 sums (int64), 3 counts (int64), 4 group offsets (int64, nbatches + 1).
Array* batched_result(Batched* b, int which) {
  if (which == 0 && !b->ran_sort) fail("batched: the last run had no sort keys");
  if (which >= 1 && which <= 4 && !b->ran_group) fail("batched: the last run had no group-by");
  hipStream_t s = stream();
  auto wrap = [&](const void* src, int64_t len, const char* fmt, int type) {
    Array* a = new_array(fmt, type, len);
    a->values = device_alloc(size_t(std::max<int64_t>(len, 1)) * std::max(type_width(type), 1));
    if (len > 0)
      AM_HIP(hipMemcpyAsync(a->values->ptr, src, size_t(len) * type_width(type), hipMemcpyDeviceToDevice, s));
    a->null_count = 0;
    return a;
  };
  Array* a = nullptr;
  const int64_t total = int64_t(b->groups);
  switch (which) {
    case 0: a = wrap(b->out_idx->ptr, b->n, "I", T_U32); break;
    case 1: {
      // Keys come back in the key column's own type.
      a = new_array(b->key_format, b->key_type, total);
      int w = type_width(b->key_type);
      a->values = device_alloc(size_t(std::max<int64_t>(total, 1)) * w);
      a->null_count = 0;
      if (total && w == 8)
        AM_HIP(hipMemcpyAsync(a->values->ptr, b->okey->ptr, size_t(total) * 8, hipMemcpyDeviceToDevice, s));
      else if (total)
        AM_HIP(hipMemcpy2DAsync(a->values->ptr, w, b->okey->ptr, 8, w, size_t(total), hipMemcpyDeviceToDevice, s));
      break;
    }
    case 2: a = wrap(b->osum->ptr, total, "l", T_I64); break;
    case 3: a = wrap(b->ocnt->ptr, total, "l", T_I64); break;
    case 4: a = wrap(b->goff->ptr, int64_t(b->nseg) + 1, "l", T_I64); break;
    default: fail("batched: unknown result " + std::to_string(which));
  }
  try {
    sync();
  } catch (...) {
    release_array(a);
    throw;
  }
  return a;
}

}  // namespace am
