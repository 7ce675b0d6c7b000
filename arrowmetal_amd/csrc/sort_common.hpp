// Device building blocks shared by the plain radix argsort (sort.hip) and the batched, segmented
// one (batched.hip): key maps, block scans, wave64 tile ranking, LDS-staged scatter, decoupled
// look-back with epoch-tagged status words, and the per-thread look-back status table.
#pragma once

#include "am_internal.hpp"

// Predecessor status words a look-back step reads at once (a build-time tuning constant).
#ifndef AM_LB_WINDOW
#define AM_LB_WINDOW 4
#endif

#include <algorithm>

namespace am {
namespace {

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;
constexpr int RADIX = 256;
constexpr int BIN_NAN = 256;
constexpr int BIN_NULL = 257;
constexpr int BIN_NONE = 258;     // lanes past the end of the tile
constexpr int NBINS_FIRST = 258;
constexpr int WC_STRIDE = 260;    // per-wave counter row in LDS
constexpr int STATUS_STRIDE = 260;
constexpr uint32_t FLAG_AGG = 1, FLAG_PREFIX = 2;


// ---------------------------------------------------------------- key maps

constexpr uint64_t SIGN64 = 0x8000000000000000ull;
constexpr uint64_t ABS64 = 0x7FFFFFFFFFFFFFFFull;
constexpr uint64_t INF64 = 0x7FF0000000000000ull;

// Rows are read in two steps so that a thread's loads are all issued before any is used: the raw
// bits of a W-byte element and its validity byte are loaded unconditionally (the byte comes from a
// one-byte all-valid constant when the column has no bitmap), and the type-dependent mapping below
// runs on registers only. A type switch around the loads themselves would make the compiler wait on
// every load at the branch join.
__device__ const uint8_t kAllValid = 0xFFu;

template <int W>
__device__ __forceinline__ uint64_t load_raw(const void* vals, uint32_t i) {
  if constexpr (W == 8) return load_global(static_cast<const uint64_t*>(vals) + i);
  else if constexpr (W == 4) return load_global(static_cast<const uint32_t*>(vals) + i);
  else if constexpr (W == 2) return load_global(static_cast<const uint16_t*>(vals) + i);
  else return load_global(static_cast<const uint8_t*>(vals) + i);
}

// The validity byte of row i and the bit of row i in it.
__device__ __forceinline__ uint32_t load_valid_byte(const uint8_t* valid, int64_t voff, uint32_t i, uint32_t& bit) {
  const int64_t b = voff + int64_t(i);
  const uint8_t* p = valid ? valid + (b >> 3) : &kAllValid;
  bit = valid ? uint32_t(b & 7) : 0u;
  return load_global(p);
}

// Category of a row (0 value, 1 NaN in its own block, 2 null) from its raw bits and validity and,
// for a value, its ascending order-preserving key with the descending flip applied
// (SortSource.swift:131-190). float_order: 0 ieee (NaN apart, -0.0 ties +0.0), 1 total (raw sign
// flip), 2 nan_largest. Registers only.
template <typename K>
__device__ __forceinline__ int classify_raw(uint64_t raw, bool is_valid, int type, int forder, int desc, K& key) {
  if (!is_valid) return 2;
  uint64_t k;
  switch (type) {
    case T_I64: k = raw ^ SIGN64; break;
    case T_U64: k = raw; break;
    case T_F64: {
      uint64_t x = raw;
      if (forder != 1) {
        if ((x & ABS64) > INF64) {
          if (forder == 0) return 1;
          x = 0x7FF0000000000001ull;
        } else if ((x & ABS64) == 0) {
          x = 0;
        }
      }
      k = (x & SIGN64) ? ~x : (x | SIGN64);
      break;
    }
    case T_I32: k = uint32_t(raw) ^ 0x80000000u; break;
    case T_I16: k = uint32_t(int32_t(int16_t(uint16_t(raw)))) ^ 0x80000000u; break;
    case T_I8: k = uint32_t(int32_t(int8_t(uint8_t(raw)))) ^ 0x80000000u; break;
    case T_U32: case T_U16: case T_U8: k = raw; break;
    default: {   // T_F32
      uint32_t x = uint32_t(raw);
      if (forder != 1) {
        if ((x & 0x7FFFFFFFu) > 0x7F800000u) {
          if (forder == 0) return 1;
          x = 0x7F800001u;
        } else if ((x & 0x7FFFFFFFu) == 0u) {
          x = 0u;
        }
      }
      k = (x & 0x80000000u) ? uint32_t(~x) : (x | 0x80000000u);
      break;
    }
  }
  if (desc) k = ~k;
  key = K(k);
  return 0;
}

// Both steps for one row (rows whose loads need not be batched).
template <typename K, int W>
__device__ __forceinline__ int classify(const void* vals, const uint8_t* valid, int64_t voff, uint32_t i,
                                        int type, int forder, int desc, K& key) {
  uint32_t bit;
  const uint32_t vb = load_valid_byte(valid, voff, i, bit);
  return classify_raw<K>(load_raw<W>(vals, i), (vb >> bit) & 1u, type, forder, desc, key);
}

// ---------------------------------------------------------------- block helpers

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x, int lane) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  return x;
}

// Exclusive scan of one value per thread over the 256-thread block. `tmp` holds WAVES words of LDS.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* tmp, uint32_t& total) {
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t x = wave_incl_scan(v, lane);
  if (lane == 63) tmp[w] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < WAVES; i++) {
    uint32_t t = tmp[i];
    pre += (i < w) ? t : 0u;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return pre + x - v;
}

__device__ __forceinline__ unsigned long long status_load(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void status_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// A ranked element packs its bin (9 bits) and its rank inside its wave (the bits above).
constexpr uint32_t BIN_BITS = 9;
__device__ __forceinline__ uint32_t bin_of(uint32_t packed) { return packed & ((1u << BIN_BITS) - 1u); }
__device__ __forceinline__ uint32_t rank_of(uint32_t packed) { return packed >> BIN_BITS; }

// Ranks a tile held in registers. Element (w, j, lane) of the tile is tile_start + w*64*ITEMS + j*64
// + lane; its stable rank among the tile's elements of the same bin is
//   (count in earlier waves) + (count in this wave at earlier items) + (lanes below it in its row).
// The last two come from here: a 9-bit match over the 64 lanes with ballots, and one LDS counter per
// (wave, bin) that only its own wave touches. On return wc[w][bin] holds the wave's bin totals and
// each bins[j] is packed with the element's rank inside its wave (bin_of / rank_of).
template <int ITEMS, int NBITS>
__device__ __forceinline__ void rank_tile(uint32_t (&bins)[ITEMS], uint32_t* wc_row, int lane) {
  const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t b = bins[j];
    // peers = the lanes whose bin equals this lane's, bit by bit: a ballot of each bit, kept where
    // this lane's bit agrees (XNOR of the scalar ballot with the lane's 0 / -1 bit mask).
    uint32_t plo = ~0u, phi = ~0u;
#pragma unroll
    for (int bit = 0; bit < NBITS; bit++) {
      const uint32_t m = uint32_t(int32_t(b << (31 - bit)) >> 31);   // 0 or all ones
      const unsigned long long bal = __builtin_amdgcn_ballot_w64(m != 0u);
      plo &= ~(uint32_t(bal) ^ m);
      phi &= ~(uint32_t(bal >> 32) ^ m);
    }
    const unsigned long long peers = (unsigned long long)phi << 32 | plo;
    const uint32_t below = uint32_t(__popcll(peers & lt));
    const uint32_t base = wc_row[b];
    __builtin_amdgcn_wave_barrier();
    if (below == 0) wc_row[b] = base + uint32_t(__popcll(peers));
    __builtin_amdgcn_wave_barrier();
    bins[j] = b | ((base + below) << BIN_BITS);
  }
}

// The shared body of a scatter pass over one tile that is already ranked and looked back.
// Stages the tile's value elements in LDS in bin order (keys first, then rows, through one buffer)
// and writes them out as runs: element at staged position `pos` with digit d goes to
// bin_base[d] + pos - bin_start[d].
// ROWS_IOTA: the row of element j is row_base + 64 * j (a first pass over the raw column), so no
// row registers are kept; otherwise rows[j].
// POS_ROWS (with ROWS_IOTA, the ORDER BY of whole tables): the row written beside each key is its own
// destination, and pos_of_row[row] receives that destination, so that later passes carry positions
// of this pass's output and the columns can be partitioned the same way from pos_of_row.
template <typename K, int ITEMS, bool WRITE_KEYS, bool ROWS_IOTA = false, bool POS_ROWS = false>
__device__ __forceinline__ void scatter_values(const K (&keys)[ITEMS], const uint32_t* rows,
                                               const uint32_t (&packed)[ITEMS], const uint32_t* wc_row,
                                               const uint32_t* bin_start, uint32_t* bin_base, uint8_t* s_digit,
                                               K* s_stage, uint32_t vcount, K* out_keys, uint32_t* out_rows,
                                               uint32_t row_base = 0, uint32_t* pos_of_row = nullptr) {
  static_assert(!POS_ROWS || (WRITE_KEYS && ROWS_IOTA), "positions as rows need the first pass's keys and rows");
  auto row = [&](int j) { return ROWS_IOTA ? row_base + uint32_t(j) * 64u : rows[j]; };
  uint32_t pos[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    uint32_t b = bin_of(packed[j]);
    pos[j] = b < RADIX ? bin_start[b] + wc_row[b] + rank_of(packed[j]) : 0xFFFFFFFFu;
    if (b < RADIX) s_digit[pos[j]] = uint8_t(b);
  }
  // Each staged position's destination is delta[digit] + position; computed once, used by both
  // phases (bin_base is turned into the per-digit delta in place, after every reader of it above).
  bin_base[threadIdx.x] -= bin_start[threadIdx.x];
  if (WRITE_KEYS) {
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
      if (pos[j] != 0xFFFFFFFFu) s_stage[pos[j]] = keys[j];
  } else {
    uint32_t* s_rows = reinterpret_cast<uint32_t*>(s_stage);
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
      if (pos[j] != 0xFFFFFFFFu) s_rows[pos[j]] = row(j);
  }
  __syncthreads();
  uint32_t dest[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    uint32_t p = threadIdx.x + uint32_t(k) * BLOCK;
    dest[k] = p < vcount ? bin_base[s_digit[p]] + p : 0u;
  }
  if (WRITE_KEYS) {
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      uint32_t p = threadIdx.x + uint32_t(k) * BLOCK;
      if (p < vcount) out_keys[dest[k]] = s_stage[p];
    }
    if constexpr (POS_ROWS) {
      // Every delta is in place (the barrier after the key staging); rows in input order, so these
      // writes are coalesced.
#pragma unroll
      for (int j = 0; j < ITEMS; j++)
        if (pos[j] != 0xFFFFFFFFu) pos_of_row[row(j)] = bin_base[bin_of(packed[j])] + pos[j];
#pragma unroll
      for (int k = 0; k < ITEMS; k++) {
        uint32_t p = threadIdx.x + uint32_t(k) * BLOCK;
        if (p < vcount) out_rows[dest[k]] = dest[k];
      }
      return;
    }
    __syncthreads();
    uint32_t* s_rows = reinterpret_cast<uint32_t*>(s_stage);
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
      if (pos[j] != 0xFFFFFFFFu) s_rows[pos[j]] = row(j);
    __syncthreads();
  }
  const uint32_t* s_rows = reinterpret_cast<const uint32_t*>(s_stage);
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    uint32_t p = threadIdx.x + uint32_t(k) * BLOCK;
    if (p < vcount) out_rows[dest[k]] = s_rows[p];
  }
}

// Per-bin wave prefixes (in place in wc), tile totals, and the bin starts of the staged tile.
template <int NB>
__device__ __forceinline__ uint32_t tile_bin_layout(uint32_t (*wc)[WC_STRIDE], uint32_t* tile_total,
                                                   uint32_t* bin_start, uint32_t* tmp) {
  for (int b = threadIdx.x; b < NB; b += BLOCK) {
    uint32_t run = 0;
#pragma unroll
    for (int w = 0; w < WAVES; w++) {
      uint32_t c = wc[w][b];
      wc[w][b] = run;
      run += c;
    }
    tile_total[b] = run;
  }
  __syncthreads();
  uint32_t vcount;
  bin_start[threadIdx.x] = block_excl_scan(tile_total[threadIdx.x], tmp, vcount);
  return vcount;
}

// Decoupled look-back for bins [0, NB): publish this tile's counts, then accumulate earlier tiles'
// until one with an inclusive prefix is found. Writes the exclusive prefix into `excl`.
template <int NB>
__device__ __forceinline__ void look_back(unsigned long long* status, uint32_t tile, uint32_t epoch,
                                          const uint32_t* tile_total, uint32_t* excl, uint32_t first_tile = 0) {
  // Tiles [first_tile, ...) form one chain (a whole column, or one segment of a batched sort): the
  // first tile of the chain publishes its inclusive prefix at once.
  const unsigned long long ep = (unsigned long long)epoch << 34;
  unsigned long long* mine = status + size_t(tile) * STATUS_STRIDE;
  const bool head = tile == first_tile;
  for (int b = threadIdx.x; b < NB; b += BLOCK) {
    uint32_t flag = head ? FLAG_PREFIX : FLAG_AGG;
    status_store(mine + b, ep | ((unsigned long long)flag << 32) | tile_total[b]);
  }
  if (head) {
    for (int b = threadIdx.x; b < NB; b += BLOCK) excl[b] = 0;
    return;
  }
  // The walk reads LB_WINDOW predecessors per step (independent loads, one memory latency), then
  // consumes them in order: aggregates are added, the first inclusive prefix ends the walk, and a
  // predecessor that has published nothing yet is polled again from where the walk stopped. The
  // chain's head publishes its prefix at once, so no walk goes below it. A thread that owns two bins
  // (bins BLOCK.. of the first pass: NaN and null) walks both chains in the same steps.
  constexpr int LB_WINDOW = AM_LB_WINDOW;
  constexpr int PER = (NB + BLOCK - 1) / BLOCK;
  static_assert(PER <= 2, "at most two bins per thread");
  const int64_t lo = int64_t(first_tile);
  int bin[PER];
  uint32_t sum[PER];
  int64_t t[PER];
  bool done[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    bin[q] = int(threadIdx.x) + q * BLOCK;
    sum[q] = 0;
    t[q] = int64_t(tile) - 1;
    done[q] = bin[q] >= NB;
  }
  while (true) {
    bool all = true;
#pragma unroll
    for (int q = 0; q < PER; q++) all = all && done[q];
    if (all) break;
    unsigned long long sv[PER][LB_WINDOW];
#pragma unroll
    for (int q = 0; q < PER; q++)
#pragma unroll
      for (int k = 0; k < LB_WINDOW; k++)
        sv[q][k] = (!done[q] && t[q] - k >= lo) ? status_load(status + size_t(t[q] - k) * STATUS_STRIDE + bin[q]) : 0ull;
    bool progressed = false;
#pragma unroll
    for (int q = 0; q < PER; q++) {
      if (done[q]) continue;
      int used = 0;
#pragma unroll
      for (int k = 0; k < LB_WINDOW; k++) {
        if (done[q] || used != k || t[q] - k < lo || uint32_t(sv[q][k] >> 34) != epoch) continue;
        sum[q] += uint32_t(sv[q][k]);
        used = k + 1;
        if (((sv[q][k] >> 32) & 3u) == FLAG_PREFIX) done[q] = true;
      }
      t[q] -= used;
      progressed = progressed || used > 0;
    }
    if (!progressed) __builtin_amdgcn_s_sleep(1);
  }
#pragma unroll
  for (int q = 0; q < PER; q++) {
    if (bin[q] >= NB) continue;
    excl[bin[q]] = sum[q];
    status_store(mine + bin[q], ep | ((unsigned long long)FLAG_PREFIX << 32) | (sum[q] + tile_total[bin[q]]));
  }
}

// Zeroes the look-back status table (only when it is first allocated or grows).
__global__ void fill_u64(unsigned long long* p, size_t n, unsigned long long v) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// Per-thread look-back status table: epoch-tagged, so it is cleared only when it grows.
struct SortWorkspace {
  DevBuf status;
  size_t entries = 0;
  uint32_t epoch = 1;
};
thread_local SortWorkspace t_ws;

uint32_t reserve_status(size_t entries, hipStream_t s) {
  if (t_ws.entries < entries || t_ws.epoch > (1u << 29)) {
    size_t want = std::max(entries, t_ws.entries);
    // Allocated outside the pool's batch retention: the table lives as long as the thread.
    t_ws.status.reset();
    void* p = nullptr;
    AM_HIP(hipMalloc(&p, want * 8));
    t_ws.status = device_wrap(p, want * 8, [](void* q) { (void)hipFree(q); });
    t_ws.entries = want;
    int blocks = int(std::min<size_t>((want + 255) / 256, 4096));
    hipLaunchKernelGGL(fill_u64, dim3(blocks), dim3(256), 0, s, static_cast<unsigned long long*>(p), want, 0ull);
    AM_HIP(hipGetLastError());
    t_ws.epoch = 1;
  }
  uint32_t e = t_ws.epoch;
  t_ws.epoch += 16;
  return e;
}

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) return false;
  return st != hipStreamCaptureStatusNone;
}

// Pinned per-thread landing place of the plan readback.
uint32_t* host_plan_words() {
  thread_local uint32_t* p = nullptr;
  if (!p) AM_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), 64, hipHostMallocDefault));
  return p;
}

}  // namespace
}  // namespace am
