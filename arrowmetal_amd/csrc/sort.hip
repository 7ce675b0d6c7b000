// Stable LSD radix argsort for CDNA4 (gfx950, wave64).
//
// Semantics follow the reference's radix sort (Sources/ArrowMetal/Kernels/Sort.swift:190-504 and
// SortSource.swift:131-330): an order-preserving unsigned key per value, a stable three-way split of
// the rows into [values][NaN][nulls] (or [nulls][NaN][values] for nulls first), 8-bit digits, and a
// pass skipped when its digit is the same in every value key. The way it runs on the GPU is this
// build's own:
//
//   sort_init      clears the per-call control block (a fill kernel, no hipMemset, so that the
//                  sequence can be captured in a graph)
//   sort_upsweep   one read of the column: maps every key, counts values / NaN / nulls, and builds
//                  the digit histograms of ALL passes at once plus the OR / AND of the keys
//   sort_plan      one workgroup: active-pass mask, block bases of the three buckets, and the
//                  exclusive digit prefix of every pass
//   sort_first     the first active pass reads the raw column again, maps keys on the fly, and
//                  scatters values by digit while NaN and null rows go straight to their final slots
//   sort_mid       one launch per further active digit; the last active pass writes only the row
//                  numbers, into the output. Inputs of 2^18 rows and more read the plan back once and
//                  launch only the active passes; smaller ones (and graph captures) launch one kernel
//                  per digit that exits at once when its digit is inactive
//
// Each scatter pass is "onesweep": tiles of 256 threads x ITEMS keys take dynamic tile ids, rank
// their keys inside the tile with wave64 ballots (a 9-bit match per 64-lane row), publish their
// per-bin counts and look back over earlier tiles for the exclusive prefix (decoupled look-back with
// epoch-tagged 64-bit status words, so the status table is never cleared between passes), then stage
// the tile in LDS in digit order so that the global writes are runs of consecutive addresses.
#include "am_internal.hpp"

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

struct SortCtl {
  uint32_t hist[8][RADIX];          // digit histograms of the value keys, one per pass
  uint32_t m, q, z, pad0;           // value rows, NaN rows (separate block), null rows
  unsigned long long key_or, key_and;
  uint32_t tile_counter[16];        // [0] first pass, [1 + p] pass p
  uint32_t digit_base[8][RADIX];    // exclusive digit prefix inside the value block, per pass
  uint32_t active_mask, first_active, last_active, pad1;
  uint32_t value_base, nan_base, null_base, pad2;
};

// ---------------------------------------------------------------- key maps

constexpr uint64_t SIGN64 = 0x8000000000000000ull;
constexpr uint64_t ABS64 = 0x7FFFFFFFFFFFFFFFull;
constexpr uint64_t INF64 = 0x7FF0000000000000ull;

// Category of row i (0 value, 1 NaN in its own block, 2 null) and, for a value, its ascending
// order-preserving key with the descending flip applied (SortSource.swift:131-190).
// float_order: 0 ieee (NaN apart, -0.0 ties +0.0), 1 total (raw sign flip), 2 nan_largest.
template <typename K>
__device__ __forceinline__ int classify(const void* vals, const uint8_t* valid, int64_t voff, uint32_t i,
                                        int type, int forder, int desc, K& key) {
  if (valid) {
    int64_t b = voff + int64_t(i);
    if (!((valid[b >> 3] >> (b & 7)) & 1)) return 2;
  }
  uint64_t k;
  switch (type) {
    case T_I64: k = uint64_t(static_cast<const int64_t*>(vals)[i]) ^ SIGN64; break;
    case T_U64: k = static_cast<const uint64_t*>(vals)[i]; break;
    case T_F64: {
      uint64_t x = static_cast<const uint64_t*>(vals)[i];
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
    case T_I32: k = uint32_t(static_cast<const int32_t*>(vals)[i]) ^ 0x80000000u; break;
    case T_I16: k = uint32_t(int32_t(static_cast<const int16_t*>(vals)[i])) ^ 0x80000000u; break;
    case T_I8: k = uint32_t(int32_t(static_cast<const int8_t*>(vals)[i])) ^ 0x80000000u; break;
    case T_U32: k = static_cast<const uint32_t*>(vals)[i]; break;
    case T_U16: k = static_cast<const uint16_t*>(vals)[i]; break;
    case T_U8: k = static_cast<const uint8_t*>(vals)[i]; break;
    default: {   // T_F32
      uint32_t x = static_cast<const uint32_t*>(vals)[i];
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

// Ranks a tile held in registers. Element (w, j, lane) of the tile is tile_start + w*64*ITEMS + j*64
// + lane; its stable rank among the tile's elements of the same bin is
//   (count in earlier waves) + (count in this wave at earlier items) + (lanes below it in its row).
// The last two come from here: a 9-bit match over the 64 lanes with ballots, and one LDS counter per
// (wave, bin) that only its own wave touches. On return wc[w][bin] holds the wave's bin totals.
template <int ITEMS>
__device__ __forceinline__ void rank_tile(const uint32_t (&bins)[ITEMS], uint32_t (&lrank)[ITEMS],
                                          uint32_t* wc_row, int lane) {
  const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    uint32_t b = bins[j];
    unsigned long long peers = ~0ull;
#pragma unroll
    for (int bit = 0; bit < 9; bit++) {
      unsigned long long bal = __ballot((b >> bit) & 1u);
      peers &= ((b >> bit) & 1u) ? bal : ~bal;
    }
    uint32_t below = uint32_t(__popcll(peers & lt));
    uint32_t base = wc_row[b];
    __builtin_amdgcn_wave_barrier();
    if (below == 0) wc_row[b] = base + uint32_t(__popcll(peers));
    __builtin_amdgcn_wave_barrier();
    lrank[j] = base + below;
  }
}

// ---------------------------------------------------------------- kernels

__global__ void sort_init(SortCtl* ctl) {
  uint32_t* w = reinterpret_cast<uint32_t*>(ctl);
  const int words = int(sizeof(SortCtl) / 4);
  for (int i = threadIdx.x; i < words; i += blockDim.x) w[i] = 0u;
  __syncthreads();
  if (threadIdx.x == 0) ctl->key_and = ~0ull;
}

// Copies the plan (active mask, first and last active pass, value rows) to pinned host memory.
__global__ void sort_read_plan(const SortCtl* ctl, uint32_t* out) {
  out[0] = ctl->active_mask;
  out[1] = ctl->first_active;
  out[2] = ctl->last_active;
  out[3] = ctl->m;
}

// Zeroes the look-back status table (only when it is first allocated or grows).
__global__ void fill_u64(unsigned long long* p, size_t n, unsigned long long v) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

template <typename K>
__global__ __launch_bounds__(BLOCK) void sort_upsweep(const void* vals, const uint8_t* valid, int64_t voff,
                                                      uint32_t n, int type, int forder, int desc,
                                                      uint32_t rows_per_block, SortCtl* ctl) {
  constexpr int P = int(sizeof(K));
  __shared__ uint32_t h[P][RADIX];
  __shared__ uint32_t cnt[3];
  __shared__ unsigned long long s_or, s_and;
  for (int i = threadIdx.x; i < P * RADIX; i += BLOCK) (&h[0][0])[i] = 0u;
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0u;
  if (threadIdx.x == 0) { s_or = 0ull; s_and = ~0ull; }
  __syncthreads();
  uint32_t begin = blockIdx.x * rows_per_block;
  uint32_t end = min(n, begin + rows_per_block);
  K o = 0, a = K(~K(0));
  uint32_t cm = 0, cq = 0, cz = 0;
  for (uint32_t i = begin + threadIdx.x; i < end; i += BLOCK) {
    K k;
    int c = classify<K>(vals, valid, voff, i, type, forder, desc, k);
    if (c == 0) {
      cm++;
      o |= k;
      a &= k;
#pragma unroll
      for (int p = 0; p < P; p++) atomicAdd(&h[p][uint32_t(k >> (8 * p)) & 0xFFu], 1u);
    } else if (c == 1) {
      cq++;
    } else {
      cz++;
    }
  }
  // Wave reductions of the counters and the OR / AND, then one LDS atomic per wave.
  unsigned long long o64 = o, a64 = a;
  if (sizeof(K) == 4) a64 |= 0xFFFFFFFF00000000ull;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    cm += __shfl_down(cm, off, 64);
    cq += __shfl_down(cq, off, 64);
    cz += __shfl_down(cz, off, 64);
    o64 |= __shfl_down(o64, off, 64);
    a64 &= __shfl_down(a64, off, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(&cnt[0], cm);
    atomicAdd(&cnt[1], cq);
    atomicAdd(&cnt[2], cz);
    atomicOr(&s_or, o64);
    atomicAnd(&s_and, a64);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P * RADIX; i += BLOCK) {
    uint32_t v = (&h[0][0])[i];
    if (v) atomicAdd(&(&ctl->hist[0][0])[i], v);
  }
  if (threadIdx.x == 0) {
    if (cnt[0]) {
      atomicAdd(&ctl->m, cnt[0]);
      atomicOr(&ctl->key_or, s_or);
      atomicAnd(&ctl->key_and, s_and);
    }
    if (cnt[1]) atomicAdd(&ctl->q, cnt[1]);
    if (cnt[2]) atomicAdd(&ctl->z, cnt[2]);
  }
}

template <int P>
__global__ __launch_bounds__(BLOCK) void sort_plan(SortCtl* ctl, int nulls_first) {
  __shared__ uint32_t tmp[WAVES];
  if (threadIdx.x == 0) {
    unsigned long long diff = ctl->key_or ^ ctl->key_and;
    uint32_t mask = 0;
    for (int p = 0; p < P; p++)
      if ((diff >> (8 * p)) & 0xFFull) mask |= 1u << p;
    // No varying digit (a single value, or one repeated key): one pass over digit 0 is then a
    // stable compaction of the value rows, which is exactly the answer.
    if (mask == 0 || ctl->m <= 1u) mask = 1;
    ctl->active_mask = mask;
    ctl->first_active = uint32_t(__builtin_ctz(mask));
    ctl->last_active = 31u - uint32_t(__builtin_clz(mask));
    uint32_t m = ctl->m, q = ctl->q, z = ctl->z;
    if (nulls_first) {
      ctl->null_base = 0;
      ctl->nan_base = z;
      ctl->value_base = z + q;
    } else {
      ctl->value_base = 0;
      ctl->nan_base = m;
      ctl->null_base = m + q;
    }
  }
  for (int p = 0; p < P; p++) {
    uint32_t total;
    uint32_t v = ctl->hist[p][threadIdx.x];
    ctl->digit_base[p][threadIdx.x] = block_excl_scan(v, tmp, total);
  }
}

// The shared body of a scatter pass over one tile that is already ranked and looked back.
// Stages the tile's value elements in LDS in bin order (keys first, then rows, through one buffer)
// and writes them out as runs: element at staged position `pos` with digit d goes to
// bin_base[d] + pos - bin_start[d].
template <typename K, int ITEMS, bool WRITE_KEYS>
__device__ __forceinline__ void scatter_values(const K (&keys)[ITEMS], const uint32_t (&rows)[ITEMS],
                                               const uint32_t (&bins)[ITEMS], const uint32_t (&lrank)[ITEMS],
                                               const uint32_t* wc_row, const uint32_t* bin_start,
                                               const uint32_t* bin_base, uint8_t* s_digit, K* s_stage,
                                               uint32_t vcount, K* out_keys, uint32_t* out_rows) {
  uint32_t pos[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    uint32_t b = bins[j];
    pos[j] = b < RADIX ? bin_start[b] + wc_row[b] + lrank[j] : 0xFFFFFFFFu;
    if (b < RADIX) s_digit[pos[j]] = uint8_t(b);
  }
  if (WRITE_KEYS) {
#pragma unroll
    for (int j = 0; j < ITEMS; j++)
      if (pos[j] != 0xFFFFFFFFu) s_stage[pos[j]] = keys[j];
    __syncthreads();
    for (uint32_t p = threadIdx.x; p < vcount; p += BLOCK) {
      uint32_t d = s_digit[p];
      out_keys[bin_base[d] + p - bin_start[d]] = s_stage[p];
    }
    __syncthreads();
  }
  uint32_t* s_rows = reinterpret_cast<uint32_t*>(s_stage);
#pragma unroll
  for (int j = 0; j < ITEMS; j++)
    if (pos[j] != 0xFFFFFFFFu) s_rows[pos[j]] = rows[j];
  __syncthreads();
  for (uint32_t p = threadIdx.x; p < vcount; p += BLOCK) {
    uint32_t d = s_digit[p];
    out_rows[bin_base[d] + p - bin_start[d]] = s_rows[p];
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
                                          const uint32_t* tile_total, uint32_t* excl) {
  const unsigned long long ep = (unsigned long long)epoch << 34;
  unsigned long long* mine = status + size_t(tile) * STATUS_STRIDE;
  for (int b = threadIdx.x; b < NB; b += BLOCK) {
    uint32_t flag = tile == 0 ? FLAG_PREFIX : FLAG_AGG;
    status_store(mine + b, ep | ((unsigned long long)flag << 32) | tile_total[b]);
  }
  if (tile == 0) {
    for (int b = threadIdx.x; b < NB; b += BLOCK) excl[b] = 0;
    return;
  }
  for (int b = threadIdx.x; b < NB; b += BLOCK) {
    uint32_t sum = 0;
    int64_t t = int64_t(tile) - 1;
    while (true) {
      unsigned long long s = status_load(status + size_t(t) * STATUS_STRIDE + b);
      if (uint32_t(s >> 34) != epoch) {
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      sum += uint32_t(s);
      if (((s >> 32) & 3u) == FLAG_PREFIX) break;
      t--;
    }
    excl[b] = sum;
    status_store(mine + b, ep | ((unsigned long long)FLAG_PREFIX << 32) | (sum + tile_total[b]));
  }
}

template <typename K, int ITEMS>
__global__ __launch_bounds__(BLOCK) void sort_first(const void* vals, const uint8_t* valid, int64_t voff,
                                                    uint32_t n, int type, int forder, int desc,
                                                    SortCtl* ctl, unsigned long long* status, uint32_t epoch,
                                                    K* out_keys, uint32_t* out_rows, uint32_t* out_idx) {
  constexpr int TILE = BLOCK * ITEMS;
  __shared__ uint32_t wc[WAVES][WC_STRIDE];
  __shared__ uint32_t tile_total[WC_STRIDE];
  __shared__ uint32_t bin_start[RADIX];
  __shared__ uint32_t bin_base[WC_STRIDE];
  __shared__ uint32_t tmp[WAVES];
  __shared__ uint32_t s_tile;
  __shared__ uint8_t s_digit[TILE];
  __shared__ K s_stage[TILE];

  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < WAVES * WC_STRIDE; i += BLOCK) (&wc[0][0])[i] = 0u;
  if (threadIdx.x == 0) s_tile = atomicAdd(&ctl->tile_counter[0], 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t first = ctl->first_active;
  const bool last = first == ctl->last_active;
  const int shift = int(8 * first);

  K keys[ITEMS];
  uint32_t rows[ITEMS], bins[ITEMS], lrank[ITEMS];
  const uint32_t base_e = tile * uint32_t(TILE) + uint32_t(w) * 64u * ITEMS + uint32_t(lane);
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    uint32_t e = base_e + uint32_t(j) * 64u;
    rows[j] = e;
    keys[j] = 0;
    uint32_t b = BIN_NONE;
    if (e < n) {
      K k;
      int c = classify<K>(vals, valid, voff, e, type, forder, desc, k);
      if (c == 0) {
        keys[j] = k;
        b = uint32_t(k >> shift) & 0xFFu;
      } else {
        b = c == 1 ? BIN_NAN : BIN_NULL;
      }
    }
    bins[j] = b;
  }
  rank_tile<ITEMS>(bins, lrank, wc[w], lane);
  __syncthreads();
  uint32_t vcount = tile_bin_layout<NBINS_FIRST>(wc, tile_total, bin_start, tmp);
  look_back<NBINS_FIRST>(status, tile, epoch, tile_total, bin_base);
  __syncthreads();
  // Turn the exclusive prefixes into absolute destinations.
  for (int b = threadIdx.x; b < NBINS_FIRST; b += BLOCK) {
    uint32_t base;
    if (b < RADIX) base = ctl->digit_base[first][b] + (last ? ctl->value_base : 0u);
    else base = b == BIN_NAN ? ctl->nan_base : ctl->null_base;
    bin_base[b] += base;
  }
  __syncthreads();
  // NaN and null rows go straight to their final slots, in input order.
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    uint32_t b = bins[j];
    if (b == BIN_NAN || b == BIN_NULL) out_idx[bin_base[b] + wc[w][b] + lrank[j]] = rows[j];
  }
  if (last) {
    scatter_values<K, ITEMS, false>(keys, rows, bins, lrank, wc[w], bin_start, bin_base, s_digit, s_stage,
                                    vcount, nullptr, out_idx);
  } else {
    scatter_values<K, ITEMS, true>(keys, rows, bins, lrank, wc[w], bin_start, bin_base, s_digit, s_stage,
                                   vcount, out_keys, out_rows);
  }
}

// MODE: PASS_MID and PASS_LAST are launched by a host that read the plan back (only active passes,
// the last one known); PASS_DEVICE is launched for every digit and decides everything on the GPU.
constexpr int PASS_MID = 0, PASS_LAST = 1, PASS_DEVICE = 2;

template <typename K, int ITEMS, int MODE>
__global__ __launch_bounds__(BLOCK) void sort_mid(int pass, SortCtl* ctl, unsigned long long* status,
                                                  uint32_t epoch, K* keys0, uint32_t* rows0, K* keys1,
                                                  uint32_t* rows1, uint32_t* out_idx) {
  constexpr int TILE = BLOCK * ITEMS;
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
  if (!((mask >> pass) & 1u) || uint32_t(pass) == first) return;   // uniform over the grid
  const uint32_t m = ctl->m;
  const uint32_t ntiles = (m + TILE - 1) / TILE;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < WAVES * WC_STRIDE; i += BLOCK) (&wc[0][0])[i] = 0u;
  if (threadIdx.x == 0) s_tile = atomicAdd(&ctl->tile_counter[1 + pass], 1u);
  __syncthreads();
  const uint32_t tile = s_tile;
  if (tile >= ntiles) return;   // surplus block: the grid is sized for every row, the pass for values
  const uint32_t rank = uint32_t(__builtin_popcount(mask & ((1u << pass) - 1u)));
  const bool last = MODE == PASS_DEVICE ? uint32_t(pass) == ctl->last_active : MODE == PASS_LAST;
  const K* in_keys = (rank & 1u) ? keys0 : keys1;
  const uint32_t* in_rows = (rank & 1u) ? rows0 : rows1;
  K* out_keys = (rank & 1u) ? keys1 : keys0;
  uint32_t* out_rows = (rank & 1u) ? rows1 : rows0;
  const int shift = 8 * pass;

  K keys[ITEMS];
  uint32_t rows[ITEMS], bins[ITEMS], lrank[ITEMS];
  const uint32_t base_e = tile * uint32_t(TILE) + uint32_t(w) * 64u * ITEMS + uint32_t(lane);
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    uint32_t e = base_e + uint32_t(j) * 64u;
    if (e < m) {
      keys[j] = in_keys[e];
      rows[j] = in_rows[e];
      bins[j] = uint32_t(keys[j] >> shift) & 0xFFu;
    } else {
      keys[j] = 0;
      rows[j] = 0;
      bins[j] = BIN_NONE;
    }
  }
  rank_tile<ITEMS>(bins, lrank, wc[w], lane);
  __syncthreads();
  uint32_t vcount = tile_bin_layout<RADIX>(wc, tile_total, bin_start, tmp);
  look_back<RADIX>(status, tile, epoch, tile_total, bin_base);
  __syncthreads();
  const uint32_t vb = last ? ctl->value_base : 0u;
  bin_base[threadIdx.x] += ctl->digit_base[pass][threadIdx.x] + vb;
  __syncthreads();
  if (last) {
    scatter_values<K, ITEMS, false>(keys, rows, bins, lrank, wc[w], bin_start, bin_base, s_digit, s_stage,
                                    vcount, nullptr, out_idx);
  } else {
    scatter_values<K, ITEMS, true>(keys, rows, bins, lrank, wc[w], bin_start, bin_base, s_digit, s_stage,
                                   vcount, out_keys, out_rows);
  }
}

// ---------------------------------------------------------------- host side

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

template <typename K, int ITEMS>
void run_sort(const void* vals, int type, const uint8_t* valid, int64_t voff, uint32_t n, bool desc,
              int placement, int forder, uint32_t* out_idx, hipStream_t s) {
  constexpr int P = int(sizeof(K));
  constexpr int TILE = BLOCK * ITEMS;
  const uint32_t tiles = (n + TILE - 1) / TILE;
  DevBuf ctl_buf = device_alloc(sizeof(SortCtl));
  SortCtl* ctl = static_cast<SortCtl*>(ctl_buf->ptr);
  uint32_t epoch = reserve_status(size_t(tiles) * STATUS_STRIDE, s);
  unsigned long long* status = static_cast<unsigned long long*>(t_ws.status->ptr);
  retain(t_ws.status);
  DevBuf k0 = device_alloc(size_t(n) * sizeof(K)), k1 = device_alloc(size_t(n) * sizeof(K));
  DevBuf r0 = device_alloc(size_t(n) * 4), r1 = device_alloc(size_t(n) * 4);
  const int dsc = desc ? 1 : 0;
  {
    ProfScope ps("sort_init");
    hipLaunchKernelGGL(sort_init, dim3(1), dim3(BLOCK), 0, s, ctl);
    AM_HIP(hipGetLastError());
  }
  {
    uint32_t max_blocks = uint32_t(device().cu_count) * 2u;
    uint32_t blocks = std::max(1u, std::min(max_blocks, (n + 4095u) / 4096u));
    uint32_t rpb = ((n + blocks - 1) / blocks + 255u) / 256u * 256u;
    blocks = (n + rpb - 1) / rpb;
    ProfScope ps("sort_upsweep");
    hipLaunchKernelGGL(sort_upsweep<K>, dim3(blocks), dim3(BLOCK), 0, s, vals, valid, voff, n, type, forder,
                       dsc, rpb, ctl);
    AM_HIP(hipGetLastError());
  }
  {
    ProfScope ps("sort_plan");
    hipLaunchKernelGGL(sort_plan<P>, dim3(1), dim3(BLOCK), 0, s, ctl, placement == 1 ? 1 : 0);
    AM_HIP(hipGetLastError());
  }
  // Large inputs read the plan back once (one sync, as the reference's analysis readback does), so
  // that only active passes are launched, over the value rows only.
  struct Plan { uint32_t mask, first, last, m; } plan{0xFFu, 0, uint32_t(P - 1), n};
  const bool analyse = n >= (1u << 18) && !capturing(s);
  if (analyse) {
    Plan* hp = reinterpret_cast<Plan*>(host_plan_words());
    hipLaunchKernelGGL(sort_read_plan, dim3(1), dim3(1), 0, s, ctl, reinterpret_cast<uint32_t*>(hp));
    AM_HIP(hipGetLastError());
    AM_HIP(hipStreamSynchronize(s));
    plan = *hp;
  }
  {
    ProfScope ps("sort_first");
    hipLaunchKernelGGL((sort_first<K, ITEMS>), dim3(tiles), dim3(BLOCK), 0, s, vals, valid, voff, n, type, forder,
                       dsc, ctl, status, epoch, static_cast<K*>(k0->ptr), static_cast<uint32_t*>(r0->ptr),
                       out_idx);
    AM_HIP(hipGetLastError());
  }
  K* kp0 = static_cast<K*>(k0->ptr);
  K* kp1 = static_cast<K*>(k1->ptr);
  uint32_t* rp0 = static_cast<uint32_t*>(r0->ptr);
  uint32_t* rp1 = static_cast<uint32_t*>(r1->ptr);
  if (analyse) {
    // The host read the plan back: launch only the active passes, each over the value rows, with
    // the last one's kernel known (no key writes).
    const uint32_t mtiles = std::max<uint32_t>(1, (plan.m + TILE - 1) / TILE);
    for (int p = 0; p < P; p++) {
      if (!((plan.mask >> p) & 1u) || uint32_t(p) == plan.first) continue;
      if (uint32_t(p) == plan.last) {
        ProfScope ps("sort_scatter_last");
        hipLaunchKernelGGL((sort_mid<K, ITEMS, PASS_LAST>), dim3(mtiles), dim3(BLOCK), 0, s, p, ctl, status,
                           epoch + 1u + uint32_t(p), kp0, rp0, kp1, rp1, out_idx);
      } else {
        ProfScope ps("sort_scatter");
        hipLaunchKernelGGL((sort_mid<K, ITEMS, PASS_MID>), dim3(mtiles), dim3(BLOCK), 0, s, p, ctl, status,
                           epoch + 1u + uint32_t(p), kp0, rp0, kp1, rp1, out_idx);
      }
      AM_HIP(hipGetLastError());
    }
  } else {
    for (int p = 0; p < P; p++) {
      ProfScope ps("sort_pass_dev");
      hipLaunchKernelGGL((sort_mid<K, ITEMS, PASS_DEVICE>), dim3(tiles), dim3(BLOCK), 0, s, p, ctl, status,
                         epoch + 1u + uint32_t(p), kp0, rp0, kp1, rp1, out_idx);
      AM_HIP(hipGetLastError());
    }
  }
}

template <typename K>
void dispatch_items(int items, const void* vals, int type, const uint8_t* valid, int64_t voff, uint32_t n,
                    bool desc, int placement, int forder, uint32_t* out_idx, hipStream_t s) {
  switch (items) {
    case 4: run_sort<K, 4>(vals, type, valid, voff, n, desc, placement, forder, out_idx, s); break;
    case 8: run_sort<K, 8>(vals, type, valid, voff, n, desc, placement, forder, out_idx, s); break;
    case 12: run_sort<K, 12>(vals, type, valid, voff, n, desc, placement, forder, out_idx, s); break;
    default: run_sort<K, 16>(vals, type, valid, voff, n, desc, placement, forder, out_idx, s); break;
  }
}

}  // namespace

void argsort_device(const void* vals, int type, const uint8_t* validity, int64_t vbit_off, int64_t n,
                    bool descending, int null_placement, int float_order, uint32_t* out_idx, hipStream_t s) {
  if (n <= 0) return;
  if (n > int64_t(UINT32_MAX)) fail("arrays above 2^32 elements are not supported yet");
  int w = type_width(type);
  if (w == 0) fail("argsort: unsupported element type");
  int kb = w == 8 ? 8 : 4;
  SortTune t = sort_tune(n, kb);
  if (kb == 8)
    dispatch_items<unsigned long long>(t.items, vals, type, validity, vbit_off, uint32_t(n), descending,
                                       null_placement, float_order, out_idx, s);
  else
    dispatch_items<uint32_t>(t.items, vals, type, validity, vbit_off, uint32_t(n), descending, null_placement,
                             float_order, out_idx, s);
}

Array* argsort(const Array& a, bool descending, int null_placement, int float_order) {
  check_length(a.length);
  check_index_rows(a.length, "argsort");
  if (a.type == T_UNSUPPORTED || a.type == T_BOOL) fail("argsort: unsupported type " + a.format);
  Array* out = new_array("I", T_U32, a.length);
  try {
    out->values = device_alloc(size_t(std::max<int64_t>(a.length, 1)) * 4);
    out->null_count = 0;
    if (a.length > 0) {
      argsort_device(a.data(), a.type, a.has_nulls() ? a.valid_bits() : nullptr, a.offset, a.length, descending,
                     null_placement, float_order, static_cast<uint32_t*>(out->values->ptr), stream());
    }
    sync();
  } catch (...) {
    release_array(out);
    throw;
  }
  return out;
}

}  // namespace am
