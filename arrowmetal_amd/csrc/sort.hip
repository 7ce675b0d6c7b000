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
//                  numbers, into the output. Inputs of 2^16 rows and more read the plan back once and
//                  launch only the active passes; smaller ones (and graph captures) launch one kernel
//                  per digit that exits at once when its digit is inactive
//
// Each scatter pass is "onesweep": tiles of 256 threads x ITEMS keys take dynamic tile ids, rank
// their keys inside the tile with wave64 ballots (a 9-bit match per 64-lane row), publish their
// per-bin counts and look back over earlier tiles for the exclusive prefix (decoupled look-back with
// epoch-tagged 64-bit status words, so the status table is never cleared between passes), then stage
// the tile in LDS in digit order so that the global writes are runs of consecutive addresses.
#include "msd_levels.hpp"

namespace am {
namespace {

struct SortCtl {
  uint32_t hist[8][RADIX];          // digit histograms of the value keys, one per pass
  uint32_t m, q, z, pad0;           // value rows, NaN rows (separate block), null rows
  unsigned long long key_or, key_and;
  // Bumped by every workgroup of a pass, so on cache lines of their own: a line that also holds
  // words the same workgroups read slows the pass down.
  alignas(128) uint32_t tile_counter[16];   // [0] first pass, [1 + p] pass p, [11] units
  alignas(128) uint32_t digit_base[8][RADIX];   // exclusive digit prefix inside the value block, per pass
  uint32_t active_mask, first_active, last_active, pad1;
  uint32_t value_base, nan_base, null_base, pad2;
  // Per-XCD partials of the upsweep (updated by workgroup-scope atomics in each XCD's own L2), folded
  // into the fields above by sort_plan.
  uint32_t xhist[NXCD][8][RADIX];
  uint32_t xm[NXCD], xq[NXCD], xz[NXCD], pad3[NXCD];
  unsigned long long xor_[NXCD], xand[NXCD];
};

// ---------------------------------------------------------------- kernels

__global__ void sort_init(SortCtl* ctl) {
  uint32_t* w = reinterpret_cast<uint32_t*>(ctl);
  const int words = int(sizeof(SortCtl) / 4);
  for (int i = threadIdx.x; i < words; i += blockDim.x) w[i] = 0u;
  __syncthreads();
  if (threadIdx.x == 0) ctl->key_and = ~0ull;
  if (threadIdx.x < NXCD) ctl->xand[threadIdx.x] = ~0ull;
}

// Copies the plan (active mask, first and last active pass, value rows) to pinned host memory.
__global__ void sort_read_plan(const SortCtl* ctl, uint32_t* out) {
  out[0] = ctl->active_mask;
  out[1] = ctl->first_active;
  out[2] = ctl->last_active;
  out[3] = ctl->m;
}


// Rows per unit of the top-digit partition (a power of two): the upsweep counts the top digit, NaN
// and null rows per unit when given unit_hist ([NBINS_FIRST][nunits], bin-major), so that the
// partition knows every unit's destinations before it starts.
// Units are 2^unit_shift rows: 16K for 4M rows and more, 4K (one tile of the partition) below, so
// that a 1M-row input still has ~250 units to spread over the CUs.
constexpr uint32_t UNIT_SHIFT_LARGE = 14, UNIT_SHIFT_SMALL = 12;
constexpr uint32_t UNITS_PER_BLOCK = 8;   // most units one upsweep workgroup counts

template <typename K, int W>
__global__ __launch_bounds__(BLOCK) void sort_upsweep(const void* vals, const uint8_t* valid, int64_t voff,
                                                      uint32_t n, int type, int forder, int desc,
                                                      uint32_t rows_per_block, SortCtl* ctl, uint32_t* unit_hist,
                                                      uint32_t nunits, uint32_t unit_shift, int all_digits) {
  constexpr int P = int(sizeof(K));
  __shared__ uint32_t h[P][RADIX];
  __shared__ uint32_t hu[UNITS_PER_BLOCK][NBINS_FIRST];
  __shared__ uint32_t cnt[3];
  __shared__ unsigned long long s_or, s_and;
  const bool units = unit_hist != nullptr;   // uniform
  for (int i = threadIdx.x; i < P * RADIX; i += BLOCK) (&h[0][0])[i] = 0u;
  if (units)
    for (int i = threadIdx.x; i < int(UNITS_PER_BLOCK * NBINS_FIRST); i += BLOCK) (&hu[0][0])[i] = 0u;
  if (threadIdx.x < 3) cnt[threadIdx.x] = 0u;
  if (threadIdx.x == 0) { s_or = 0ull; s_and = ~0ull; }
  __syncthreads();
  uint32_t begin = blockIdx.x * rows_per_block;
  uint32_t end = min(n, begin + rows_per_block);
  K o = 0, a = K(~K(0));
  uint32_t cm = 0, cq = 0, cz = 0;
  // UP_ITEMS rows per thread per step, loaded before any is counted; two steps' rows take turns, so that
  // the next step's loads are in flight while this step is counted, and every loaded register is taken
  // at the top of a step (a load the compiler moved into one of the classification's branches would be
  // waited for there alone). The step count is the same for every thread of the block.
  constexpr int UP_ITEMS = 8;
  constexpr uint32_t UP_STEP = BLOCK * UP_ITEMS;
  const int steps = end > begin ? int((end - begin + UP_STEP - 1) / UP_STEP) : 0;
  uint64_t raw_a[UP_ITEMS], raw_b[UP_ITEMS];
  uint32_t vb_a[UP_ITEMS], vb_b[UP_ITEMS];
  auto load_step = [&](uint64_t base, uint64_t (&raw)[UP_ITEMS], uint32_t (&vb)[UP_ITEMS]) {
#pragma unroll
    for (int u = 0; u < UP_ITEMS; u++) {   // rows past the end read the last row
      const uint32_t i = uint32_t(std::min<uint64_t>(base + uint64_t(u) * BLOCK + threadIdx.x, uint64_t(end) - 1u));
      uint32_t bit;
      raw[u] = load_raw<W>(vals, i);
      vb[u] = load_valid_byte(valid, voff, i, bit);
    }
  };
  auto count_step = [&](uint64_t base, const uint64_t (&raw)[UP_ITEMS], const uint32_t (&vb)[UP_ITEMS]) {
#pragma unroll
    for (int u = 0; u < UP_ITEMS; u++) asm volatile("" ::"v"(raw[u]), "v"(vb[u]));
    K kk[UP_ITEMS];
    int cc[UP_ITEMS];
#pragma unroll
    for (int u = 0; u < UP_ITEMS; u++) {
      const uint64_t i = base + uint64_t(u) * BLOCK + threadIdx.x;
      const uint32_t ic = uint32_t(std::min<uint64_t>(i, uint64_t(end) - 1u));
      const uint32_t bit = valid ? uint32_t((voff + int64_t(ic)) & 7) : 0u;
      kk[u] = 0;
      cc[u] = i < end ? classify_raw<K>(raw[u], (vb[u] >> bit) & 1u, type, forder, desc, kk[u]) : 3;
    }
#pragma unroll
    for (int u = 0; u < UP_ITEMS; u++) {
      const uint32_t ui = uint32_t(base + uint64_t(u) * BLOCK + threadIdx.x - begin) >> unit_shift;
      if (cc[u] == 0) {
        K k = kk[u];
        cm++;
        o |= k;
        a &= k;
        // all_digits == 0 (the hybrid sort's analysis): only the top digit's histogram, per unit (the
        // levels count the lower digits per bucket; the LSD passes, if they run instead, analyse again).
        if (all_digits) {
#pragma unroll
          for (int p = 0; p < P - 1; p++) atomicAdd(&h[p][uint32_t(k >> (8 * p)) & 0xFFu], 1u);
        }
        const uint32_t top = uint32_t(k >> (8 * (P - 1))) & 0xFFu;
        atomicAdd(units ? &hu[ui & (UNITS_PER_BLOCK - 1)][top] : &h[P - 1][top], 1u);
      } else if (cc[u] == 1) {
        cq++;
        if (units) atomicAdd(&hu[ui & (UNITS_PER_BLOCK - 1)][BIN_NAN], 1u);
      } else if (cc[u] == 2) {
        cz++;
        if (units) atomicAdd(&hu[ui & (UNITS_PER_BLOCK - 1)][BIN_NULL], 1u);
      }
    }
  };
  if (steps > 0) {
    load_step(begin, raw_a, vb_a);
    for (int t = 0; t < steps; t += 2) {
      const uint64_t base = uint64_t(begin) + uint64_t(t) * UP_STEP;
      load_step(base + UP_STEP, raw_b, vb_b);
      count_step(base, raw_a, vb_a);
      if (t + 1 >= steps) break;
      load_step(base + 2 * UP_STEP, raw_a, vb_a);
      count_step(base + UP_STEP, raw_b, vb_b);
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
  if (units) {
    // The block's top-digit histogram is the sum of its units'; every unit's counts go out bin-major.
    const uint32_t u0 = begin >> unit_shift;
    const uint32_t nu = (end - begin + (1u << unit_shift) - 1u) >> unit_shift;
    for (uint32_t b = threadIdx.x; b < uint32_t(NBINS_FIRST); b += BLOCK) {
      uint32_t sum = 0;
      for (uint32_t ui = 0; ui < nu; ui++) {
        sum += hu[ui][b];
        unit_hist[size_t(b) * nunits + u0 + ui] = hu[ui][b];
      }
      if (b < uint32_t(RADIX)) h[P - 1][b] = sum;
    }
    __syncthreads();
  }
  const uint32_t x = xcd_id();
  uint32_t* xh = &ctl->xhist[x][0][0];
  for (int i = threadIdx.x; i < P * RADIX; i += BLOCK) {
    uint32_t v = (&h[0][0])[i];
    if (v) __hip_atomic_fetch_add(xh + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  if (threadIdx.x == 0) {
    if (cnt[0]) {
      __hip_atomic_fetch_add(&ctl->xm[x], cnt[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_or(&ctl->xor_[x], s_or, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_fetch_and(&ctl->xand[x], s_and, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (cnt[1]) __hip_atomic_fetch_add(&ctl->xq[x], cnt[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (cnt[2]) __hip_atomic_fetch_add(&ctl->xz[x], cnt[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

template <int P>
__global__ __launch_bounds__(BLOCK) void sort_plan(SortCtl* ctl, int nulls_first) {
  __shared__ uint32_t tmp[WAVES];
  // Fold the per-XCD partials of the upsweep.
  for (int p = 0; p < P; p++) {
    uint32_t v = 0;
#pragma unroll
    for (int x = 0; x < NXCD; x++) v += ctl->xhist[x][p][threadIdx.x];
    ctl->hist[p][threadIdx.x] = v;
  }
  if (threadIdx.x == 0) {
    uint32_t m = 0, q = 0, z = 0;
    unsigned long long o = 0ull, a = ~0ull;
    for (int x = 0; x < NXCD; x++) {
      m += ctl->xm[x];
      q += ctl->xq[x];
      z += ctl->xz[x];
      o |= ctl->xor_[x];
      a &= ctl->xand[x];
    }
    ctl->m = m;
    ctl->q = q;
    ctl->z = z;
    ctl->key_or = o;
    ctl->key_and = a;
  }
  __syncthreads();
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

template <typename K, int W, int ITEMS>
__global__ __launch_bounds__(BLOCK, 3) void sort_first(const void* vals, const uint8_t* valid, int64_t voff,
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
  uint32_t bins[ITEMS];
  const uint32_t rows[1] = {0};   // rows are base_e + 64 * j (ROWS_IOTA)
  const uint32_t base_e = tile * uint32_t(TILE) + uint32_t(w) * 64u * ITEMS + uint32_t(lane);
  // Every load first (lanes past the end read the last row), then the mapping on registers.
  uint32_t vbit[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t e = min(base_e + uint32_t(j) * 64u, n - 1u);
    uint32_t bit;
    const uint32_t vb = load_valid_byte(valid, voff, e, bit);
    vbit[j] = (vb >> bit) & 1u;
    keys[j] = K(load_raw<W>(vals, e));
  }
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    uint32_t e = base_e + uint32_t(j) * 64u;
    uint32_t b = BIN_NONE;
    const uint64_t raw = uint64_t(keys[j]);
    keys[j] = 0;
    if (e < n) {
      K k;
      int c = classify_raw<K>(raw, vbit[j] != 0u, type, forder, desc, k);
      if (c == 0) {
        keys[j] = k;
        b = uint32_t(k >> shift) & 0xFFu;
      } else {
        b = c == 1 ? BIN_NAN : BIN_NULL;
      }
    }
    bins[j] = b;
  }
  rank_tile<ITEMS, 9>(bins, wc[w], lane);
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
    uint32_t b = bin_of(bins[j]);
    if (b == BIN_NAN || b == BIN_NULL) out_idx[bin_base[b] + wc[w][b] + rank_of(bins[j])] = base_e + uint32_t(j) * 64u;
  }
  if (last) {
    scatter_values<K, ITEMS, false, true>(keys, rows, bins, wc[w], bin_start, bin_base, s_digit, s_stage, vcount,
                                          nullptr, out_idx, base_e);
  } else {
    scatter_values<K, ITEMS, true, true>(keys, rows, bins, wc[w], bin_start, bin_base, s_digit, s_stage, vcount,
                                         out_keys, out_rows, base_e);
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
  uint32_t rows[ITEMS], bins[ITEMS];
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
  // A full tile has no lanes past the end, so its bins need 8 bits; the last tile needs the 9th.
  if ((tile + 1u) * uint32_t(TILE) <= m) rank_tile<ITEMS, 8>(bins, wc[w], lane);
  else rank_tile<ITEMS, 9>(bins, wc[w], lane);
  __syncthreads();
  uint32_t vcount = tile_bin_layout<RADIX>(wc, tile_total, bin_start, tmp);
  look_back<RADIX>(status, tile, epoch, tile_total, bin_base);
  __syncthreads();
  const uint32_t vb = last ? ctl->value_base : 0u;
  bin_base[threadIdx.x] += ctl->digit_base[pass][threadIdx.x] + vb;
  __syncthreads();
  if (last) {
    scatter_values<K, ITEMS, false>(keys, rows, bins, wc[w], bin_start, bin_base, s_digit, s_stage, vcount,
                                    nullptr, out_idx);
  } else {
    scatter_values<K, ITEMS, true>(keys, rows, bins, wc[w], bin_start, bin_base, s_digit, s_stage, vcount,
                                   out_keys, out_rows);
  }
}

// ---------------------------------------------------------------- top-digit partition by units
//
// The hybrid sort's first pass, without look-back: the upsweep counted every unit's top digits, NaN
// and null rows, so one workgroup per bin scans those counts into every unit's destinations
// (msd_unit_scan), and persistent workgroups then partition whole units, tile after tile, keeping a
// running destination per bin in LDS (msd_part). Units are taken from a counter, in any order.

// unit_base[b][u] = the first destination of bin b's rows of unit u: inside the value block for
// digits (positions in the partitioned key array), absolute output slots for NaN and null rows.
template <int P>
__global__ __launch_bounds__(BLOCK) void msd_unit_scan(const SortCtl* ctl, const uint32_t* unit_hist,
                                                       uint32_t nunits, uint32_t* unit_base) {
  __shared__ uint32_t tmp[WAVES];
  const uint32_t b = blockIdx.x;
  uint32_t carry = b < uint32_t(RADIX) ? ctl->digit_base[P - 1][b] : (b == uint32_t(BIN_NAN) ? ctl->nan_base
                                                                                                 : ctl->null_base);
  for (uint32_t u0 = 0; u0 < nunits; u0 += BLOCK) {
    const uint32_t u = u0 + threadIdx.x;
    const uint32_t v = u < nunits ? unit_hist[size_t(b) * nunits + u] : 0u;
    uint32_t total;
    const uint32_t ex = block_excl_scan(v, tmp, total);
    if (u < nunits) unit_base[size_t(b) * nunits + u] = carry + ex;
    carry += total;
  }
}

// POS_ROWS: the rows written beside the keys are their own positions in this pass's output, and
// pos_of_row[row] gets that position (all ones for NaN and null rows), for the ORDER BY of whole
// tables (order_by_columns).
template <typename K, int W, int ITEMS, bool POS_ROWS = false>
__global__ __launch_bounds__(BLOCK, 3) void msd_part(const void* vals, const uint8_t* valid, int64_t voff,
                                                     uint32_t n, int type, int forder, int desc, SortCtl* ctl,
                                                     const uint32_t* unit_base, uint32_t nunits,
                                                     uint32_t unit_shift, K* out_keys, uint32_t* out_rows,
                                                     uint32_t* out_idx, uint32_t* pos_of_row = nullptr) {
  constexpr int P = int(sizeof(K));
  constexpr int TILE = BLOCK * ITEMS;
  static_assert((1u << UNIT_SHIFT_SMALL) % TILE == 0, "units hold whole tiles");
  __shared__ uint32_t wc[WAVES][WC_STRIDE];
  __shared__ uint32_t tile_total[WC_STRIDE];
  __shared__ uint32_t bin_start[RADIX];
  __shared__ uint32_t bin_base[WC_STRIDE];
  __shared__ uint32_t running[WC_STRIDE];
  __shared__ uint32_t tmp[WAVES];
  __shared__ uint32_t s_unit;
  __shared__ uint8_t s_digit[TILE];
  __shared__ K s_stage[TILE];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int shift = 8 * (P - 1);
  const uint32_t rows[1] = {0};   // rows are base_e + 64 * j (ROWS_IOTA)
  while (true) {
    if (threadIdx.x == 0) s_unit = atomicAdd(&ctl->tile_counter[11], 1u);
    __syncthreads();
    const uint32_t unit = s_unit;
    if (unit >= nunits) return;   // uniform: every workgroup ends here once the units run out
    for (int b = threadIdx.x; b < NBINS_FIRST; b += BLOCK) running[b] = unit_base[size_t(b) * nunits + unit];
    const uint32_t begin = unit << unit_shift;
    const uint32_t end = min(n, begin + (1u << unit_shift));
    for (uint32_t t0 = begin; t0 < end; t0 += TILE) {
      for (int i = threadIdx.x; i < WAVES * WC_STRIDE; i += BLOCK) (&wc[0][0])[i] = 0u;
      __syncthreads();
      K keys[ITEMS];
      uint32_t bins[ITEMS], vbit[ITEMS];
      const uint32_t base_e = t0 + uint32_t(w) * 64u * ITEMS + uint32_t(lane);
#pragma unroll
      for (int j = 0; j < ITEMS; j++) {   // every load first (lanes past the unit read its last row)
        const uint32_t e = min(base_e + uint32_t(j) * 64u, end - 1u);
        uint32_t bit;
        const uint32_t vb = load_valid_byte(valid, voff, e, bit);
        vbit[j] = (vb >> bit) & 1u;
        keys[j] = K(load_raw<W>(vals, e));
      }
#pragma unroll
      for (int j = 0; j < ITEMS; j++) {
        const uint32_t e = base_e + uint32_t(j) * 64u;
        uint32_t b = BIN_NONE;
        const uint64_t raw = uint64_t(keys[j]);
        keys[j] = 0;
        if (e < end) {
          K k;
          const int c = classify_raw<K>(raw, vbit[j] != 0u, type, forder, desc, k);
          if (c == 0) {
            keys[j] = k;
            b = uint32_t(k >> shift) & 0xFFu;
          } else {
            b = c == 1 ? BIN_NAN : BIN_NULL;
          }
        }
        bins[j] = b;
      }
      rank_tile<ITEMS, 9>(bins, wc[w], lane);
      __syncthreads();
      const uint32_t vcount = tile_bin_layout<NBINS_FIRST>(wc, tile_total, bin_start, tmp);
      for (int b = threadIdx.x; b < NBINS_FIRST; b += BLOCK) {
        bin_base[b] = running[b];
        running[b] += tile_total[b];
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < ITEMS; j++) {   // NaN and null rows go straight to their final slots
        const uint32_t b = bin_of(bins[j]);
        if (b == BIN_NAN || b == BIN_NULL) {
          out_idx[bin_base[b] + wc[w][b] + rank_of(bins[j])] = base_e + uint32_t(j) * 64u;
          if constexpr (POS_ROWS) pos_of_row[base_e + uint32_t(j) * 64u] = 0xFFFFFFFFu;
        }
      }
      scatter_values<K, ITEMS, true, true, POS_ROWS>(keys, rows, bins, wc[w], bin_start, bin_base, s_digit, s_stage,
                                                     vcount, out_keys, out_rows, base_e, pos_of_row);
      __syncthreads();
    }
  }
}

// ---------------------------------------------------------------- hybrid MSD sort (64-bit keys)
//
// For large inputs of 64-bit keys whose top digit varies: the first pass partitions the value rows
// stably by the top digit (msd_part, above). Then, level by level down the active digits, every
// bucket still larger than LOCAL_MAX rows is partitioned stably by the next digit (onesweep tiles
// whose look-back stops at their bucket's first tile), into the other key / row buffer at the same
// positions; a bucket whose digit is the same on all its rows is carried to the next level without
// moving. Buckets of at most LOCAL_MAX rows leave for the bucket sort, which sorts each in LDS by its
// remaining digits and writes its rows to their final slots. After the lowest active digit every
// bucket holds equal keys, in input order, and is copied. Uniform keys take three passes over the
// rows (top digit, second digit, bucket sort) instead of eight.

#ifndef AM_MSD_FITEMS
#define AM_MSD_FITEMS 16   // keys per thread of the top-digit partition
#endif
// One workgroup: the level kernels' MsdCtl; the top-digit buckets msd_part made, into the first
// large-bucket list or the bucket sort's list; the counters cleared. `low` = the active digits below
// the top one.
template <int P>
__global__ __launch_bounds__(BLOCK) void msd_init_lists(const SortCtl* ctl, MsdCtl* mc, uint32_t epoch,
                                                        Bucket* large, Bucket* local, uint32_t local_cap,
                                                        uint32_t* lc, unsigned long long* key_out,
                                                        unsigned long long key_xor) {
  if (threadIdx.x < LC_WORDS) lc[threadIdx.x] = 0u;
  if (threadIdx.x == 0) *mc = {ctl->active_mask, ctl->value_base, epoch, 0u, key_out, key_xor};
  __syncthreads();
  const uint32_t b = threadIdx.x;
  const uint32_t size = ctl->hist[P - 1][b];
  if (size == 0u) return;
  const Bucket bk{ctl->digit_base[P - 1][b], size, 0u, uint32_t(P - 1)};
  const uint32_t low = ctl->active_mask & ((1u << (P - 1)) - 1u);
  if (size > LOCAL_MAX && low != 0u) large[atomicAdd(&lc[LC_LIST], 1u)] = bk;
  else push_local(local, local_cap, lc, bk, low == 0u);
}

// ---------------------------------------------------------------- host side

// The hybrid sort of 64-bit keys (the analysed plan says the top digit varies; `mask` = the active
// digits). Returns false, with nothing written to out_idx that the LSD passes do not write again, when
// the bucket sort's list runs out of room.
template <typename K, int W, int FITEMS, int ITEMS>
bool run_hybrid(const void* vals, int type, const uint8_t* valid, int64_t voff, uint32_t n, int dsc, int forder,
                SortCtl* ctl, unsigned long long* status, uint32_t epoch, uint32_t m, uint32_t mask, K* k0,
                uint32_t* r0, K* k1, uint32_t* r1, uint32_t* out_idx, const uint32_t* unit_hist, uint32_t* unit_base,
                uint32_t nunits, uint32_t unit_shift, hipStream_t s, uint32_t* pos_of_row,
                unsigned long long* key_out, unsigned long long key_xor) {
  constexpr int P = int(sizeof(K));
  const MsdWork w = msd_work(m, ITEMS, device_alloc);
  {
    ProfScope ps("sort_msd_unit_scan");
    hipLaunchKernelGGL(msd_unit_scan<P>, dim3(NBINS_FIRST), dim3(BLOCK), 0, s, ctl, unit_hist, nunits, unit_base);
    AM_HIP(hipGetLastError());
  }
  {
    ProfScope ps("sort_msd_first");
    const uint32_t workers = std::min(nunits, uint32_t(device().cu_count) * 3u);
    if (pos_of_row)
      hipLaunchKernelGGL((msd_part<K, W, FITEMS, true>), dim3(workers), dim3(BLOCK), 0, s, vals, valid, voff, n, type,
                         forder, dsc, ctl, unit_base, nunits, unit_shift, k0, r0, out_idx, pos_of_row);
    else
      hipLaunchKernelGGL((msd_part<K, W, FITEMS>), dim3(workers), dim3(BLOCK), 0, s, vals, valid, voff, n, type, forder,
                         dsc, ctl, unit_base, nunits, unit_shift, k0, r0, out_idx, nullptr);
    AM_HIP(hipGetLastError());
  }
  {
    ProfScope ps("sort_msd_lists");
    hipLaunchKernelGGL(msd_init_lists<P>, dim3(1), dim3(BLOCK), 0, s, ctl, w.mc, epoch, w.large[0], w.local,
                       w.local_cap, w.lc, key_out, key_xor);
    AM_HIP(hipGetLastError());
  }
  return msd_levels<K, ITEMS>(w, m, P - 1, mask, false, LOCAL_MAX, status, k0, r0, k1, r1, out_idx, s,
                              key_out != nullptr);
}

// With pos_of_row (the ORDER BY of whole tables): only the hybrid sort runs, with positions for rows
// (msd_part's POS_ROWS form); when it cannot (too few rows, a constant top digit, a full bucket list)
// the call returns false and out_idx holds nothing to use. *value_rows receives m. With key_out (an
// integer key) the bucket sorts also write the value block of the sorted key column there.
#ifndef AM_HYBRID_MIN_LOG
#define AM_HYBRID_MIN_LOG 16
#endif
// Rows from which a 64-bit key takes the hybrid sort (and the plan is read back).
constexpr uint32_t HYBRID_MIN = 1u << AM_HYBRID_MIN_LOG;

template <typename K, int W, int ITEMS>
bool run_sort(const void* vals, int type, const uint8_t* valid, int64_t voff, uint32_t n, bool desc,
              int placement, int forder, uint32_t* out_idx, hipStream_t s, uint32_t* pos_of_row = nullptr,
              uint32_t* value_rows = nullptr, uint32_t* bucket_base = nullptr, unsigned long long* key_out = nullptr,
              unsigned long long key_xor = 0) {
  constexpr int P = int(sizeof(K));
  if (pos_of_row && !(P == 8 && n >= HYBRID_MIN && !capturing(s))) return false;
  constexpr int TILE = BLOCK * ITEMS;
  // The first pass keeps at most 16 keys per thread: its raw-column loads and 258-bin classification
  // need more registers, and 16 keeps it at three workgroups per CU.
  constexpr int FITEMS = ITEMS > 16 ? 16 : ITEMS;
  constexpr int FTILE = BLOCK * FITEMS;
  const uint32_t tiles = (n + TILE - 1) / TILE;
  const uint32_t ftiles = (n + FTILE - 1) / FTILE;
  DevBuf ctl_buf = device_alloc(sizeof(SortCtl));
  SortCtl* ctl = static_cast<SortCtl*>(ctl_buf->ptr);
  // Status rows for every tile of a pass; a level of the hybrid sort has up to one more tile per large
  // bucket (fewer than the levels' bucket capacity, m / 2048 + 4096).
  const uint32_t hyb_tiles = tiles + n / 2048u + 4096u + 2u;
  uint32_t epoch = reserve_status(size_t(std::max(hyb_tiles, ftiles)) * STATUS_STRIDE, s);
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
  // Large 64-bit inputs may take the hybrid sort: the upsweep then also counts its units.
  const bool hybrid_ok = P == 8 && n >= HYBRID_MIN && !capturing(s);
  const uint32_t unit_shift = n >= (1u << 22) ? UNIT_SHIFT_LARGE : UNIT_SHIFT_SMALL;
  const uint32_t unit_rows = 1u << unit_shift;
  const uint32_t nunits = hybrid_ok ? (n + unit_rows - 1) / unit_rows : 0u;
  DevBuf units_buf;
  uint32_t* unit_hist = nullptr;
  if (hybrid_ok) {
    units_buf = device_alloc(size_t(nunits) * NBINS_FIRST * 4 * 2);
    unit_hist = static_cast<uint32_t*>(units_buf->ptr);
  }
  // The upsweep and the plan. With units (the hybrid sort's analysis) the upsweep counts only the top
  // digit; if the LSD passes run after all, the analysis is redone with every digit.
  auto analyse_column = [&](bool units) {
    uint32_t max_blocks = uint32_t(device().cu_count) * 4u;
    uint32_t blocks = std::max(1u, std::min(max_blocks, (n + 8191u) / 8192u));
    // Block ranges are whole units when units are counted, at most UNITS_PER_BLOCK of them.
    const uint32_t align = units ? unit_rows : 2048u;
    uint32_t rpb = ((n + blocks - 1) / blocks + align - 1) / align * align;
    if (units) rpb = std::min(rpb, UNITS_PER_BLOCK * unit_rows);
    blocks = (n + rpb - 1) / rpb;
    {
      ProfScope ps("sort_upsweep");
      hipLaunchKernelGGL((sort_upsweep<K, W>), dim3(blocks), dim3(BLOCK), 0, s, vals, valid, voff, n, type, forder,
                         dsc, rpb, ctl, units ? unit_hist : nullptr, nunits, unit_shift, units ? 0 : 1);
      AM_HIP(hipGetLastError());
    }
    ProfScope ps("sort_plan");
    hipLaunchKernelGGL(sort_plan<P>, dim3(1), dim3(BLOCK), 0, s, ctl, placement == 1 ? 1 : 0);
    AM_HIP(hipGetLastError());
  };
  analyse_column(hybrid_ok);
  if (bucket_base) {   // where each top-digit bucket starts in the value block (the partition's layout)
    AM_HIP(hipMemcpyAsync(bucket_base, &ctl->digit_base[P - 1][0], RADIX * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
  // Large inputs read the plan back once (one sync, as the reference's analysis readback does), so
  // that only active passes are launched, over the value rows only.
  struct Plan { uint32_t mask, first, last, m; } plan{0xFFu, 0, uint32_t(P - 1), n};
  const bool analyse = n >= HYBRID_MIN && !capturing(s);
  if (analyse) {
    Plan* hp = reinterpret_cast<Plan*>(host_plan_words());
    hipLaunchKernelGGL(sort_read_plan, dim3(1), dim3(1), 0, s, ctl, reinterpret_cast<uint32_t*>(hp));
    AM_HIP(hipGetLastError());
    AM_HIP(hipStreamSynchronize(s));
    plan = *hp;
  }
  K* kp0 = static_cast<K*>(k0->ptr);
  K* kp1 = static_cast<K*>(k1->ptr);
  uint32_t* rp0 = static_cast<uint32_t*>(r0->ptr);
  uint32_t* rp1 = static_cast<uint32_t*>(r1->ptr);
  if constexpr (P == 8) {
    // The hybrid sort when the top digit varies; the LSD passes below run instead if its bucket list
    // runs out of room.
    if (value_rows) *value_rows = plan.m;
    if (hybrid_ok && analyse && ((plan.mask >> (P - 1)) & 1u) &&
        run_hybrid<K, W, AM_MSD_FITEMS, ITEMS>(
            vals, type, valid, voff, n, dsc, forder, ctl, status, epoch, plan.m, plan.mask, kp0, rp0, kp1, rp1,
            out_idx, unit_hist, unit_hist + size_t(nunits) * NBINS_FIRST, nunits, unit_shift, s, pos_of_row,
            key_out, key_xor))
      return true;
  }
  if (pos_of_row) return false;
  if (hybrid_ok) {
    // The LSD passes need every digit's histogram: the analysis again, from a cleared control block
    // (the plan read back above stays the same).
    {
      ProfScope ps("sort_init");
      hipLaunchKernelGGL(sort_init, dim3(1), dim3(BLOCK), 0, s, ctl);
      AM_HIP(hipGetLastError());
    }
    analyse_column(false);
  }
  {
    ProfScope ps("sort_first");
    hipLaunchKernelGGL((sort_first<K, W, FITEMS>), dim3(ftiles), dim3(BLOCK), 0, s, vals, valid, voff, n, type, forder,
                       dsc, ctl, status, epoch, kp0, rp0, out_idx);
    AM_HIP(hipGetLastError());
  }
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
  return true;
}

template <typename K, int W>
void dispatch_items(int items, const void* vals, int type, const uint8_t* valid, int64_t voff, uint32_t n,
                    bool desc, int placement, int forder, uint32_t* out_idx, hipStream_t s) {
  switch (items) {
    case 4: run_sort<K, W, 4>(vals, type, valid, voff, n, desc, placement, forder, out_idx, s); break;
    case 8: run_sort<K, W, 8>(vals, type, valid, voff, n, desc, placement, forder, out_idx, s); break;
    case 12: run_sort<K, W, 12>(vals, type, valid, voff, n, desc, placement, forder, out_idx, s); break;
    case 20: run_sort<K, W, 20>(vals, type, valid, voff, n, desc, placement, forder, out_idx, s); break;
    default: run_sort<K, W, 16>(vals, type, valid, voff, n, desc, placement, forder, out_idx, s); break;
  }
}

// ---------------------------------------------------------------- ORDER BY of whole tables
//
// order_by_columns sorts by one 64-bit key with the hybrid sort in its POS_ROWS form: rows carried
// through the levels are positions p in the top-digit partition's output (the value block), and
// pos_of_row[r] = p for every value row r. A column then moves in two row passes:
//   ob_partition  part[pos_of_row[r]] = col[r]: the column laid out as the partition laid out its keys
//                 (reads in row order; writes in runs, as the partition's own)
//   ob_gather     out[i] = part[out_idx[i]] in the value block, col[out_idx[i]] in the NaN and null
//                 blocks (whose rows keep input order)
// Output i of the value block and its position p lie in the same top-digit bucket, so the gather of an
// output range reads one bucket's part of the column. Each XCD takes its own eighth of the outputs
// (XCD-private chunk counters; a workgroup that finds its eighth done takes chunks of the others), so a
// bucket's lines are fetched into one L2, not eight.

#ifndef AM_OB_CHUNK
#define AM_OB_CHUNK 4096
#endif
#ifndef AM_OB_GATHER_WG
#define AM_OB_GATHER_WG 4
#endif
constexpr uint32_t OB_CHUNK = AM_OB_CHUNK;   // outputs per chunk of the gather (64-row words)
constexpr uint32_t OB_GROUPS = OB_CHUNK / 64 / WAVES;   // 64-row groups per wave per chunk

// Each workgroup takes one contiguous range of rows, tile by tile. The partition is stable, so the
// rows of one tile that belong to one top-digit bucket have consecutive positions p: the tile is
// staged in LDS in position order (bucket by bucket, found by a binary search of p in the buckets'
// starts) and written out as runs of consecutive addresses, from this XCD's L2 only.
constexpr uint32_t OBP_ITEMS = 8;
constexpr uint32_t OBP_TILE = BLOCK * OBP_ITEMS;
template <int W>
__global__ __launch_bounds__(BLOCK) void ob_partition(const void* col, const uint8_t* valid, int64_t voff,
                                                      uint32_t n, uint32_t rows_per_block, const uint32_t* pos_of_row,
                                                      const uint32_t* bucket_base, void* part, uint8_t* vpart) {
  __shared__ uint32_t s_base[RADIX];
  __shared__ uint32_t s_minp[RADIX], s_cnt[RADIX], s_off[RADIX];
  __shared__ uint32_t tmp[WAVES];
  __shared__ uint64_t s_val[OBP_TILE];
  __shared__ uint8_t s_bkt[OBP_TILE];
  __shared__ uint8_t s_vb[OBP_TILE];
  s_base[threadIdx.x] = bucket_base[threadIdx.x];
  s_minp[threadIdx.x] = 0xFFFFFFFFu;
  s_cnt[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t first = blockIdx.x * rows_per_block;
  const uint32_t last = min(n, first + rows_per_block);
  for (uint32_t t0 = first; t0 < last; t0 += OBP_TILE) {
    uint32_t p[OBP_ITEMS], bk[OBP_ITEMS], vb[OBP_ITEMS], bit[OBP_ITEMS];
    uint64_t v[OBP_ITEMS];
#pragma unroll
    for (uint32_t j = 0; j < OBP_ITEMS; j++) {   // every load first (rows past the range read its last row)
      const uint32_t r = min(t0 + j * BLOCK + threadIdx.x, last - 1u);
      p[j] = pos_of_row[r];
      v[j] = load_raw<W>(col, r);
      vb[j] = load_valid_byte(valid, voff, r, bit[j]);
    }
#pragma unroll
    for (uint32_t j = 0; j < OBP_ITEMS; j++) {
      bk[j] = 0xFFFFFFFFu;
      if (t0 + j * BLOCK + threadIdx.x >= last || p[j] == 0xFFFFFFFFu) continue;
      uint32_t lo = 0, hi = RADIX;   // the last bucket whose start is <= p
      while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_base[mid] <= p[j]) lo = mid;
        else hi = mid;
      }
      bk[j] = lo;
      atomicMin(&s_minp[lo], p[j]);
      atomicAdd(&s_cnt[lo], 1u);
    }
    __syncthreads();
    uint32_t total;
    s_off[threadIdx.x] = block_excl_scan(s_cnt[threadIdx.x], tmp, total);
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < OBP_ITEMS; j++) {
      if (bk[j] == 0xFFFFFFFFu) continue;
      const uint32_t k = s_off[bk[j]] + (p[j] - s_minp[bk[j]]);
      if (k >= OBP_TILE) continue;   // cannot happen for a stable partition; keeps LDS writes in bounds
      s_val[k] = v[j];
      s_bkt[k] = uint8_t(bk[j]);
      s_vb[k] = uint8_t((vb[j] >> bit[j]) & 1u);
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < total; k += BLOCK) {
      const uint32_t b = s_bkt[k];
      const uint32_t dest = s_minp[b] + (k - s_off[b]);
      if constexpr (W == 8) static_cast<uint64_t*>(part)[dest] = s_val[k];
      else if constexpr (W == 4) static_cast<uint32_t*>(part)[dest] = uint32_t(s_val[k]);
      else if constexpr (W == 2) static_cast<uint16_t*>(part)[dest] = uint16_t(s_val[k]);
      else static_cast<uint8_t*>(part)[dest] = uint8_t(s_val[k]);
      if (vpart) vpart[dest] = s_vb[k];
    }
    __syncthreads();
    s_minp[threadIdx.x] = 0xFFFFFFFFu;
    s_cnt[threadIdx.x] = 0u;
    __syncthreads();
  }
}

// Counters of the gather: [x * 32] the next chunk of XCD x's eighth (each on its own 128-byte line).
constexpr uint32_t OB_CTR_STRIDE = 32;

template <int W>
__global__ __launch_bounds__(BLOCK) void ob_gather(const uint32_t* out_idx, uint32_t n, uint32_t vbase, uint32_t m,
                                                   const void* part, const uint8_t* vpart, const void* col,
                                                   const uint8_t* valid, int64_t voff, void* out,
                                                   unsigned long long* out_valid, uint32_t* ctr) {
  __shared__ uint32_t s_chunk;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint32_t words = (n + 63u) / 64u;
  // Eighth x of the outputs: 64-row words [x * words / 8, (x + 1) * words / 8).
  const uint32_t home = xcd_id();
  for (uint32_t k = 0; k < uint32_t(NXCD); k++) {
    const uint32_t x = (home + k) & uint32_t(NXCD - 1);
    const uint32_t w0 = uint32_t(uint64_t(words) * x / NXCD), w1 = uint32_t(uint64_t(words) * (x + 1) / NXCD);
    const uint32_t chunks = (w1 - w0 + OB_CHUNK / 64 - 1) / (OB_CHUNK / 64);
    while (true) {
      if (threadIdx.x == 0) s_chunk = atomicAdd(&ctr[x * OB_CTR_STRIDE], 1u);
      __syncthreads();
      const uint32_t c = s_chunk;
      __syncthreads();
      if (c >= chunks) break;   // uniform: this eighth has no chunk left
      uint32_t idx[OB_GROUPS];
      bool in_val[OB_GROUPS];
#pragma unroll
      for (uint32_t g = 0; g < OB_GROUPS; g++) {
        const uint32_t word = w0 + c * (OB_CHUNK / 64) + uint32_t(w) * OB_GROUPS + g;
        const uint32_t i = min(word * 64u + uint32_t(lane), n - 1u);
        in_val[g] = i - vbase < m;
        // Positions index the m-row partition, rows the n-row column (clamped: reads stay in bounds).
        idx[g] = min(out_idx[i], (in_val[g] ? m : n) - 1u);
      }
      uint64_t v[OB_GROUPS];
      uint32_t vb[OB_GROUPS], bit[OB_GROUPS];
#pragma unroll
      for (uint32_t g = 0; g < OB_GROUPS; g++) {
        // A value-block output reads the partitioned column at its position, any other the column at
        // its row.
        const void* src = in_val[g] ? part : col;
        v[g] = load_raw<W>(src, idx[g]);
        if (in_val[g]) {
          vb[g] = vpart ? load_global(vpart + idx[g]) : 1u;
          bit[g] = 0u;
        } else {
          vb[g] = load_valid_byte(valid, voff, idx[g], bit[g]);
        }
      }
#pragma unroll
      for (uint32_t g = 0; g < OB_GROUPS; g++) {
        const uint32_t word = w0 + c * (OB_CHUNK / 64) + uint32_t(w) * OB_GROUPS + g;
        const uint32_t i = word * 64u + uint32_t(lane);
        const bool live = word < w1 && i < n;
        if (live) {
          if constexpr (W == 8) static_cast<uint64_t*>(out)[i] = v[g];
          else if constexpr (W == 4) static_cast<uint32_t*>(out)[i] = uint32_t(v[g]);
          else if constexpr (W == 2) static_cast<uint16_t*>(out)[i] = uint16_t(v[g]);
          else static_cast<uint8_t*>(out)[i] = uint8_t(v[g]);
        }
        if (out_valid) {
          const unsigned long long bits = __ballot(live && ((vb[g] >> bit[g]) & 1u));
          if (lane == 0 && word < w1) out_valid[word] = bits;
        }
      }
    }
  }
}

__global__ void ob_clear(uint32_t* ctr) {
  if (threadIdx.x < uint32_t(NXCD * OB_CTR_STRIDE)) ctr[threadIdx.x] = 0u;
}

template <int W>
void ob_move_column(const Array& c, bool nulls, const uint32_t* pos_of_row, const uint32_t* bucket_base,
                    const uint32_t* out_idx, uint32_t n, uint32_t vbase, uint32_t m, Array* out, uint32_t* ctr,
                    hipStream_t s) {
  DevBuf part = device_alloc(size_t(std::max<uint32_t>(m, 1)) * W);
  DevBuf vpart;
  if (nulls) vpart = device_alloc(std::max<uint32_t>(m, 1));
  retain(part);
  if (vpart) retain(vpart);
  const uint8_t* valid = nulls ? c.valid_bits() : nullptr;
  {
    ProfScope ps("order_by_partition");
    const uint32_t tiles = (n + OBP_TILE - 1) / OBP_TILE;
    const uint32_t blocks = std::max(1u, std::min(uint32_t(device().cu_count) * 4u, tiles));
    const uint32_t rpb = (tiles + blocks - 1) / blocks * OBP_TILE;   // whole tiles per workgroup
    hipLaunchKernelGGL((ob_partition<W>), dim3(blocks), dim3(BLOCK), 0, s, c.data(), valid, c.offset, n, rpb,
                       pos_of_row, bucket_base, part->ptr, nulls ? static_cast<uint8_t*>(vpart->ptr) : nullptr);
    AM_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(ob_clear, dim3(1), dim3(BLOCK), 0, s, ctr);
  AM_HIP(hipGetLastError());
  {
    ProfScope ps("order_by_gather");
    const uint32_t blocks = uint32_t(device().cu_count) * AM_OB_GATHER_WG;   // workgroups per CU
    hipLaunchKernelGGL((ob_gather<W>), dim3(blocks), dim3(BLOCK), 0, s, out_idx, n, vbase, m, part->ptr,
                       nulls ? static_cast<const uint8_t*>(vpart->ptr) : nullptr, c.data(), valid, c.offset,
                       out->values->ptr, nulls ? static_cast<unsigned long long*>(out->validity->ptr) : nullptr, ctr);
    AM_HIP(hipGetLastError());
  }
}

template <int ITEMS>
bool ob_sort(const Array& key, bool desc, int placement, int forder, uint32_t* out_idx, uint32_t* pos_of_row,
             uint32_t* m, uint32_t* bucket_base, hipStream_t s, unsigned long long* key_out,
             unsigned long long key_xor) {
  return run_sort<unsigned long long, 8, ITEMS>(key.data(), key.type, key.has_nulls() ? key.valid_bits() : nullptr,
                                                key.offset, uint32_t(key.length), desc, placement, forder, out_idx,
                                                s, pos_of_row, m, bucket_base, key_out, key_xor);
}

// The rest of a sorted integer key column whose value block the bucket sorts wrote: zeros outside the
// value block [vbase, vbase + m), and the validity bitmap (the value block valid, the rest null).
__global__ void ob_key_rest(unsigned long long* out, uint32_t n, uint32_t vbase, uint32_t m,
                            unsigned long long* out_valid) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in_val = i - vbase < m;
  if (i < n && !in_val) out[i] = 0ull;
  if (out_valid) {
    const unsigned long long bits = __ballot(i < n && in_val);
    if ((threadIdx.x & 63) == 0 && i < n) out_valid[i / 64u] = bits;
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
  switch (w) {
    case 8:
      dispatch_items<unsigned long long, 8>(t.items, vals, type, validity, vbit_off, uint32_t(n), descending,
                                            null_placement, float_order, out_idx, s);
      break;
    case 4:
      dispatch_items<uint32_t, 4>(t.items, vals, type, validity, vbit_off, uint32_t(n), descending, null_placement,
                                  float_order, out_idx, s);
      break;
    case 2:
      dispatch_items<uint32_t, 2>(t.items, vals, type, validity, vbit_off, uint32_t(n), descending, null_placement,
                                  float_order, out_idx, s);
      break;
    default:
      dispatch_items<uint32_t, 1>(t.items, vals, type, validity, vbit_off, uint32_t(n), descending, null_placement,
                                  float_order, out_idx, s);
      break;
  }
}

Array* argsort(const Array& a, bool descending, int null_placement, int float_order) {
  check_length(a.length);
  check_index_rows(a.length, "argsort");
  if (!type_is_primitive(a.type)) fail_unsupported("sort by " + a.format + " is not implemented");
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

bool order_by_columns(const Array& key, bool descending, int null_placement, int float_order,
                      const std::vector<const Array*>& cols, std::vector<Array*>& out) {
  out.clear();
  const int64_t n = key.length;
  if (key.width() != 8 || !type_is_primitive(key.type) || n < (int64_t(1) << 18) || n > int64_t(UINT32_MAX) ||
      capturing(stream()))
    return false;
  for (const Array* c : cols)
    if (c->length != n || !type_is_primitive(c->type)) return false;
  check_index_rows(n, "argsort");
  hipStream_t s = stream();
  DevBuf idx = device_alloc(size_t(n) * 4), pos = device_alloc(size_t(n) * 4), ctr = device_alloc(NXCD * OB_CTR_STRIDE * 4);
  DevBuf bb = device_alloc(RADIX * 4);
  retain(idx);
  retain(pos);
  retain(ctr);
  retain(bb);
  auto* out_idx = static_cast<uint32_t*>(idx->ptr);
  auto* pos_of_row = static_cast<uint32_t*>(pos->ptr);
  auto* bbase = static_cast<uint32_t*>(bb->ptr);
  uint32_t m = 0;
  const int place = null_placement == 1 ? 1 : 0;
  // An int64 / uint64 key column among the columns comes out of the bucket sorts themselves: its
  // mapped keys are a bijection of its values (x ^ 2^63 for int64, then ~ when descending).
  DevBuf key_vals;
  unsigned long long key_xor = 0;
  const bool key_direct = (key.type == T_I64 || key.type == T_U64) &&
                          std::find(cols.begin(), cols.end(), &key) != cols.end();
  if (key_direct) {
    key_vals = device_alloc(size_t(n) * 8);
    key_xor = (key.type == T_I64 ? 0x8000000000000000ull : 0ull) ^ (descending ? ~0ull : 0ull);
  }
  unsigned long long* kout = key_direct ? static_cast<unsigned long long*>(key_vals->ptr) : nullptr;
  bool ok;
  switch (sort_tune(n, 8).items) {
    case 4: ok = ob_sort<4>(key, descending, place, float_order, out_idx, pos_of_row, &m, bbase, s, kout, key_xor); break;
    case 8: ok = ob_sort<8>(key, descending, place, float_order, out_idx, pos_of_row, &m, bbase, s, kout, key_xor); break;
    case 12: ok = ob_sort<12>(key, descending, place, float_order, out_idx, pos_of_row, &m, bbase, s, kout, key_xor); break;
    case 20: ok = ob_sort<20>(key, descending, place, float_order, out_idx, pos_of_row, &m, bbase, s, kout, key_xor); break;
    default: ok = ob_sort<16>(key, descending, place, float_order, out_idx, pos_of_row, &m, bbase, s, kout, key_xor); break;
  }
  if (!ok) return false;
  // The value block: [0, m) with nulls last, [n - m, n) with nulls first.
  const uint32_t vbase = place == 1 ? uint32_t(n) - m : 0u;
  try {
    for (const Array* c : cols) {
      Array* o = new_array(c->format, c->type, n);
      out.push_back(o);
      int64_t nulls = c->validity ? c->null_count : 0;
      if (nulls < 0) nulls = count_nulls(*c);
      o->null_count = nulls;   // a permutation keeps the null count
      if (nulls > 0) o->validity = device_alloc(size_t((n + 63) / 64) * 8);
      if (c == &key && key_direct) {
        o->values = key_vals;
        hipLaunchKernelGGL(ob_key_rest, dim3((uint32_t(n) + 255u) / 256u), dim3(256), 0, s,
                           static_cast<unsigned long long*>(key_vals->ptr), uint32_t(n), vbase, m,
                           nulls > 0 ? static_cast<unsigned long long*>(o->validity->ptr) : nullptr);
        AM_HIP(hipGetLastError());
        continue;
      }
      o->values = device_alloc(size_t(n) * c->width());
      auto* ctrp = static_cast<uint32_t*>(ctr->ptr);
      switch (c->width()) {
        case 8: ob_move_column<8>(*c, nulls > 0, pos_of_row, bbase, out_idx, uint32_t(n), vbase, m, o, ctrp, s); break;
        case 4: ob_move_column<4>(*c, nulls > 0, pos_of_row, bbase, out_idx, uint32_t(n), vbase, m, o, ctrp, s); break;
        case 2: ob_move_column<2>(*c, nulls > 0, pos_of_row, bbase, out_idx, uint32_t(n), vbase, m, o, ctrp, s); break;
        default: ob_move_column<1>(*c, nulls > 0, pos_of_row, bbase, out_idx, uint32_t(n), vbase, m, o, ctrp, s); break;
      }
    }
    sync();
  } catch (...) {
    for (Array* a : out) release_array(a);
    out.clear();
    throw;
  }
  return true;
}

}  // namespace am
