// Top-k by radix select, sorted values, and the multi-key (lexicographic) argsort.
//
//   am_top_k_ex   the first k indices am_argsort_ex2 gives with the same options (reference
//                 include/arrowmetal.h:1342-1351, Sources/ArrowMetalC/ArrowMetalC_Options.swift:188-203,
//                 Kernels/Sort.swift:95-133).
//   am_sort_ex2   the values in am_argsort_ex2's order, same type (ArrowMetalC_Options.swift:177-186,
//                 Kernels/Sort.swift:67-93): -0.0 and NaN payloads come out as they went in.
//   am_lexsort_ex2  stable multi-key argsort, each key with its own direction, null placement and float
//                 order (ArrowMetalC_Options.swift:205-228, Kernels/MultiSort.swift:28-65).
//
// Top-k, this file's method. The rows of the answer are the value rows whose order key is below a
// threshold T, plus the first `need` rows (in row order) whose key equals T. T is found digit by digit
// from the most significant: one pass histograms the top digit of every row (and counts NaN and null
// rows), the rows in the bin that holds the k-th key are copied out as candidates, and the lower digits
// are histogrammed over the candidates only. A last pass over the column writes the chosen rows in row
// order (per-tile counts, one scan, per-tile offsets), and the k chosen values are sorted by the
// ordinary argsort, whose stability keeps equal keys in row order. The work is a few reads of the key
// column plus a sort of k values, against a full sort of n.
#include "sort_common.hpp"

#include <vector>

namespace am {
namespace {

constexpr int TK_ITEMS = 8;
constexpr uint32_t TK_TILE = BLOCK * TK_ITEMS;
constexpr int TK_BINS = 258;   // 256 digit bins of value rows, NaN (ieee order), null

// Histogram of the top digit of every value row, and the NaN and null counts.
template <typename K, int W>
__global__ __launch_bounds__(BLOCK) void tk_hist(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n,
                                                 int type, int forder, int desc, uint32_t* hist) {
  __shared__ uint32_t h[TK_BINS];
  for (int b = threadIdx.x; b < TK_BINS; b += BLOCK) h[b] = 0;
  __syncthreads();
  constexpr int SH = int(sizeof(K)) * 8 - 8;
  const uint32_t stride = gridDim.x * BLOCK;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < n; i += stride) {
    K key;
    const int c = classify<K, W>(vals, valid, voff, i, type, forder, desc, key);
    const uint32_t bin = c == 0 ? uint32_t(key >> SH) : (c == 1 ? 256u : 257u);
    atomicAdd(&h[bin], 1u);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < TK_BINS; b += BLOCK)
    if (h[b]) atomicAdd(&hist[b], h[b]);
}

// The order keys of the value rows whose top digit is `bin`, in any order.
template <typename K, int W>
__global__ __launch_bounds__(BLOCK) void tk_extract(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n,
                                                    int type, int forder, int desc, uint32_t bin, K* cand,
                                                    uint32_t* count) {
  constexpr int SH = int(sizeof(K)) * 8 - 8;
  const int lane = threadIdx.x & 63;
  const uint32_t stride = gridDim.x * BLOCK;
  for (uint32_t base = blockIdx.x * BLOCK; base < n; base += stride) {
    const uint32_t i = base + threadIdx.x;
    K key = 0;
    bool take = false;
    if (i < n) take = classify<K, W>(vals, valid, voff, i, type, forder, desc, key) == 0 && uint32_t(key >> SH) == bin;
    // One atomic per wave for its candidates.
    const unsigned long long bal = __ballot(take);
    uint32_t first = 0;
    if (lane == 0 && bal) first = atomicAdd(count, uint32_t(__popcll(bal)));
    first = __shfl(first, 0, 64);
    if (take) cand[first + uint32_t(__popcll(bal & ((1ull << lane) - 1ull)))] = key;
  }
}

// Histogram of digit `shift` over the candidates whose bits above it equal `prefix`.
template <typename K>
__global__ __launch_bounds__(BLOCK) void tk_hist_cand(const K* cand, uint32_t c, K prefix, int shift, uint32_t* hist) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t stride = gridDim.x * BLOCK;
  const int hi = shift + 8;
  for (uint32_t i = blockIdx.x * BLOCK + threadIdx.x; i < c; i += stride) {
    const K k = cand[i];
    const bool match = hi >= int(sizeof(K)) * 8 ? true : (k >> hi) == (prefix >> hi);
    if (match) atomicAdd(&h[uint32_t(k >> shift) & 255u], 1u);
  }
  __syncthreads();
  if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

// Row j of tile `tile` is tile * TK_TILE + j * BLOCK + threadIdx.x; its flags: key < T, key == T.
template <typename K, int W>
__device__ __forceinline__ void tk_flags(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n, int type,
                                         int forder, int desc, K T, uint32_t tile, bool (&lt)[TK_ITEMS],
                                         bool (&eq)[TK_ITEMS]) {
#pragma unroll
  for (int j = 0; j < TK_ITEMS; j++) {
    const uint32_t i = tile * TK_TILE + uint32_t(j) * BLOCK + threadIdx.x;
    K key = 0;
    const bool v = i < n && classify<K, W>(vals, valid, voff, i, type, forder, desc, key) == 0;
    lt[j] = v && key < T;
    eq[j] = v && key == T;
  }
}

// Per-tile counts of the rows below and at the threshold.
template <typename K, int W>
__global__ __launch_bounds__(BLOCK) void tk_count(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n,
                                                  int type, int forder, int desc, K T, uint32_t* tile_lt,
                                                  uint32_t* tile_eq) {
  __shared__ uint32_t s[2];
  if (threadIdx.x < 2) s[threadIdx.x] = 0;
  __syncthreads();
  bool lt[TK_ITEMS], eq[TK_ITEMS];
  tk_flags<K, W>(vals, valid, voff, n, type, forder, desc, T, blockIdx.x, lt, eq);
  uint32_t a = 0, b = 0;
#pragma unroll
  for (int j = 0; j < TK_ITEMS; j++) {
    a += uint32_t(__popcll(__ballot(lt[j])));
    b += uint32_t(__popcll(__ballot(eq[j])));
  }
  if ((threadIdx.x & 63) == 0) {
    if (a) atomicAdd(&s[0], a);
    if (b) atomicAdd(&s[1], b);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    tile_lt[blockIdx.x] = s[0];
    tile_eq[blockIdx.x] = s[1];
  }
}

// Exclusive scan, in place, of the two per-tile count arrays (one workgroup).
__global__ __launch_bounds__(BLOCK) void tk_scan(uint32_t* tile_lt, uint32_t* tile_eq, uint32_t ntiles) {
  __shared__ uint32_t tmp[WAVES];
  uint32_t carry_lt = 0, carry_eq = 0;
  for (uint32_t base = 0; base < ntiles; base += BLOCK) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t a = i < ntiles ? tile_lt[i] : 0u, b = i < ntiles ? tile_eq[i] : 0u;
    uint32_t ta, tb;
    const uint32_t ea = block_excl_scan(a, tmp, ta);
    const uint32_t eb = block_excl_scan(b, tmp, tb);
    if (i < ntiles) {
      tile_lt[i] = carry_lt + ea;
      tile_eq[i] = carry_eq + eb;
    }
    carry_lt += ta;
    carry_eq += tb;
  }
}

// Writes the chosen rows in row order: rows below T at [tile_lt[tile] + rank), rows at T at
// lt_total + (tile_eq[tile] + rank) while that is below `need`. Also writes their raw values.
template <typename K, int W>
__global__ __launch_bounds__(BLOCK) void tk_write(const void* vals, const uint8_t* valid, int64_t voff, uint32_t n,
                                                  int type, int forder, int desc, K T, const uint32_t* tile_lt,
                                                  const uint32_t* tile_eq, uint32_t lt_total, uint32_t need,
                                                  uint32_t* out_rows, void* out_vals) {
  __shared__ uint32_t wl[TK_ITEMS][WAVES], we[TK_ITEMS][WAVES];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  bool lt[TK_ITEMS], eq[TK_ITEMS];
  tk_flags<K, W>(vals, valid, voff, n, type, forder, desc, T, blockIdx.x, lt, eq);
  uint32_t rl[TK_ITEMS], re[TK_ITEMS];
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < TK_ITEMS; j++) {
    const unsigned long long bl = __ballot(lt[j]), be = __ballot(eq[j]);
    rl[j] = uint32_t(__popcll(bl & below));
    re[j] = uint32_t(__popcll(be & below));
    if (lane == 0) {
      wl[j][w] = uint32_t(__popcll(bl));
      we[j][w] = uint32_t(__popcll(be));
    }
  }
  __syncthreads();
  // Rows are ordered (j, wave, lane): the offset of (j, w) sums every (j', w') before it.
  uint32_t ol = tile_lt[blockIdx.x], oe = tile_eq[blockIdx.x];
#pragma unroll
  for (int j = 0; j < TK_ITEMS; j++) {
    uint32_t pl = ol, pe = oe;
    for (int v = 0; v < WAVES; v++) {
      if (v < w) {
        pl += wl[j][v];
        pe += we[j][v];
      }
      ol += wl[j][v];
      oe += we[j][v];
    }
    const uint32_t i = blockIdx.x * TK_TILE + uint32_t(j) * BLOCK + threadIdx.x;
    uint32_t dst = 0xFFFFFFFFu;
    if (lt[j]) dst = pl + rl[j];
    else if (eq[j] && pe + re[j] < need) dst = lt_total + pe + re[j];
    if (dst != 0xFFFFFFFFu) {
      out_rows[dst] = i;
      if constexpr (W == 8) static_cast<uint64_t*>(out_vals)[dst] = static_cast<const uint64_t*>(vals)[i];
      else if constexpr (W == 4) static_cast<uint32_t*>(out_vals)[dst] = static_cast<const uint32_t*>(vals)[i];
      else if constexpr (W == 2) static_cast<uint16_t*>(out_vals)[dst] = static_cast<const uint16_t*>(vals)[i];
      else static_cast<uint8_t*>(out_vals)[dst] = static_cast<const uint8_t*>(vals)[i];
    }
  }
}

__global__ void tk_compose(const uint32_t* rows, const uint32_t* perm, uint32_t k, uint32_t* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) out[i] = rows[perm[i]];
}

__global__ void tk_zero(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

int tk_grid(uint32_t n) {
  return int(std::max<int64_t>(1, std::min<int64_t>((int64_t(n) + BLOCK - 1) / BLOCK, int64_t(device().cu_count) * 8)));
}

// The first k rows of the argsort, k < the number of value rows, which come first in the order.
template <typename K, int W>
void top_k_select(const Array& a, uint32_t k, bool desc, int forder, const std::vector<uint32_t>& top_hist,
                  uint32_t* out, hipStream_t s) {
  const uint32_t n = uint32_t(a.length);
  const uint8_t* valid = a.has_nulls() ? a.valid_bits() : nullptr;
  const void* vals = a.data();
  constexpr int NDIG = int(sizeof(K));
  // Top digit: the bin holding the k-th key (1-based) and the rows before it.
  uint32_t before = 0, bin = 0;
  for (bin = 0; bin < 256; bin++) {
    if (before + top_hist[bin] >= k) break;
    before += top_hist[bin];
  }
  K prefix = K(bin) << (8 * (NDIG - 1));
  uint32_t rank = k - before;                 // 1-based rank of T inside the bin
  const uint32_t cand_n = top_hist[bin];
  DevBuf cand = device_alloc(size_t(std::max<uint32_t>(cand_n, 1)) * sizeof(K));
  DevBuf hist = device_alloc(256 * 4 + 64);
  auto* hp = static_cast<uint32_t*>(hist->ptr);
  uint32_t* cnt = hp + 256;
  std::vector<uint32_t> h(256);
  {
    ProfScope ps("topk_extract");
    hipLaunchKernelGGL(tk_zero, dim3(2), dim3(256), 0, s, hp, 257u);
    hipLaunchKernelGGL((tk_extract<K, W>), dim3(tk_grid(n)), dim3(BLOCK), 0, s, vals, valid, a.offset, n, a.type,
                       forder, int(desc), bin, static_cast<K*>(cand->ptr), cnt);
    AM_HIP(hipGetLastError());
  }
  for (int d = NDIG - 2; d >= 0; d--) {
    const int shift = 8 * d;
    hipLaunchKernelGGL(tk_zero, dim3(1), dim3(256), 0, s, hp, 256u);
    hipLaunchKernelGGL((tk_hist_cand<K>), dim3(tk_grid(cand_n)), dim3(BLOCK), 0, s,
                       static_cast<const K*>(cand->ptr), cand_n, prefix, shift, hp);
    AM_HIP(hipGetLastError());
    AM_HIP(hipMemcpyAsync(h.data(), hp, 256 * 4, hipMemcpyDeviceToHost, s));
    AM_HIP(hipStreamSynchronize(s));
    uint32_t b = 0, acc = 0;
    for (b = 0; b < 256; b++) {
      if (acc + h[b] >= rank) break;
      acc += h[b];
    }
    rank -= acc;
    prefix |= K(b) << shift;
  }
  const K T = prefix;                          // the k-th smallest order key
  const uint32_t ntiles = (n + TK_TILE - 1) / TK_TILE;
  DevBuf tiles = device_alloc(size_t(ntiles) * 8);
  auto* tl = static_cast<uint32_t*>(tiles->ptr);
  uint32_t* te = tl + ntiles;
  ProfScope ps("topk_select");
  hipLaunchKernelGGL((tk_count<K, W>), dim3(ntiles), dim3(BLOCK), 0, s, vals, valid, a.offset, n, a.type, forder,
                     int(desc), T, tl, te);
  hipLaunchKernelGGL(tk_scan, dim3(1), dim3(BLOCK), 0, s, tl, te, ntiles);
  // Rows strictly below T: k minus the ties taken (rank of T among its equals, 1-based).
  const uint32_t need = rank, lt_total = k - rank;
  DevBuf rows = device_alloc(size_t(k) * 4), sel = device_alloc(size_t(k) * W);
  hipLaunchKernelGGL((tk_write<K, W>), dim3(ntiles), dim3(BLOCK), 0, s, vals, valid, a.offset, n, a.type, forder,
                     int(desc), T, tl, te, lt_total, need, static_cast<uint32_t*>(rows->ptr), sel->ptr);
  AM_HIP(hipGetLastError());
  // The k chosen values in the argsort's order; equal keys keep row order (stable sort, rows written in
  // row order).
  DevBuf perm = device_alloc(size_t(k) * 4);
  argsort_device(sel->ptr, a.type, nullptr, 0, k, desc, 0, forder, static_cast<uint32_t*>(perm->ptr), s);
  hipLaunchKernelGGL(tk_compose, dim3((k + 255) / 256), dim3(256), 0, s, static_cast<const uint32_t*>(rows->ptr),
                     static_cast<const uint32_t*>(perm->ptr), k, out);
  AM_HIP(hipGetLastError());
  AM_HIP(hipStreamSynchronize(s));
}

template <typename K, int W>
bool top_k_try(const Array& a, uint32_t k, bool desc, int placement, int forder, uint32_t* out, hipStream_t s) {
  const uint32_t n = uint32_t(a.length);
  DevBuf hist = device_alloc(TK_BINS * 4);
  auto* hp = static_cast<uint32_t*>(hist->ptr);
  std::vector<uint32_t> h(TK_BINS);
  {
    ProfScope ps("topk_hist");
    hipLaunchKernelGGL(tk_zero, dim3(2), dim3(256), 0, s, hp, uint32_t(TK_BINS));
    hipLaunchKernelGGL((tk_hist<K, W>), dim3(tk_grid(n)), dim3(BLOCK), 0, s, a.data(),
                       a.has_nulls() ? a.valid_bits() : nullptr, a.offset, n, a.type, forder, int(desc), hp);
    AM_HIP(hipGetLastError());
  }
  AM_HIP(hipMemcpyAsync(h.data(), hp, TK_BINS * 4, hipMemcpyDeviceToHost, s));
  AM_HIP(hipStreamSynchronize(s));
  uint64_t values = 0;
  for (int b = 0; b < 256; b++) values += h[b];
  const uint32_t nan = h[256], nulls = h[257];
  // The value rows come first at_end, and at_start when no null or NaN block precedes them.
  const bool values_first = placement == 0 || (nulls == 0 && nan == 0);
  if (!values_first || uint64_t(k) > values) return false;
  std::vector<uint32_t> top(h.begin(), h.begin() + 256);
  top_k_select<K, W>(a, k, desc, forder, top, out, s);
  return true;
}

}  // namespace

Array* top_k(const Array& a, int64_t k, bool largest, int placement, int forder) {
  if (k < 0) fail_invalid("top_k needs k >= 0");
  if (!type_is_primitive(a.type)) fail_unsupported("sort by " + a.format + " is not implemented");
  if (k == 0) {
    Array* out = new_array("I", T_U32, 0);
    out->values = device_alloc(4);
    out->null_count = 0;
    return out;
  }
  check_length(a.length);
  check_index_rows(a.length, "top_k");
  const int64_t kk = std::min<int64_t>(k, a.length);
  Array* out = new_array("I", T_U32, kk);
  try {
    hipStream_t s = stream();
    out->null_count = 0;
    bool done = false;
    if (kk > 0) {
      out->values = device_alloc(size_t(kk) * 4);
      auto* o = static_cast<uint32_t*>(out->values->ptr);
      switch (a.width()) {
        case 8: done = top_k_try<unsigned long long, 8>(a, uint32_t(kk), largest, placement, forder, o, s); break;
        case 4: done = top_k_try<uint32_t, 4>(a, uint32_t(kk), largest, placement, forder, o, s); break;
        case 2: done = top_k_try<uint32_t, 2>(a, uint32_t(kk), largest, placement, forder, o, s); break;
        default: done = top_k_try<uint32_t, 1>(a, uint32_t(kk), largest, placement, forder, o, s); break;
      }
    } else {
      out->values = device_alloc(4);
      done = true;
    }
    if (!done) {
      // The answer reaches into the NaN or null block (or starts there): the first k of the full sort.
      DevBuf all = device_alloc(size_t(a.length) * 4);
      argsort_device(a.data(), a.type, a.has_nulls() ? a.valid_bits() : nullptr, a.offset, a.length, largest,
                     placement, forder, static_cast<uint32_t*>(all->ptr), s);
      AM_HIP(hipMemcpyAsync(out->values->ptr, all->ptr, size_t(kk) * 4, hipMemcpyDeviceToDevice, s));
      AM_HIP(hipStreamSynchronize(s));
    }
  } catch (...) {
    release_array(out);
    throw;
  }
  return out;
}

Array* sort_values(const Array& a, bool descending, int placement, int forder) {
  Array* idx = argsort(a, descending, placement, forder);
  Array* out = nullptr;
  try {
    out = take(a, *idx);
  } catch (...) {
    release_array(idx);
    throw;
  }
  release_array(idx);
  return out;
}

Array* lexsort(Array* const* cols, int64_t count, const int* desc, const int* placement, const int* forder) {
  if (count < 1) fail_invalid("lexsort needs at least one column");
  const int64_t n = cols[0]->length;
  for (int64_t c = 1; c < count; c++)
    if (cols[c]->length != n) fail_length_mismatch(n, cols[c]->length);
  check_index_rows(n, "lexsort");
  auto opt = [](const int* p, int64_t c) { return p ? p[c] : 0; };
  if (count == 1) return argsort(*cols[0], opt(desc, 0) != 0, opt(placement, 0) == 1 ? 1 : 0, opt(forder, 0));
  if (n == 0) {
    Array* out = new_array("I", T_U32, 0);
    out->values = device_alloc(4);
    out->null_count = 0;
    return out;
  }
  // Least significant key first; every later (more significant) key is sorted stably in the order the
  // earlier passes left, so ties fall back to the keys after it (Kernels/MultiSort.swift:3-15).
  Array* perm = nullptr;
  try {
    for (int64_t c = count - 1; c >= 0; c--) {
      Array* keys = perm ? take(*cols[c], *perm) : cols[c];
      Array* idx = nullptr;
      try {
        idx = argsort(*keys, opt(desc, c) != 0, opt(placement, c) == 1 ? 1 : 0, opt(forder, c));
      } catch (...) {
        if (perm) release_array(keys);
        throw;
      }
      if (perm) {
        release_array(keys);
        Array* next = take(*perm, *idx);
        release_array(idx);
        release_array(perm);
        perm = next;
      } else {
        perm = idx;
      }
    }
  } catch (...) {
    if (perm) release_array(perm);
    throw;
  }
  return perm;
}

}  // namespace am
