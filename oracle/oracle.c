/*
 * CPU oracle for the argsort / take / group-by path. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and only
 * as the checker (or the timed CPU baseline); the product library never links or calls it.
 *
 * Each function restates the reference's semantics from the cited files (paths relative to the
 * reference checkout) in plain, obviously-correct C:
 *
 *   or_argsort        Sources/ArrowMetal/Kernels/SortSource.swift:131-190 (key maps),
 *                     Sources/ArrowMetal/Kernels/Sort.swift:32-43, 204-233 (NaN block, placement),
 *                     Tests/ArrowMetalTests/SortOptionsTests.swift:25-38 (the reference permutation).
 *                     The sort itself is an 8-bit-digit stable LSD radix sort, the reference's own
 *                     algorithm (SortSource.swift:3-17), run on the CPU.
 *   or_lexsort        Sources/ArrowMetal/Kernels/MultiSort.swift:28-65 (stable argsort of the least
 *                     significant key first, each more significant key sorted stably in the order the
 *                     earlier passes left; options per key).
 *   or_take           Sources/ArrowMetal/Kernels/Take.swift:41-80, KernelSource.swift:209-227.
 *   or_group_keys     Sources/ArrowMetal/Kernels/GroupByKeys.swift:49-58, 170-236, 387-413
 *                     (ascending key order, null group last), HashTable.swift:290-302 (key
 *                     normalisation: -0.0 == 0.0, one NaN), GroupByKeys.swift:126-149 (lowest row
 *                     per group is the representative).
 *   or_group_agg      ArrowMetalC_GroupByExtra.swift:126-181 (op table), GroupBy.swift:46-182
 *                     (wrapping int64 sums, counts, null where no valid value),
 *                     GroupBySource.swift:28-40 (integer mean = double(sum)/double(count)),
 *                     GroupByExtremaSource.swift:50-55 (float64 min/max on the raw sign-flip key:
 *                     -0.0 < +0.0, NaN skipped), GroupSumExactSource.swift:3-32 and
 *                     GroupSumExact.swift:172-279 (correctly rounded float64 sum/mean and the
 *                     special-value rules).
 *   or_group_by_dense Sources/ArrowMetalC/ArrowMetalC.swift:416-469, GroupBy.swift:4-5
 *                     (dense keys; out-of-range and null keys skipped).
 *   or_unique_rows    Sources/ArrowMetal/Engine/Executor.swift:134-148 (keep-first unique rows).
 *   or_unique_values  Sources/ArrowMetal/Kernels/UniqueOrder.swift:117-125, Kernels/Unique.swift:46-51.
 *
 * Every function here is pinned by tests/test_oracle.py against the reference's literal golden
 * vectors, pyarrow/numpy fixtures under tests/golden, math.fsum / fractions, and brute force.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Element type codes shared with tests/oracle_py.py. */
enum {
    OR_I8 = 0, OR_I16 = 1, OR_I32 = 2, OR_I64 = 3,
    OR_U8 = 4, OR_U16 = 5, OR_U32 = 6, OR_U64 = 7,
    OR_F32 = 8, OR_F64 = 9
};

static int or_width(int t) {
    switch (t) {
    case OR_I8: case OR_U8: return 1;
    case OR_I16: case OR_U16: return 2;
    case OR_I32: case OR_U32: case OR_F32: return 4;
    default: return 8;
    }
}
static int or_is_float(int t) { return t == OR_F32 || t == OR_F64; }
static int or_is_signed(int t) { return t <= OR_I64; }

/* SplitMix64 (Tests/ArrowMetalTests/RadixSortScaleTests.swift:14-24). */
uint64_t or_splitmix64_next(uint64_t* s) {
    *s += 0x9E3779B97F4A7C15ull;
    uint64_t z = *s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
void or_splitmix64_fill(uint64_t seed, uint64_t* out, int64_t n) {
    uint64_t s = seed;
    for (int64_t i = 0; i < n; i++) out[i] = or_splitmix64_next(&s);
}

static int bit_get(const uint8_t* v, int64_t i) { return (v[i >> 3] >> (i & 7)) & 1; }

/* Raw bits of element i, sign- or zero-extended to 64 bits for integers, raw for floats. */
static uint64_t raw64(const void* vals, int t, int64_t i) {
    switch (t) {
    case OR_I8: return (uint64_t)(int64_t)((const int8_t*)vals)[i];
    case OR_I16: return (uint64_t)(int64_t)((const int16_t*)vals)[i];
    case OR_I32: return (uint64_t)(int64_t)((const int32_t*)vals)[i];
    case OR_I64: return (uint64_t)((const int64_t*)vals)[i];
    case OR_U8: return ((const uint8_t*)vals)[i];
    case OR_U16: return ((const uint16_t*)vals)[i];
    case OR_U32: case OR_F32: return ((const uint32_t*)vals)[i];
    default: return ((const uint64_t*)vals)[i];
    }
}

/* ------------------------------------------------------------------ stable LSD radix sort */

typedef struct { uint64_t key; uint32_t row; } kr_t;

/* Stable ascending sort of (key, row) by key over `bytes` low bytes, 8-bit digits, LSD first. */
static void lsd_sort(kr_t* a, int64_t m, int bytes) {
    if (m <= 1) return;
    kr_t* tmp = (kr_t*)malloc((size_t)m * sizeof(kr_t));
    kr_t *src = a, *dst = tmp;
    for (int p = 0; p < bytes; p++) {
        int64_t cnt[257];
        memset(cnt, 0, sizeof cnt);
        int sh = 8 * p;
        for (int64_t i = 0; i < m; i++) cnt[((src[i].key >> sh) & 0xFF) + 1]++;
        int constant = 0;
        for (int d = 0; d < 256; d++) if (cnt[d + 1] == m) constant = 1;
        if (constant) continue;                 /* identity permutation: nothing to move */
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (int64_t i = 0; i < m; i++) dst[cnt[(src[i].key >> sh) & 0xFF]++] = src[i];
        kr_t* t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, (size_t)m * sizeof(kr_t));
    free(tmp);
}

/* ------------------------------------------------------------------ sort keys */

/*
 * The ascending order-preserving key of a non-null element (SortSource.swift:131-190) and whether it
 * belongs in the separate NaN block (ieee only). Narrow integers widen to 32-bit keys.
 * float_order: 0 ieee, 1 total, 2 nan_largest.
 */
static uint64_t sort_key(const void* vals, int t, int64_t i, int forder, int* is_nan) {
    *is_nan = 0;
    uint64_t b = raw64(vals, t, i);
    switch (t) {
    case OR_I8: case OR_I16: case OR_I32: return (uint64_t)(((uint32_t)b) ^ 0x80000000u);
    case OR_U8: case OR_U16: case OR_U32: return (uint32_t)b;
    case OR_I64: return b ^ 0x8000000000000000ull;
    case OR_U64: return b;
    case OR_F32: {
        uint32_t x = (uint32_t)b;
        if (forder != 1) {
            if ((x & 0x7FFFFFFFu) > 0x7F800000u) { *is_nan = (forder == 0); x = 0x7F800001u; }
            else if ((x & 0x7FFFFFFFu) == 0u) x = 0u;
        }
        return (x & 0x80000000u) ? (uint32_t)~x : (x | 0x80000000u);
    }
    default: {
        uint64_t x = b;
        if (forder != 1) {
            if ((x & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) { *is_nan = (forder == 0); x = 0x7FF0000000000001ull; }
            else if ((x & 0x7FFFFFFFFFFFFFFFull) == 0ull) x = 0ull;
        }
        return (x & 0x8000000000000000ull) ? ~x : (x | 0x8000000000000000ull);
    }
    }
}

/*
 * Stable argsort (Arrow array_sort_indices with the reference's options).
 * validity: Arrow bitmap or NULL; bit i of the bitmap describes element i (the caller applies any
 * Arrow offset beforehand). Output: n uint32 row numbers. Returns 0.
 */
int or_argsort(const void* vals, int t, const uint8_t* validity, int64_t n,
               int descending, int null_placement, int float_order, uint32_t* out) {
    int kbytes = or_width(t) <= 4 ? 4 : 8;
    uint64_t kmask = kbytes == 4 ? 0xFFFFFFFFull : ~0ull;
    kr_t* v = (kr_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(kr_t));
    uint32_t* nan_rows = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * 4);
    uint32_t* null_rows = (uint32_t*)malloc((size_t)(n > 0 ? n : 1) * 4);
    int64_t m = 0, q = 0, z = 0;
    for (int64_t i = 0; i < n; i++) {
        if (validity && !bit_get(validity, i)) { null_rows[z++] = (uint32_t)i; continue; }
        int is_nan;
        uint64_t k = sort_key(vals, t, i, or_is_float(t) ? float_order : 1, &is_nan);
        if (is_nan) { nan_rows[q++] = (uint32_t)i; continue; }
        if (descending) k = ~k & kmask;
        v[m].key = k; v[m].row = (uint32_t)i; m++;
    }
    lsd_sort(v, m, kbytes);
    int64_t o = 0;
    if (null_placement == 1) {
        for (int64_t i = 0; i < z; i++) out[o++] = null_rows[i];
        for (int64_t i = 0; i < q; i++) out[o++] = nan_rows[i];
        for (int64_t i = 0; i < m; i++) out[o++] = v[i].row;
    } else {
        for (int64_t i = 0; i < m; i++) out[o++] = v[i].row;
        for (int64_t i = 0; i < q; i++) out[o++] = nan_rows[i];
        for (int64_t i = 0; i < z; i++) out[o++] = null_rows[i];
    }
    free(v); free(nan_rows); free(null_rows);
    return 0;
}

/*
 * Multi-key stable argsort (MultiSort.swift:28-65). cols[c] / validity[c] / types[c] describe key c
 * (validity bitmaps or NULL), desc / placement / forder hold one option per key. Output: n rows.
 */
int or_lexsort(int ncols, const int* types, const void* const* cols, const uint8_t* const* validity,
               int64_t n, const int* desc, const int* placement, const int* forder, uint32_t* out) {
    if (n == 0) return 0;
    uint32_t* perm = (uint32_t*)malloc((size_t)n * 4);
    uint32_t* idx = (uint32_t*)malloc((size_t)n * 4);
    uint8_t* gathered = (uint8_t*)malloc((size_t)n * 8);
    uint8_t* gvalid = (uint8_t*)malloc((size_t)(n + 7) / 8);
    for (int64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
    for (int c = ncols - 1; c >= 0; c--) {
        int w = or_width(types[c]);
        /* the key column in the order the passes so far left (perm), with its validity */
        memset(gvalid, 0, (size_t)(n + 7) / 8);
        for (int64_t i = 0; i < n; i++) {
            memcpy(gathered + i * w, (const char*)cols[c] + (int64_t)perm[i] * w, (size_t)w);
            if (!validity[c] || bit_get(validity[c], perm[i])) gvalid[i >> 3] |= (uint8_t)(1u << (i & 7));
        }
        or_argsort(gathered, types[c], validity[c] ? gvalid : NULL, n, desc[c], placement[c], forder[c], idx);
        for (int64_t i = 0; i < n; i++) out[i] = perm[idx[i]];
        memcpy(perm, out, (size_t)n * 4);
    }
    free(perm); free(idx); free(gathered); free(gvalid);
    return 0;
}

/*
 * take: out[i] = vals[idx[i]]; a null index (or a null source element) gives a null output.
 * idx_type: OR_I32, OR_I64 or OR_U32. Returns 0, or 1 when an index is out of range [0, len).
 * out_valid receives one byte per output element (1 valid, 0 null).
 */
int or_take(const void* vals, int width, const uint8_t* validity, int64_t len,
            const void* idx, int idx_type, const uint8_t* idx_validity, int64_t n,
            void* out, uint8_t* out_valid) {
    for (int64_t i = 0; i < n; i++) {
        if (idx_validity && !bit_get(idx_validity, i)) { memset((char*)out + i * width, 0, (size_t)width); out_valid[i] = 0; continue; }
        int64_t j = idx_type == OR_I32 ? (int64_t)((const int32_t*)idx)[i]
                  : idx_type == OR_U32 ? (int64_t)((const uint32_t*)idx)[i]
                  : ((const int64_t*)idx)[i];
        if (j < 0 || j >= len) return 1;
        memcpy((char*)out + i * width, (const char*)vals + j * width, (size_t)width);
        out_valid[i] = (validity && !bit_get(validity, j)) ? 0 : 1;
    }
    return 0;
}

/* ------------------------------------------------------------------ group keys */

/*
 * The grouping key of a non-null element: integers widen (sign-extended when signed), floats are
 * normalised so that bit equality is Arrow value equality (HashTable.swift:290-302). The returned
 * `order` key sorts the groups ascending by value: signed order for signed integers, unsigned for
 * unsigned ones, and the float value order with the one NaN after +inf.
 */
static uint64_t group_order_key(const void* vals, int t, int64_t i) {
    uint64_t b = raw64(vals, t, i);
    if (t == OR_F64) {
        if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) b = 0x7FF8000000000000ull;
        else if ((b & 0x7FFFFFFFFFFFFFFFull) == 0ull) b = 0ull;
        return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
    }
    if (t == OR_F32) {
        uint32_t x = (uint32_t)b;
        if ((x & 0x7FFFFFFFu) > 0x7F800000u) x = 0x7FC00000u;
        else if ((x & 0x7FFFFFFFu) == 0u) x = 0u;
        return (x & 0x80000000u) ? (uint32_t)~x : (x | 0x80000000u);
    }
    if (or_is_signed(t)) return b ^ 0x8000000000000000ull;
    return b;
}

/*
 * Dense group ids for `ncols` key columns (all of length n), groups ordered lexicographically in
 * column order, each column ascending with its null last (GroupByKeys.swift:49-58; one column is the
 * common case). types[c], vals[c], valid[c] (bitmap or NULL) describe column c.
 * Writes ids[n] (int32) and rep_rows[K] (the lowest row of each group); returns K.
 */
int64_t or_group_keys(int ncols, const int* types, const void* const* vals, const uint8_t* const* valid,
                      int64_t n, int32_t* ids, uint32_t* rep_rows) {
    if (n == 0) return 0;
    /* Lexicographic order = LSD over columns: stable sort by the last column first. */
    kr_t* a = (kr_t*)malloc((size_t)n * sizeof(kr_t));
    uint32_t* perm = (uint32_t*)malloc((size_t)n * 4);
    for (int64_t i = 0; i < n; i++) perm[i] = (uint32_t)i;
    for (int c = ncols - 1; c >= 0; c--) {
        /* Null sorts after every value: rows are split, values sorted, nulls appended stably. */
        int64_t m = 0, z = 0;
        uint32_t* nulls = (uint32_t*)malloc((size_t)n * 4);
        for (int64_t j = 0; j < n; j++) {
            uint32_t r = perm[j];
            if (valid[c] && !bit_get(valid[c], r)) { nulls[z++] = r; continue; }
            a[m].key = group_order_key(vals[c], types[c], r); a[m].row = r; m++;
        }
        lsd_sort(a, m, 8);
        for (int64_t j = 0; j < m; j++) perm[j] = a[j].row;
        for (int64_t j = 0; j < z; j++) perm[m + j] = nulls[j];
        free(nulls);
    }
    /* Walk the sorted rows; a new group starts where any column's (null, key) differs. */
    int64_t K = 0;
    for (int64_t j = 0; j < n; j++) {
        uint32_t r = perm[j];
        int same = j > 0;
        if (same) {
            uint32_t p = perm[j - 1];
            for (int c = 0; c < ncols && same; c++) {
                int nr = valid[c] && !bit_get(valid[c], r), np = valid[c] && !bit_get(valid[c], p);
                if (nr != np) same = 0;
                else if (!nr && group_order_key(vals[c], types[c], r) != group_order_key(vals[c], types[c], p)) same = 0;
            }
        }
        if (!same) { rep_rows[K] = r; K++; }   /* stable: the first row of a run is the lowest */
        ids[r] = (int32_t)(K - 1);
    }
    free(a); free(perm);
    return K;
}

/* ------------------------------------------------------------------ exact float64 sums */

#define XL 36   /* 2,304-bit two's-complement fixed point, bit 0 weighs 2^-1074 */

typedef struct {
    uint64_t limb[XL];
    int64_t count;
    int nonzero, pos_zero, pos_inf, neg_inf;
    int64_t nan_row;
    uint64_t nan_bits;
} xsum_t;

static void xsum_init(xsum_t* s) { memset(s, 0, sizeof *s); s->nan_row = -1; }

static void xsum_add(xsum_t* s, uint64_t bits, int64_t row) {
    s->count++;
    uint32_t e = (uint32_t)((bits >> 52) & 0x7FF);
    uint64_t mag = bits & 0x7FFFFFFFFFFFFFFFull;
    if (e == 0x7FF) {
        if (mag == 0x7FF0000000000000ull) { if (bits >> 63) s->neg_inf = 1; else s->pos_inf = 1; }
        else if (s->nan_row < 0 || row < s->nan_row) { s->nan_row = row; s->nan_bits = bits; }
        return;
    }
    if (mag == 0) { if (bits == 0) s->pos_zero = 1; return; }
    s->nonzero = 1;
    uint64_t m = (bits & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0);
    uint32_t off = (e ? e : 1) - 1;            /* bit position of the mantissa's bit 0 */
    uint32_t li = off / 64, sh = off % 64;
    uint64_t lo = m << sh, hi = sh ? (m >> (64 - sh)) : 0;
    if (!(bits >> 63)) {
        uint64_t c = 0;
        for (uint32_t j = li; j < XL; j++) {
            uint64_t add = (j == li) ? lo : (j == li + 1 ? hi : 0);
            uint64_t t = s->limb[j] + add; uint64_t c1 = t < add;
            uint64_t u = t + c; uint64_t c2 = u < c;
            s->limb[j] = u; c = c1 | c2;
            if (j > li + 1 && !c) break;
        }
    } else {
        uint64_t b = 0;
        for (uint32_t j = li; j < XL; j++) {
            uint64_t sub = (j == li) ? lo : (j == li + 1 ? hi : 0);
            uint64_t t = s->limb[j] - sub; uint64_t b1 = s->limb[j] < sub;
            uint64_t u = t - b; uint64_t b2 = t < b;
            s->limb[j] = u; b = b1 | b2;
            if (j > li + 1 && !b) break;
        }
    }
}

/* Round-to-nearest-even of (w + sticky) * 2^(lsb - 1074) to binary64 bits, w != 0, nw limbs. */
static uint64_t round_mag(const uint64_t* w, int nw, int lsb, int sticky, int neg) {
    int top = -1;
    for (int j = nw - 1; j >= 0; j--) if (w[j]) { top = j * 64 + 63 - __builtin_clzll(w[j]); break; }
    uint64_t sgn = neg ? 0x8000000000000000ull : 0;
    int P = top + lsb;                          /* position of the top bit in 2^-1074 units */
    if (P >= 2098) return sgn | 0x7FF0000000000000ull;   /* >= 2^1024 */
    int s = (P >= 52) ? top - 52 : -lsb;        /* low bits that do not fit (normal / subnormal) */
    uint64_t r;
    if (s <= 0) {
        r = w[0] << (-s);                       /* exact: every bit fits */
    } else {
        /* r = w >> s, the round bit is bit s-1, sticky is anything below it. */
        int j = s / 64, b = s % 64;
        r = (j < nw ? w[j] >> b : 0);
        if (b && j + 1 < nw) r |= w[j + 1] << (64 - b);
        int rb = (int)((s - 1) / 64 < nw ? (w[(s - 1) / 64] >> ((s - 1) % 64)) & 1 : 0);
        int st = sticky;
        for (int t = 0; t < (s - 1) / 64 && !st; t++) if (w[t]) st = 1;
        int bb = (s - 1) % 64;
        if (!st && bb && (w[(s - 1) / 64] & ((1ull << bb) - 1))) st = 1;
        if (rb && (st || (r & 1))) r++;
    }
    uint64_t bitsv = (P >= 52) ? r + ((uint64_t)(P - 52) << 52) : r;
    if (bitsv >= 0x7FF0000000000000ull) bitsv = 0x7FF0000000000000ull;
    return sgn | bitsv;
}

/* The correctly rounded sum (mean = 0) or mean (mean = 1) with the reference's special-value rules. */
static uint64_t xsum_value(const xsum_t* s, int mean) {
    if (s->nan_row >= 0) return s->nan_bits | (1ull << 51);            /* first NaN in row order, quiet */
    if (s->pos_inf && s->neg_inf) return 0x7FF8000000000000ull;
    if (s->pos_inf) return 0x7FF0000000000000ull;
    if (s->neg_inf) return 0xFFF0000000000000ull;
    uint64_t mag[XL];
    int neg = (int)(s->limb[XL - 1] >> 63);
    uint64_t c = 1;
    for (int j = 0; j < XL; j++) {
        if (neg) { uint64_t x = ~s->limb[j] + c; c = (c && x == 0); mag[j] = x; }
        else mag[j] = s->limb[j];
    }
    int zero = 1;
    for (int j = 0; j < XL; j++) if (mag[j]) zero = 0;
    if (zero) return (s->nonzero || s->pos_zero) ? 0ull : 0x8000000000000000ull;
    if (!mean) return round_mag(mag, XL, 0, 0, neg);
    /* (mag << 128) / count by long division over 32-bit digits; the remainder is sticky. */
    uint32_t d[2 * XL + 4], q[2 * XL + 4];
    int nd = 2 * XL + 4;
    memset(d, 0, sizeof d);
    for (int j = 0; j < XL; j++) { d[4 + 2 * j] = (uint32_t)mag[j]; d[4 + 2 * j + 1] = (uint32_t)(mag[j] >> 32); }
    uint64_t rem = 0, cnt = (uint64_t)s->count;
    for (int j = nd - 1; j >= 0; j--) {
        unsigned __int128 cur = ((unsigned __int128)rem << 32) | d[j];
        q[j] = (uint32_t)(cur / cnt);
        rem = (uint64_t)(cur % cnt);
    }
    uint64_t qw[XL + 2];
    for (int j = 0; j < XL + 2; j++) qw[j] = (uint64_t)q[2 * j] | ((uint64_t)q[2 * j + 1] << 32);
    return round_mag(qw, XL + 2, -128, rem != 0, neg);
}

/* ------------------------------------------------------------------ aggregates */

/*
 * One grouped aggregate over dense ids (am_group_agg_ex's op table, ArrowMetalC_GroupByExtra.swift).
 * ids[i] in [0, K) for every row; a row whose id is negative or >= K is skipped (the dense-key form).
 * vals/t/valid describe the value column (vals may be NULL for op 1).
 * op: 0 sum, 1 count_all, 2 count, 3 mean, 4 min, 5 max.
 * out_bits[K]: the result's bits (int64 / uint64 / double bits, or the min/max element widened to
 * 64 bits: sign-extended for signed integers); out_valid[K]: 1 where the result is not null.
 * Returns 0, or 2 for an op this oracle does not define for the type.
 */
int or_group_agg(const int32_t* ids, int64_t n, int64_t K, const void* vals, int t,
                 const uint8_t* valid, int op, uint64_t* out_bits, uint8_t* out_valid) {
    int64_t* cnt = (int64_t*)calloc((size_t)(K > 0 ? K : 1), 8);
    if (op == 1) {
        for (int64_t i = 0; i < n; i++) { int64_t k = ids[i]; if (k >= 0 && k < K) cnt[k]++; }
        for (int64_t k = 0; k < K; k++) { out_bits[k] = (uint64_t)cnt[k]; out_valid[k] = 1; }
        free(cnt); return 0;
    }
    if (op == 2) {
        for (int64_t i = 0; i < n; i++) {
            int64_t k = ids[i];
            if (k < 0 || k >= K) continue;
            if (valid && !bit_get(valid, i)) continue;
            cnt[k]++;
        }
        for (int64_t k = 0; k < K; k++) { out_bits[k] = (uint64_t)cnt[k]; out_valid[k] = 1; }
        free(cnt); return 0;
    }
    if (or_is_float(t) && (op == 0 || op == 3)) {
        if (t != OR_F64) { free(cnt); return 2; }
        xsum_t* xs = (xsum_t*)malloc((size_t)(K > 0 ? K : 1) * sizeof(xsum_t));
        for (int64_t k = 0; k < K; k++) xsum_init(&xs[k]);
        for (int64_t i = 0; i < n; i++) {
            int64_t k = ids[i];
            if (k < 0 || k >= K || (valid && !bit_get(valid, i))) continue;
            xsum_add(&xs[k], ((const uint64_t*)vals)[i], i);
        }
        for (int64_t k = 0; k < K; k++) {
            out_valid[k] = xs[k].count > 0;
            out_bits[k] = xs[k].count > 0 ? xsum_value(&xs[k], op == 3) : 0;
        }
        free(xs); free(cnt); return 0;
    }
    if (op == 0 || op == 3) {
        uint64_t* acc = (uint64_t*)calloc((size_t)(K > 0 ? K : 1), 8);
        for (int64_t i = 0; i < n; i++) {
            int64_t k = ids[i];
            if (k < 0 || k >= K || (valid && !bit_get(valid, i))) continue;
            acc[k] += raw64(vals, t, i);           /* wraps modulo 2^64 */
            cnt[k]++;
        }
        for (int64_t k = 0; k < K; k++) {
            out_valid[k] = cnt[k] > 0;
            if (op == 0) out_bits[k] = cnt[k] > 0 ? acc[k] : 0;
            else {
                double s = or_is_signed(t) ? (double)(int64_t)acc[k] : (double)acc[k];
                double m = cnt[k] > 0 ? s / (double)cnt[k] : 0.0;
                memcpy(&out_bits[k], &m, 8);
                if (!cnt[k]) out_bits[k] = 0;
            }
        }
        free(acc); free(cnt); return 0;
    }
    if (op == 4 || op == 5) {
        /* An order-preserving unsigned key per value; NaN is skipped for floats. */
        uint64_t* best = (uint64_t*)malloc((size_t)(K > 0 ? K : 1) * 8);
        for (int64_t i = 0; i < n; i++) {
            int64_t k = ids[i];
            if (k < 0 || k >= K || (valid && !bit_get(valid, i))) continue;
            uint64_t b = raw64(vals, t, i), key;
            if (t == OR_F64) {
                if ((b & 0x7FFFFFFFFFFFFFFFull) > 0x7FF0000000000000ull) continue;
                key = (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
            } else if (t == OR_F32) {
                uint32_t x = (uint32_t)b;
                if ((x & 0x7FFFFFFFu) > 0x7F800000u) continue;
                key = (x & 0x80000000u) ? (uint32_t)~x : (x | 0x80000000u);
            } else if (or_is_signed(t)) key = b ^ 0x8000000000000000ull;
            else key = b;
            if (!cnt[k] || (op == 4 ? key < best[k] : key > best[k])) best[k] = key;
            cnt[k]++;
        }
        for (int64_t k = 0; k < K; k++) {
            out_valid[k] = cnt[k] > 0;
            if (!cnt[k]) { out_bits[k] = 0; continue; }
            uint64_t key = best[k];
            if (t == OR_F64) out_bits[k] = (key & 0x8000000000000000ull) ? (key & 0x7FFFFFFFFFFFFFFFull) : ~key;
            else if (t == OR_F32) { uint32_t u = (uint32_t)key; out_bits[k] = (u & 0x80000000u) ? (u & 0x7FFFFFFFu) : (uint32_t)~u; }
            else if (or_is_signed(t)) out_bits[k] = key ^ 0x8000000000000000ull;
            else out_bits[k] = key;
        }
        free(best); free(cnt); return 0;
    }
    free(cnt);
    return 2;
}

/*
 * am_group_by over dense keys (ArrowMetalC.swift:416-469): keys of type OR_I32 / OR_I64 / OR_U32,
 * null or out-of-range keys skipped. agg: 0 sum, 1 count rows, 2 min, 3 max, 4 mean, 5 count values.
 */
int or_group_by_dense(const void* keys, int kt, const uint8_t* kvalid, int64_t n, int64_t K,
                      const void* vals, int t, const uint8_t* valid, int agg,
                      uint64_t* out_bits, uint8_t* out_valid) {
    int32_t* ids = (int32_t*)malloc((size_t)(n > 0 ? n : 1) * 4);
    for (int64_t i = 0; i < n; i++) {
        int64_t k = kt == OR_I32 ? (int64_t)((const int32_t*)keys)[i]
                  : kt == OR_U32 ? (int64_t)((const uint32_t*)keys)[i] : ((const int64_t*)keys)[i];
        if (kvalid && !bit_get(kvalid, i)) k = -1;
        ids[i] = (k < 0 || k >= K) ? -1 : (int32_t)k;
    }
    static const int map[6] = {0, 1, 4, 5, 3, 2};
    int rc = or_group_agg(ids, n, K, vals, t, valid, map[agg], out_bits, out_valid);
    free(ids);
    return rc;
}

/*
 * Keep-first unique rows over `ncols` key columns (Executor.swift:134-148): the lowest row of each
 * group, ascending. Returns the number of rows written to out_rows.
 */
int64_t or_unique_rows(int ncols, const int* types, const void* const* vals, const uint8_t* const* valid,
                       int64_t n, uint32_t* out_rows) {
    if (n == 0) return 0;
    int32_t* ids = (int32_t*)malloc((size_t)n * 4);
    uint32_t* rep = (uint32_t*)malloc((size_t)n * 4);
    int64_t K = or_group_keys(ncols, types, vals, valid, n, ids, rep);
    uint8_t* seen = (uint8_t*)calloc((size_t)K, 1);
    int64_t o = 0;
    for (int64_t i = 0; i < n; i++) if (!seen[ids[i]]) { seen[ids[i]] = 1; out_rows[o++] = (uint32_t)i; }
    free(ids); free(rep); free(seen);
    return o;
}

/*
 * Arrow `unique` of one column (ArrowMetalC_Options.swift:284-297): the rows whose values form the
 * result, to be gathered. order 0 (Kernels/UniqueOrder.swift:117-125): the first row of every
 * distinct value in row order, the first null row included. order 1 (Kernels/Unique.swift:46-51):
 * one row per distinct non-null value, ascending by value (the group order), nulls dropped.
 * Returns the number of rows written.
 */
int64_t or_unique_values(int type, const void* vals, const uint8_t* valid, int64_t n, int order,
                         uint32_t* out_rows) {
    if (n == 0) return 0;
    const void* const v1[1] = {vals};
    const uint8_t* const b1[1] = {valid};
    if (order != 1) return or_unique_rows(1, &type, v1, b1, n, out_rows);
    int32_t* ids = (int32_t*)malloc((size_t)n * 4);
    uint32_t* rep = (uint32_t*)malloc((size_t)n * 4);
    int64_t K = or_group_keys(1, &type, v1, b1, n, ids, rep);
    int has_null = 0;
    for (int64_t i = 0; i < n && valid; i++) if (!((valid[i >> 3] >> (i & 7)) & 1)) { has_null = 1; break; }
    int64_t k = K - has_null;
    for (int64_t g = 0; g < k; g++) out_rows[g] = rep[g];
    free(ids); free(rep);
    return k;
}
