// Streaming group-by over a producer of record batches: am_stream_from_c_stream, am_stream_group_by and
// the stream result accessors.
//
// Reference: include/arrowmetal.h:1515-1600 (the stream handle, the aggregate op codes 0 sum, 1 count,
// 2 min, 3 max, 4 mean, the result accessors), Sources/ArrowMetalC/ArrowMetalC_Stream.swift:53-130,
// 259-281, 419-496 (the C conventions), Sources/ArrowMetal/Stream/StreamingExecutor.swift:508-1420 (what
// the group-by terminal computes: groups in ascending key order with the null key last; sum null for a
// group without a value; count of non-null values, or of rows for a NULL column; mean = sum / count;
// the dense form keeps one integer key inside [0, dense_key_count) and returns it as int32, dropping
// keys no row landed on).
//
// How it runs here: every batch is imported into HBM and reduced to its groups' partial states (the
// hash group-by and its aggregates: keys, sums, counts, minima, maxima). The partials of many batches
// are concatenated and, once they hold more rows than the running answer can justify, grouped again by
// key with each state combined by its own rule (sums and counts added, minima and maxima taken again).
// Memory is one batch plus the partials, which shrink back to the running answer at each fold. Integer
// sums, counts, minima and maxima are exact in any batch order. A float sum is the correctly rounded
// sum of the batches' correctly rounded sums (the reference adds the batches' float sums in float64
// too, so neither is the exact sum of all rows); variance, stddev and count_distinct_approx are not
// taken.
#include <algorithm>
#include <chrono>
#include <map>
#include <memory>

#include "am_internal.hpp"

namespace am {

namespace {

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------- device helpers

__global__ void sc_expand_bits(const uint8_t* bits, int64_t off, int64_t n, uint8_t* out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = bits ? uint8_t((load_global(bits + ((off + i) >> 3)) >> ((off + i) & 7)) & 1u) : uint8_t(1);
}

// Bytes -> bitmap, one ballot per 64 rows, and the count of zero bytes.
__global__ void sc_pack(const uint8_t* bytes, int64_t n, unsigned long long* words, unsigned long long* nulls) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool v = i < n && bytes[i] != 0;
  const unsigned long long bal = __ballot(v);
  if (lane == 0 && i < n) {
    words[i >> 6] = bal;
    const int64_t rows = std::min<int64_t>(64, n - i);
    atomicAdd(nulls, (unsigned long long)(rows - __popcll(bal)));
  }
}

__global__ void sc_zero(unsigned long long* p) { *p = 0ull; }

// out = sum / count where count > 0 (integer sums converted as signed or unsigned), validity bytes.
__global__ void sc_mean(const void* sum, int stype, const unsigned long long* count, int64_t n, double* out,
                        uint8_t* vb) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned long long c = count[i];
  double s;
  if (stype == T_F64) s = static_cast<const double*>(sum)[i];
  else if (stype == T_U64) s = double(static_cast<const unsigned long long*>(sum)[i]);
  else s = double(static_cast<const long long*>(sum)[i]);
  out[i] = c ? s / double(c) : 0.0;
  vb[i] = c != 0;
}

__global__ void sc_i64_to_i32(const long long* in, int64_t n, int32_t* out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = int32_t(in[i]);
}

inline uint32_t blocks_for(int64_t n) { return uint32_t(std::max<int64_t>((n + 255) / 256, 1)); }

// Packs validity bytes into a's bitmap (and its null count), dropping the bitmap when nothing is null.
void set_validity(Array* a, const uint8_t* bytes, int64_t n) {
  hipStream_t s = stream();
  a->validity = device_alloc(size_t((std::max<int64_t>(n, 1) + 63) / 64) * 8);
  DevBuf nz = device_alloc(8);
  hipLaunchKernelGGL(sc_zero, dim3(1), dim3(1), 0, s, static_cast<unsigned long long*>(nz->ptr));
  hipLaunchKernelGGL(sc_pack, dim3(blocks_for(n)), dim3(256), 0, s, bytes, n,
                     static_cast<unsigned long long*>(a->validity->ptr), static_cast<unsigned long long*>(nz->ptr));
  AM_HIP(hipGetLastError());
  const unsigned long long nulls = read_back(static_cast<const unsigned long long*>(nz->ptr));
  a->null_count = int64_t(nulls);
  if (nulls == 0) a->validity.reset();
}

// One array of the parts' rows in order (all of one type).
Array* concat(const std::vector<Array*>& parts) {
  int64_t total = 0;
  bool nulls = false;
  for (Array* p : parts) {
    total += p->length;
    nulls = nulls || p->has_nulls();
  }
  const Array& f = *parts.front();
  const int w = f.width();
  Array* out = new_array(f.format, f.type, total);
  try {
    hipStream_t s = stream();
    out->values = device_alloc(size_t(std::max<int64_t>(total, 1)) * w);
    out->null_count = 0;
    DevBuf bytes;
    if (nulls) bytes = device_alloc(size_t(std::max<int64_t>(total, 1)));
    int64_t at = 0;
    for (Array* p : parts) {
      if (p->length > 0) {
        AM_HIP(hipMemcpyAsync(static_cast<char*>(out->values->ptr) + at * w, p->data(), size_t(p->length) * w,
                              hipMemcpyDeviceToDevice, s));
        if (nulls) {
          hipLaunchKernelGGL(sc_expand_bits, dim3(blocks_for(p->length)), dim3(256), 0, s,
                             p->has_nulls() ? p->valid_bits() : nullptr, p->offset, p->length,
                             static_cast<uint8_t*>(bytes->ptr) + at);
          AM_HIP(hipGetLastError());
        }
      }
      at += p->length;
    }
    if (nulls) set_validity(out, static_cast<const uint8_t*>(bytes->ptr), total);
    else AM_HIP(hipStreamSynchronize(s));
  } catch (...) {
    release_array(out);
    throw;
  }
  return out;
}

// ---------------------------------------------------------------- the stream

enum StreamOp { SO_SUM = 0, SO_COUNT = 1, SO_MIN = 2, SO_MAX = 3, SO_MEAN = 4 };

struct AggSpec {
  int op;
  int col = -1;   // index among the needed columns, -1: count of rows
  std::string name;
};

// The partial states of one aggregate: for sum and mean a sum column and a count column, for count a
// count column, for min / max the extreme column.
struct Partial {
  std::vector<Array*> keys;                 // one per key column
  std::vector<std::vector<Array*>> state;   // [agg][0..1]
  int64_t rows() const { return keys.empty() ? 0 : keys[0]->length; }
};

void release_partial(Partial& p) {
  for (Array* a : p.keys) release_array(a);
  for (auto& v : p.state)
    for (Array* a : v) release_array(a);
  p.keys.clear();
  p.state.clear();
}

}  // namespace

struct Stream {
  ArrowArrayStream src{};
  ArrowSchema schema{};
  bool consumed = false;
  std::vector<std::string> names;
  ~Stream() {
    if (schema.release) schema.release(&schema);
    if (src.release) src.release(&src);
  }
};

struct StreamResult {
  std::vector<std::string> names;
  std::vector<Array*> cols;
  int64_t rows_out = 0, batches = 0, rows = 0, bytes = 0;
  double wall = 0, read = 0, gpu = 0, merge = 0;
  ~StreamResult() {
    for (Array* a : cols) release_array(a);
  }
};

Stream* stream_from_c(ArrowArrayStream* in) {
  if (!in->release) fail_invalid("am_stream_from_c_stream: the stream has already been released");
  auto s = std::make_unique<Stream>();
  s->src = *in;
  in->release = nullptr;   // moved
  if (s->src.get_schema(&s->src, &s->schema) != 0) {
    const char* e = s->src.get_last_error ? s->src.get_last_error(&s->src) : nullptr;
    fail_invalid(std::string("am_stream_from_c_stream: get_schema failed: ") + (e ? e : "unknown error"));
  }
  if (!s->schema.format || std::string(s->schema.format) != "+s")
    fail_unsupported("am_stream_from_c_stream: a stream of record batches (struct \"+s\") is needed");
  for (int64_t i = 0; i < s->schema.n_children; i++)
    s->names.push_back(s->schema.children[i]->name ? s->schema.children[i]->name : "");
  return s.release();
}

void stream_release(Stream* s) { delete s; }

namespace {

// The partial states of one batch's needed columns.
Partial batch_partial(const std::vector<Array*>& keys, const std::vector<Array*>& cols, const std::vector<AggSpec>& aggs) {
  Partial p;
  GroupBy* g = group_by_keys(const_cast<Array* const*>(keys.data()), int64_t(keys.size()));
  try {
    for (size_t i = 0; i < keys.size(); i++) p.keys.push_back(group_by_keys_result(g, int64_t(i)));
    for (const AggSpec& a : aggs) {
      std::vector<Array*> st;
      const Array* v = a.col >= 0 ? cols[size_t(a.col)] : nullptr;
      switch (a.op) {
        case SO_SUM:
        case SO_MEAN: {
          Array* sum = group_agg(g, v, 0, 0.0);
          // Arrow's sum of an unsigned column is uint64; the wrapped 64-bit total has the same bits.
          if (sum->type == T_I64 && type_is_int(v->type) && !type_is_signed(v->type)) {
            sum->format = "L";
            sum->type = T_U64;
          }
          st.push_back(sum);
          st.push_back(group_agg(g, v, 2, 0.0));
          break;
        }
        case SO_COUNT:
          st.push_back(group_agg(g, v, v ? 2 : 1, 0.0));
          break;
        default:
          st.push_back(group_agg(g, v, a.op == SO_MIN ? 4 : 5, 0.0));
          break;
      }
      p.state.push_back(std::move(st));
    }
  } catch (...) {
    group_by_release(g);
    release_partial(p);
    throw;
  }
  group_by_release(g);
  return p;
}

// The partials of several batches as one: concatenated, grouped by key again, states combined.
Partial fold(std::vector<Partial>& parts, const std::vector<AggSpec>& aggs) {
  const size_t nk = parts.front().keys.size();
  std::vector<Array*> keys, tmp;
  Partial out;
  auto cat = [&](auto pick) {
    std::vector<Array*> v;
    for (Partial& p : parts) v.push_back(pick(p));
    return concat(v);
  };
  try {
    for (size_t i = 0; i < nk; i++) keys.push_back(cat([&](Partial& p) { return p.keys[i]; }));
    GroupBy* g = group_by_keys(keys.data(), int64_t(nk));
    try {
      for (size_t i = 0; i < nk; i++) out.keys.push_back(group_by_keys_result(g, int64_t(i)));
      for (size_t a = 0; a < aggs.size(); a++) {
        std::vector<Array*> st;
        for (size_t w = 0; w < parts.front().state[a].size(); w++) {
          Array* c = cat([&](Partial& p) { return p.state[a][w]; });
          tmp.push_back(c);
          const bool extreme = aggs[a].op == SO_MIN || aggs[a].op == SO_MAX;
          st.push_back(group_agg(g, c, extreme ? (aggs[a].op == SO_MIN ? 4 : 5) : 0, 0.0));
        }
        out.state.push_back(std::move(st));
      }
    } catch (...) {
      group_by_release(g);
      throw;
    }
    group_by_release(g);
  } catch (...) {
    for (Array* a : keys) release_array(a);
    for (Array* a : tmp) release_array(a);
    release_partial(out);
    throw;
  }
  for (Array* a : keys) release_array(a);
  for (Array* a : tmp) release_array(a);
  for (Partial& p : parts) release_partial(p);
  parts.clear();
  return out;
}

Array* empty_of(const std::string& fmt) {
  Array* a = new_array(fmt, type_of_format(fmt), 0);
  a->values = device_alloc(8);
  a->null_count = 0;
  return a;
}

}  // namespace

StreamResult* stream_group_by(Stream* s, const char* const* keys, int64_t n_keys, const int* ops,
                              const char* const* columns, const char* const* names, int64_t n_aggs,
                              int64_t dense_key_count) {
  if (s->consumed) fail_invalid("am_stream_group_by: the stream has already been consumed by a terminal");
  const auto t0 = std::chrono::steady_clock::now();
  // The columns every batch must provide: the keys, then the aggregates' value columns.
  std::vector<int> need;   // child index per needed column
  auto child_of = [&](const std::string& n) {
    for (size_t i = 0; i < s->names.size(); i++)
      if (s->names[i] == n) return int(i);
    fail_invalid("no column named " + n);
  };
  auto need_index = [&](int child) {
    for (size_t i = 0; i < need.size(); i++)
      if (need[i] == child) return int(i);
    need.push_back(child);
    return int(need.size() - 1);
  };
  for (int64_t i = 0; i < n_keys; i++) need_index(child_of(keys[i]));
  std::vector<AggSpec> aggs;
  for (int64_t i = 0; i < n_aggs; i++) {
    AggSpec a;
    a.op = ops[i];
    a.name = names[i];
    if (a.op == 5 || a.op == 6 || a.op == 7)
      fail_unsupported("streaming aggregate op " + std::to_string(a.op) + " is not supported by this build");
    if (a.op < 0 || a.op > 7) fail_unsupported("unknown streaming aggregate op " + std::to_string(a.op));
    if (columns && columns[i]) a.col = need_index(child_of(columns[i]));
    else if (a.op != SO_COUNT) fail_invalid("a column of NULL is only valid for count");
    aggs.push_back(a);
  }
  for (int c : need) {
    const ArrowSchema* cs = s->schema.children[c];
    if (!type_is_primitive(type_of_format(cs->format ? cs->format : "")))
      fail_unsupported("streaming group-by over a " + std::string(cs->format ? cs->format : "?") +
                       " column is not supported by this build");
  }
  if (dense_key_count > 0) {
    if (n_keys != 1) fail_invalid("a dense streaming group-by takes one key column");
    const int kt = type_of_format(s->schema.children[need[0]]->format);
    if (!type_is_int(kt)) fail_invalid("a dense streaming group-by needs an integer key column");
  }
  for (const AggSpec& a : aggs)
    if (a.col >= 0 && (a.op == SO_SUM || a.op == SO_MEAN)) {
      const int t = type_of_format(s->schema.children[need[size_t(a.col)]]->format);
      if (!type_is_primitive(t)) fail_unsupported("streaming sum over a non-numeric column");
    }
  s->consumed = true;
  auto r = std::make_unique<StreamResult>();
  std::vector<Partial> parts;
  int64_t part_rows = 0, answer_rows = 0;
  struct Cleanup {
    std::vector<Partial>& p;
    ~Cleanup() {
      for (Partial& x : p) release_partial(x);
    }
  } cleanup{parts};
  while (true) {
    auto tr = std::chrono::steady_clock::now();
    ArrowArray batch{};
    if (s->src.get_next(&s->src, &batch) != 0) {
      const char* e = s->src.get_last_error ? s->src.get_last_error(&s->src) : nullptr;
      fail_invalid(std::string("am_stream_group_by: get_next failed: ") + (e ? e : "unknown error"));
    }
    if (!batch.release) break;   // end of stream
    struct BatchGuard {
      ArrowArray& b;
      ~BatchGuard() {
        if (b.release) b.release(&b);
      }
    } guard{batch};
    std::vector<Array*> cols;
    struct ColsGuard {
      std::vector<Array*>& c;
      ~ColsGuard() {
        for (Array* a : c) release_array(a);
      }
    } cguard{cols};
    if (batch.n_children != int64_t(s->names.size())) fail_invalid("am_stream_group_by: a batch does not match the schema");
    for (int c : need) {
      ArrowArray* child = batch.children[c];
      if (batch.offset != 0) fail_unsupported("am_stream_group_by: a record batch with an offset");
      r->bytes += child->length * type_width(type_of_format(s->schema.children[c]->format));
      cols.push_back(import_host(s->schema.children[c], child));
    }
    r->batches++;
    r->rows += batch.length;
    r->read += seconds_since(tr);
    if (batch.length == 0) continue;
    auto tg = std::chrono::steady_clock::now();
    std::vector<Array*> key_cols(cols.begin(), cols.begin() + n_keys);
    parts.push_back(batch_partial(key_cols, cols, aggs));
    part_rows += parts.back().rows();
    sync();
    r->gpu += seconds_since(tg);
    // Fold when the partials outgrow the running answer.
    if (parts.size() > 1 && part_rows > std::max<int64_t>(int64_t(1) << 22, 2 * answer_rows)) {
      auto tm = std::chrono::steady_clock::now();
      Partial f = fold(parts, aggs);
      answer_rows = f.rows();
      part_rows = answer_rows;
      parts.push_back(f);
      r->merge += seconds_since(tm);
    }
  }
  auto tm = std::chrono::steady_clock::now();
  // The answer: keys, then one column per aggregate.
  for (int64_t i = 0; i < n_keys; i++) r->names.push_back(keys[i]);
  for (const AggSpec& a : aggs) r->names.push_back(a.name);
  if (parts.empty()) {
    for (int64_t i = 0; i < n_keys; i++)
      r->cols.push_back(empty_of(dense_key_count > 0 ? "i" : s->schema.children[need[size_t(i)]]->format));
    for (const AggSpec& a : aggs) {
      std::string f = "l";
      if (a.op == SO_MEAN) f = "g";
      if (a.col >= 0 && a.op != SO_COUNT && a.op != SO_MEAN) {
        const int t = type_of_format(s->schema.children[need[size_t(a.col)]]->format);
        f = a.op == SO_SUM ? (type_is_float(t) ? "g" : type_is_signed(t) ? "l" : "L")
                           : std::string(s->schema.children[need[size_t(a.col)]]->format);
      }
      r->cols.push_back(empty_of(f));
    }
  } else {
    Partial ans = parts.size() == 1 ? parts.front() : fold(parts, aggs);
    if (parts.size() == 1) parts.clear();
    struct AnsGuard {
      Partial& p;
      ~AnsGuard() { release_partial(p); }
    } aguard{ans};
    hipStream_t st = stream();
    const int64_t K = ans.rows();
    for (Array* k : ans.keys) {
      retain_array(k);
      r->cols.push_back(k);
    }
    for (size_t a = 0; a < aggs.size(); a++) {
      if (aggs[a].op == SO_MEAN) {
        Array* sum = ans.state[a][0];
        Array* cnt = ans.state[a][1];
        Array* m = new_array("g", T_F64, K);
        r->cols.push_back(m);
        m->values = device_alloc(size_t(std::max<int64_t>(K, 1)) * 8);
        DevBuf vb = device_alloc(size_t(std::max<int64_t>(K, 1)));
        hipLaunchKernelGGL(sc_mean, dim3(blocks_for(K)), dim3(256), 0, st, sum->data(), sum->type,
                           static_cast<const unsigned long long*>(cnt->data()), K, static_cast<double*>(m->values->ptr),
                           static_cast<uint8_t*>(vb->ptr));
        AM_HIP(hipGetLastError());
        set_validity(m, static_cast<const uint8_t*>(vb->ptr), K);
      } else {
        Array* c = ans.state[a][0];
        retain_array(c);
        r->cols.push_back(c);
      }
    }
    if (dense_key_count > 0) {
      // One integer key inside [0, K), returned as int32; groups of other keys (and the null key) go.
      Array* key = r->cols[0];
      std::vector<uint8_t> vbits;
      std::vector<int64_t> kv(static_cast<size_t>(K));
      if (K > 0) {
        // The key column widened on the device, then read back (one row per group).
        Array* wide = key;
        std::vector<uint8_t> raw(size_t(K) * key->width());
        AM_HIP(hipMemcpyAsync(raw.data(), wide->data(), raw.size(), hipMemcpyDeviceToHost, st));
        if (key->has_nulls()) {
          vbits.resize(size_t((key->offset + K + 7) / 8));
          AM_HIP(hipMemcpyAsync(vbits.data(), key->valid_bits(), vbits.size(), hipMemcpyDeviceToHost, st));
        }
        AM_HIP(hipStreamSynchronize(st));
        for (int64_t i = 0; i < K; i++) {
          switch (key->type) {
            case T_I8: kv[size_t(i)] = reinterpret_cast<const int8_t*>(raw.data())[i]; break;
            case T_U8: kv[size_t(i)] = raw[size_t(i)]; break;
            case T_I16: kv[size_t(i)] = reinterpret_cast<const int16_t*>(raw.data())[i]; break;
            case T_U16: kv[size_t(i)] = reinterpret_cast<const uint16_t*>(raw.data())[i]; break;
            case T_I32: kv[size_t(i)] = reinterpret_cast<const int32_t*>(raw.data())[i]; break;
            case T_U32: kv[size_t(i)] = reinterpret_cast<const uint32_t*>(raw.data())[i]; break;
            default: kv[size_t(i)] = reinterpret_cast<const int64_t*>(raw.data())[i]; break;
          }
        }
      }
      std::vector<uint32_t> keep;
      std::vector<long long> kept_keys;
      for (int64_t i = 0; i < K; i++) {
        const int64_t b = key->offset + i;
        const bool valid = vbits.empty() || ((vbits[size_t(b >> 3)] >> (b & 7)) & 1);
        if (valid && kv[size_t(i)] >= 0 && kv[size_t(i)] < dense_key_count) {
          keep.push_back(uint32_t(i));
          kept_keys.push_back(kv[size_t(i)]);
        }
      }
      Array* idx = new_array("I", T_U32, int64_t(keep.size()));
      idx->values = device_alloc(std::max<size_t>(keep.size(), 1) * 4);
      idx->null_count = 0;
      Array* k64 = new_array("l", T_I64, int64_t(kept_keys.size()));
      k64->values = device_alloc(std::max<size_t>(kept_keys.size(), 1) * 8);
      k64->null_count = 0;
      if (!keep.empty()) {
        AM_HIP(hipMemcpyAsync(idx->values->ptr, keep.data(), keep.size() * 4, hipMemcpyHostToDevice, st));
        AM_HIP(hipMemcpyAsync(k64->values->ptr, kept_keys.data(), kept_keys.size() * 8, hipMemcpyHostToDevice, st));
      }
      Array* k32 = new_array("i", T_I32, int64_t(keep.size()));
      k32->values = device_alloc(std::max<size_t>(keep.size(), 1) * 4);
      k32->null_count = 0;
      hipLaunchKernelGGL(sc_i64_to_i32, dim3(blocks_for(int64_t(keep.size()))), dim3(256), 0, st,
                         static_cast<const long long*>(k64->values->ptr), int64_t(keep.size()),
                         static_cast<int32_t*>(k32->values->ptr));
      AM_HIP(hipGetLastError());
      AM_HIP(hipStreamSynchronize(st));
      release_array(k64);
      std::vector<Array*> kept{k32};
      for (size_t c = 1; c < r->cols.size(); c++) kept.push_back(take(*r->cols[c], *idx));
      release_array(idx);
      for (Array* a : r->cols) release_array(a);
      r->cols = kept;
    }
  }
  sync();
  r->merge += seconds_since(tm);
  r->rows_out = r->cols.empty() ? 0 : r->cols[0]->length;
  r->wall = seconds_since(t0);
  return r.release();
}

int64_t stream_result_column_count(const StreamResult* r) { return int64_t(r->cols.size()); }
const char* stream_result_column_name(const StreamResult* r, int64_t i) { return r->names[size_t(i)].c_str(); }
Array* stream_result_column(const StreamResult* r, int64_t i) {
  retain_array(r->cols[size_t(i)]);
  return r->cols[size_t(i)];
}
int64_t stream_result_rows_out(const StreamResult* r) { return r->rows_out; }
void stream_result_stats(const StreamResult* r, int64_t* batches, int64_t* rows, int64_t* bytes, double* wall,
                         double* read, double* gpu, double* merge, double* overlap) {
  if (batches) *batches = r->batches;
  if (rows) *rows = r->rows;
  if (bytes) *bytes = r->bytes;
  if (wall) *wall = r->wall;
  if (read) *read = r->read;
  if (gpu) *gpu = r->gpu;
  if (merge) *merge = r->merge;
  if (overlap) *overlap = r->wall > 0 ? (r->read + r->gpu + r->merge) / r->wall : 0.0;
}
void stream_result_release(StreamResult* r) { delete r; }

}  // namespace am
