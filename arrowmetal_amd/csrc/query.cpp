// am_query's fused dense group-by, the form the DuckDB extension sends for small key spans.
//
// Reference: include/arrowmetal.h:1178-1236 (the query grammar and result accessors),
// duckdb-extension/src/arrowmetal_rewrite.cpp:1245-1257, 1319-1350 (the queries it builds: a key of
// (sub (cast (col "c0") i64) (i64 MIN)) over a span, and count / sum / min / max needs, sums split as
// (shr c 32) and (bit_and c 0xFFFFFFFF)), Sources/ArrowMetal/Expr/ExprGroupBy.swift:48-66, 360-421
// (the aggregate kinds, the refusals, and the result columns: the key slot 0..K-1 as int32, then one
// column per aggregate, null where no value contributed, counts never null).
//
// Taken: "(query (group_by K \"name\" KEY) (aggregate AGG+))" where KEY is a column reference, optionally
// cast to an integer type, optionally with an integer literal subtracted or added, and AGG is
// (count "n"), (count "n" V), (sum "n" V), (mean "n" V), (min "n" V), (max "n" V) with V a column
// reference, (shr V N) or (bit_and V N). Filters, projections, whole-input aggregates and every other
// expression return "Unsupported Arrow type: am_query: ... is not supported by this build".
// Each aggregate is one row pass of the dense group-by kernels (group_by_dense_ex).
#include <cstdlib>
#include <map>
#include <memory>
#include <sstream>

#include "am_internal.hpp"

namespace am {

namespace {

// ---------------------------------------------------------------- s-expressions

struct SNode {
  bool list = false;
  bool text = false;       // a "..." literal (atoms otherwise)
  std::string atom;
  std::vector<SNode> kids;
};

class SReader {
 public:
  explicit SReader(const std::string& t) : s_(t) {}
  SNode read() {
    SNode n = node();
    skip();
    if (i_ != s_.size()) fail_invalid("am_query: text after the query");
    return n;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;
  static bool blank(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r'; }
  void skip() {
    while (i_ < s_.size() && blank(s_[i_])) i_++;
  }
  SNode node() {
    skip();
    if (i_ >= s_.size()) fail_invalid("am_query: unexpected end of expression text");
    SNode n;
    if (s_[i_] == '(') {
      i_++;
      n.list = true;
      while (true) {
        skip();
        if (i_ >= s_.size()) fail_invalid("am_query: missing ')' in expression text");
        if (s_[i_] == ')') {
          i_++;
          return n;
        }
        n.kids.push_back(node());
      }
    }
    if (s_[i_] == ')') fail_invalid("am_query: unexpected ')' in expression text");
    if (s_[i_] == '"') {
      i_++;
      n.text = true;
      while (i_ < s_.size() && s_[i_] != '"') {
        char c = s_[i_++];
        if (c == '\\' && i_ < s_.size()) {
          char x = s_[i_++];
          c = x == 'n' ? '\n' : x == 't' ? '\t' : x;
        }
        n.atom.push_back(c);
      }
      if (i_ >= s_.size()) fail_invalid("am_query: unterminated string in expression text");
      i_++;
      return n;
    }
    while (i_ < s_.size() && !blank(s_[i_]) && s_[i_] != '(' && s_[i_] != ')') n.atom.push_back(s_[i_++]);
    return n;
  }
};

const std::string& head(const SNode& n) {
  static const std::string none;
  return n.list && !n.kids.empty() && !n.kids[0].list ? n.kids[0].atom : none;
}

std::string show(const SNode& n) {
  if (!n.list) {
    if (!n.text) return n.atom;
    std::string o = "\"";
    for (char c : n.atom) {
      if (c == '"' || c == '\\') o.push_back('\\');
      o.push_back(c);
    }
    return o + "\"";
  }
  std::string o = "(";
  for (size_t i = 0; i < n.kids.size(); i++) o += (i ? " " : "") + show(n.kids[i]);
  return o + ")";
}

[[noreturn]] void unsupported(const SNode& n, const char* what) {
  fail_unsupported(std::string("am_query: ") + what + " " + show(n) + " is not supported by this build");
}

bool is_int_type(const std::string& t) {
  return t == "i8" || t == "i16" || t == "i32" || t == "i64" || t == "u8" || t == "u16" || t == "u32" || t == "u64";
}

// An integer literal: (int N) or (i64 N) and the other typed integer literals.
bool int_literal(const SNode& n, long long& v) {
  const std::string& h = head(n);
  if (!(h == "int" || is_int_type(h)) || n.kids.size() != 2 || n.kids[1].list) return false;
  char* end = nullptr;
  v = strtoll(n.kids[1].atom.c_str(), &end, 10);
  return end && *end == 0 && !n.kids[1].atom.empty();
}

std::string column_ref(const SNode& n) {
  if (head(n) != "col" || n.kids.size() < 2 || n.kids[1].list) return {};
  return n.kids[1].atom;
}

struct Key {
  std::string column;
  long long bias = 0;
};

// KEY := (col c) | (cast KEY INT-TYPE) | (sub KEY LIT) | (add KEY LIT)
bool parse_key(const SNode& n, Key& k) {
  const std::string c = column_ref(n);
  if (!c.empty()) {
    k.column = c;
    return true;
  }
  const std::string& h = head(n);
  long long lit;
  if (h == "cast" && n.kids.size() == 3 && !n.kids[2].list && is_int_type(n.kids[2].atom)) return parse_key(n.kids[1], k);
  if ((h == "sub" || h == "add") && n.kids.size() == 3 && int_literal(n.kids[2], lit)) {
    if (!parse_key(n.kids[1], k)) return false;
    k.bias += h == "sub" ? -lit : lit;
    return true;
  }
  return false;
}

struct Agg {
  std::string op, name, column;   // column empty: count of rows
  int xop = 0;
  long long xarg = 0;
};

// V := (col c) | (shr (col c) LIT) | (bit_and (col c) LIT)
bool parse_value(const SNode& n, Agg& a) {
  a.column = column_ref(n);
  if (!a.column.empty()) return true;
  const std::string& h = head(n);
  long long lit;
  if ((h == "shr" || h == "bit_and") && n.kids.size() == 3 && int_literal(n.kids[2], lit)) {
    a.column = column_ref(n.kids[1]);
    if (a.column.empty()) return false;
    a.xop = h == "shr" ? 1 : 2;
    a.xarg = lit;
    return true;
  }
  return false;
}

struct Query {
  long long K = 0;
  std::string key_name;
  Key key;
  std::vector<Agg> aggs;
};

Query parse_query(const char* text) {
  const std::string t(text);
  SReader r(t);
  const SNode root = r.read();
  if (head(root) != "query") fail_invalid("am_query: expected (query ...)");
  Query q;
  bool grouped = false, terminal = false;
  for (size_t i = 1; i < root.kids.size(); i++) {
    const SNode& part = root.kids[i];
    const std::string& h = head(part);
    if (h == "group_by") {
      long long K;
      if (part.kids.size() != 4 || part.kids[1].list || part.kids[2].list)
        fail_invalid("am_query: group_by takes a key count, a name and a key expression");
      char* end = nullptr;
      K = strtoll(part.kids[1].atom.c_str(), &end, 10);
      if (!end || *end != 0) fail_invalid("am_query: group_by takes a key count, a name and a key expression");
      if (K <= 0 || K > int64_t(UINT32_MAX) / 8)
        fail_invalid("group_by keyCount out of range: " + std::to_string(K));
      q.K = K;
      q.key_name = part.kids[2].atom;
      if (!parse_key(part.kids[3], q.key)) unsupported(part.kids[3], "the group_by key");
      grouped = true;
    } else if (h == "aggregate") {
      for (size_t j = 1; j < part.kids.size(); j++) {
        const SNode& a = part.kids[j];
        const std::string& op = head(a);
        if (op != "sum" && op != "min" && op != "max" && op != "mean" && op != "count")
          fail_invalid("am_query: unknown aggregate " + show(a));
        if (a.kids.size() < 2 || a.kids[1].list) fail_invalid("am_query: an aggregate needs a name: " + show(a));
        Agg g;
        g.op = op;
        g.name = a.kids[1].atom;
        if (a.kids.size() > 2) {
          if (!parse_value(a.kids[2], g)) unsupported(a.kids[2], "the aggregate expression");
        } else if (op != "count") {
          fail_invalid(op + " \"" + g.name + "\" needs an expression");
        }
        q.aggs.push_back(std::move(g));
      }
      terminal = true;
    } else if (h == "filter" || h == "project") {
      unsupported(part, h == "filter" ? "the filter" : "the project terminal");
    } else {
      fail_invalid("am_query: unknown query part " + show(part));
    }
  }
  if (!terminal) fail_invalid("am_query: a query needs a terminal");
  if (!grouped) unsupported(root, "a whole-input aggregate");
  return q;
}

}  // namespace

// The result: named columns (reference ExprQueryResult; scalars are never produced here).
struct QueryResult {
  std::vector<std::string> names;
  std::vector<Array*> cols;
  ~QueryResult() {
    for (Array* a : cols) release_array(a);
  }
};

namespace {

// int32 0 .. K-1, the key slot of each group.
__global__ void q_iota(int32_t* out, long long K) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < K) out[i] = int32_t(i);
}

}  // namespace

QueryResult* query_run(Array* const* cols, const char* const* names, int64_t n, const char* text) {
  const Query q = parse_query(text);
  std::map<std::string, Array*> by_name;
  for (int64_t i = 0; i < n; i++) by_name[names[i]] = cols[i];
  auto column = [&](const std::string& c) -> Array* {
    auto it = by_name.find(c);
    if (it == by_name.end()) fail_invalid("no column named " + c);
    return it->second;
  };
  Array* key = column(q.key.column);
  if (!type_is_int(key->type)) fail_unsupported("group_by key must be an integer expression, got " + key->format);
  if (key->type != T_I32 && key->type != T_I64 && key->type != T_U32)
    fail_unsupported("am_query: a group_by key column of type " + key->format + " is not supported by this build");
  auto r = std::make_unique<QueryResult>();
  Array* keys_out = new_array("i", T_I32, q.K);
  r->cols.push_back(keys_out);
  r->names.push_back(q.key_name);
  keys_out->values = device_alloc(size_t(q.K) * 4);
  keys_out->null_count = 0;
  hipLaunchKernelGGL(q_iota, dim3(uint32_t((q.K + 255) / 256)), dim3(256), 0, stream(),
                     static_cast<int32_t*>(keys_out->values->ptr), q.K);
  AM_HIP(hipGetLastError());
  for (const Agg& a : q.aggs) {
    Array* v = a.column.empty() ? nullptr : column(a.column);
    int agg;
    if (a.op == "count") {
      agg = v ? 5 : 1;
    } else {
      const int t = v->type;
      const bool integral = type_is_int(t) || (a.xop != 0);
      if (a.xop != 0 && !type_is_int(t))
        fail_unsupported("am_query: shr / bit_and of a " + v->format + " column");
      if (a.op == "sum" || a.op == "mean") {
        if (t == T_F64)
          fail_unsupported("group_by " + a.op + " \"" + a.name + "\" over float64; cast the expression to float32");
        if (!integral && t != T_F32) fail_unsupported("group_by " + a.op + " \"" + a.name + "\" over " + v->format);
        if (a.op == "mean" && t == T_F32)
          fail_unsupported("am_query: mean of a float32 expression is not supported by this build");
        agg = a.op == "sum" ? 0 : 4;
      } else {
        const bool narrow = (type_is_int(t) && type_width(t) <= 4 && a.xop == 0) || t == T_F32;
        if (!narrow)
          fail_unsupported("group_by " + a.op + " \"" + a.name +
                           "\" needs a 32-bit or narrower integer or a float32 expression (Metal has no 64-bit "
                           "atomics); cast first");
        agg = a.op == "min" ? 2 : 3;
      }
    }
    Array* out = group_by_dense_ex(*key, q.K, q.key.bias, agg, v, a.xop, a.xarg);
    // Sums of unsigned columns are uint64 (ExprGroupBy.swift's sumU); the wrapped total has the same bits.
    if (agg == 0 && out->type == T_I64 && v && !type_is_signed(v->type) && type_is_int(v->type)) {
      out->format = "L";
      out->type = T_U64;
    }
    r->cols.push_back(out);
    r->names.push_back(a.name);
  }
  sync();
  return r.release();
}

const char* query_canonical(const char* text) {
  thread_local std::string t;
  const std::string s(text);
  SReader r(s);
  const SNode root = r.read();
  parse_query(text);   // refuses what this build does not run
  t = show(root);
  return t.c_str();
}

int64_t query_column_count(const QueryResult* r) { return int64_t(r->cols.size()); }
const char* query_column_name(const QueryResult* r, int64_t i) { return r->names[size_t(i)].c_str(); }
Array* query_column(const QueryResult* r, int64_t i) {
  retain_array(r->cols[size_t(i)]);
  return r->cols[size_t(i)];
}
void query_result_release(QueryResult* r) { delete r; }

}  // namespace am
