// The plan engine's subset on this path: JSON plans of scan -> sort / group_by / unique / limit over
// registered tables, in one call (am_plan_source_create, am_plan_run, am_plan_explain and the result
// accessors).
//
// Reference: include/arrowmetal.h:1362-1433 (the grammar), Sources/ArrowMetalC/ArrowMetalC_Plan.swift
// (the C entry points, their NULL-handle and error conventions), Sources/ArrowMetal/Engine/PlanJSON.swift
// (parsing and its error texts), Engine/Executor.swift:46-160, 243-350 (what each operator computes)
// and Engine/PhysicalPlan.swift:118-128 (a one-key sort under a limit of 1..1024 rows, no offset, is a
// top-k selection).
//
// Operators and expressions outside the subset (filter, select, with_columns, aggregate, join,
// join_asof, concat, window, explode; any expression other than a column reference) are refused with
// "Unsupported Arrow type: plan: ... is not supported by this build". The `optimize` flag is accepted:
// on this subset the reference's optimizer rules (pushdown, pruning, folding) do not change a result.
//
// Execution is this build's own. A sort by one 64-bit key of a table of at least 2^18 rows moves
// every column with the sort (order_by_columns in sort.hip): the columns are partitioned along the
// sort's top-digit partition, and the final gather reads from the partition of its own output range,
// so no column is gathered across the whole table; an int64 / uint64 key column comes out of the
// sort's bucket sorts directly. Other sorts are a lexsort and a take per column,
// as the reference does.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <memory>
#include <sstream>

#include "am_internal.hpp"

namespace am {

Array* slice(const Array& a, int64_t offset, int64_t length);

namespace {

// ---------------------------------------------------------------- a small JSON reader

struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double d = 0.0;
  bool integral = false;
  int64_t i = 0;
  std::string s;
  std::vector<JVal> arr;
  std::vector<std::pair<std::string, JVal>> obj;
  const JVal* get(const char* key) const {
    if (kind != OBJ) return nullptr;
    for (auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

class JsonReader {
 public:
  explicit JsonReader(const std::string& t) : p_(t.data()), e_(t.data() + t.size()) {}
  JsonReader(std::string&&) = delete;   // the text must outlive the reader
  bool document(JVal& v) {
    if (!value(v, 0)) return false;
    ws();
    return p_ == e_;
  }

 private:
  const char* p_;
  const char* e_;
  void ws() {
    while (p_ < e_ && (*p_ == ' ' || *p_ == '\t' || *p_ == '\n' || *p_ == '\r')) p_++;
  }
  bool lit(const char* w) {
    size_t n = strlen(w);
    if (size_t(e_ - p_) < n || strncmp(p_, w, n) != 0) return false;
    p_ += n;
    return true;
  }
  static void utf8(std::string& o, uint32_t cp) {
    if (cp < 0x80) {
      o.push_back(char(cp));
    } else if (cp < 0x800) {
      o.push_back(char(0xC0 | (cp >> 6)));
      o.push_back(char(0x80 | (cp & 0x3F)));
    } else if (cp < 0x10000) {
      o.push_back(char(0xE0 | (cp >> 12)));
      o.push_back(char(0x80 | ((cp >> 6) & 0x3F)));
      o.push_back(char(0x80 | (cp & 0x3F)));
    } else {
      o.push_back(char(0xF0 | (cp >> 18)));
      o.push_back(char(0x80 | ((cp >> 12) & 0x3F)));
      o.push_back(char(0x80 | ((cp >> 6) & 0x3F)));
      o.push_back(char(0x80 | (cp & 0x3F)));
    }
  }
  bool hex4(uint32_t& cp) {
    if (e_ - p_ < 4) return false;
    cp = 0;
    for (int k = 0; k < 4; k++) {
      char c = *p_++;
      cp <<= 4;
      if (c >= '0' && c <= '9') cp |= uint32_t(c - '0');
      else if (c >= 'a' && c <= 'f') cp |= uint32_t(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F') cp |= uint32_t(c - 'A' + 10);
      else return false;
    }
    return true;
  }
  bool str(std::string& o) {
    if (p_ >= e_ || *p_ != '"') return false;
    p_++;
    while (p_ < e_ && *p_ != '"') {
      char c = *p_++;
      if (c != '\\') {
        o.push_back(c);
        continue;
      }
      if (p_ >= e_) return false;
      char x = *p_++;
      switch (x) {
        case '"': o.push_back('"'); break;
        case '\\': o.push_back('\\'); break;
        case '/': o.push_back('/'); break;
        case 'b': o.push_back('\b'); break;
        case 'f': o.push_back('\f'); break;
        case 'n': o.push_back('\n'); break;
        case 'r': o.push_back('\r'); break;
        case 't': o.push_back('\t'); break;
        case 'u': {
          uint32_t cp;
          if (!hex4(cp)) return false;
          if (cp >= 0xD800 && cp < 0xDC00 && e_ - p_ >= 6 && p_[0] == '\\' && p_[1] == 'u') {
            p_ += 2;
            uint32_t lo;
            if (!hex4(lo) || lo < 0xDC00 || lo >= 0xE000) return false;
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          }
          utf8(o, cp);
          break;
        }
        default: return false;
      }
    }
    if (p_ >= e_) return false;
    p_++;
    return true;
  }
  bool num(JVal& v) {
    const char* s = p_;
    if (p_ < e_ && *p_ == '-') p_++;
    if (p_ >= e_ || !isdigit(uint8_t(*p_))) return false;
    while (p_ < e_ && isdigit(uint8_t(*p_))) p_++;
    bool integral = true;
    if (p_ < e_ && *p_ == '.') {
      integral = false;
      p_++;
      if (p_ >= e_ || !isdigit(uint8_t(*p_))) return false;
      while (p_ < e_ && isdigit(uint8_t(*p_))) p_++;
    }
    if (p_ < e_ && (*p_ == 'e' || *p_ == 'E')) {
      integral = false;
      p_++;
      if (p_ < e_ && (*p_ == '+' || *p_ == '-')) p_++;
      if (p_ >= e_ || !isdigit(uint8_t(*p_))) return false;
      while (p_ < e_ && isdigit(uint8_t(*p_))) p_++;
    }
    std::string t(s, p_);
    v.kind = JVal::NUM;
    v.d = strtod(t.c_str(), nullptr);
    v.integral = integral;
    if (integral) {
      errno = 0;
      long long x = strtoll(t.c_str(), nullptr, 10);
      v.i = x;   // strtoll saturates out of range
    } else {
      v.i = std::isnan(v.d) ? 0 : v.d >= 9.2e18 ? INT64_MAX : v.d <= -9.2e18 ? INT64_MIN : int64_t(v.d);
    }
    return true;
  }
  bool value(JVal& v, int depth) {
    if (depth > 256) return false;
    ws();
    if (p_ >= e_) return false;
    char c = *p_;
    if (c == '{') {
      p_++;
      v.kind = JVal::OBJ;
      ws();
      if (p_ < e_ && *p_ == '}') {
        p_++;
        return true;
      }
      while (true) {
        ws();
        std::string k;
        if (!str(k)) return false;
        ws();
        if (p_ >= e_ || *p_ != ':') return false;
        p_++;
        JVal x;
        if (!value(x, depth + 1)) return false;
        v.obj.emplace_back(std::move(k), std::move(x));
        ws();
        if (p_ < e_ && *p_ == ',') {
          p_++;
          continue;
        }
        if (p_ < e_ && *p_ == '}') {
          p_++;
          return true;
        }
        return false;
      }
    }
    if (c == '[') {
      p_++;
      v.kind = JVal::ARR;
      ws();
      if (p_ < e_ && *p_ == ']') {
        p_++;
        return true;
      }
      while (true) {
        JVal x;
        if (!value(x, depth + 1)) return false;
        v.arr.push_back(std::move(x));
        ws();
        if (p_ < e_ && *p_ == ',') {
          p_++;
          continue;
        }
        if (p_ < e_ && *p_ == ']') {
          p_++;
          return true;
        }
        return false;
      }
    }
    if (c == '"') {
      v.kind = JVal::STR;
      return str(v.s);
    }
    if (lit("true")) {
      v.kind = JVal::BOOL;
      v.b = true;
      return true;
    }
    if (lit("false")) {
      v.kind = JVal::BOOL;
      return true;
    }
    if (lit("null")) return true;
    return num(v);
  }
};

// ---------------------------------------------------------------- expressions

// A column reference `(col "name")` or `(col name)` (the s-expression lexer of
// Sources/ArrowMetal/Expr/ExprText.swift:14-45: blanks separate atoms, "..." is text with \n, \t and
// \<c> escapes). Any other expression is outside this build's subset.
std::string column_of_expr(const std::string& text) {
  size_t i = 0;
  auto blank = [&](char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r'; };
  auto skip = [&] {
    while (i < text.size() && blank(text[i])) i++;
  };
  auto unsupported = [&]() -> std::string {
    fail_unsupported("plan: the expression " + text + " is not supported by this build (only (col \"name\"))");
  };
  skip();
  if (i >= text.size() || text[i] != '(') unsupported();
  i++;
  skip();
  size_t h = i;
  while (i < text.size() && !blank(text[i]) && text[i] != '(' && text[i] != ')' && text[i] != '"') i++;
  if (text.compare(h, i - h, "col") != 0) unsupported();
  skip();
  std::string name;
  if (i < text.size() && text[i] == '"') {
    i++;
    while (i < text.size() && text[i] != '"') {
      if (text[i] == '\\' && i + 1 < text.size()) {
        i++;
        name.push_back(text[i] == 'n' ? '\n' : text[i] == 't' ? '\t' : text[i]);
      } else {
        name.push_back(text[i]);
      }
      i++;
    }
    if (i >= text.size()) fail_invalid("plan: unterminated string in expression text");
    i++;
  } else {
    size_t a = i;
    while (i < text.size() && !blank(text[i]) && text[i] != '(' && text[i] != ')') i++;
    if (i == a) fail_invalid("plan: col is missing an argument");
    name = text.substr(a, i - a);
  }
  skip();
  if (i >= text.size() || text[i] != ')') unsupported();
  i++;
  skip();
  if (i != text.size()) unsupported();
  return name;
}

// ---------------------------------------------------------------- tables

// One counted reference to an Array.
struct Ref {
  Array* a = nullptr;
  Ref() = default;
  explicit Ref(Array* x) : a(x) {}   // adopts a +1 reference
  Ref(const Ref& o) : a(o.a) {
    if (a) retain_array(a);
  }
  Ref& operator=(const Ref& o) {
    if (this != &o) {
      if (o.a) retain_array(o.a);
      if (a) release_array(a);
      a = o.a;
    }
    return *this;
  }
  ~Ref() {
    if (a) release_array(a);
  }
  static Ref share(Array* x) {
    retain_array(x);
    return Ref(x);
  }
};

struct Table {
  std::vector<std::string> names;
  std::vector<Ref> cols;
  int64_t length() const { return cols.empty() ? 0 : cols[0].a->length; }
  const Array* find(const std::string& n) const {
    for (size_t c = 0; c < names.size(); c++)
      if (names[c] == n) return cols[c].a;
    return nullptr;
  }
};

Table take_table(const Table& t, const Array& idx) {
  Table o;
  o.names = t.names;
  for (auto& c : t.cols) o.cols.emplace_back(take(*c.a, idx));
  return o;
}

}  // namespace

// A registered table (reference ArrowMetalC_Plan.swift:22-26, 52-74).
struct PlanSource {
  std::string name;
  Table table;
};

// A finished plan's columns (reference ArrowMetalC_Plan.swift:28-40).
struct PlanResult {
  Table table;
};

namespace {

// ---------------------------------------------------------------- plan nodes

struct SortKey {
  std::string column;
  bool descending = false;
  bool nulls_first = false;
  int float_order = 0;
};
struct Agg {
  std::string op, name, column;   // column empty: no expression
};
struct Node {
  std::string op;
  std::unique_ptr<Node> input;
  const PlanSource* source = nullptr;   // scan
  bool has_columns = false;
  std::vector<std::string> columns;     // scan columns, unique subset
  bool has_subset = false;
  std::vector<SortKey> keys;            // sort
  int64_t count = INT64_MAX, offset = 0;   // limit
  std::vector<std::pair<std::string, std::string>> gkeys;   // group_by: (output name, column)
  std::vector<Agg> aggs;
};

bool flag(const JVal* v) {
  if (!v) return false;
  if (v->kind == JVal::BOOL) return v->b;
  if (v->kind == JVal::NUM) return v->d != 0.0;
  return false;
}

bool nulls_first_of(const JVal* v, bool fallback) {
  if (!v || v->kind == JVal::NUL) return fallback;
  if (v->kind == JVal::STR && v->s == "first") return true;
  if (v->kind == JVal::STR && v->s == "last") return false;
  fail_invalid("plan: sort \"nulls\" must be \"first\" or \"last\"");
}

int float_order_of(const JVal* v, int fallback) {
  if (!v || v->kind == JVal::NUL) return fallback;
  if (v->kind == JVal::STR) {
    if (v->s == "ieee") return 0;
    if (v->s == "total") return 1;
    if (v->s == "nan_largest") return 2;
  }
  fail_invalid("plan: sort \"float_order\" must be \"ieee\", \"total\" or \"nan_largest\"");
}

// The keys of a sort's "by", with the sort-level "nulls" / "float_order" as defaults
// (PlanJSON.swift:96-118).
std::vector<SortKey> sort_keys(const JVal& o) {
  std::vector<SortKey> out;
  const JVal* by = o.get("by");
  if (!by || by->kind != JVal::ARR) return out;
  const bool dnf = nulls_first_of(o.get("nulls"), false);
  const int dfo = float_order_of(o.get("float_order"), 0);
  for (const JVal& x : by->arr) {
    SortKey k;
    if (x.kind == JVal::OBJ) {
      const JVal* c = x.get("column");
      if (!c || c->kind != JVal::STR) fail_invalid("plan: a sort key object needs a \"column\"");
      k.column = c->s;
      k.descending = flag(x.get("descending"));
      k.nulls_first = nulls_first_of(x.get("nulls"), dnf);
      k.float_order = float_order_of(x.get("float_order"), dfo);
    } else if (x.kind == JVal::ARR && !x.arr.empty() && x.arr[0].kind == JVal::STR) {
      k.column = x.arr[0].s;
      k.descending = x.arr.size() > 1 ? flag(&x.arr[1]) : false;
      const JVal* opts = x.arr.size() > 2 && x.arr[2].kind == JVal::OBJ ? &x.arr[2] : nullptr;
      k.nulls_first = nulls_first_of(opts ? opts->get("nulls") : nullptr, dnf);
      k.float_order = float_order_of(opts ? opts->get("float_order") : nullptr, dfo);
    } else {
      continue;   // entries of another shape are skipped, as the reference's compactMap does
    }
    out.push_back(std::move(k));
  }
  return out;
}

std::vector<std::string> strings_of(const JVal* v) {
  std::vector<std::string> out;
  if (!v || v->kind != JVal::ARR) return out;
  for (auto& x : v->arr)
    if (x.kind == JVal::STR) out.push_back(x.s);
  return out;
}

std::unique_ptr<Node> parse_node(const JVal& o, const std::map<std::string, const PlanSource*>& sources) {
  const JVal* opv = o.get("op");
  if (!opv || opv->kind != JVal::STR) fail_invalid("plan: node has no \"op\"");
  auto n = std::make_unique<Node>();
  n->op = opv->s;
  auto input = [&](const char* key = "input") {
    const JVal* d = o.get(key);
    if (!d || d->kind != JVal::OBJ) fail_invalid("plan: " + n->op + " has no \"" + key + "\"");
    return parse_node(*d, sources);
  };
  const std::string& op = n->op;
  if (op == "scan") {
    const JVal* s = o.get("source");
    const std::string name = s && s->kind == JVal::STR ? s->s : "?";
    auto it = s && s->kind == JVal::STR ? sources.find(s->s) : sources.end();
    if (it == sources.end()) fail_invalid("plan: unknown source \"" + name + "\"");
    n->source = it->second;
    const JVal* cols = o.get("columns");
    if (cols && cols->kind == JVal::ARR) {
      n->has_columns = true;
      n->columns = strings_of(cols);
    }
  } else if (op == "sort") {
    n->input = input();
    n->keys = sort_keys(o);
  } else if (op == "limit") {
    const JVal* c = o.get("count");
    const JVal* f = o.get("offset");
    n->count = c && c->kind == JVal::NUM ? c->i : INT64_MAX;
    n->offset = f && f->kind == JVal::NUM ? f->i : 0;
    n->input = input();
  } else if (op == "unique") {
    const JVal* s = o.get("subset");
    if (s && s->kind == JVal::ARR) {
      n->has_subset = true;
      n->columns = strings_of(s);
    }
    n->input = input();
  } else if (op == "group_by") {
    n->input = input();
    if (const JVal* ks = o.get("keys"); ks && ks->kind == JVal::ARR) {
      for (auto& x : ks->arr) {
        if (x.kind != JVal::ARR || x.arr.size() != 2 || x.arr[0].kind != JVal::STR || x.arr[1].kind != JVal::STR)
          fail_invalid("plan: group_by.keys entries are [name, sexpr] pairs");
        n->gkeys.emplace_back(x.arr[0].s, column_of_expr(x.arr[1].s));
      }
    }
    if (const JVal* as = o.get("aggs"); as && as->kind == JVal::ARR) {
      for (auto& x : as->arr) {
        if (x.kind != JVal::ARR || x.arr.size() < 2 || x.arr[0].kind != JVal::STR || x.arr[1].kind != JVal::STR)
          fail_invalid("plan: aggs entries are [op, name, sexpr?]");
        const std::string& aop = x.arr[0].s;
        if (aop != "sum" && aop != "min" && aop != "max" && aop != "mean" && aop != "count")
          fail_invalid("plan: aggs entries are [op, name, sexpr?]");
        Agg a{aop, x.arr[1].s, ""};
        if (x.arr.size() > 2 && x.arr[2].kind == JVal::STR && !x.arr[2].s.empty())
          a.column = column_of_expr(x.arr[2].s);
        n->aggs.push_back(std::move(a));
      }
    }
  } else if (op == "filter" || op == "select" || op == "with_columns" || op == "aggregate" || op == "join" ||
             op == "join_asof" || op == "concat" || op == "window" || op == "explode") {
    fail_unsupported("plan: operator \"" + op + "\" is not supported by this build");
  } else {
    fail_invalid("plan: unknown operator \"" + op + "\"");
  }
  return n;
}

std::unique_ptr<Node> parse_plan(const char* text, const std::map<std::string, const PlanSource*>& sources) {
  JVal root;
  const std::string t(text);
  JsonReader r(t);   // reads `t` in place
  if (!r.document(root) || root.kind != JVal::OBJ) fail_invalid("plan: not a JSON object");
  return parse_node(root, sources);
}

// ---------------------------------------------------------------- execution

Table run_node(const Node& n);

// The result type of a group-by aggregate over a column of `type` (Executor.swift:300-350).
const char* agg_format(const std::string& op, const Array* values) {
  if (op == "count") return "l";
  if (op == "mean") return "g";
  const int t = values->type;
  if (op == "sum") {
    if (t == T_U8 || t == T_U16 || t == T_U32 || t == T_U64) return "L";
    if (type_is_float(t)) return "g";
    return "l";
  }
  return values->format.c_str();
}

Array* empty_array(const std::string& format) {
  const int t = type_of_format(format);
  Array* out = new_array(format, t, 0);
  out->values = device_alloc(8);
  out->null_count = 0;
  return out;
}

Table sort_table(Table in, const std::vector<SortKey>& keys);

Table run_sort(const Node& n) { return sort_table(run_node(*n.input), n.keys); }

// A stable sort of every column by `keys` (Executor.swift:118-121, Kernels/MultiSort.swift:201-213).
Table sort_table(Table in, const std::vector<SortKey>& keys) {
  if (keys.empty()) return in;
  std::vector<Array*> kc;
  for (auto& k : keys) {
    const Array* c = in.find(k.column);
    if (!c) fail_invalid("no column named " + k.column);
    kc.push_back(const_cast<Array*>(c));
  }
  if (in.length() == 0) return in;
  if (keys.size() == 1) {
    std::vector<const Array*> cols;
    for (auto& c : in.cols) cols.push_back(c.a);
    std::vector<Array*> sorted;
    const SortKey& k = keys[0];
    if (order_by_columns(*kc[0], k.descending, k.nulls_first ? 1 : 0, k.float_order, cols, sorted)) {
      Table o;
      o.names = in.names;
      for (Array* a : sorted) o.cols.emplace_back(a);
      return o;
    }
  }
  std::vector<int> desc, place, forder;
  for (auto& k : keys) {
    desc.push_back(k.descending ? 1 : 0);
    place.push_back(k.nulls_first ? 1 : 0);
    forder.push_back(k.float_order);
  }
  Ref idx(lexsort(kc.data(), int64_t(kc.size()), desc.data(), place.data(), forder.data()));
  return take_table(in, *idx.a);
}

Table run_limit(const Node& n) {
  // `sort` then `limit` of 1..1024 rows with no offset and one numeric key is a top-k selection
  // (PhysicalPlan.swift:118-126): the first k rows of the stable sort, without sorting the rest.
  Table in;
  if (n.offset == 0 && n.input->op == "sort" && n.input->keys.size() == 1 && n.count > 0 && n.count <= 1024) {
    Table base = run_node(*n.input->input);
    const SortKey& k = n.input->keys[0];
    const Array* c = base.find(k.column);
    if (c && type_is_primitive(c->type)) {
      if (base.length() == 0) return base;
      Ref idx(top_k(*c, n.count, k.descending, k.nulls_first ? 1 : 0, k.float_order));
      return take_table(base, *idx.a);
    }
    in = sort_table(std::move(base), n.input->keys);   // not numeric (or missing): sort, then slice
  } else {
    in = run_node(*n.input);
  }
  const int64_t len_in = in.length();
  const int64_t lo = std::min(std::max<int64_t>(n.offset, 0), len_in);
  const int64_t len = std::max<int64_t>(std::min<int64_t>(n.count, len_in - lo), 0);
  if (lo == 0 && len == len_in) return in;
  Table o;
  o.names = in.names;
  for (auto& c : in.cols) o.cols.emplace_back(slice(*c.a, lo, len));
  return o;
}

Table run_unique(const Node& n) {
  Table in = run_node(*n.input);
  std::vector<Array*> kc;
  const std::vector<std::string>& names = n.has_subset ? n.columns : in.names;
  for (auto& nm : names) {
    const Array* c = in.find(nm);
    if (!c) fail_invalid("unique: no column named " + nm);
    kc.push_back(const_cast<Array*>(c));
  }
  if (in.length() == 0 || kc.empty()) return in;
  Ref keep(unique_rows(kc.data(), int64_t(kc.size())));
  return take_table(in, *keep.a);
}

Table run_group_by(const Node& n) {
  Table in = run_node(*n.input);
  std::vector<Array*> kc;
  for (auto& k : n.gkeys) {
    const Array* c = in.find(k.second);
    if (!c) fail_invalid("no column named " + k.second);
    kc.push_back(const_cast<Array*>(c));
  }
  std::vector<const Array*> vals;
  for (auto& a : n.aggs) {
    const Array* v = nullptr;
    if (!a.column.empty()) {
      v = in.find(a.column);
      if (!v) fail_invalid("no column named " + a.column);
    } else if (a.op != "count") {
      fail_invalid(a.op + " needs an expression");
    }
    vals.push_back(v);
  }
  Table o;
  for (auto& k : n.gkeys) o.names.push_back(k.first);
  for (auto& a : n.aggs) o.names.push_back(a.name);
  if (in.length() == 0) {
    // No groups; the output still has its columns (Executor.swift:248-259).
    for (Array* c : kc) o.cols.emplace_back(slice(*c, 0, 0));
    for (size_t i = 0; i < n.aggs.size(); i++)
      o.cols.emplace_back(empty_array(vals[i] ? agg_format(n.aggs[i].op, vals[i]) : "l"));
    return o;
  }
  if (kc.empty()) fail_invalid("group-by needs at least one key column");
  GroupBy* g = group_by_keys(kc.data(), int64_t(kc.size()));
  try {
    for (size_t i = 0; i < kc.size(); i++) o.cols.emplace_back(group_by_keys_result(g, int64_t(i)));
    for (size_t i = 0; i < n.aggs.size(); i++) {
      const std::string& op = n.aggs[i].op;
      int code;
      if (op == "count") code = vals[i] ? 2 : 1;   // non-null values of the expression, or rows
      else if (op == "sum") code = 0;
      else if (op == "mean") code = 3;
      else if (op == "min") code = 4;
      else code = 5;
      Array* r = group_agg(g, vals[i], code, 0.0);
      // An unsigned sum is uint64 on both of the reference's plan paths (Executor.swift:308-311, and
      // ExprGroupBy.swift's sumU); the wrapped 64-bit total has the same bits either way.
      if (code == 0 && r->type == T_I64 && (vals[i]->type == T_U8 || vals[i]->type == T_U16 || vals[i]->type == T_U32)) {
        r->format = "L";
        r->type = T_U64;
      }
      o.cols.emplace_back(r);
    }
  } catch (...) {
    group_by_release(g);
    throw;
  }
  group_by_release(g);
  return o;
}

Table run_node(const Node& n) {
  if (n.op == "scan") {
    const Table& t = n.source->table;
    if (!n.has_columns) return t;
    Table o;
    for (auto& nm : n.columns) {
      const Array* c = t.find(nm);
      if (!c) fail_invalid("no column named " + nm);
      o.names.push_back(nm);
      o.cols.push_back(Ref::share(const_cast<Array*>(c)));
    }
    return o;
  }
  if (n.op == "sort") return run_sort(n);
  if (n.op == "limit") return run_limit(n);
  if (n.op == "unique") return run_unique(n);
  return run_group_by(n);
}

// The physical plan as text (this build's own wording).
void describe(const Node& n, int depth, std::ostringstream& os) {
  std::string pad(size_t(2 * depth), ' ');
  if (n.op == "scan") {
    os << pad << "SOURCE " << n.source->name;
    if (n.has_columns) {
      os << " [";
      for (size_t i = 0; i < n.columns.size(); i++) os << (i ? ", " : "") << n.columns[i];
      os << "]";
    }
    os << "\n";
    return;
  }
  const Node* child = n.input.get();
  if (n.op == "sort") {
    os << pad << "SORT [";
    for (size_t i = 0; i < n.keys.size(); i++) {
      const SortKey& k = n.keys[i];
      os << (i ? ", " : "") << k.column << (k.descending ? " DESC" : "") << (k.nulls_first ? " NULLS FIRST" : "")
         << (k.float_order == 1 ? " TOTAL_ORDER" : k.float_order == 2 ? " NAN_LARGEST" : "");
    }
    os << "]" << (n.keys.size() == 1 ? " (one-key MSD sort moving every column when the key is 64-bit)" : " (lexsort + take)")
       << "\n";
  } else if (n.op == "limit") {
    if (n.offset == 0 && child->op == "sort" && child->keys.size() == 1 && n.count > 0 && n.count <= 1024) {
      os << pad << "SLICE 0, " << n.count << "\n" << pad << "  TOP_K " << n.count << " BY " << child->keys[0].column
         << (child->keys[0].descending ? " DESC" : "") << "\n";
      describe(*child->input, depth + 2, os);
      return;
    }
    os << pad << "SLICE " << n.offset << ", " << n.count << "\n";
  } else if (n.op == "unique") {
    os << pad << "DISTINCT";
    if (n.has_subset) {
      os << " [";
      for (size_t i = 0; i < n.columns.size(); i++) os << (i ? ", " : "") << n.columns[i];
      os << "]";
    }
    os << " (keep first)\n";
  } else {
    os << pad << "HASH_AGGREGATE keys [";
    for (size_t i = 0; i < n.gkeys.size(); i++) os << (i ? ", " : "") << n.gkeys[i].first;
    os << "] aggs [";
    for (size_t i = 0; i < n.aggs.size(); i++)
      os << (i ? ", " : "") << n.aggs[i].op << "(" << n.aggs[i].column << ") AS " << n.aggs[i].name;
    os << "]\n";
  }
  describe(*child, depth + 1, os);
}

thread_local std::string t_explain;

}  // namespace

PlanSource* plan_source_create(const char* name, Array* const* cols, const char* const* names, int64_t n) {
  auto* s = new PlanSource;
  s->name = name;
  try {
    for (int64_t i = 0; i < n; i++) {
      s->table.names.push_back(names[i]);
      s->table.cols.push_back(Ref::share(cols[i]));
    }
    for (int64_t i = 1; i < n; i++)
      if (cols[i]->length != cols[0]->length) fail_length_mismatch(cols[0]->length, cols[i]->length);
  } catch (...) {
    delete s;
    throw;
  }
  return s;
}

void plan_source_release(PlanSource* s) { delete s; }

static std::map<std::string, const PlanSource*> source_map(PlanSource* const* sources, int64_t n) {
  std::map<std::string, const PlanSource*> m;
  for (int64_t i = 0; i < n; i++) m[sources[i]->name] = sources[i];   // a later source of a name wins
  return m;
}

PlanResult* plan_run(const char* text, PlanSource* const* sources, int64_t n, int /*optimize*/) {
  const auto m = source_map(sources, n);
  std::unique_ptr<Node> plan = parse_plan(text, m);
  auto* r = new PlanResult;
  try {
    r->table = run_node(*plan);
    sync();
  } catch (...) {
    delete r;
    throw;
  }
  return r;
}

const char* plan_explain(const char* text, PlanSource* const* sources, int64_t n, int /*optimize*/) {
  const auto m = source_map(sources, n);
  std::unique_ptr<Node> plan = parse_plan(text, m);
  std::ostringstream os;
  os << "PHYSICAL PLAN\n";
  describe(*plan, 1, os);
  t_explain = os.str();
  return t_explain.c_str();
}

int64_t plan_column_count(const PlanResult* r) { return int64_t(r->table.names.size()); }
int64_t plan_row_count(const PlanResult* r) { return r->table.length(); }
const char* plan_column_name(const PlanResult* r, int64_t i) { return r->table.names[size_t(i)].c_str(); }
Array* plan_column(const PlanResult* r, int64_t i) {
  Array* a = r->table.cols[size_t(i)].a;
  retain_array(a);
  return a;
}
void plan_result_release(PlanResult* r) { delete r; }

}  // namespace am
