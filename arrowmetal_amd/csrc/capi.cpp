// The extern "C" entry points of include/arrowmetal.h.
//
// Each one mirrors the reference's @_cdecl function (Sources/ArrowMetalC/ArrowMetalC.swift,
// ArrowMetalC_Options.swift, ArrowMetalC_GroupByExtra.swift): a NULL handle or out-pointer returns 2
// with no message; an operation error returns 1 and leaves its message in the thread-local
// am_last_error(); success returns 0 and writes a new reference to `*out`.
#include <cstring>
#include <vector>

#include "../../include/arrowmetal.h"
#include "am_internal.hpp"

namespace am {
const char* last_error_cstr();
Array* import_host(const ArrowSchema* schema, ArrowArray* array);
Array* import_device(const ArrowSchema* schema, ArrowDeviceArray* array);
bool chunks_supported(const ArrowSchema* schema);
Array* import_chunks(const ArrowSchema* schema, ArrowArray* arrays, int64_t n);
void export_host(Array* a, ArrowSchema* schema, ArrowArray* array);
void export_device(Array* a, ArrowSchema* schema, ArrowDeviceArray* array);
void batch_begin();
void batch_end();
void prof_enable(bool on);
void prof_reset();
std::string prof_report();
Array* slice(const Array& a, int64_t offset, int64_t length);
struct Batched;
Batched* batched_create(const int64_t* offsets, int64_t nbatches, int hash_capacity, int items);
void batched_release(Batched* b);
void batched_run(Batched* b, const Array* sort_keys, const Array* group_keys, const Array* values, int desc,
                 int forder, bool use_graph);
void batched_tune(Batched* b, const Array* sort_keys, const Array* group_keys, const Array* values, int desc,
                  int forder);
void batched_config(const Batched* b, int64_t* out, int count);
Array* batched_result(Batched* b, int which);
struct Server;
Server* server_create(int slots, int64_t budget_bytes);
void server_release(Server* s);
int64_t server_submit(Server* s, int op, Array* const* cols, int64_t ncols, Array* values, int desc, int placement,
                      int forder);
int64_t server_wait(Server* s, int64_t ticket, Array** out, int64_t cap, double* times);
void server_stats(Server* s, int64_t* out, int count);
}  // namespace am

using am::Array;

namespace {

inline Array* A(am_array* p) { return reinterpret_cast<Array*>(p); }
inline am_array* H(Array* p) { return reinterpret_cast<am_array*>(p); }
inline am::GroupBy* G(am_groupby* p) { return reinterpret_cast<am::GroupBy*>(p); }
inline am::Batched* B(am_batched* p) { return reinterpret_cast<am::Batched*>(p); }
inline const Array* OA(am_array* p) { return p ? reinterpret_cast<Array*>(p) : nullptr; }

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    am::set_last_error(e.what());
    return 1;
  } catch (...) {
    am::set_last_error("unknown error");
    return 1;
  }
}

thread_local std::string t_text;

}  // namespace

extern "C" {

const char* am_version(void) { return "0.1.0-mi355x"; }

const char* am_device_name(void) {
  try {
    t_text = am::device().name;
  } catch (const std::exception& e) {
    t_text = std::string("no ROCm device: ") + e.what();
  }
  return t_text.c_str();
}

const char* am_last_error(void) { return am::last_error_cstr(); }

int am_import(const struct ArrowSchema* schema, struct ArrowArray* array, am_array** out) {
  if (!schema || !array || !out) return 2;
  return guarded([&] { *out = H(am::import_host(schema, array)); });
}

int am_import_chunks(const struct ArrowSchema* schema, struct ArrowArray* arrays, int64_t n, am_array** out) {
  if (!schema || n < 0 || (n > 0 && !arrays) || !out) return 2;
  if (!am::chunks_supported(schema)) {
    am::set_last_error(std::string("Unsupported Arrow type: am_import_chunks does not take ") +
                       (schema->format ? schema->format : "(null format)") + " arrays; import their concatenation");
    return 3;
  }
  return guarded([&] { *out = H(am::import_chunks(schema, arrays, n)); });
}

int am_import_chunks_supported(const struct ArrowSchema* schema) { return am::chunks_supported(schema) ? 1 : 0; }

int am_import_device(const struct ArrowSchema* schema, struct ArrowDeviceArray* array, am_array** out) {
  if (!schema || !array || !out) return 2;
  return guarded([&] { *out = H(am::import_device(schema, array)); });
}

int am_export(am_array* a, struct ArrowSchema* schema, struct ArrowArray* array) {
  if (!a || !schema || !array) return 2;
  return guarded([&] { am::export_host(A(a), schema, array); });
}

int am_export_device(am_array* a, struct ArrowSchema* schema, struct ArrowDeviceArray* array) {
  if (!a || !schema || !array) return 2;
  return guarded([&] { am::export_device(A(a), schema, array); });
}

void am_release(am_array* a) {
  if (a) am::release_array(A(a));
}

int64_t am_length(am_array* a) { return a ? A(a)->length : -1; }

int64_t am_null_count(am_array* a) {
  if (!a) return -1;
  Array* x = A(a);
  if (x->null_count < 0) {
    try {
      am::sync();
      if (x->null_count < 0) x->null_count = am::count_nulls(*x);
    } catch (const std::exception& e) {
      am::set_last_error(e.what());
      return -1;
    }
  }
  return x->null_count;
}

const char* am_format(am_array* a) { return a ? A(a)->format.c_str() : nullptr; }
const char* am_compute_format(am_array* a) { return a ? A(a)->format.c_str() : nullptr; }

int am_take(am_array* a, am_array* indices, am_array** out) {
  if (!a || !indices || !out) return 2;
  return guarded([&] { *out = H(am::take(*A(a), *A(indices))); });
}

int am_argsort(am_array* a, int descending, am_array** out) {
  if (!a || !out) return 2;
  return guarded([&] { *out = H(am::argsort(*A(a), descending != 0, 0, 0)); });
}

int am_argsort_ex(am_array* a, int descending, int null_placement, am_array** out) {
  if (!a || !out) return 2;
  return guarded([&] { *out = H(am::argsort(*A(a), descending != 0, null_placement == 1 ? 1 : 0, 0)); });
}

int am_argsort_ex2(am_array* a, int descending, int null_placement, int float_order, am_array** out) {
  if (!a || !out) return 2;
  return guarded([&] {
    if (float_order < 0 || float_order > 2) am::fail_invalid("unknown float_order " + std::to_string(float_order));
    *out = H(am::argsort(*A(a), descending != 0, null_placement == 1 ? 1 : 0, float_order));
  });
}

namespace {
void check_float_order(int f) {
  if (f < 0 || f > 2) am::fail_invalid("unknown float_order " + std::to_string(f));
}
// Handles of `count` key columns, or nullptr when one is NULL.
bool all_handles(am_array** columns, int64_t count) {
  if (!columns || count <= 0) return false;
  for (int64_t i = 0; i < count; i++)
    if (!columns[i]) return false;
  return true;
}
}  // namespace

int am_sort(am_array* a, int descending, am_array** out) {
  if (!a || !out) return 2;
  return guarded([&] { *out = H(am::sort_values(*A(a), descending != 0, 0, 0)); });
}

int am_sort_ex(am_array* a, int descending, int null_placement, am_array** out) {
  if (!a || !out) return 2;
  return guarded([&] { *out = H(am::sort_values(*A(a), descending != 0, null_placement == 1 ? 1 : 0, 0)); });
}

int am_sort_ex2(am_array* a, int descending, int null_placement, int float_order, am_array** out) {
  if (!a || !out) return 2;
  return guarded([&] {
    check_float_order(float_order);
    *out = H(am::sort_values(*A(a), descending != 0, null_placement == 1 ? 1 : 0, float_order));
  });
}

int am_top_k(am_array* a, int64_t k, int largest, am_array** out) {
  if (!a || !out) return 2;
  return guarded([&] { *out = H(am::top_k(*A(a), k, largest != 0, 0, 0)); });
}

int am_top_k_ex(am_array* a, int64_t k, int largest, int null_placement, int float_order, am_array** out) {
  if (!a || !out) return 2;
  return guarded([&] {
    if (k < 0) am::fail_invalid("top_k needs k >= 0");
    check_float_order(float_order);
    *out = H(am::top_k(*A(a), k, largest != 0, null_placement == 1 ? 1 : 0, float_order));
  });
}

int am_lexsort(am_array** columns, const int* descending, int64_t count, am_array** out) {
  if (!all_handles(columns, count) || !out) return 2;
  return guarded([&] {
    *out = H(am::lexsort(reinterpret_cast<Array* const*>(columns), count, descending, nullptr, nullptr));
  });
}

int am_lexsort_ex(am_array** columns, const int* descending, int64_t count, int null_placement, am_array** out) {
  if (!all_handles(columns, count) || !out) return 2;
  return guarded([&] {
    std::vector<int> place(size_t(count), null_placement == 1 ? 1 : 0);
    *out = H(am::lexsort(reinterpret_cast<Array* const*>(columns), count, descending, place.data(), nullptr));
  });
}

int am_lexsort_ex2(am_array** columns, const int* descending, const int* null_placement, const int* float_order,
                   int64_t count, am_array** out) {
  if (!all_handles(columns, count) || !out) return 2;
  return guarded([&] {
    if (float_order)
      for (int64_t i = 0; i < count; i++) check_float_order(float_order[i]);
    *out = H(am::lexsort(reinterpret_cast<Array* const*>(columns), count, descending, null_placement, float_order));
  });
}

int am_group_by(am_array* keys, int64_t key_count, int agg, am_array* values, am_array** out) {
  if (!keys || !out) return 2;
  return guarded([&] { *out = H(am::group_by_dense(*A(keys), key_count, agg, values ? A(values) : nullptr)); });
}

int am_batch_begin(void) {
  return guarded([] { am::batch_begin(); });
}

int am_batch_end(void) {
  return guarded([] { am::batch_end(); });
}

int am_group_by_keys(am_array** columns, int64_t count, am_groupby** out) {
  if (!columns || !out || count <= 0) return 2;
  for (int64_t i = 0; i < count; i++)
    if (!columns[i]) return 2;
  return guarded([&] {
    *out = reinterpret_cast<am_groupby*>(am::group_by_keys(reinterpret_cast<Array* const*>(columns), count));
  });
}

int64_t am_group_by_group_count(am_groupby* gb) { return gb ? am::group_by_group_count(G(gb)) : -1; }

int am_group_by_keys_result(am_groupby* gb, int64_t i, am_array** out) {
  if (!gb || !out) return 2;
  return guarded([&] { *out = H(am::group_by_keys_result(G(gb), i)); });
}

int am_group_by_ids(am_groupby* gb, am_array** out) {
  if (!gb || !out) return 2;
  return guarded([&] { *out = H(am::group_by_ids(G(gb))); });
}

void am_group_by_release(am_groupby* gb) {
  if (gb) am::group_by_release(G(gb));
}

int am_group_agg_ex(am_groupby* gb, am_array* values, int op, double p1, am_array** out) {
  if (!gb || !out) return 2;
  return guarded([&] { *out = H(am::group_agg(G(gb), values ? A(values) : nullptr, op, p1)); });
}

int am_slice(am_array* a, int64_t offset, int64_t length, am_array** out) {
  if (!a || !out) return 2;
  return guarded([&] { *out = H(am::slice(*A(a), offset, length)); });
}

int am_batched_create(const int64_t* offsets, int64_t nbatches, int hash_capacity, int tile_items,
                      am_batched** out) {
  if (!offsets || !out) return 2;
  return guarded([&] {
    *out = reinterpret_cast<am_batched*>(am::batched_create(offsets, nbatches, hash_capacity, tile_items));
  });
}

int am_batched_tune(am_batched* b, am_array* sort_keys, am_array* group_keys, am_array* values, int descending,
                    int float_order) {
  if (!b) return 2;
  return guarded([&] { am::batched_tune(B(b), OA(sort_keys), OA(group_keys), OA(values), descending, float_order); });
}

int am_batched_run(am_batched* b, am_array* sort_keys, am_array* group_keys, am_array* values, int descending,
                   int float_order, int use_graph) {
  if (!b) return 2;
  return guarded([&] {
    am::batched_run(B(b), OA(sort_keys), OA(group_keys), OA(values), descending, float_order, use_graph != 0);
  });
}

int am_batched_config(am_batched* b, int64_t* out, int count) {
  if (!b || !out) return 2;
  return guarded([&] { am::batched_config(B(b), out, count); });
}

int am_batched_result(am_batched* b, int which, am_array** out) {
  if (!b || !out) return 2;
  return guarded([&] { *out = H(am::batched_result(B(b), which)); });
}

void am_batched_release(am_batched* b) {
  try {
    am::batched_release(B(b));
  } catch (...) {
  }
}

int am_server_create(int slots, int64_t budget_bytes, am_server** out) {
  if (!out) return 2;
  return guarded([&] { *out = reinterpret_cast<am_server*>(am::server_create(slots, budget_bytes)); });
}

int am_server_submit(am_server* s, int op, am_array** columns, int64_t count, am_array* values, int descending,
                     int null_placement, int float_order, int64_t* ticket) {
  if (!s || !columns || !ticket || count < 1) return 2;
  for (int64_t i = 0; i < count; i++)
    if (!columns[i]) return 2;
  return guarded([&] {
    *ticket = am::server_submit(reinterpret_cast<am::Server*>(s), op, reinterpret_cast<Array* const*>(columns),
                                count, values ? A(values) : nullptr, descending, null_placement, float_order);
  });
}

int am_server_wait(am_server* s, int64_t ticket, am_array** out, int64_t cap, double* times_ms, int64_t* count) {
  if (!s || (!out && cap > 0)) return 2;
  return guarded([&] {
    int64_t k = am::server_wait(reinterpret_cast<am::Server*>(s), ticket, reinterpret_cast<Array**>(out), cap,
                                times_ms);
    if (count) *count = k;
  });
}

int am_server_stats(am_server* s, int64_t* out, int count) {
  if (!s || !out) return 2;
  return guarded([&] { am::server_stats(reinterpret_cast<am::Server*>(s), out, count); });
}

void am_server_release(am_server* s) {
  try {
    am::server_release(reinterpret_cast<am::Server*>(s));
  } catch (...) {
  }
}

int am_unique_ex(am_array* a, int order, am_array** out) {
  if (!a || !out) return 2;
  return guarded([&] { *out = H(am::unique_values(*A(a), order == 1 ? 1 : 0)); });
}

int am_unique_rows(am_array** columns, int64_t count, am_array** out) {
  if (!columns || !out || count <= 0) return 2;
  for (int64_t i = 0; i < count; i++)
    if (!columns[i]) return 2;
  return guarded([&] { *out = H(am::unique_rows(reinterpret_cast<Array* const*>(columns), count)); });
}

// ---- the streaming group-by (reference Sources/ArrowMetalC/ArrowMetalC_Stream.swift:53-130, 259-281,
// 419-496). Errors go to the same thread-local text as am_last_error().

am_stream* am_stream_from_c_stream(struct ArrowArrayStream* in, int /*prefetch_depth*/) {
  if (!in) return nullptr;
  am_stream* s = nullptr;
  if (guarded([&] { s = reinterpret_cast<am_stream*>(am::stream_from_c(in)); }) != 0) return nullptr;
  return s;
}

void am_stream_release(am_stream* s) {
  if (s) am::stream_release(reinterpret_cast<am::Stream*>(s));
}

const char* am_stream_last_error(void) { return am::last_error_cstr(); }

int am_stream_group_by(am_stream* s, const char** keys, int64_t n_keys, const int* ops, const char** columns,
                       const char** names, int64_t n_aggs, int64_t dense_key_count, int /*ddof*/,
                       am_stream_result** out) {
  if (!s || !ops || !names || !out || n_keys <= 0 || n_aggs <= 0 || !keys) return 2;
  for (int64_t i = 0; i < n_keys; i++)
    if (!keys[i]) return 2;
  for (int64_t i = 0; i < n_aggs; i++)
    if (!names[i]) return 2;
  return guarded([&] {
    *out = reinterpret_cast<am_stream_result*>(am::stream_group_by(
        reinterpret_cast<am::Stream*>(s), keys, n_keys, ops, columns, names, n_aggs, dense_key_count));
  });
}

int64_t am_stream_result_column_count(am_stream_result* r) {
  return r ? am::stream_result_column_count(reinterpret_cast<am::StreamResult*>(r)) : 0;
}

const char* am_stream_result_column_name(am_stream_result* r, int64_t i) {
  if (!r || i < 0 || i >= am::stream_result_column_count(reinterpret_cast<am::StreamResult*>(r))) return nullptr;
  return am::stream_result_column_name(reinterpret_cast<am::StreamResult*>(r), i);
}

int am_stream_result_column(am_stream_result* r, int64_t i, am_array** out) {
  if (!r || !out || i < 0 || i >= am::stream_result_column_count(reinterpret_cast<am::StreamResult*>(r))) return 2;
  *out = H(am::stream_result_column(reinterpret_cast<am::StreamResult*>(r), i));
  return 0;
}

// The group-by terminal returns columns only.
int64_t am_stream_result_scalar_count(am_stream_result*) { return 0; }
const char* am_stream_result_scalar_name(am_stream_result*, int64_t) { return nullptr; }
int am_stream_result_scalar(am_stream_result*, int64_t, int64_t*, double*, int*, int*) { return 2; }

int64_t am_stream_result_rows_out(am_stream_result* r) {
  return r ? am::stream_result_rows_out(reinterpret_cast<am::StreamResult*>(r)) : 0;
}

int am_stream_result_stats(am_stream_result* r, int64_t* batches, int64_t* rows, int64_t* bytes_read, double* wall_s,
                           double* read_s, double* gpu_s, double* merge_s, double* overlap) {
  if (!r) return 2;
  am::stream_result_stats(reinterpret_cast<am::StreamResult*>(r), batches, rows, bytes_read, wall_s, read_s, gpu_s,
                          merge_s, overlap);
  return 0;
}

void am_stream_result_release(am_stream_result* r) {
  if (r) am::stream_result_release(reinterpret_cast<am::StreamResult*>(r));
}

// ---- am_query (reference Sources/ArrowMetalC/ArrowMetalC_Expr.swift:36-123)

int am_query(am_array** columns, const char** names, int64_t n_columns, const char* expr_text,
             am_query_result** out) {
  if (!columns || !names || !expr_text || !out || n_columns < 0) return 2;
  for (int64_t i = 0; i < n_columns; i++)
    if (!columns[i] || !names[i]) return 2;
  return guarded([&] {
    *out = reinterpret_cast<am_query_result*>(
        am::query_run(reinterpret_cast<Array* const*>(columns), names, n_columns, expr_text));
  });
}

int64_t am_query_column_count(am_query_result* r) {
  return r ? am::query_column_count(reinterpret_cast<am::QueryResult*>(r)) : -1;
}

const char* am_query_column_name(am_query_result* r, int64_t i) {
  if (!r || i < 0 || i >= am::query_column_count(reinterpret_cast<am::QueryResult*>(r))) return nullptr;
  return am::query_column_name(reinterpret_cast<am::QueryResult*>(r), i);
}

int am_query_column(am_query_result* r, int64_t i, am_array** out) {
  if (!r || !out || i < 0 || i >= am::query_column_count(reinterpret_cast<am::QueryResult*>(r))) return 2;
  *out = H(am::query_column(reinterpret_cast<am::QueryResult*>(r), i));
  return 0;
}

// The queries taken here end in a group_by terminal: no scalars.
int64_t am_query_scalar_count(am_query_result* r) { return r ? 0 : -1; }
const char* am_query_scalar_name(am_query_result*, int64_t) { return nullptr; }
int am_query_scalar(am_query_result*, int64_t, int64_t*, double*, int*, int*) { return 2; }

void am_query_result_release(am_query_result* r) {
  if (r) am::query_result_release(reinterpret_cast<am::QueryResult*>(r));
}

const char* am_query_canonical(const char* expr_text) {
  if (!expr_text) return nullptr;
  const char* text = nullptr;
  if (guarded([&] { text = am::query_canonical(expr_text); }) != 0) return nullptr;
  return text;
}

// ---- the plan engine (reference Sources/ArrowMetalC/ArrowMetalC_Plan.swift)

int am_plan_source_create(const char* name, am_array** columns, const char** names, int64_t n_columns,
                          am_plan_source** out) {
  if (!name || !columns || !names || !out || n_columns < 0) return 2;
  for (int64_t i = 0; i < n_columns; i++)
    if (!columns[i] || !names[i]) return 2;
  return guarded([&] {
    *out = reinterpret_cast<am_plan_source*>(
        am::plan_source_create(name, reinterpret_cast<Array* const*>(columns), names, n_columns));
  });
}

void am_plan_source_release(am_plan_source* s) {
  if (s) am::plan_source_release(reinterpret_cast<am::PlanSource*>(s));
}

int am_plan_run(const char* plan_json, am_plan_source** sources, int64_t n_sources, int optimize,
                am_plan_result** out) {
  if (!plan_json || !out || n_sources < 0 || (!sources && n_sources > 0)) return 2;
  for (int64_t i = 0; i < n_sources; i++)
    if (!sources[i]) return 2;
  return guarded([&] {
    *out = reinterpret_cast<am_plan_result*>(am::plan_run(
        plan_json, reinterpret_cast<am::PlanSource* const*>(sources), n_sources, optimize));
  });
}

const char* am_plan_explain(const char* plan_json, am_plan_source** sources, int64_t n_sources, int optimize) {
  if (!plan_json || n_sources < 0 || (!sources && n_sources > 0)) return nullptr;
  for (int64_t i = 0; i < n_sources; i++)
    if (!sources[i]) return nullptr;
  const char* text = nullptr;
  if (guarded([&] {
        text = am::plan_explain(plan_json, reinterpret_cast<am::PlanSource* const*>(sources), n_sources, optimize);
      }) != 0)
    return nullptr;
  return text;
}

int64_t am_plan_column_count(am_plan_result* r) {
  return r ? am::plan_column_count(reinterpret_cast<am::PlanResult*>(r)) : -1;
}

int64_t am_plan_row_count(am_plan_result* r) {
  return r ? am::plan_row_count(reinterpret_cast<am::PlanResult*>(r)) : -1;
}

const char* am_plan_column_name(am_plan_result* r, int64_t i) {
  if (!r || i < 0 || i >= am::plan_column_count(reinterpret_cast<am::PlanResult*>(r))) return nullptr;
  return am::plan_column_name(reinterpret_cast<am::PlanResult*>(r), i);
}

int am_plan_column(am_plan_result* r, int64_t i, am_array** out) {
  if (!r || !out || i < 0 || i >= am::plan_column_count(reinterpret_cast<am::PlanResult*>(r))) return 2;
  *out = H(am::plan_column(reinterpret_cast<am::PlanResult*>(r), i));
  return 0;
}

void am_plan_result_release(am_plan_result* r) {
  if (r) am::plan_result_release(reinterpret_cast<am::PlanResult*>(r));
}

int am_group_by_noalloc(am_array* keys, int64_t key_count, int agg, am_array* values, void* out_values,
                        uint64_t* out_validity, int64_t* out_null_count) {
  if (!keys || !out_values || !out_validity || !out_null_count) return 2;
  return guarded([&] {
    am::group_by_noalloc(*A(keys), key_count, agg, OA(values), out_values, out_validity, out_null_count);
  });
}

int64_t am_set_max_row_index(int64_t max_row) { return am::set_max_row_index(max_row); }

int am_prof_enable(int on) {
  am::prof_enable(on != 0);
  return 0;
}

int am_prof_reset(void) {
  return guarded([] { am::prof_reset(); });
}

int64_t am_prof_report(char* buf, int64_t cap) {
  std::string r;
  try {
    r = am::prof_report();
  } catch (...) {
    return -1;
  }
  if (buf && cap > 0) {
    size_t k = std::min<size_t>(size_t(cap - 1), r.size());
    std::memcpy(buf, r.data(), k);
    buf[k] = 0;
  }
  return int64_t(r.size() + 1);
}

}  // extern "C"
