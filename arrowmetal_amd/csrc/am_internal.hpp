// Internal declarations of libArrowMetalC (MI355X / gfx950 build).
//
// Layout of the library:
//   runtime.cpp   device context, per-thread HIP streams, the caching device allocator, the am_array
//                 handle, Arrow C Data / C Device Data import and export, errors, batching, profiling
//   sort.hip      the stable LSD radix argsort (onesweep: one histogram sweep, then one scatter
//                 kernel per 8-bit digit with a decoupled look-back) and its key maps
//   take.hip      the gather behind am_take
//   groupby.hip   the hash group-by: key table build, ascending relabel, grouped aggregates
//   select.hip    sorted values, top-k by radix select, lexsort
//   plan.cpp      the plan engine's scan / sort / group_by / unique / limit subset (am_plan_*)
//   query.cpp     am_query's fused dense group-by subset
//   stream.cpp    the streaming group-by (am_stream_from_c_stream, am_stream_group_by)
//   batched.hip   the batched plan (am_batched_*); serve.cpp the request server (am_server_*)
//   capi.cpp      the extern "C" am_* entry points of include/arrowmetal.h
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/arrow_abi.h"

namespace am {

// ---------------------------------------------------------------- errors

// An operation error: the entry point returns 1 and the message goes to am_last_error().
struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};
[[noreturn]] void fail(const std::string& msg);
// The reference's error texts (ArrowMetalError.description, Sources/ArrowMetal/MetalContext.swift:19-31):
// "Invalid ArrowArray: <msg>", "Unsupported Arrow type: <msg>", "Array length mismatch: <a> vs <b>".
[[noreturn]] void fail_invalid(const std::string& msg);
[[noreturn]] void fail_unsupported(const std::string& msg);
[[noreturn]] void fail_length_mismatch(int64_t a, int64_t b);
void hip_check(hipError_t e, const char* what);
#define AM_HIP(x) ::am::hip_check((x), #x)

void set_last_error(const std::string& msg);

// ---------------------------------------------------------------- element types

// Type codes shared by every kernel family (and by the CPU oracle's tests).
enum TypeCode : int {
  T_I8 = 0, T_I16 = 1, T_I32 = 2, T_I64 = 3,
  T_U8 = 4, T_U16 = 5, T_U32 = 6, T_U64 = 7,
  T_F32 = 8, T_F64 = 9,
  T_BOOL = 10,
  T_STRUCT = 11,   // a result-only struct of child arrays (hash_min_max)
  T_UNSUPPORTED = -1
};
__host__ __device__ inline int type_width(int t) {
  switch (t) {
    case T_I8: case T_U8: return 1;
    case T_I16: case T_U16: return 2;
    case T_I32: case T_U32: case T_F32: return 4;
    case T_I64: case T_U64: case T_F64: return 8;
    default: return 0;
  }
}
// The XCD (accelerator complex die) a wave runs on, from the XCC_ID hardware register. Each XCD has
// its own L2: memory private to one XCD can be updated with workgroup-scope atomics, which the XCD's
// L2 executes, instead of device-scope ones, which must go to memory.
constexpr int NXCD = 8;
__device__ __forceinline__ uint32_t xcd_id() {
  return uint32_t(__builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20)) & (NXCD - 1);
}

// A load through a pointer known to address global memory (device buffers, __device__ variables).
// A pointer chosen between two buffers (a column or a constant, a partition or a column) is a generic
// pointer to the compiler, and a generic (flat) load counts against both the vector-memory and the
// LDS counters, so waiting for it also waits for every LDS operation in flight and the reverse.
template <typename T>
__device__ __forceinline__ T load_global(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

__host__ __device__ inline bool type_is_float(int t) { return t == T_F32 || t == T_F64; }
__host__ __device__ inline bool type_is_signed(int t) { return t >= T_I8 && t <= T_I64; }
__host__ __device__ inline bool type_is_int(int t) { return t >= T_I8 && t <= T_U64; }
__host__ __device__ inline bool type_is_primitive(int t) { return t >= T_I8 && t <= T_F64; }
// The storage type of an Arrow format string: primitives, and temporal types as their integers.
int type_of_format(const std::string& fmt);

// ---------------------------------------------------------------- device memory

// One device allocation, returned to the caching pool (or to its producer) when the last reference
// goes. `bytes` is the usable size.
struct DevBlock {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipStream_t owner = nullptr;          // stream of the thread that took the block from the pool
  std::function<void(void*)> on_free;   // empty: the block belongs to the pool
  ~DevBlock();
};
using DevBuf = std::shared_ptr<DevBlock>;

// Device memory from the library's caching pool (size classes of 2^k and 1.5 * 2^k above 4 KiB).
// The pool never hands a block back to the driver while the process runs, so steady-state calls
// allocate nothing from HIP.
DevBuf device_alloc(size_t bytes);
// A view of memory owned elsewhere (zero-copy device import); `release` runs when the view dies.
DevBuf device_wrap(void* ptr, size_t bytes, std::function<void(void*)> release);
// Bytes the pool holds (in use + cached), for admission control and tests.
size_t pool_bytes_in_use();
size_t pool_bytes_cached();
void pool_trim();

// ---------------------------------------------------------------- streams and devices

struct DeviceInfo {
  int id = 0;
  int cu_count = 256;
  size_t lds_per_block = 65536;
  size_t lds_per_cu = 163840;
  std::string name;
};
const DeviceInfo& device();

// The calling thread's stream (created on first use, default priority, non-blocking).
hipStream_t stream();
// Every public call is synchronous unless the thread is inside am_batch_begin / am_batch_end:
// `sync()` waits for the thread's stream unless a batch is open.
void sync();
bool in_batch();
// Keeps `b` alive until the thread's open batch ends (no-op outside a batch).
void retain(const DevBuf& b);
// Runs `f` after the stream has drained at the next sync point (deferred GPU-side error flags).
void after_sync(std::function<void()> f);

// Small read-backs (status words, counts: at most PINNED_SLOT bytes, a multiple of 4, 4-byte aligned).
// A one-wave kernel copies the words into pinned host memory, which the host reads once the stream has
// drained: a copy into pageable memory would be staged and costs ~10 us more per read-back.
constexpr size_t PINNED_SLOT = 256;
// Copies `bytes` of device memory at `dev` to `host` and waits for the thread's stream.
void read_back(void* host, const void* dev, size_t bytes);
template <class T>
T read_back(const T* dev) {
  T v;
  read_back(&v, dev, sizeof(T));
  return v;
}
// A pinned slot of PINNED_SLOT bytes (returned to a shared free list when released), for a read-back
// whose check is deferred to the next sync point.
std::shared_ptr<void> pinned_slot();
// Queues the copy of `bytes` of device memory at `dev` into the pinned slot `dst` on the thread's stream.
void copy_to_pinned(void* dst, const void* dev, size_t bytes);

// Kernel-time profiling of the library's own launches (extension API am_prof_*): when enabled,
// each named launch is bracketed with HIP events on the launching stream.
struct ProfScope {
  ProfScope(const char* name);
  ~ProfScope();
  const char* name_;
  hipEvent_t start_ = nullptr;
};

// ---------------------------------------------------------------- arrays

struct Array {
  std::atomic<int> refs{1};
  uint64_t uid = 0;        // unique per array for the life of the process (cache keys)
  std::string format;      // Arrow C data format string, as imported
  int type = T_UNSUPPORTED;
  int64_t length = 0;
  int64_t offset = 0;      // Arrow offset, in elements, carried rather than applied
  int64_t null_count = 0;
  DevBuf values;           // device memory of element 0 (before the offset)
  DevBuf validity;         // device bitmap, bit (offset + i) for element i; null when no nulls
  // Cached host copy of the exported buffers (so repeated am_export calls copy once).
  std::shared_ptr<std::vector<uint8_t>> host_values, host_validity;
  // A struct ("+s") holds one reference to each child; it has no values buffer of its own.
  std::vector<Array*> children;
  std::vector<std::string> child_names;
  ~Array();

  int width() const { return type_width(type); }
  const void* data() const {   // element `offset`
    return values ? static_cast<const char*>(values->ptr) + offset * width() : nullptr;
  }
  void* mutable_data() { return values ? static_cast<char*>(values->ptr) + offset * width() : nullptr; }
  const uint8_t* valid_bits() const { return validity ? static_cast<const uint8_t*>(validity->ptr) : nullptr; }
  bool has_nulls() const { return validity && null_count != 0; }
};

Array* new_array(const std::string& format, int type, int64_t length);   // refs = 1
void retain_array(Array* a);
void release_array(Array* a);
// Counts the unset bits of the array's validity, on the device.
int64_t count_nulls(const Array& a);

// ---------------------------------------------------------------- limits

// Row numbers are uint32 ("I"): every index-returning call refuses more than 2^32 rows
// (Sources/ArrowMetal/Kernels/Dispatch.swift:186-203). The largest row number an index array may hold
// is 2^32 - 1 unless a test lowers it (the reference's Dispatch.maxRowIndex, :190-196), so that the
// refusal is reachable at small sizes.
void check_length(int64_t n);
void check_index_rows(int64_t rows, const char* op);
int64_t max_row_index();
int64_t set_max_row_index(int64_t v);   // returns the previous limit; out-of-range v restores 2^32 - 1

// ---------------------------------------------------------------- operations

// sort.hip
Array* argsort(const Array& a, bool descending, int null_placement, int float_order);
// Device-level argsort on raw buffers, shared by the group-by relabel and the batched path.
// `vals` points at element 0 of `n` elements of `type`; validity bit (vbit_off + i) or null.
void argsort_device(const void* vals, int type, const uint8_t* validity, int64_t vbit_off, int64_t n,
                    bool descending, int null_placement, int float_order, uint32_t* out_idx,
                    hipStream_t s);

// ORDER BY of a whole table by one 64-bit key (int64 / uint64 / float64): every column of `cols`
// (each of n rows) in the key's stable order with am_argsort_ex2's options, as new arrays in `out`.
// The columns move with the hybrid sort's top-digit partition instead of one gather across the table.
// Returns false, with nothing in `out`, when the key does not take the hybrid sort (fewer than 2^18
// rows, a constant top digit, a full bucket list, a graph capture); the caller then sorts and takes.
bool order_by_columns(const Array& key, bool descending, int null_placement, int float_order,
                      const std::vector<const Array*>& cols, std::vector<Array*>& out);

// take.hip
Array* take(const Array& a, const Array& indices);

// select.hip
Array* top_k(const Array& a, int64_t k, bool largest, int placement, int forder);
Array* sort_values(const Array& a, bool descending, int placement, int forder);
Array* lexsort(Array* const* cols, int64_t count, const int* desc, const int* placement, const int* forder);

// groupby.hip
struct GroupBy;
GroupBy* group_by_keys(Array* const* columns, int64_t count);
int64_t group_by_group_count(const GroupBy* g);
Array* group_by_keys_result(GroupBy* g, int64_t i);
Array* group_by_ids(GroupBy* g);
Array* group_agg(GroupBy* g, const Array* values, int op, double p1);
void group_by_release(GroupBy* g);
Array* group_by_dense(const Array& keys, int64_t key_count, int agg, const Array* values);
// group_by_dense with every key offset by `bias` before the range check, and the integer values
// transformed before they are aggregated: xop 0 none, 1 shift right by xarg (arithmetic for signed
// columns), 2 bitwise and with xarg.
Array* group_by_dense_ex(const Array& keys, int64_t key_count, long long bias, int agg, const Array* values, int xop,
                         long long xarg);
// group_by_dense's integer aggregates into caller device buffers, allocating nothing (a fixed static
// device workspace; key_count <= 65536).
void group_by_noalloc(const Array& keys, int64_t key_count, int agg, const Array* values, void* out_values,
                      uint64_t* out_validity, int64_t* out_null_count);
Array* unique_rows(Array* const* columns, int64_t count);
Array* unique_values(const Array& a, int order);

// plan.cpp: the plan engine's scan / sort / group_by / unique / limit subset.
struct PlanSource;
struct PlanResult;
PlanSource* plan_source_create(const char* name, Array* const* cols, const char* const* names, int64_t n);
void plan_source_release(PlanSource* s);
PlanResult* plan_run(const char* text, PlanSource* const* sources, int64_t n, int optimize);
const char* plan_explain(const char* text, PlanSource* const* sources, int64_t n, int optimize);
int64_t plan_column_count(const PlanResult* r);
int64_t plan_row_count(const PlanResult* r);
const char* plan_column_name(const PlanResult* r, int64_t i);
Array* plan_column(const PlanResult* r, int64_t i);
void plan_result_release(PlanResult* r);

// query.cpp: am_query's fused dense group-by subset.
struct QueryResult;
QueryResult* query_run(Array* const* cols, const char* const* names, int64_t n, const char* text);
const char* query_canonical(const char* text);
int64_t query_column_count(const QueryResult* r);
const char* query_column_name(const QueryResult* r, int64_t i);
Array* query_column(const QueryResult* r, int64_t i);
void query_result_release(QueryResult* r);

// stream.cpp: the streaming group-by over a producer of record batches.
struct Stream;
struct StreamResult;
Stream* stream_from_c(ArrowArrayStream* in);
void stream_release(Stream* s);
StreamResult* stream_group_by(Stream* s, const char* const* keys, int64_t n_keys, const int* ops,
                              const char* const* columns, const char* const* names, int64_t n_aggs,
                              int64_t dense_key_count);
int64_t stream_result_column_count(const StreamResult* r);
const char* stream_result_column_name(const StreamResult* r, int64_t i);
Array* stream_result_column(const StreamResult* r, int64_t i);
int64_t stream_result_rows_out(const StreamResult* r);
void stream_result_stats(const StreamResult* r, int64_t* batches, int64_t* rows, int64_t* bytes, double* wall,
                         double* read, double* gpu, double* merge, double* overlap);
void stream_result_release(StreamResult* r);
// Imports a host array (moving it), as am_import does.
Array* import_host(const ArrowSchema* schema, ArrowArray* array);

// Launch-shape parameters the autotuner chooses (runtime.cpp keeps the table).
struct SortTune {
  int items = 16;        // keys per thread per scatter tile (tile = 256 * items)
};
SortTune sort_tune(int64_t n, int key_bytes);

}  // namespace am
