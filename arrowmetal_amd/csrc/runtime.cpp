// Runtime of libArrowMetalC: device, streams, caching allocator, arrays, Arrow C Data import/export,
// errors, batching and kernel-time profiling.
//
// The reference's runtime (Sources/ArrowMetal/MetalContext.swift, MetalArrowBuffer.swift,
// CInterop.swift) wraps Apple unified memory, where an imported array is visible to the GPU without a
// copy. Here host memory and HBM are separate: am_import copies a CPU array into HBM once, and
// am_import_device takes ROCm device buffers without a copy. Work runs on one HIP stream per calling
// thread; every call that returns arrays is synchronous unless the thread has a batch open
// (MetalContext.swift:245-312). am_group_by_keys returns a handle while its last kernels still run:
// the calls that use the handle are ordered after them.
#include "am_internal.hpp"

#include <algorithm>
#include <cstdio>
#include <map>
#include <set>
#include <unordered_map>

namespace am {

// ---------------------------------------------------------------- errors

namespace {
thread_local std::string t_last_error;
}

void set_last_error(const std::string& msg) { t_last_error = msg; }
const char* last_error_cstr() { return t_last_error.c_str(); }

[[noreturn]] void fail(const std::string& msg) { throw Error(msg); }
[[noreturn]] void fail_invalid(const std::string& msg) { throw Error("Invalid ArrowArray: " + msg); }
[[noreturn]] void fail_unsupported(const std::string& msg) { throw Error("Unsupported Arrow type: " + msg); }
[[noreturn]] void fail_length_mismatch(int64_t a, int64_t b) {
  throw Error("Array length mismatch: " + std::to_string(a) + " vs " + std::to_string(b));
}

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(std::string("HIP error ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ") in " + what);
  }
}

// ---------------------------------------------------------------- types

int type_of_format(const std::string& f) {
  if (f == "c") return T_I8;
  if (f == "C") return T_U8;
  if (f == "s") return T_I16;
  if (f == "S") return T_U16;
  if (f == "i") return T_I32;
  if (f == "I") return T_U32;
  if (f == "l") return T_I64;
  if (f == "L") return T_U64;
  if (f == "f") return T_F32;
  if (f == "g") return T_F64;
  // Temporal types compute on their storage integers (Sources/ArrowMetal/Kernels/MultiSort.swift:76-113).
  if (f == "tdD" || f == "tts" || f == "ttm" || f == "tiM") return T_I32;
  if (f == "tdm" || f == "ttu" || f == "ttn") return T_I64;
  if (f.size() >= 4 && f[0] == 't' && f[1] == 's' && f[3] == ':') return T_I64;   // timestamps
  if (f.size() == 3 && f[0] == 't' && f[1] == 'D') return T_I64;                  // durations
  return T_UNSUPPORTED;
}

// ---------------------------------------------------------------- device

namespace {
std::once_flag g_dev_once;
DeviceInfo g_dev;
}

const DeviceInfo& device() {
  std::call_once(g_dev_once, [] {
    int id = 0;
    AM_HIP(hipGetDevice(&id));
    hipDeviceProp_t p;
    AM_HIP(hipGetDeviceProperties(&p, id));
    g_dev.id = id;
    g_dev.cu_count = p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
    g_dev.lds_per_block = p.sharedMemPerBlock;
    if (p.maxSharedMemoryPerMultiProcessor > 0) g_dev.lds_per_cu = p.maxSharedMemoryPerMultiProcessor;
    g_dev.name = std::string(p.name) + " (" + p.gcnArchName + ", " + std::to_string(g_dev.cu_count) + " CUs)";
  });
  return g_dev;
}

// ---------------------------------------------------------------- caching allocator

namespace {

// Free blocks are kept per stream. A block freed while kernels queued on its stream may still use
// it (a temporary dropped before the call's final sync) goes back to that stream's list, where only
// later work of the same stream, ordered after those kernels, can take it. Another stream takes it
// only once the owner stream has drained (hipStreamQuery), so concurrent threads never share a block
// that is still in use.
struct Pool {
  std::mutex mu;
  std::map<std::pair<hipStream_t, size_t>, std::vector<void*>> free_lists;   // (stream, class) -> blocks
  std::set<hipStream_t> live;   // streams of running threads; only these are ever queried
  size_t in_use = 0, cached = 0;
};
Pool& pool() {
  static Pool* p = new Pool();   // never destroyed: blocks may be freed during process exit
  return *p;
}

size_t size_class(size_t bytes) {
  if (bytes <= 4096) return (bytes + 255) & ~size_t(255);
  size_t p = 4096;
  while (p < bytes) {
    if (p + p / 2 >= bytes) return p + p / 2;
    p <<= 1;
  }
  return p;
}

void pool_free(void* ptr, size_t cls, hipStream_t owner) {
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  // A block that outlives its thread (a result released after the thread ended) is idle: its stream
  // was drained before it was destroyed.
  if (owner && !P.live.count(owner)) owner = nullptr;
  P.free_lists[{owner, cls}].push_back(ptr);
  P.in_use -= cls;
  P.cached += cls;
}

// A cached block of class `cls` that stream `s` may use now, or null. Caller holds the lock.
void* pool_take(Pool& P, hipStream_t s, size_t cls) {
  auto it = P.free_lists.find({s, cls});
  if (it == P.free_lists.end() || it->second.empty()) {
    it = P.free_lists.end();
    for (auto jt = P.free_lists.begin(); jt != P.free_lists.end(); ++jt) {
      if (jt->first.second != cls || jt->second.empty()) continue;
      hipStream_t other = jt->first.first;
      if (other && P.live.count(other)) {   // otherwise the stream of a finished thread, drained before it ended
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(other, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) continue;
        if (hipStreamQuery(other) != hipSuccess) {   // still running work that may use the block
          (void)hipGetLastError();
          continue;
        }
      }
      it = jt;
      break;
    }
    if (it == P.free_lists.end()) return nullptr;
  }
  void* ptr = it->second.back();
  it->second.pop_back();
  P.cached -= cls;
  P.in_use += cls;
  return ptr;
}

}  // namespace

void pool_register_stream(hipStream_t s) {
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  P.live.insert(s);
}

// A thread's stream is about to be destroyed: drain it and hand its free blocks to every stream.
void pool_retire_stream(hipStream_t s) {
  (void)hipStreamSynchronize(s);
  Pool& P = pool();
  std::lock_guard<std::mutex> g(P.mu);
  P.live.erase(s);
  for (auto it = P.free_lists.begin(); it != P.free_lists.end();) {
    if (it->first.first != s) {
      ++it;
      continue;
    }
    auto& dst = P.free_lists[{nullptr, it->first.second}];
    dst.insert(dst.end(), it->second.begin(), it->second.end());
    it = P.free_lists.erase(it);
  }
}

DevBlock::~DevBlock() {
  if (!ptr) return;
  if (on_free) on_free(ptr);
  else pool_free(ptr, bytes, owner);
}

DevBuf device_alloc(size_t bytes) {
  size_t cls = size_class(bytes ? bytes : 1);
  Pool& P = pool();
  hipStream_t s = stream();
  void* ptr = nullptr;
  {
    std::lock_guard<std::mutex> g(P.mu);
    ptr = pool_take(P, s, cls);
  }
  if (!ptr) {
    device();
    hipError_t e = hipMalloc(&ptr, cls);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      pool_trim();
      e = hipMalloc(&ptr, cls);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        fail("out of device memory allocating " + std::to_string(bytes) + " bytes");
      }
    }
    std::lock_guard<std::mutex> g(P.mu);
    P.in_use += cls;
  }
  auto b = std::make_shared<DevBlock>();
  b->ptr = ptr;
  b->bytes = cls;
  b->owner = s;
  retain(b);   // inside a batch, scratch outlives the queued kernels that use it
  return b;
}

DevBuf device_wrap(void* ptr, size_t bytes, std::function<void(void*)> release) {
  auto b = std::make_shared<DevBlock>();
  b->ptr = ptr;
  b->bytes = bytes;
  b->on_free = release ? std::move(release) : [](void*) {};
  return b;
}

size_t pool_bytes_in_use() {
  std::lock_guard<std::mutex> g(pool().mu);
  return pool().in_use;
}
size_t pool_bytes_cached() {
  std::lock_guard<std::mutex> g(pool().mu);
  return pool().cached;
}
void pool_trim() {
  Pool& P = pool();
  (void)hipDeviceSynchronize();   // blocks of other streams may still be read by queued kernels
  std::map<std::pair<hipStream_t, size_t>, std::vector<void*>> lists;
  {
    std::lock_guard<std::mutex> g(P.mu);
    lists.swap(P.free_lists);
    P.cached = 0;
  }
  for (auto& kv : lists)
    for (void* p : kv.second) (void)hipFree(p);
}

// ---------------------------------------------------------------- per-thread stream and batch

namespace {

struct ThreadState {
  hipStream_t s = nullptr;
  int batch_depth = 0;
  std::vector<DevBuf> retained;
  std::vector<std::function<void()>> deferred;
  ~ThreadState() {
    // Streams are left to the driver at process exit; a live thread's stream is reused.
    if (s && batch_depth == 0) {
      pool_retire_stream(s);
      (void)hipStreamDestroy(s);
    }
  }
};
thread_local ThreadState t_state;

// Profiling state.
struct ProfEntry { int64_t count = 0; double ms = 0; };
std::atomic<bool> g_prof_on{false};
std::mutex g_prof_mu;
std::map<std::string, ProfEntry> g_prof;
struct PendingTiming { const char* name; hipEvent_t a, b; };
thread_local std::vector<PendingTiming> t_pending;

void drain_prof() {
  for (auto& p : t_pending) {
    float ms = 0;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      std::lock_guard<std::mutex> g(g_prof_mu);
      auto& e = g_prof[p.name];
      e.count++;
      e.ms += ms;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  t_pending.clear();
}

}  // namespace

hipStream_t stream() {
  if (!t_state.s) {
    device();
    AM_HIP(hipStreamCreateWithFlags(&t_state.s, hipStreamNonBlocking));
    pool_register_stream(t_state.s);
  }
  return t_state.s;
}

bool in_batch() { return t_state.batch_depth > 0; }

void retain(const DevBuf& b) {
  if (t_state.batch_depth > 0 && b) t_state.retained.push_back(b);
}

void after_sync(std::function<void()> f) { t_state.deferred.push_back(std::move(f)); }

static void run_deferred() {
  auto fs = std::move(t_state.deferred);
  t_state.deferred.clear();
  std::string first_error;
  for (auto& f : fs) {
    try {
      f();
    } catch (const Error& e) {
      if (first_error.empty()) first_error = e.what();
    }
  }
  if (!first_error.empty()) fail(first_error);
}

void sync() {
  if (t_state.batch_depth > 0) return;
  AM_HIP(hipStreamSynchronize(stream()));
  if (!t_pending.empty()) drain_prof();
  run_deferred();
}

// ---------------------------------------------------------------- pinned read-backs

namespace {

__global__ void copy_words(const uint32_t* src, uint32_t* dst, uint32_t words) {
  for (uint32_t i = threadIdx.x; i < words; i += blockDim.x) dst[i] = src[i];
}

// Pinned slots are carved from chunks of 256 and never freed (a few KiB per thread that reads back).
struct PinnedSlots {
  std::mutex mu;
  std::vector<void*> free;
};
PinnedSlots& pinned_slots() {
  static PinnedSlots* p = new PinnedSlots();   // outlives every thread's slots
  return *p;
}

void check_read_back(const void* dev, size_t bytes) {
  if (bytes > PINNED_SLOT || bytes % 4 != 0 || reinterpret_cast<uintptr_t>(dev) % 4 != 0)
    fail("internal: read-back of " + std::to_string(bytes) + " bytes is not a pinned-slot copy");
}

}  // namespace

std::shared_ptr<void> pinned_slot() {
  PinnedSlots& P = pinned_slots();
  void* p = nullptr;
  {
    std::lock_guard<std::mutex> g(P.mu);
    if (P.free.empty()) {
      constexpr size_t CHUNK = 256;
      char* c = nullptr;
      AM_HIP(hipHostMalloc(reinterpret_cast<void**>(&c), CHUNK * PINNED_SLOT, hipHostMallocDefault));
      for (size_t i = 0; i < CHUNK; i++) P.free.push_back(c + i * PINNED_SLOT);
    }
    p = P.free.back();
    P.free.pop_back();
  }
  return std::shared_ptr<void>(p, [](void* q) {
    PinnedSlots& S = pinned_slots();
    std::lock_guard<std::mutex> g(S.mu);
    S.free.push_back(q);
  });
}

void copy_to_pinned(void* dst, const void* dev, size_t bytes) {
  check_read_back(dev, bytes);
  hipLaunchKernelGGL(copy_words, dim3(1), dim3(64), 0, stream(), static_cast<const uint32_t*>(dev),
                     static_cast<uint32_t*>(dst), uint32_t(bytes / 4));
  AM_HIP(hipGetLastError());
}

void read_back(void* host, const void* dev, size_t bytes) {
  thread_local std::shared_ptr<void> landing;   // one per thread: each read-back waits before returning
  if (!landing) landing = pinned_slot();
  copy_to_pinned(landing.get(), dev, bytes);
  AM_HIP(hipStreamSynchronize(stream()));
  std::memcpy(host, landing.get(), bytes);
}

void batch_begin() { t_state.batch_depth++; }

void batch_end() {
  if (t_state.batch_depth == 0) fail("am_batch_end without am_batch_begin");
  if (--t_state.batch_depth > 0) return;
  hipError_t e = hipStreamSynchronize(stream());
  t_state.retained.clear();
  if (!t_pending.empty()) drain_prof();
  AM_HIP(e);
  run_deferred();
}

ProfScope::ProfScope(const char* name) : name_(name) {
  if (!g_prof_on.load(std::memory_order_relaxed)) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(stream(), &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
  if (hipEventCreate(&start_) != hipSuccess) { start_ = nullptr; return; }
  (void)hipEventRecord(start_, stream());
}
ProfScope::~ProfScope() {
  if (!start_) return;
  hipEvent_t end = nullptr;
  if (hipEventCreate(&end) != hipSuccess) { (void)hipEventDestroy(start_); return; }
  (void)hipEventRecord(end, stream());
  t_pending.push_back({name_, start_, end});
}

void prof_enable(bool on) { g_prof_on.store(on); }
void prof_reset() {
  drain_prof();
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof.clear();
}
std::string prof_report() {
  drain_prof();
  std::lock_guard<std::mutex> g(g_prof_mu);
  std::string out;
  char line[256];
  for (auto& kv : g_prof) {
    snprintf(line, sizeof line, "%s %lld %.6f\n", kv.first.c_str(), (long long)kv.second.count, kv.second.ms);
    out += line;
  }
  return out;
}

// ---------------------------------------------------------------- arrays

Array* new_array(const std::string& format, int type, int64_t length) {
  static std::atomic<uint64_t> next_uid{1};
  Array* a = new Array();
  a->uid = next_uid.fetch_add(1, std::memory_order_relaxed);
  a->format = format;
  a->type = type;
  a->length = length;
  return a;
}
Array::~Array() {
  for (Array* c : children) release_array(c);
}
void retain_array(Array* a) { a->refs.fetch_add(1, std::memory_order_relaxed); }
void release_array(Array* a) {
  if (a && a->refs.fetch_sub(1, std::memory_order_acq_rel) == 1) delete a;
}

// A zero-copy window onto `a` (ArrowMetalC.swift:384-389): the buffers are shared, the offset moves.
Array* slice(const Array& a, int64_t offset, int64_t length) {
  if (offset < 0 || length < 0 || offset + length > a.length) fail_invalid("slice out of range");
  if (a.type == T_STRUCT) fail("am_slice of a struct array is not supported by this build");
  Array* out = new_array(a.format, a.type, length);
  out->values = a.values;
  out->offset = a.offset + offset;
  out->null_count = 0;
  if (a.has_nulls() && length > 0) {
    out->validity = a.validity;
    try {
      out->null_count = count_nulls(*out);
    } catch (...) {
      release_array(out);
      throw;
    }
    if (out->null_count == 0) out->validity.reset();
  }
  return out;
}

void check_length(int64_t n) {
  if (n > int64_t(UINT32_MAX)) fail_invalid("arrays above 2^32 elements are not supported yet");
}

namespace {
std::atomic<int64_t> g_max_row_index{int64_t(UINT32_MAX)};
}
int64_t max_row_index() { return g_max_row_index.load(std::memory_order_relaxed); }
int64_t set_max_row_index(int64_t v) {
  if (v < 0 || v > int64_t(UINT32_MAX)) v = int64_t(UINT32_MAX);
  return g_max_row_index.exchange(v);
}
void check_index_rows(int64_t rows, const char* op) {
  const int64_t last = rows - 1, top = max_row_index();
  if (last > top)
    fail_invalid(std::string(op) + ": row " + std::to_string(last) +
                 " does not fit the UInt32 index type (row numbers go up to " + std::to_string(top) + ")");
}

namespace {

__global__ void popcount_kernel(const uint8_t* bits, int64_t bit_off, int64_t n, unsigned long long* zeros) {
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  int64_t stride = int64_t(gridDim.x) * blockDim.x;
  unsigned long long z = 0;
  for (; i < n; i += stride) {
    int64_t b = bit_off + i;
    z += ((bits[b >> 3] >> (b & 7)) & 1) ? 0 : 1;
  }
  // Wave reduction, then one atomic per wave.
  for (int off = 32; off > 0; off >>= 1) z += __shfl_down(z, off, 64);
  if ((threadIdx.x & 63) == 0 && z) atomicAdd(zeros, z);
}

}  // namespace

int64_t count_nulls(const Array& a) {
  if (!a.validity || a.length == 0) return 0;
  auto cnt = device_alloc(8);
  hipStream_t s = stream();
  AM_HIP(hipMemsetAsync(cnt->ptr, 0, 8, s));
  int blocks = int(std::min<int64_t>((a.length + 255) / 256, 4096));
  {
    ProfScope ps("count_nulls");
    hipLaunchKernelGGL(popcount_kernel, dim3(blocks), dim3(256), 0, s, a.valid_bits(), a.offset, a.length,
                       static_cast<unsigned long long*>(cnt->ptr));
    AM_HIP(hipGetLastError());
  }
  return int64_t(read_back(static_cast<const unsigned long long*>(cnt->ptr)));
}

// ---------------------------------------------------------------- import / export

namespace {

// Copies bits [off, off + n) of `src` into a new LSB-first bitmap starting at bit 0.
std::vector<uint8_t> shift_bitmap(const uint8_t* src, int64_t off, int64_t n, int64_t* zeros) {
  std::vector<uint8_t> out(size_t((n + 7) / 8), 0);
  int64_t z = 0;
  if (off % 8 == 0) {
    std::memcpy(out.data(), src + off / 8, out.size());
    if (n % 8) out.back() &= uint8_t((1u << (n % 8)) - 1);
    for (uint8_t b : out) z += __builtin_popcount(b);
    z = n - z;
  } else {
    for (int64_t i = 0; i < n; i++) {
      int64_t b = off + i;
      int v = (src[b >> 3] >> (b & 7)) & 1;
      out[size_t(i >> 3)] |= uint8_t(v << (i & 7));
      z += !v;
    }
  }
  if (zeros) *zeros = z;
  return out;
}

struct ImportedArray {
  ArrowArray arr;   // the producer's array, moved here
  ~ImportedArray() {
    if (arr.release) arr.release(&arr);
  }
};

}  // namespace

Array* import_host(const ArrowSchema* schema, ArrowArray* array) {
  if (!schema || !array || !schema->format) fail("am_import: NULL schema or array");
  if (!array->release) fail("am_import: the array has already been released");
  std::string fmt = schema->format;
  int t = type_of_format(fmt);
  if (t == T_UNSUPPORTED) fail("am_import: Arrow format \"" + fmt + "\" is not supported by this build");
  if (array->n_buffers != 2) fail("am_import: a primitive array needs 2 buffers");
  // Move the producer's struct: its release is ours to call from here on.
  auto holder = std::make_shared<ImportedArray>();
  holder->arr = *array;
  array->release = nullptr;
  ArrowArray& A = holder->arr;
  int64_t n = A.length, off = A.offset;
  check_length(n);
  int w = type_width(t);
  Array* out = new_array(fmt, t, n);
  try {
    hipStream_t s = stream();
    out->values = device_alloc(size_t(std::max<int64_t>(n, 1)) * w);
    const uint8_t* vals = static_cast<const uint8_t*>(A.buffers[1]);
    if (n > 0) {
      if (!vals) fail("am_import: NULL values buffer");
      AM_HIP(hipMemcpyAsync(out->values->ptr, vals + off * w, size_t(n) * w, hipMemcpyHostToDevice, s));
    }
    const uint8_t* bits = static_cast<const uint8_t*>(A.buffers[0]);
    if (bits && A.null_count != 0 && n > 0) {
      int64_t zeros = 0;
      auto bm = shift_bitmap(bits, off, n, &zeros);
      out->null_count = zeros;
      if (zeros > 0) {
        out->validity = device_alloc(bm.size());
        AM_HIP(hipMemcpyAsync(out->validity->ptr, bm.data(), bm.size(), hipMemcpyHostToDevice, s));
      }
      AM_HIP(hipStreamSynchronize(s));   // bm is a host temporary
    } else {
      out->null_count = 0;
      AM_HIP(hipStreamSynchronize(s));
    }
  } catch (...) {
    release_array(out);
    throw;
  }
  return out;   // holder goes out of scope: the producer's release runs now, the data is copied
}

// ---- chunked import (reference include/arrowmetal.h:22-34, Sources/ArrowMetal/ChunkedImport.swift)

bool chunks_supported(const ArrowSchema* schema) {
  return schema && schema->format && type_is_primitive(type_of_format(schema->format));
}

namespace {
// ORs `len` bits of an aligned bitmap (bit 0 first, trailing bits clear) into `dst` from bit `pos`.
void or_bits_at(std::vector<uint8_t>& dst, int64_t pos, const std::vector<uint8_t>& src, int64_t len) {
  const int sh = int(pos & 7);
  const size_t at = size_t(pos >> 3), nbytes = size_t((len + 7) / 8);
  for (size_t i = 0; i < nbytes; i++) {
    const unsigned v = unsigned(src[i]) << sh;
    dst[at + i] |= uint8_t(v);
    if (v >> 8) dst[at + i + 1] |= uint8_t(v >> 8);
  }
}
}  // namespace

// n chunks of one primitive type as one array of their total length: each chunk's values (from its
// offset) are copied host to device into their place in one buffer, and the validity bitmap is built
// on the host at the chunks' bit positions (a chunk without a bitmap, or whose null_count is 0, is all
// valid; a null_count of -1 is counted) and copied once. Everything is checked before anything is
// moved; on success every chunk's release has run and is NULL.
Array* import_chunks(const ArrowSchema* schema, ArrowArray* arrays, int64_t n) {
  if (n == 1) return import_host(schema, &arrays[0]);
  const std::string fmt = schema->format;
  const int t = type_of_format(fmt);
  const int w = type_width(t);
  int64_t total = 0;
  bool any_nulls = false;
  for (int64_t c = 0; c < n; c++) {
    const ArrowArray& A = arrays[c];
    if (!A.release) fail_invalid("am_import_chunks: chunk " + std::to_string(c) + " has already been released");
    if (A.n_buffers != 2) fail_invalid("am_import_chunks: a primitive array needs 2 buffers");
    if (A.length < 0 || A.offset < 0) fail_invalid("am_import_chunks: negative length or offset");
    if (A.length > 0 && !A.buffers[1]) fail_invalid("am_import_chunks: NULL values buffer");
    total += A.length;
    any_nulls = any_nulls || (A.buffers[0] && A.null_count != 0 && A.length > 0);
  }
  check_length(total);
  Array* out = new_array(fmt, t, total);
  try {
    hipStream_t s = stream();
    out->values = device_alloc(size_t(std::max<int64_t>(total, 1)) * w);
    int64_t pos = 0;
    for (int64_t c = 0; c < n; c++) {
      const ArrowArray& A = arrays[c];
      if (A.length > 0)
        AM_HIP(hipMemcpyAsync(static_cast<uint8_t*>(out->values->ptr) + pos * w,
                              static_cast<const uint8_t*>(A.buffers[1]) + A.offset * w, size_t(A.length) * w,
                              hipMemcpyHostToDevice, s));
      pos += A.length;
    }
    out->null_count = 0;
    std::vector<uint8_t> bm;
    if (any_nulls) {
      bm.assign(size_t((total + 7) / 8) + 1, 0);
      int64_t zeros = 0;
      pos = 0;
      for (int64_t c = 0; c < n; c++) {
        const ArrowArray& A = arrays[c];
        if (A.length == 0) continue;
        std::vector<uint8_t> bits;
        if (A.buffers[0] && A.null_count != 0) {
          int64_t z = 0;
          bits = shift_bitmap(static_cast<const uint8_t*>(A.buffers[0]), A.offset, A.length, &z);
          zeros += z;
        } else {
          bits.assign(size_t((A.length + 7) / 8), 0xFF);
          if (A.length % 8) bits.back() = uint8_t((1u << (A.length % 8)) - 1u);
        }
        or_bits_at(bm, pos, bits, A.length);
        pos += A.length;
      }
      out->null_count = zeros;
      if (zeros > 0) {
        out->validity = device_alloc(size_t((total + 7) / 8));
        AM_HIP(hipMemcpyAsync(out->validity->ptr, bm.data(), size_t((total + 7) / 8), hipMemcpyHostToDevice, s));
      }
    }
    AM_HIP(hipStreamSynchronize(s));   // the chunks' buffers and bm are read until here
  } catch (...) {
    release_array(out);
    throw;
  }
  for (int64_t c = 0; c < n; c++) {   // moved: the data is copied, so the producers' buffers go now
    ArrowArray& A = arrays[c];
    A.release(&A);
    A.release = nullptr;
  }
  return out;
}

Array* import_device(const ArrowSchema* schema, ArrowDeviceArray* darr) {
  if (!schema || !darr || !schema->format) fail("am_import_device: NULL schema or array");
  if (darr->device_type == ARROW_DEVICE_CPU || darr->device_type == ARROW_DEVICE_ROCM_HOST)
    return import_host(schema, &darr->array);
  if (darr->device_type != ARROW_DEVICE_ROCM)
    fail("am_import_device: device type " + std::to_string(darr->device_type) +
         " is not ROCm (10) or CPU (1)");
  if (darr->device_id != device().id)
    fail("am_import_device: device " + std::to_string(darr->device_id) + " is not the current device " +
         std::to_string(device().id));
  ArrowArray* array = &darr->array;
  if (!array->release) fail("am_import_device: the array has already been released");
  std::string fmt = schema->format;
  int t = type_of_format(fmt);
  if (t == T_UNSUPPORTED) fail("am_import_device: Arrow format \"" + fmt + "\" is not supported by this build");
  if (array->n_buffers != 2) fail("am_import_device: a primitive array needs 2 buffers");
  if (darr->sync_event) AM_HIP(hipStreamWaitEvent(stream(), *static_cast<hipEvent_t*>(darr->sync_event), 0));
  auto holder = std::make_shared<ImportedArray>();
  holder->arr = *array;
  array->release = nullptr;
  ArrowArray& A = holder->arr;
  check_length(A.length);
  int w = type_width(t);
  Array* out = new_array(fmt, t, A.length);
  out->offset = A.offset;
  void* vals = const_cast<void*>(A.buffers[1]);
  out->values = device_wrap(vals, size_t(A.offset + A.length) * w, [holder](void*) {});
  void* bits = const_cast<void*>(A.buffers[0]);
  if (bits && A.null_count != 0 && A.length > 0) {
    out->validity = device_wrap(bits, size_t((A.offset + A.length + 7) / 8), [holder](void*) {});
    try {
      out->null_count = A.null_count >= 0 ? A.null_count : count_nulls(*out);
    } catch (...) {
      release_array(out);
      throw;
    }
    if (out->null_count == 0) out->validity.reset();
  } else {
    out->null_count = 0;
  }
  return out;
}

namespace {

struct ExportHolder {
  std::shared_ptr<std::vector<uint8_t>> values, validity;
  const void* buffers[2] = {nullptr, nullptr};
  Array* keep = nullptr;   // device export: the array whose buffers are handed out
  std::vector<ArrowArray> kids;
  std::vector<ArrowArray*> kid_ptrs;
  ~ExportHolder() {
    if (keep) release_array(keep);
  }
};

void release_export_array(ArrowArray* a) {
  if (!a || !a->release) return;
  auto* h = static_cast<ExportHolder*>(a->private_data);
  for (ArrowArray& k : h->kids)   // children not moved out by the consumer
    if (k.release) k.release(&k);
  delete h;
  a->release = nullptr;
}

struct SchemaHolder {
  std::string format, name;
  std::vector<ArrowSchema> kids;
  std::vector<ArrowSchema*> kid_ptrs;
};

void release_export_schema(ArrowSchema* s) {
  if (!s || !s->release) return;
  auto* h = static_cast<SchemaHolder*>(s->private_data);
  for (ArrowSchema& k : h->kids)
    if (k.release) k.release(&k);
  delete h;
  s->release = nullptr;
}

void fill_schema(const Array& a, const std::string& name, ArrowSchema* schema) {
  auto* h = new SchemaHolder();
  h->format = a.format;
  h->name = name;
  h->kids.resize(a.children.size());
  for (size_t i = 0; i < a.children.size(); i++) {
    fill_schema(*a.children[i], i < a.child_names.size() ? a.child_names[i] : std::string(), &h->kids[i]);
    h->kid_ptrs.push_back(&h->kids[i]);
  }
  std::memset(schema, 0, sizeof *schema);
  schema->format = h->format.c_str();
  schema->name = h->name.c_str();
  schema->metadata = nullptr;
  schema->flags = ARROW_FLAG_NULLABLE;
  schema->n_children = int64_t(h->kids.size());
  schema->children = h->kids.empty() ? nullptr : h->kid_ptrs.data();
  schema->dictionary = nullptr;
  schema->release = release_export_schema;
  schema->private_data = h;
}

// Copies the array's buffers to host memory once and keeps them with the array.
void cache_host_copy(Array* a) {
  if (a->host_values || a->type == T_STRUCT) return;
  int w = a->width();
  auto hv = std::make_shared<std::vector<uint8_t>>(size_t(std::max<int64_t>(a->length, 1)) * w);
  hipStream_t s = stream();
  if (a->length > 0) AM_HIP(hipMemcpyAsync(hv->data(), a->data(), size_t(a->length) * w, hipMemcpyDeviceToHost, s));
  std::shared_ptr<std::vector<uint8_t>> hb;
  if (a->has_nulls() && a->length > 0) {
    int64_t first_byte = a->offset / 8;
    int64_t nbytes = (a->offset + a->length + 7) / 8 - first_byte;
    std::vector<uint8_t> raw(static_cast<size_t>(nbytes));
    AM_HIP(hipMemcpyAsync(raw.data(), a->valid_bits() + first_byte, size_t(nbytes), hipMemcpyDeviceToHost, s));
    AM_HIP(hipStreamSynchronize(s));
    hb = std::make_shared<std::vector<uint8_t>>(shift_bitmap(raw.data(), a->offset % 8, a->length, nullptr));
  }
  AM_HIP(hipStreamSynchronize(s));
  a->host_values = hv;
  a->host_validity = hb;
}

// Fills `array` with the host copy (device == false) or the device buffers (device == true) of `a`,
// children included.
void fill_array(Array* a, bool device, ArrowArray* array) {
  auto* h = new ExportHolder();
  h->kids.resize(a->children.size());
  try {
    for (size_t i = 0; i < a->children.size(); i++) {
      fill_array(a->children[i], device, &h->kids[i]);
      h->kid_ptrs.push_back(&h->kids[i]);
    }
  } catch (...) {
    for (ArrowArray& k : h->kids)
      if (k.release) k.release(&k);
    delete h;
    throw;
  }
  const bool is_struct = a->type == T_STRUCT;
  std::memset(array, 0, sizeof *array);
  if (device) {
    retain_array(a);
    h->keep = a;
    h->buffers[0] = a->has_nulls() ? a->validity->ptr : nullptr;
    h->buffers[1] = a->values ? a->values->ptr : nullptr;
    array->null_count = a->has_nulls() ? a->null_count : 0;
    array->offset = a->offset;
  } else {
    cache_host_copy(a);
    h->values = a->host_values;
    h->validity = a->host_validity;
    h->buffers[0] = h->validity ? h->validity->data() : nullptr;
    h->buffers[1] = h->values ? h->values->data() : nullptr;
    array->null_count = h->validity ? a->null_count : 0;
    array->offset = 0;
  }
  array->length = a->length;
  array->n_buffers = is_struct ? 1 : 2;
  array->buffers = h->buffers;
  array->n_children = int64_t(h->kids.size());
  array->children = h->kids.empty() ? nullptr : h->kid_ptrs.data();
  array->dictionary = nullptr;
  array->release = release_export_array;
  array->private_data = h;
}

}  // namespace

void export_host(Array* a, ArrowSchema* schema, ArrowArray* array) {
  sync();   // a batch open on this thread completes before its results are read
  fill_array(a, false, array);
  fill_schema(*a, "", schema);
}

void export_device(Array* a, ArrowSchema* schema, ArrowDeviceArray* darr) {
  sync();
  std::memset(darr, 0, sizeof *darr);
  fill_array(a, true, &darr->array);
  fill_schema(*a, "", schema);
  darr->device_id = device().id;
  darr->device_type = ARROW_DEVICE_ROCM;
  darr->sync_event = nullptr;
}

// ---------------------------------------------------------------- tuning table

SortTune sort_tune(int64_t n, int key_bytes) {
  // Tiles of 256 threads x `items` keys. Small inputs take short tiles so that the grid still covers
  // every CU; large ones take long tiles, which amortise the per-tile look-back over more keys.
  // ARROWMETAL_SORT_ITEMS (4, 8, 12, 16 or 20) overrides the choice, for tuning sweeps.
  SortTune t;
  static const int forced = [] {
    const char* e = getenv("ARROWMETAL_SORT_ITEMS");
    int v = e ? atoi(e) : 0;
    return (v == 4 || v == 8 || v == 12 || v == 16 || v == 20) ? v : 0;
  }();
  if (forced) {
    t.items = forced;
    return t;
  }
  int64_t cus = device().cu_count;
  if (n <= cus * 256 * 4) t.items = 4;
  else if (n <= cus * 256 * 8 * 2) t.items = 8;
  else if (n <= cus * 256 * 16 * 8) t.items = 16;
  else t.items = 20;
  (void)key_bytes;
  return t;
}

}  // namespace am
