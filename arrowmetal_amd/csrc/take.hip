// Gather (Arrow `take`) for CDNA4: out[i] = values[indices[i]].
//
// Semantics follow the reference (Sources/ArrowMetal/Kernels/Take.swift:41-80,
// KernelSource.swift:209-227): a null index gives a null output, a null source element gives a null,
// and an index outside [0, length) makes the call fail with "take: index out of range (array length
// N)" once the kernel has run. Each wave handles 64 consecutive outputs, so the output validity is one
// 64-bit ballot per wave, written as one word, and the null count is one atomic per wave.
#include "am_internal.hpp"

#include <algorithm>

namespace am {
namespace {

struct TakeStatus {
  unsigned int out_of_range;
  unsigned int pad;
  unsigned long long nulls;
};

template <typename V, typename I>
__global__ __launch_bounds__(256) void take_kernel(const V* __restrict__ vals, const uint8_t* __restrict__ vvalid,
                                                   int64_t voff, int64_t len, const I* __restrict__ idx,
                                                   const uint8_t* __restrict__ ivalid, int64_t ioff, int64_t n,
                                                   V* __restrict__ out, unsigned long long* __restrict__ out_valid,
                                                   TakeStatus* st) {
  const int lane = threadIdx.x & 63;
  int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t stride = int64_t(gridDim.x) * blockDim.x;
  const bool want_validity = out_valid != nullptr;
  // Every wave walks whole 64-row groups, so a group's ballot is one word of the output bitmap.
  for (; i - lane < n; i += stride) {
    bool valid = false;
    if (i < n) {
      bool iv = true;
      if (ivalid) {
        int64_t b = ioff + i;
        iv = (ivalid[b >> 3] >> (b & 7)) & 1;
      }
      V v = V(0);
      if (iv) {
        int64_t j = int64_t(idx[i]);
        if (j < 0 || j >= len) {
          atomicOr(&st->out_of_range, 1u);
        } else {
          v = vals[j];
          valid = true;
          if (vvalid) {
            int64_t b = voff + j;
            valid = (vvalid[b >> 3] >> (b & 7)) & 1;
          }
        }
      }
      out[i] = v;
    }
    if (want_validity) {
      unsigned long long bal = __ballot(valid);
      if (lane == 0) {
        int64_t rows = std::min<int64_t>(64, n - (i - lane));
        unsigned long long mask = rows >= 64 ? ~0ull : ((1ull << rows) - 1ull);
        out_valid[(i - lane) >> 6] = bal & mask;
        unsigned long long z = uint64_t(rows) - uint64_t(__popcll(bal & mask));
        if (z) atomicAdd(&st->nulls, z);
      }
    }
  }
}

template <typename V>
void launch_take(const Array& a, const Array& ix, Array* out, TakeStatus* st, hipStream_t s) {
  int64_t n = ix.length;
  int blocks = int(std::min<int64_t>((n + 255) / 256, int64_t(device().cu_count) * 16));
  blocks = std::max(blocks, 1);
  const V* vals = static_cast<const V*>(a.data());
  const uint8_t* vv = a.has_nulls() ? a.valid_bits() : nullptr;
  const uint8_t* iv = ix.has_nulls() ? ix.valid_bits() : nullptr;
  unsigned long long* ov = out->validity ? static_cast<unsigned long long*>(out->validity->ptr) : nullptr;
  V* o = static_cast<V*>(out->values->ptr);
  ProfScope ps("take");
  switch (ix.type) {
    case T_I32:
      hipLaunchKernelGGL((take_kernel<V, int32_t>), dim3(blocks), dim3(256), 0, s, vals, vv, a.offset, a.length,
                         static_cast<const int32_t*>(ix.data()), iv, ix.offset, n, o, ov, st);
      break;
    case T_U32:
      hipLaunchKernelGGL((take_kernel<V, uint32_t>), dim3(blocks), dim3(256), 0, s, vals, vv, a.offset, a.length,
                         static_cast<const uint32_t*>(ix.data()), iv, ix.offset, n, o, ov, st);
      break;
    default:
      hipLaunchKernelGGL((take_kernel<V, int64_t>), dim3(blocks), dim3(256), 0, s, vals, vv, a.offset, a.length,
                         static_cast<const int64_t*>(ix.data()), iv, ix.offset, n, o, ov, st);
      break;
  }
  AM_HIP(hipGetLastError());
}

__global__ void take_status_init(TakeStatus* st) {
  st->out_of_range = 0;
  st->pad = 0;
  st->nulls = 0;
}

}  // namespace

Array* take(const Array& a, const Array& ix) {
  check_length(a.length);
  check_length(ix.length);
  if (ix.type != T_I32 && ix.type != T_I64 && ix.type != T_U32)
    fail_unsupported("take indices must be int32, int64 or uint32");
  int w = a.width();
  if (w == 0) fail_unsupported("take of " + a.format + " is not supported by this build");
  int64_t n = ix.length;
  Array* out = new_array(a.format, a.type, n);
  try {
    hipStream_t s = stream();
    out->values = device_alloc(size_t(std::max<int64_t>(n, 1)) * w);
    bool need_validity = a.has_nulls() || ix.has_nulls();
    if (need_validity) out->validity = device_alloc(size_t((std::max<int64_t>(n, 1) + 63) / 64) * 8);
    DevBuf stb = device_alloc(sizeof(TakeStatus));
    auto* st = static_cast<TakeStatus*>(stb->ptr);
    hipLaunchKernelGGL(take_status_init, dim3(1), dim3(1), 0, s, st);
    if (n > 0) {
      switch (w) {
        case 1: launch_take<uint8_t>(a, ix, out, st, s); break;
        case 2: launch_take<uint16_t>(a, ix, out, st, s); break;
        case 4: launch_take<uint32_t>(a, ix, out, st, s); break;
        default: launch_take<unsigned long long>(a, ix, out, st, s); break;
      }
    }
    std::shared_ptr<void> slot = pinned_slot();   // the status lands in pinned memory
    copy_to_pinned(slot.get(), st, sizeof(TakeStatus));
    int64_t len = a.length;
    Array* res = out;
    retain_array(res);   // the check below may run at the end of a batch
    after_sync([slot, len, res, need_validity] {
      struct Drop {
        Array* a;
        ~Drop() { release_array(a); }
      } drop{res};
      const TakeStatus* host = static_cast<const TakeStatus*>(slot.get());
      if (host->out_of_range) fail_invalid("take: index out of range (array length " + std::to_string(len) + ")");
      res->null_count = need_validity ? int64_t(host->nulls) : 0;
      if (res->null_count == 0) res->validity.reset();
    });
    if (!in_batch()) {
      sync();
    } else {
      // Inside a batch the null count is not known until the batch ends; the validity is kept.
      out->null_count = need_validity ? -1 : 0;
    }
  } catch (...) {
    release_array(out);
    throw;
  }
  return out;
}

}  // namespace am
