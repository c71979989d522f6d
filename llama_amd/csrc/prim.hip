// Device-wide scan and sort plumbing (rocPRIM through hipCUB).  These carry no
// graph logic: they turn per-vertex / per-block counts into offsets and order
// key arrays for the snapshot and triangle-counting preprocessing.
#include <hipcub/hipcub.hpp>

#include "ll_gpu_internal.h"

namespace llgpu {

namespace {

struct WidenU32 {
  __host__ __device__ uint64_t operator()(uint32_t x) const { return uint64_t(x); }
};

__global__ void k_write_total_u32(const uint32_t* in, uint64_t* out, uint64_t n) {
  out[n] = n ? out[n - 1] + uint64_t(in[n - 1]) : 0;
}
__global__ void k_write_total_u64(const uint64_t* in, uint64_t* out, uint64_t n) {
  out[n] = n ? out[n - 1] + in[n - 1] : 0;
}

}  // namespace

void scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint64_t n, hipStream_t s) {
  if (n > 0) {
    hipcub::TransformInputIterator<uint64_t, WidenU32, const uint32_t*> it(in, WidenU32());
    size_t tmp = 0;
    LLG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, it, out, n, s));
    DevBuf t;
    t.alloc(tmp, s);
    LLG_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, it, out, n, s));
  }
  k_write_total_u32<<<1, 1, 0, s>>>(in, out, n);
  LLG_HIP(hipGetLastError());
}

void scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t s) {
  if (n > 0) {
    size_t tmp = 0;
    LLG_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n, s));
    DevBuf t;
    t.alloc(tmp, s);
    LLG_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, tmp, in, out, n, s));
  }
  k_write_total_u64<<<1, 1, 0, s>>>(in, out, n);
  LLG_HIP(hipGetLastError());
}

void sort_pairs_u64_u32(const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in,
                        uint32_t* vals_out, uint64_t n, int end_bit, hipStream_t s) {
  if (n == 0) return;
  size_t tmp = 0;
  LLG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys_in, keys_out, vals_in, vals_out,
                                             n, 0, end_bit, s));
  DevBuf t;
  t.alloc(tmp, s);
  LLG_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, keys_in, keys_out, vals_in, vals_out, n,
                                             0, end_bit, s));
}

void sort_pairs_u32_u32(const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in,
                        uint32_t* vals_out, uint64_t n, int end_bit, hipStream_t s) {
  if (n == 0) return;
  size_t tmp = 0;
  LLG_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys_in, keys_out, vals_in, vals_out,
                                             n, 0, end_bit, s));
  DevBuf t;
  t.alloc(tmp, s);
  LLG_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, tmp, keys_in, keys_out, vals_in, vals_out, n,
                                             0, end_bit, s));
}

void sort_keys_u64(const uint64_t* keys_in, uint64_t* keys_out, uint64_t n, int end_bit,
                   hipStream_t s) {
  if (n == 0) return;
  size_t tmp = 0;
  LLG_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp, keys_in, keys_out, n, 0, end_bit, s));
  DevBuf t;
  t.alloc(tmp, s);
  LLG_HIP(hipcub::DeviceRadixSort::SortKeys(t.p, tmp, keys_in, keys_out, n, 0, end_bit, s));
}

uint64_t unique_u64(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t s) {
  if (n == 0) return 0;
  DevBuf cnt;
  cnt.alloc(sizeof(uint64_t), s);
  size_t tmp = 0;
  LLG_HIP(hipcub::DeviceSelect::Unique(nullptr, tmp, in, out, cnt.as<uint64_t>(), n, s));
  DevBuf t;
  t.alloc(tmp, s);
  LLG_HIP(hipcub::DeviceSelect::Unique(t.p, tmp, in, out, cnt.as<uint64_t>(), n, s));
  uint64_t h = 0;
  LLG_HIP(hipMemcpyAsync(&h, cnt.p, sizeof h, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace llgpu
