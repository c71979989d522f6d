// Device helper shared by the sweeps over the binned out-edge layout
// (pr_bins.hip builds it): vertices in bins of 2^bb consecutive IDs, every
// out-edge (u, v) owning one slot, the slots of destination bin b' grouped by
// the source's bin b into chunks (b', b), contiguous, and a chunk table by
// source bin, ct[b * nb + b'] = {first slot, slots}.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace llgpu {

// Writes the values of one source bin's vertices into the bin's chunks for
// every destination bin: slot k of a chunk gets cl[sidx[k]], where cl holds
// the bin's values by offset (LDS) and sidx names each slot's source offset.
// A wave takes 64 table entries at a time, then kGroup chunks at a time, the
// first 128 slots of each requested together (two per lane) before any is
// written; longer chunks finish 64 slots at a time.  Every slot of the bin's
// chunks is written.
template <int kBlock, int kGroup, typename T>
__device__ __forceinline__ void scatter_chunks(const uint2* row, uint32_t nb,
                                               const uint16_t* sidx, const T* cl, T* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint32_t c0 = uint32_t(wave) * 64; c0 < nb; c0 += (kBlock / 64) * 64) {
    const uint32_t mine = c0 + lane;
    const uint2 t = mine < nb ? row[mine] : make_uint2(0u, 0u);
    const uint32_t n = min(64u, nb - c0);
    for (uint32_t j = 0; j < n; j += kGroup) {
      uint32_t st[kGroup], len[kGroup];
      uint16_t s0[kGroup], s1[kGroup];
#pragma unroll
      for (int q = 0; q < kGroup; q++) {
        const int src = int(min(j + uint32_t(q), n - 1));
        st[q] = __shfl(t.x, src, 64);
        len[q] = j + uint32_t(q) < n ? __shfl(t.y, src, 64) : 0u;
      }
#pragma unroll
      for (int q = 0; q < kGroup; q++) {
        s0[q] = uint32_t(lane) < len[q] ? sidx[st[q] + lane] : uint16_t(0);
        s1[q] = uint32_t(lane) + 64 < len[q] ? sidx[st[q] + lane + 64] : uint16_t(0);
      }
#pragma unroll
      for (int q = 0; q < kGroup; q++) {
        if (uint32_t(lane) < len[q]) out[st[q] + lane] = cl[s0[q]];
        if (uint32_t(lane) + 64 < len[q]) out[st[q] + lane + 64] = cl[s1[q]];
      }
#pragma unroll
      for (int q = 0; q < kGroup; q++)
        for (uint32_t k = 128 + uint32_t(lane); k < len[q]; k += 64) {
          const uint32_t slot = st[q] + k;
          out[slot] = cl[sidx[slot]];
        }
    }
  }
}

}  // namespace llgpu
