// Device-side helpers for kernels that consume row blocks (see rowblocks.hip).
#pragma once

#include "ll_gpu_internal.h"

namespace llgpu {

constexpr uint32_t kRowsPerItem = 256;   // rows of a stream item (one thread each)
constexpr uint32_t kPiece = 8192;        // entries of one piece of a long row (PR, BFS)
constexpr uint32_t kMaxCap = 4096;       // largest stream capacity (LDS: 1.5*cap values)

// Neumaier compensated summation in FP32.
struct Neu {
  float s, c;
  __device__ static inline Neu zero() { return Neu{0.f, 0.f}; }
  __device__ inline void add(float x) {
    float t = s + x;
    if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  __device__ inline void merge(const Neu& o) {
    add(o.s);
    c += o.c;
  }
  __device__ inline float value() const { return s + c; }
};

__device__ inline Neu wave_reduce(Neu v) {
  for (int o = 32; o > 0; o >>= 1) {
    Neu w;
    w.s = __shfl_down(v.s, o, 64);
    w.c = __shfl_down(v.c, o, 64);
    v.merge(w);
  }
  return v;
}

}  // namespace llgpu
