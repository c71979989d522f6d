// Strongly connected components of the snapshot (replaces task tarjan_scc,
// benchmark/benchmarks/tarjan_scc.h:121-206, whose Green-Marl transcription
// is not a correct SCC algorithm).  Output label = smallest vertex ID of the
// component, so results are canonical and comparable bit for bit.
//
//   1. trim: an active vertex with no active in-neighbour or no active
//      out-neighbour (self-loops ignored) is a singleton component; a few
//      sweeps peel most of R-MAT's low-degree fringe;
//   2. forward-backward from a pivot (largest (out+1)*(in+1) among active
//      vertices): the intersection of the forward and backward reach sets,
//      each a direction-optimising BFS restricted to active vertices, is the
//      pivot's component (the giant SCC on R-MAT);
//   3. colouring rounds until nothing is active: propagate the maximum vertex
//      ID forward (colour[v] = max ID that reaches v); every vertex whose
//      colour is its own ID is a root; a backward sweep inside each colour
//      marks the vertices that reach their root: those form the root's
//      component.  Each round retires at least the component of the largest
//      active ID, so the loop terminates.
#include "ll_gpu_internal.h"

namespace llgpu {

namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr int kReduceBlock = 256;

__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t x) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = x;
}

__global__ void k_zero(unsigned long long* p) { *p = 0; }

// Active (unlabelled) vertices as a bitmap; wave w builds word w.
__global__ void k_active_bitmap(const uint32_t* label, uint64_t V, uint64_t* act) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t words = (V + 63) / 64;
  if ((v >> 6) >= words) return;
  const uint64_t m = __ballot(v < V && label[v] == kNone);
  if ((threadIdx.x & 63) == 0) act[v >> 6] = m;
}

__device__ inline bool has_active_bm(const uint64_t* ofs, const uint32_t* adj, uint64_t v,
                                     const uint64_t* act) {
  for (uint64_t e = ofs[v], end = ofs[v + 1]; e < end; e++) {
    const uint32_t u = adj[e];
    if (u != v && ((act[u >> 6] >> (u & 63)) & 1)) return true;
  }
  return false;
}

// Neighbours are tested in the active bitmap (V/8 bytes, L2-resident) rather
// than in the labels; a wave's vertices share one bitmap word, whose trimmed
// bits the wave clears with one atomic.  Vertices trimmed earlier in the same
// sweep may or may not be seen as inactive: either way a trimmed vertex is a
// singleton component, and the sweeps repeat until none trims.
__global__ __launch_bounds__(256) void k_trim(const uint64_t* iofs, const uint32_t* iadj,
                                              const uint64_t* oofs, const uint32_t* oadj,
                                              uint32_t* label, uint64_t* act, uint64_t V,
                                              unsigned long long* changed) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool t = false;
  if (v < V && ((act[v >> 6] >> (v & 63)) & 1))
    t = !has_active_bm(iofs, iadj, v, act) || !has_active_bm(oofs, oadj, v, act);
  if (t) label[v] = uint32_t(v);
  const uint64_t m = __ballot(t);  // the wave's vertices: word v >> 6 (blocks of 256)
  if ((threadIdx.x & 63) == 0 && m) atomicAnd((unsigned long long*)&act[v >> 6], ~m);
  block_atomic_add(changed, t ? 1ull : 0ull);
}

// The workgroup's largest (kMax) or smallest x, in thread 0.  Every thread
// calls it.  The reductions below run a fixed grid (grid-stride), so the
// single result word takes one atomic per workgroup, not one per wave.
template <bool kMax>
__device__ inline unsigned long long block_extreme(unsigned long long x) {
  __shared__ unsigned long long part[kReduceBlock / 64];
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_down(x, o, 64);
    x = kMax ? (t > x ? t : x) : (t < x ? t : x);
  }
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 1; i < kReduceBlock / 64; i++)
      x = kMax ? (part[i] > x ? part[i] : x) : (part[i] < x ? part[i] : x);
  return x;
}

__global__ __launch_bounds__(kReduceBlock) void k_pivot(const uint32_t* label,
                                                        const uint32_t* dout,
                                                        const uint32_t* din, uint64_t V,
                                                        unsigned long long* best) {
  unsigned long long key = 0;
  for (uint64_t v = uint64_t(blockIdx.x) * kReduceBlock + threadIdx.x; v < V;
       v += uint64_t(gridDim.x) * kReduceBlock) {
    if (label[v] == kNone) {
      unsigned long long p = (unsigned long long)(dout[v] + 1ull) * (din[v] + 1ull);
      if (p > 0x7fffffffull) p = 0x7fffffffull;
      const unsigned long long k = ((p + 1) << 32) | v;
      key = k > key ? k : key;
    }
  }
  key = block_extreme<true>(key);
  if (threadIdx.x == 0 && key) atomicMax(best, key);
}

// inactive (labelled) vertices as a bitmap; wave w builds word w
__global__ void k_inactive_bitmap(const uint32_t* label, uint64_t V, uint64_t* bm) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t words = (V + 63) / 64;
  if ((v >> 6) >= words) return;
  const uint64_t m = __ballot(v < V && label[v] != kNone);
  if ((threadIdx.x & 63) == 0) bm[v >> 6] = m;
}

__global__ __launch_bounds__(kReduceBlock) void k_fwbw_min(const uint64_t* fw, const uint64_t* bw,
                                                           const uint64_t* excl, uint64_t V,
                                                           unsigned long long* mn) {
  unsigned long long m = ~0ull;
  for (uint64_t v = uint64_t(blockIdx.x) * kReduceBlock + threadIdx.x; v < V;
       v += uint64_t(gridDim.x) * kReduceBlock) {
    const uint64_t w = v >> 6, b = uint64_t(1) << (v & 63);
    if (((fw[w] & bw[w] & ~excl[w]) & b) && v < m) m = v;
  }
  m = block_extreme<false>(m);
  if (threadIdx.x == 0 && m != ~0ull) atomicMin(mn, m);
}

__global__ void k_fwbw_assign(const uint64_t* fw, const uint64_t* bw, const uint64_t* excl,
                              uint64_t V, const unsigned long long* mn, uint32_t* label) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= V) return;
  const uint64_t w = v >> 6, b = uint64_t(1) << (v & 63);
  if ((fw[w] & bw[w] & ~excl[w]) & b) label[v] = uint32_t(*mn);
}

__global__ void k_color_init(const uint32_t* label, uint32_t* color, uint32_t* mark,
                             uint32_t* minlab, uint64_t V, unsigned long long* active) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool a = false;
  if (v < V) {
    a = label[v] == kNone;
    color[v] = a ? uint32_t(v) : kNone;
    mark[v] = 0;
    minlab[v] = kNone;
  }
  block_atomic_add(active, a ? 1ull : 0ull);
}

// colour[v] = max(colour[v], colour[u]) over active in-neighbours u
__global__ void k_color_prop(const uint64_t* iofs, const uint32_t* iadj, const uint32_t* label,
                             uint32_t* color, uint64_t V, unsigned long long* changed) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool ch = false;
  if (v < V && label[v] == kNone) {
    uint32_t c = color[v];
    for (uint64_t e = iofs[v], end = iofs[v + 1]; e < end; e++) {
      const uint32_t u = iadj[e];
      if (label[u] == kNone) {
        const uint32_t cu = color[u];
        c = cu > c ? cu : c;
      }
    }
    if (c > color[v]) {
      color[v] = c;
      ch = true;
    }
  }
  block_atomic_add(changed, ch ? 1ull : 0ull);
}

__global__ void k_root_mark(const uint32_t* label, const uint32_t* color, uint32_t* mark,
                            uint64_t V) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V && label[v] == kNone && color[v] == uint32_t(v)) mark[v] = 1;
}

// v joins its colour's component if an active out-neighbour of the same colour is marked
__global__ void k_bw_prop(const uint64_t* oofs, const uint32_t* oadj, const uint32_t* label,
                          const uint32_t* color, uint32_t* mark, uint64_t V,
                          unsigned long long* changed) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  bool ch = false;
  if (v < V && label[v] == kNone && !mark[v]) {
    const uint32_t c = color[v];
    for (uint64_t e = oofs[v], end = oofs[v + 1]; e < end; e++) {
      const uint32_t w = oadj[e];
      if (mark[w] && color[w] == c && label[w] == kNone) {
        mark[v] = 1;
        ch = true;
        break;
      }
    }
  }
  block_atomic_add(changed, ch ? 1ull : 0ull);
}

__global__ void k_comp_min(const uint32_t* label, const uint32_t* color, const uint32_t* mark,
                           uint32_t* minlab, uint64_t V) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V && label[v] == kNone && mark[v]) atomicMin(&minlab[color[v]], uint32_t(v));
}

__global__ void k_comp_assign(uint32_t* label, const uint32_t* color, const uint32_t* mark,
                              const uint32_t* minlab, uint64_t V) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V && label[v] == kNone && mark[v]) label[v] = minlab[color[v]];
}

__global__ __launch_bounds__(kReduceBlock) void k_labels_out(const uint32_t* label, int64_t* out,
                                                             uint64_t V,
                                                             unsigned long long* nscc) {
  unsigned long long roots = 0;
  for (uint64_t v = uint64_t(blockIdx.x) * kReduceBlock + threadIdx.x; v < V;
       v += uint64_t(gridDim.x) * kReduceBlock) {
    const uint32_t l = label[v];
    out[v] = int64_t(l);
    roots += l == uint32_t(v) ? 1ull : 0ull;
  }
  block_atomic_add(nscc, roots);
}

// One counter through the context's pinned read-back buffer.
unsigned long long read_ull(ll_gpu_ctx* ctx, const DevBuf& b) {
  LLG_HIP(hipMemcpyAsync(ctx->readback, b.p, 8, hipMemcpyDeviceToHost, ctx->stream));
  LLG_HIP(hipStreamSynchronize(ctx->stream));
  return *static_cast<const unsigned long long*>(ctx->readback);
}

}  // namespace

struct SccResult {
  uint64_t num_scc = 0;
  uint64_t sweeps = 0;
};

SccResult scc_device(ll_gpu_graph* g, int level, DevBuf& labels64) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V;
  const uint64_t words = (V + 63) / 64;
  const Csr& out = ensure_csr(g, LL_GPU_DIR_OUT, level);
  const Csr& in = ensure_csr(g, LL_GPU_DIR_IN, level);
  const uint32_t* dout = ensure_degrees(g, LL_GPU_DIR_OUT, level);
  const uint32_t* din = ensure_degrees(g, LL_GPU_DIR_IN, level);
  const uint64_t* io = in.ofs.as<uint64_t>();
  const uint32_t* ia = in.adj.as<uint32_t>();
  const uint64_t* oo = out.ofs.as<uint64_t>();
  const uint32_t* oa = out.adj.as<uint32_t>();
  DevBuf label, cnt, excl, fw, bw, color, mark, minlab;
  label.alloc(V * 4, s);
  cnt.alloc(8, s);
  k_fill_u32<<<grid_for(V, 256), 256, 0, s>>>(label.as<uint32_t>(), V, kNone);
  SccResult r;
  const unsigned gV = grid_for(V, 256);
  DevBuf act;
  act.alloc(words * 8, s);
  auto trim = [&](int rounds) {
    // the active bitmap from the labels as they stand (fw-bw and colouring
    // label whole components)
    k_active_bitmap<<<grid_for(words * 64, 256), 256, 0, s>>>(label.as<uint32_t>(), V,
                                                             act.as<uint64_t>());
    for (int i = 0; i < rounds; i++) {
      k_zero<<<1, 1, 0, s>>>(cnt.as<unsigned long long>());
      k_trim<<<gV, 256, 0, s>>>(io, ia, oo, oa, label.as<uint32_t>(), act.as<uint64_t>(), V,
                                cnt.as<unsigned long long>());
      LLG_HIP(hipGetLastError());
      r.sweeps++;
      if (read_ull(g->ctx, cnt) == 0) break;
    }
  };
  trim(4);
  // forward-backward from the pivot
  k_zero<<<1, 1, 0, s>>>(cnt.as<unsigned long long>());
  // reductions over every vertex: a fixed grid, one atomic per workgroup
  const unsigned gR = std::min(gV, 4u * unsigned(ctx->num_cus));
  k_pivot<<<gR, kReduceBlock, 0, s>>>(label.as<uint32_t>(), dout, din, V,
                                      cnt.as<unsigned long long>());
  LLG_HIP(hipGetLastError());
  const unsigned long long best = read_ull(g->ctx, cnt);
  if (best) {
    const uint32_t pivot = uint32_t(best & 0xffffffffull);
    excl.alloc(words * 8, s);
    k_inactive_bitmap<<<grid_for(words * 64, 256), 256, 0, s>>>(label.as<uint32_t>(), V,
                                                               excl.as<uint64_t>());
    LLG_HIP(hipGetLastError());
    bfs_core(g, out, in, dout, pivot, excl.as<uint64_t>(), fw, nullptr);
    bfs_core(g, in, out, din, pivot, excl.as<uint64_t>(), bw, nullptr);
    r.sweeps += 2;
    DevBuf mn;
    mn.alloc(8, s);
    k_fill_u32<<<1, 2, 0, s>>>(mn.as<uint32_t>(), 2, kNone);
    k_fwbw_min<<<gR, kReduceBlock, 0, s>>>(fw.as<uint64_t>(), bw.as<uint64_t>(),
                                           excl.as<uint64_t>(), V, mn.as<unsigned long long>());
    k_fwbw_assign<<<gV, 256, 0, s>>>(fw.as<uint64_t>(), bw.as<uint64_t>(), excl.as<uint64_t>(),
                                     V, mn.as<unsigned long long>(), label.as<uint32_t>());
    LLG_HIP(hipGetLastError());
    fw.release();
    bw.release();
    excl.release();
    trim(4);
  }
  // colouring rounds
  color.alloc(V * 4, s);
  mark.alloc(V * 4, s);
  minlab.alloc(V * 4, s);
  for (int round = 0;; round++) {
    LLG_CHECK(round < 1000000, LL_GPU_ESTATE, "SCC colouring did not converge");
    k_zero<<<1, 1, 0, s>>>(cnt.as<unsigned long long>());
    k_color_init<<<gV, 256, 0, s>>>(label.as<uint32_t>(), color.as<uint32_t>(),
                                    mark.as<uint32_t>(), minlab.as<uint32_t>(), V,
                                    cnt.as<unsigned long long>());
    LLG_HIP(hipGetLastError());
    if (read_ull(g->ctx, cnt) == 0) break;
    for (;;) {
      k_zero<<<1, 1, 0, s>>>(cnt.as<unsigned long long>());
      k_color_prop<<<gV, 256, 0, s>>>(io, ia, label.as<uint32_t>(), color.as<uint32_t>(), V,
                                      cnt.as<unsigned long long>());
      LLG_HIP(hipGetLastError());
      r.sweeps++;
      if (read_ull(g->ctx, cnt) == 0) break;
    }
    k_root_mark<<<gV, 256, 0, s>>>(label.as<uint32_t>(), color.as<uint32_t>(),
                                   mark.as<uint32_t>(), V);
    for (;;) {
      k_zero<<<1, 1, 0, s>>>(cnt.as<unsigned long long>());
      k_bw_prop<<<gV, 256, 0, s>>>(oo, oa, label.as<uint32_t>(), color.as<uint32_t>(),
                                   mark.as<uint32_t>(), V, cnt.as<unsigned long long>());
      LLG_HIP(hipGetLastError());
      r.sweeps++;
      if (read_ull(g->ctx, cnt) == 0) break;
    }
    k_comp_min<<<gV, 256, 0, s>>>(label.as<uint32_t>(), color.as<uint32_t>(),
                                  mark.as<uint32_t>(), minlab.as<uint32_t>(), V);
    k_comp_assign<<<gV, 256, 0, s>>>(label.as<uint32_t>(), color.as<uint32_t>(),
                                     mark.as<uint32_t>(), minlab.as<uint32_t>(), V);
    LLG_HIP(hipGetLastError());
    trim(2);
  }
  labels64.alloc(V * 8, s);
  k_zero<<<1, 1, 0, s>>>(cnt.as<unsigned long long>());
  k_labels_out<<<gR, kReduceBlock, 0, s>>>(label.as<uint32_t>(), labels64.as<int64_t>(), V,
                                           cnt.as<unsigned long long>());
  LLG_HIP(hipGetLastError());
  r.num_scc = read_ull(g->ctx, cnt);
  return r;
}

}  // namespace llgpu

extern "C" {

int ll_gpu_scc_ex(ll_gpu_graph* g, int level, int64_t* labels_out, uint64_t* num_scc,
                  uint64_t* sweeps) {
  if (!g) return LL_GPU_EINVAL;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    llgpu::DevBuf lab;
    llgpu::SccResult r = llgpu::scc_device(g, level, lab);
    if (labels_out) {
      LLG_HIP(hipMemcpyAsync(labels_out, lab.p, g->V * 8, hipMemcpyDeviceToHost, g->ctx->stream));
      LLG_HIP(hipStreamSynchronize(g->ctx->stream));
    }
    if (num_scc) *num_scc = r.num_scc;
    if (sweeps) *sweeps = r.sweeps;
    return LL_GPU_OK;
  } catch (const llgpu::Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

int ll_gpu_scc(ll_gpu_graph* g, int level, int64_t* labels_out, uint64_t* num_scc) {
  return ll_gpu_scc_ex(g, level, labels_out, num_scc, nullptr);
}

}  // extern "C"
