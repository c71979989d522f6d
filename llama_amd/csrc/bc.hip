// Betweenness centrality (tasks bc_adj / bc_random,
// benchmark/benchmarks/bc_adj.h, bc_random.h): Brandes from each source.
//
// Per source the direction-optimising BFS (bfs_core) gives the levels; then
//   forward, levels 1..D:  sigma[r] = sum of sigma[u] over in-neighbours u one
//                          level up (the reference adds sigma[v] to each
//                          down-neighbour atomically; the terms are the same),
//   reverse, levels D-1..0: delta[r] = sum over out-edges (r, w) to the next
//                          level, in iterator order, of
//                          sigma[r] / sigma[w] * (1 + delta[w]); BC[r] += delta[r]
//                          for r != source,
// in FP32 as the generated task code computes them.  Each pass runs over the
// row blocks of the in- (forward) or out-CSR (reverse): a workgroup with no
// row on the pass's level returns at once; otherwise only the edges of those
// rows are loaded (an LDS map from entry to row masks the rest), their terms
// staged in LDS, and each row sums its own terms sequentially in edge order,
// so short rows reproduce the reference's sums bit for bit.  Rows split into
// pieces (long rows) are combined in piece order by a fix-up kernel.
#include <algorithm>

#include "ll_gpu_internal.h"
#include "rowblocks.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kCap = kBfsRowCap;  // row blocks shared with the bottom-up BFS
constexpr uint32_t kMaxEdges = kCap + kCap / 2;

struct BcArgs {
  const RowItem* items;
  const uint64_t* ofs;
  const uint32_t* adj;
  const int32_t* lv;
  float* sigma;
  float* delta;
  float* bc;
  float* part;      // partials of split rows, per item
  uint32_t source;
  int32_t d;        // the pass's level
};

constexpr uint32_t kPer = kMaxEdges / kBlock;  // entries per thread of a chunk

// Term of edge (r, nb) whose far end is on level l: forward, sigma of an
// in-neighbour one level up; reverse, the reference's
// sigma[r] / sigma[w] * (1 + delta[w]) for an out-neighbour one level down.
template <bool kFwd>
__device__ inline float term(const BcArgs& a, uint32_t nb, int32_t l, float srow) {
  if (kFwd) return l == a.d - 1 ? a.sigma[nb] : 0.0f;
  return l == a.d + 1 ? srow / a.sigma[nb] * (1 + a.delta[nb]) : 0.0f;
}

template <bool kFwd>
__device__ inline void finish(const BcArgs& a, uint32_t r, float S) {
  if (kFwd) {
    a.sigma[r] = S;
  } else {
    a.delta[r] = S;
    if (r != a.source) a.bc[r] = a.bc[r] + S;
  }
}

template <bool kFwd>
__global__ __launch_bounds__(kBlock) void k_bc_pass(BcArgs a) {
  __shared__ uint64_t rofs[kRowsPerItem + 1];
  __shared__ float vals[kMaxEdges];
  __shared__ uint8_t rowof[kMaxEdges];
  __shared__ float srow[kRowsPerItem];
  __shared__ uint8_t onf[kRowsPerItem];
  __shared__ float red[kBlock / 64];
  const RowItem it = a.items[blockIdx.x];
  if (it.npieces == 0) {
    const uint32_t nrows = it.row_end - it.row_begin;
    const uint32_t r = it.row_begin + threadIdx.x;
    const bool on = threadIdx.x < nrows && a.lv[r] == a.d;
    if (!__syncthreads_or(on)) return;
    for (uint32_t i = threadIdx.x; i <= nrows; i += kBlock) rofs[i] = a.ofs[it.row_begin + i];
    __syncthreads();
    const uint64_t e0 = rofs[0];
    const uint32_t ne = uint32_t(rofs[nrows] - e0);
    const uint32_t lo = threadIdx.x < nrows ? uint32_t(rofs[threadIdx.x] - e0) : 0u;
    const uint32_t hi = threadIdx.x < nrows ? uint32_t(rofs[threadIdx.x + 1] - e0) : 0u;
    if (threadIdx.x < nrows) {
      onf[threadIdx.x] = on ? 1 : 0;
      srow[threadIdx.x] = on && !kFwd ? a.sigma[r] : 0.0f;
      if (on)
        for (uint32_t j = lo; j < hi; j++) rowof[j] = uint8_t(threadIdx.x);
    }
    __syncthreads();
    // entries of rows off the level are neither loaded nor staged (rowof is
    // written for rows on the level only, so the row's range is checked too);
    // all neighbour IDs first, then their levels, so the loads overlap
    uint32_t nb[kPer], rr[kPer];
    int32_t l[kPer];
    bool need[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      const uint32_t i = k * kBlock + threadIdx.x;
      need[k] = false;
      rr[k] = 0;
      if (i < ne) {
        rr[k] = rowof[i];
        need[k] = rr[k] < nrows && onf[rr[k]] && i >= uint32_t(rofs[rr[k]] - e0) &&
                  i < uint32_t(rofs[rr[k] + 1] - e0);
      }
      nb[k] = need[k] ? a.adj[e0 + i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) l[k] = need[k] ? a.lv[nb[k]] : -1;
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++)
      if (need[k]) vals[k * kBlock + threadIdx.x] = term<kFwd>(a, nb[k], l[k], srow[rr[k]]);
    __syncthreads();
    if (on) {
      // the row's terms in edge order, as the reference's loop adds them
      float S = 0.0f;
      for (uint32_t j = lo; j < hi; j++) S = S + vals[j];
      finish<kFwd>(a, r, S);
    }
  } else if (!kFwd) {
    // Reverse pass, long row: the first piece's workgroup walks the whole row
    // in chunks; the chunk's terms are staged in LDS and one lane adds them in
    // edge order, so the dependency is the reference's sequential sum.
    const uint32_t r = it.row_begin;
    if (it.piece != 0 || a.lv[r] != a.d) return;  // uniform over the workgroup
    const uint64_t rb = a.ofs[r], re = a.ofs[r + 1];
    const float sr = a.sigma[r];
    float S = 0.0f;
    for (uint64_t c0 = rb; c0 < re; c0 += kMaxEdges) {
      const uint32_t n = uint32_t(min<uint64_t>(re - c0, kMaxEdges));
      uint32_t nb[kPer];
      int32_t l[kPer];
#pragma unroll
      for (uint32_t k = 0; k < kPer; k++) {
        const uint32_t i = k * kBlock + threadIdx.x;
        nb[k] = i < n ? a.adj[c0 + i] : 0u;
      }
#pragma unroll
      for (uint32_t k = 0; k < kPer; k++) l[k] = k * kBlock + threadIdx.x < n ? a.lv[nb[k]] : -1;
#pragma unroll
      for (uint32_t k = 0; k < kPer; k++) {
        const uint32_t i = k * kBlock + threadIdx.x;
        if (i < n) vals[i] = term<false>(a, nb[k], l[k], sr);
      }
      __syncthreads();
      if (threadIdx.x == 0)
        for (uint32_t j = 0; j < n; j++) S = S + vals[j];
      __syncthreads();
    }
    if (threadIdx.x == 0) finish<false>(a, r, S);
  } else {
    // Forward pass, long row: path counts are whole numbers, exact in any
    // order below 2^24; pieces reduce in parallel, k_bc_fixup joins them.
    const uint32_t r = it.row_begin;
    if (a.lv[r] != a.d) return;  // uniform over the workgroup
    const uint64_t lo = a.ofs[r] + uint64_t(it.piece) * kPiece;
    const uint64_t hi = min(a.ofs[r + 1], lo + kPiece);
    float S = 0.0f;
    for (uint64_t e = lo + threadIdx.x; e < hi; e += kBlock) {
      const uint32_t u = a.adj[e];
      S = S + term<true>(a, u, a.lv[u], 0.0f);
    }
    for (int o = 32; o > 0; o >>= 1) S = S + __shfl_down(S, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = S;
    __syncthreads();
    if (threadIdx.x == 0) {
      float t = red[0];
      for (int i = 1; i < kBlock / 64; i++) t = t + red[i];
      if (it.npieces == 1) finish<true>(a, r, t);
      else a.part[blockIdx.x] = t;
    }
  }
}

// Forward pass, rows split over several pieces: combine the partials in
// piece order.
__global__ void k_bc_fixup(BcArgs a, uint64_t nitems) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nitems) return;
  const RowItem it = a.items[i];
  if (it.npieces > 1 && it.piece == 0 && a.lv[it.row_begin] == a.d) {
    float t = a.part[i];
    for (uint32_t p = 1; p < it.npieces; p++) t = t + a.part[i + p];
    finish<true>(a, it.row_begin, t);
  }
}

__global__ void k_bc_source_init(float* sigma, float* delta, uint64_t V, uint32_t s) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V) {
    sigma[v] = v == s ? 1.0f : 0.0f;
    delta[v] = 0.0f;
  }
}

__global__ void k_fill_f32(float* p, uint64_t n, float x) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = x;
}

}  // namespace

}  // namespace llgpu

extern "C" {

int ll_gpu_bc(ll_gpu_graph* g, int level, const int64_t* sources, int64_t n, float* bc_out) {
  if (!g) return LL_GPU_EINVAL;
  using namespace llgpu;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    ll_gpu_ctx* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint64_t V = g->V;
    const int64_t ns = sources ? n : int64_t(V);
    LLG_CHECK(!sources || n >= 0, LL_GPU_EINVAL, "bad source count");
    for (int64_t i = 0; sources && i < n; i++)
      LLG_CHECK(sources[i] >= 0 && uint64_t(sources[i]) < V, LL_GPU_EINVAL,
                "source outside [0, max_nodes)");
    const Csr& out = ensure_csr(g, LL_GPU_DIR_OUT, level);
    const Csr& in = ensure_csr(g, LL_GPU_DIR_IN, level);
    const RowBlocks& in_rb = ensure_rowblocks(g, LL_GPU_DIR_IN, level, kCap, kPiece);
    const RowBlocks& out_rb = ensure_rowblocks(g, LL_GPU_DIR_OUT, level, kCap, kPiece);
    const uint32_t* outdeg = ensure_degrees(g, LL_GPU_DIR_OUT, level);
    DevBuf lv, visited, sigma, delta, bc, part;
    lv.alloc(V * 4, s);
    sigma.alloc(V * 4, s);
    delta.alloc(V * 4, s);
    bc.alloc(V * 4, s);
    part.alloc(std::max<uint64_t>(std::max(in_rb.n, out_rb.n), 1) * 4, s);
    k_fill_f32<<<grid_for(V, 256), 256, 0, s>>>(bc.as<float>(), V, 0.0f);
    LLG_HIP(hipGetLastError());
    for (int64_t i = 0; i < ns; i++) {
      const uint32_t src = uint32_t(sources ? sources[i] : i);
      // the search's step count bounds the deepest level (a last step may
      // find nothing): passes over an empty level return at once
      const int32_t D =
          bfs_core(g, out, in, in_rb, outdeg, src, nullptr, visited, lv.as<int32_t>()).levels;
      k_bc_source_init<<<grid_for(V, 256), 256, 0, s>>>(sigma.as<float>(), delta.as<float>(), V,
                                                        src);
      BcArgs a;
      a.lv = lv.as<int32_t>();
      a.sigma = sigma.as<float>();
      a.delta = delta.as<float>();
      a.bc = bc.as<float>();
      a.part = part.as<float>();
      a.source = src;
      a.items = in_rb.items.as<RowItem>();
      a.ofs = in.ofs.as<uint64_t>();
      a.adj = in.adj.as<uint32_t>();
      for (int32_t d = 1; d <= D; d++) {
        a.d = d;
        {
          TimedLaunch t(ctx, "k_bc_pass");
          k_bc_pass<true><<<unsigned(in_rb.n), kBlock, 0, s>>>(a);
        }
        if (in_rb.nsplit) k_bc_fixup<<<grid_for(in_rb.n, 256), 256, 0, s>>>(a, in_rb.n);
      }
      a.items = out_rb.items.as<RowItem>();
      a.ofs = out.ofs.as<uint64_t>();
      a.adj = out.adj.as<uint32_t>();
      // level 0 is the source alone, whose dependency is never added
      for (int32_t d = D - 1; d >= 1; d--) {
        a.d = d;
        {
          TimedLaunch t(ctx, "k_bc_pass");
          k_bc_pass<false><<<unsigned(out_rb.n), kBlock, 0, s>>>(a);
        }
      }
      LLG_HIP(hipGetLastError());
    }
    if (bc_out) LLG_HIP(hipMemcpyAsync(bc_out, bc.p, V * 4, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    return LL_GPU_OK;
  } catch (const Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  }
}

}  // extern "C"
