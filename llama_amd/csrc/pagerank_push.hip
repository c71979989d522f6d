// Push PageRank (task pagerank_push / pagerank_double_push,
// benchmark/benchmarks/pagerank.h:217-306): every vertex t with out-edges adds
// rank[t] / outdeg(t) to next[w] of each out-neighbour w, then
// rank' = (1-d)/N + d * next.
//
// Over row blocks of the snapshot out-CSR: a workgroup stages its rows'
// contributions in LDS, loads its run of neighbour IDs with coalesced reads
// and adds with hardware FP64 atomics (the L2 performs them); pieces of long
// rows add one contribution to every entry.  FP64 accumulation keeps the
// result within rounding of the FP64 reference whatever order the atomics
// take.  Bytes per iteration: 4 B neighbour ID + 8 B atomic per edge, 8 B
// offset + 8 B rank + 8 B next read/clear per vertex.
#include "ll_gpu_internal.h"
#include "rowblocks.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kCap = 2048;

__global__ __launch_bounds__(kBlock) void k_push(const RowItem* items, const uint64_t* ofs,
                                                 const uint32_t* adj, const double* rank,
                                                 double* next) {
  __shared__ uint64_t rofs[kRowsPerItem + 1];
  __shared__ double contrib[kRowsPerItem];
  const RowItem it = items[blockIdx.x];
  if (it.npieces == 0) {
    const uint32_t nrows = it.row_end - it.row_begin;
    for (uint32_t i = threadIdx.x; i <= nrows; i += kBlock) rofs[i] = ofs[it.row_begin + i];
    __syncthreads();
    if (threadIdx.x < nrows) {
      const uint64_t deg = rofs[threadIdx.x + 1] - rofs[threadIdx.x];
      contrib[threadIdx.x] = deg ? rank[it.row_begin + threadIdx.x] / double(deg) : 0.0;
    }
    __syncthreads();
    const uint64_t e0 = rofs[0];
    const uint32_t ne = uint32_t(rofs[nrows] - e0);
    for (uint32_t i = threadIdx.x; i < ne; i += kBlock) {
      uint32_t a = 0, b = nrows;  // row of entry i: last row starting at or before it
      while (b - a > 1) {
        const uint32_t mid = (a + b) >> 1;
        if (rofs[mid] - e0 <= i) a = mid;
        else b = mid;
      }
      unsafeAtomicAdd(&next[adj[e0 + i]], contrib[a]);
    }
  } else {
    const uint32_t r = it.row_begin;
    const uint64_t rb = ofs[r], re = ofs[r + 1];
    const uint64_t lo = rb + uint64_t(it.piece) * kPiece;
    const uint64_t hi = min(re, lo + kPiece);
    const double c = rank[r] / double(re - rb);
    for (uint64_t e = lo + threadIdx.x; e < hi; e += kBlock) unsafeAtomicAdd(&next[adj[e]], c);
  }
}

// rank' = (1-d)/N + d*next, next cleared for the following iteration; the
// L1 change is accumulated when diff != nullptr.
__global__ void k_push_apply(double* rank, double* next, uint64_t V, double base, double d,
                             double* diff) {
  double x = 0;
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
       v += uint64_t(gridDim.x) * blockDim.x) {
    const double nv = base + d * next[v];
    if (diff) x += fabs(nv - rank[v]);
    rank[v] = nv;
    next[v] = 0.0;
  }
  if (!diff) return;
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  __shared__ double part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0;
    for (int i = 0; i < int(blockDim.x / 64); i++) t += part[i];
    unsafeAtomicAdd(diff, t);
  }
}

__global__ void k_push_init(double* rank, double* next, uint64_t V, double r0, double* diff) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V) {
    rank[v] = r0;
    next[v] = 0.0;
  }
  if (v == 0) *diff = 0.0;
}

}  // namespace

}  // namespace llgpu

extern "C" {

int ll_gpu_pagerank_push(ll_gpu_graph* g, int level, int iters, double damping,
                         double* ranks_out, double* last_diff) {
  if (!g) return LL_GPU_EINVAL;
  using namespace llgpu;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    LLG_CHECK(iters >= 1, LL_GPU_EINVAL, "iters must be >= 1 (the reference runs do/while)");
    LLG_CHECK(damping >= 0 && damping <= 1, LL_GPU_EINVAL, "damping must be in [0, 1]");
    ll_gpu_ctx* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint64_t V = g->V;
    const Csr& out = ensure_csr(g, LL_GPU_DIR_OUT, level);
    const RowBlocks& rb = ensure_rowblocks(g, LL_GPU_DIR_OUT, level, kCap, kPiece);
    DevBuf rank, next, diff;
    rank.alloc(V * 8, s);
    next.alloc(V * 8, s);
    diff.alloc(8, s);
    k_push_init<<<grid_for(V, 256), 256, 0, s>>>(rank.as<double>(), next.as<double>(), V,
                                                 1.0 / double(V), diff.as<double>());
    const unsigned apply_grid = std::min(grid_for(V, 256), 4u * unsigned(ctx->num_cus));
    for (int it = 0; it < iters; it++) {
      {
        TimedLaunch t(ctx, "k_push");
        k_push<<<unsigned(rb.n), kBlock, 0, s>>>(rb.items.as<RowItem>(), out.ofs.as<uint64_t>(),
                                                 out.adj.as<uint32_t>(), rank.as<double>(),
                                                 next.as<double>());
      }
      const bool last = it == iters - 1 && last_diff;
      k_push_apply<<<apply_grid, 256, 0, s>>>(rank.as<double>(), next.as<double>(), V,
                                              (1.0 - damping) / double(V), damping,
                                              last ? diff.as<double>() : nullptr);
      LLG_HIP(hipGetLastError());
    }
    if (ranks_out) LLG_HIP(hipMemcpyAsync(ranks_out, rank.p, V * 8, hipMemcpyDeviceToHost, s));
    if (last_diff) LLG_HIP(hipMemcpyAsync(last_diff, diff.p, 8, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    return LL_GPU_OK;
  } catch (const Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

}  // extern "C"
