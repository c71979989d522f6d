// Weighted PageRank and its adjoint over the snapshot (no reference
// counterpart; the forward is the weighted form of ll_b_pagerank_pull_ext,
// pagerank.h:98-182, the adjoint is the transposed sweep, i.e. the push
// pattern of ll_b_pagerank_push_ext, pagerank.h:217-360, computed as a pull
// over out-edges so it needs no atomics).
//
// Weights w[e] are given per live out-edge in the snapshot's out-edge order.
//   forward:  r0 = p,  r' = (1-d) p + d W^T r            (pull over in-edges)
//   adjoint:  l0 = g,  l' = g + d W l                     (pull over out-edges)
//             dL/dw[e=(u,v)] = d r[u] l[v],  dL/dp = (1-d) l
// The in-edge view carries each edge's out-edge position (radix sort of
// (head, tail) keys), so the forward gathers w into in-edge order once per
// call.  Both sweeps use one row-block kernel: a workgroup stages
// val[e] * x[col[e]] for a run of short rows in LDS with coalesced loads and
// reduces each row in one thread (FP32 Neumaier); long rows are split into
// pieces combined in order.  Bytes per sweep: 12 per edge + 16 per vertex.
#include "ll_gpu_internal.h"
#include "rowblocks.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kCap = 1024;

struct SpArgs {
  const RowItem* items;
  const uint32_t* slots;
  const uint64_t* ofs;
  const uint32_t* col;
  const float* val;
  const float* x;
  const float* base;   // y = base_scale * base[r] + d * S
  float base_scale, d;
  float* y;
  Neu* part;
};

__device__ inline void sp_finalize(const SpArgs& a, uint32_t r, float S) {
  a.y[r] = a.base_scale * a.base[r] + a.d * S;
}

__global__ __launch_bounds__(kBlock) void k_wspmv(SpArgs a) {
  __shared__ float vals[kCap + kCap / 2];
  __shared__ uint64_t rofs[kRowsPerItem + 1];
  __shared__ Neu red[kBlock / 64];
  const RowItem it = a.items[blockIdx.x];
  if (it.npieces == 0) {
    const uint32_t nrows = it.row_end - it.row_begin;
    for (uint32_t i = threadIdx.x; i <= nrows; i += kBlock) rofs[i] = a.ofs[it.row_begin + i];
    __syncthreads();
    const uint64_t e0 = rofs[0];
    const uint32_t ne = uint32_t(rofs[nrows] - e0);
    constexpr uint32_t kPer = (kCap + kCap / 2) / kBlock;
    uint32_t idx[kPer];
    float wv[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kBlock;
      idx[k] = i < ne ? a.col[e0 + i] : 0u;
      wv[k] = i < ne ? a.val[e0 + i] : 0.f;
    }
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      const uint32_t i = threadIdx.x + k * kBlock;
      if (i < ne) vals[i] = wv[k] * a.x[idx[k]];
    }
    __syncthreads();
    if (threadIdx.x < nrows) {
      const uint32_t lo = uint32_t(rofs[threadIdx.x] - e0), hi = uint32_t(rofs[threadIdx.x + 1] - e0);
      Neu s = Neu::zero();
      for (uint32_t j = lo; j < hi; j++) s.add(vals[j]);
      sp_finalize(a, it.row_begin + threadIdx.x, s.value());
    }
  } else {
    const uint32_t r = it.row_begin;
    const uint64_t lo = a.ofs[r] + uint64_t(it.piece) * kPiece;
    const uint64_t hi = min(a.ofs[r + 1], lo + kPiece);
    Neu s = Neu::zero();
    for (uint64_t e = lo + threadIdx.x; e < hi; e += kBlock) s.add(a.val[e] * a.x[a.col[e]]);
    s = wave_reduce(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      Neu t = red[0];
      for (int i = 1; i < kBlock / 64; i++) t.merge(red[i]);
      if (it.npieces == 1) sp_finalize(a, r, t.value());
      else a.part[a.slots[blockIdx.x]] = t;
    }
  }
}

__global__ void k_wspmv_fixup(SpArgs a, uint64_t nitems) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nitems) return;
  const RowItem it = a.items[i];
  if (it.npieces > 1 && it.piece == 0) {
    Neu t = a.part[a.slots[i]];
    for (uint32_t p = 1; p < it.npieces; p++) t.merge(a.part[a.slots[i + p]]);
    sp_finalize(a, it.row_begin, t.value());
  }
}

__global__ void k_wkeys(const uint64_t* ofs, const uint32_t* adj, uint64_t V, uint64_t* keys,
                        uint32_t* pos) {
  const uint64_t u = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= V) return;
  for (uint64_t e = ofs[u], end = ofs[u + 1]; e < end; e++) {
    keys[e] = (uint64_t(adj[e]) << 32) | u;
    pos[e] = uint32_t(e);
  }
}

// CSR rows from sorted (head << 32 | tail) keys.
__global__ void k_key_rows(const uint64_t* keys, uint64_t m, uint64_t V, uint64_t* ofs,
                           uint32_t* col) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i > m) return;
  const uint64_t a = i < m ? keys[i] >> 32 : V;
  const uint64_t p = i == 0 ? 0 : (keys[i - 1] >> 32) + 1;
  for (uint64_t x = p; x <= a && x <= V; x++) ofs[x] = i;
  if (i < m) col[i] = uint32_t(keys[i]);
}

__global__ void k_gather_w(const float* w, const uint32_t* perm, float* out, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = w[perm[i]];
}

__global__ void k_scale(const float* x, float* y, uint64_t n, float s) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) y[i] = s * x[i];
}

// dL/dw over out-edges: rows u, edges (u, v).
__global__ __launch_bounds__(kBlock) void k_edge_grad(const RowItem* items, const uint64_t* ofs,
                                                      const uint32_t* adj, const float* r,
                                                      const float* lam, float d, float* dw) {
  __shared__ uint64_t rofs[kRowsPerItem + 1];
  const RowItem it = items[blockIdx.x];
  if (it.npieces == 0) {
    const uint32_t nrows = it.row_end - it.row_begin;
    for (uint32_t i = threadIdx.x; i <= nrows; i += kBlock) rofs[i] = ofs[it.row_begin + i];
    __syncthreads();
    const uint64_t e0 = rofs[0], e1 = rofs[nrows];
    for (uint64_t e = e0 + threadIdx.x; e < e1; e += kBlock) {
      uint32_t lo = 0, hi = nrows;  // row k with rofs[k] <= e < rofs[k+1]
      while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (rofs[mid] <= e) lo = mid; else hi = mid;
      }
      dw[e] = d * r[it.row_begin + lo] * lam[adj[e]];
    }
  } else {
    const uint32_t u = it.row_begin;
    const uint64_t lo = ofs[u] + uint64_t(it.piece) * kPiece;
    const uint64_t hi = min(ofs[u + 1], lo + kPiece);
    const float ru = d * r[u];
    for (uint64_t e = lo + threadIdx.x; e < hi; e += kBlock) dw[e] = ru * lam[adj[e]];
  }
}

void sweep(ll_gpu_ctx* ctx, const RowBlocks& rb, const Csr& c, const float* val, const float* x,
           const float* base, float base_scale, float d, float* y, DevBuf& part, const char* name) {
  hipStream_t s = ctx->stream;
  SpArgs a;
  a.items = rb.items.as<RowItem>();
  a.slots = rb.slots.as<uint32_t>();
  a.ofs = c.ofs.as<uint64_t>();
  a.col = c.adj.as<uint32_t>();
  a.val = val;
  a.x = x;
  a.base = base;
  a.base_scale = base_scale;
  a.d = d;
  a.y = y;
  a.part = part.as<Neu>();
  {
    TimedLaunch t(ctx, name);
    k_wspmv<<<unsigned(rb.n), kBlock, 0, s>>>(a);
  }
  if (rb.nsplit) k_wspmv_fixup<<<grid_for(rb.n, 256), 256, 0, s>>>(a, rb.n);
  LLG_HIP(hipGetLastError());
}

}  // namespace

void ensure_wcsr(ll_gpu_graph* g, int level) {
  level = resolve_level(g, level);
  if (g->w_level == level) return;
  hipStream_t s = g->ctx->stream;
  const Csr& out = ensure_csr(g, LL_GPU_DIR_OUT, level);
  const uint64_t V = g->V, E = out.E;
  LLG_CHECK(E < (uint64_t(1) << 32), LL_GPU_ERANGE, "weighted PageRank needs E < 2^32");
  DevBuf keys, pos, skeys;
  keys.alloc(std::max<uint64_t>(E, 1) * 8, s);
  pos.alloc(std::max<uint64_t>(E, 1) * 4, s);
  skeys.alloc(std::max<uint64_t>(E, 1) * 8, s);
  g->wperm.alloc(std::max<uint64_t>(E, 1) * 4, s);
  k_wkeys<<<grid_for(V, 256), 256, 0, s>>>(out.ofs.as<uint64_t>(), out.adj.as<uint32_t>(), V,
                                            keys.as<uint64_t>(), pos.as<uint32_t>());
  LLG_HIP(hipGetLastError());
  sort_pairs_u64_u32(keys.as<uint64_t>(), skeys.as<uint64_t>(), pos.as<uint32_t>(),
                     g->wperm.as<uint32_t>(), E, 64, s);
  g->wcsr = Csr();
  g->wcsr.ofs.alloc((V + 1) * 8, s);
  g->wcsr.adj.alloc(std::max<uint64_t>(E, 1) * 4, s);
  k_key_rows<<<grid_for(E + 1, 256), 256, 0, s>>>(skeys.as<uint64_t>(), E, V,
                                                  g->wcsr.ofs.as<uint64_t>(),
                                                  g->wcsr.adj.as<uint32_t>());
  LLG_HIP(hipGetLastError());
  g->wcsr.valid = true;
  g->wcsr.level = level;
  g->wcsr.V = V;
  g->wcsr.E = E;
  g->wrb.clear();
  g->wrb.emplace_back();
  build_rowblocks(g->ctx, g->wcsr, kCap, kPiece, g->wrb.back());
  g->w_level = level;
}

// Forward: iters sweeps from r0 = p; result in r_out (device).
void wpr_forward_device(ll_gpu_graph* g, int level, const float* w, const float* p, int iters,
                        double d, float* r_out) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  ensure_wcsr(g, level);
  const uint64_t V = g->V, E = g->wcsr.E;
  const RowBlocks& rb = g->wrb.back();
  DevBuf w_in, buf[2], part;
  w_in.alloc(std::max<uint64_t>(E, 1) * 4, s);
  k_gather_w<<<grid_for(E, 256), 256, 0, s>>>(w, g->wperm.as<uint32_t>(), w_in.as<float>(), E);
  buf[0].alloc(V * 4, s);
  buf[1].alloc(V * 4, s);
  part.alloc(std::max<uint64_t>(rb.nsplit, 1) * sizeof(Neu), s);
  const float* x = p;
  for (int it = 0; it < iters; it++) {
    float* y = it == iters - 1 ? r_out : buf[it & 1].as<float>();
    sweep(ctx, rb, g->wcsr, w_in.as<float>(), x, p, float(1.0 - d), float(d), y, part,
          "k_wspmv_fwd");
    x = y;
  }
  if (iters == 0) LLG_HIP(hipMemcpyAsync(r_out, p, V * 4, hipMemcpyDeviceToDevice, s));
}

void wpr_adjoint_device(ll_gpu_graph* g, int level, const float* w, const float* r,
                        const float* gr, int iters, double d, float* dw, float* dp) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const Csr& out = ensure_csr(g, LL_GPU_DIR_OUT, level);
  const RowBlocks& rb = ensure_rowblocks(g, LL_GPU_DIR_OUT, level, kCap, kPiece);
  const uint64_t V = g->V;
  DevBuf buf[2], part;
  buf[0].alloc(V * 4, s);
  buf[1].alloc(V * 4, s);
  part.alloc(std::max<uint64_t>(rb.nsplit, 1) * sizeof(Neu), s);
  const float* lam = gr;
  for (int it = 0; it < iters; it++) {
    float* y = buf[it & 1].as<float>();
    sweep(ctx, rb, out, w, lam, gr, 1.0f, float(d), y, part, "k_wspmv_adj");
    lam = y;
  }
  if (dw) {
    TimedLaunch t(ctx, "k_edge_grad");
    k_edge_grad<<<unsigned(rb.n), kBlock, 0, s>>>(rb.items.as<RowItem>(), out.ofs.as<uint64_t>(),
                                                  out.adj.as<uint32_t>(), r, lam, float(d), dw);
  }
  if (dp) k_scale<<<grid_for(V, 256), 256, 0, s>>>(lam, dp, V, float(1.0 - d));
  LLG_HIP(hipGetLastError());
}

}  // namespace llgpu

extern "C" {

int ll_gpu_wpr_forward(ll_gpu_graph* g, int level, const float* edge_w, const float* teleport,
                       int iters, double damping, float* r_out) {
  if (!g) return LL_GPU_EINVAL;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    LLG_CHECK(edge_w && teleport && r_out, LL_GPU_EINVAL, "NULL device pointer");
    LLG_CHECK(iters >= 0 && damping >= 0 && damping <= 1, LL_GPU_EINVAL, "bad iters/damping");
    llgpu::wpr_forward_device(g, level, edge_w, teleport, iters, damping, r_out);
    LLG_HIP(hipStreamSynchronize(g->ctx->stream));
    return LL_GPU_OK;
  } catch (const llgpu::Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

int ll_gpu_wpr_adjoint(ll_gpu_graph* g, int level, const float* edge_w, const float* r,
                       const float* dl_dr, int iters, double damping, float* dl_dw_out,
                       float* dl_dteleport_out) {
  if (!g) return LL_GPU_EINVAL;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    LLG_CHECK(edge_w && r && dl_dr, LL_GPU_EINVAL, "NULL device pointer");
    LLG_CHECK(iters >= 0 && damping >= 0 && damping <= 1, LL_GPU_EINVAL, "bad iters/damping");
    llgpu::wpr_adjoint_device(g, level, edge_w, r, dl_dr, iters, damping, dl_dw_out,
                              dl_dteleport_out);
    LLG_HIP(hipStreamSynchronize(g->ctx->stream));
    return LL_GPU_OK;
  } catch (const llgpu::Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

}  // extern "C"
