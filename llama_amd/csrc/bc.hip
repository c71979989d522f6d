// Betweenness centrality (tasks bc_adj / bc_random,
// benchmark/benchmarks/bc_adj.h, bc_random.h): Brandes from each source.
//
// Per source the direction-optimising BFS (bfs_core) gives the levels; then
//   forward, levels 1..D:  sigma[r] = sum of sigma[u] over in-neighbours u one
//                          level up (the reference adds sigma[v] to each
//                          down-neighbour atomically; the terms are the same),
//   reverse, levels D-1..1: delta[r] = sum over out-edges (r, w) to the next
//                          level, in iterator order, of
//                          sigma[r] / sigma[w] * (1 + delta[w]); BC[r] += delta[r]
//                          for r != source,
// in FP32 as the generated task code computes them.  Per source, each row
// block item gets a mask of the levels its rows are on, and per level a list
// of the items with a row there (long rows' pieces first); a pass launches one
// workgroup per item of its level's list.  Within an item only the edges of
// rows on the level are loaded (an LDS map from entry to row masks the rest),
// their terms staged in LDS, and each row sums its own terms sequentially in
// edge order, so the reference's sums are reproduced bit for bit.  Long rows:
// the forward pass reduces pieces in parallel (path counts are whole numbers,
// exact in any order below 2^24) and joins them in piece order; for the
// reverse pass their terms are computed piece-parallel into a buffer by edge
// position first, then one lane adds them in edge order while the other waves
// stage the next chunk.
#include <algorithm>
#include <cstring>

#include "ll_gpu_internal.h"
#include "rowblocks.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kCap = 1024;  // row-block capacity of the path-count and dependency passes
constexpr uint32_t kMaxEdges = kCap + kCap / 2;

// Per-vertex Brandes state in one 16-byte record, so visiting a neighbour is
// one load (level, path count and dependency together) instead of a level
// load followed by dependent path-count and dependency loads.
struct alignas(16) VState {
  int32_t lv;    // BFS level from the source, kUnreachedLevel when unreached
  float sigma;   // shortest-path count
  float delta;   // dependency
  float pad;
};

struct BcArgs {
  const RowItem* items;
  const uint64_t* ofs;
  const uint32_t* adj;
  VState* st;
  float* bc;
  float* part;      // partials of split rows, per item
  float* hub;       // reverse pass: terms of long rows, by out-edge position
  uint32_t source;
  int32_t d;        // the pass's level
};

constexpr uint32_t kPer = kMaxEdges / kBlock;  // entries per thread of a chunk
constexpr int32_t kUnreachedLevel = 2147483646;  // INT_MAX - 1, as the BFS marks it

// Term of edge (r, nb) whose far end is on level l: forward, sigma of an
// in-neighbour one level up; reverse, the reference's
// sigma[r] / sigma[w] * (1 + delta[w]) for an out-neighbour one level down.
template <bool kFwd>
__device__ inline float term(const BcArgs& a, const VState& w, float srow) {
  if (kFwd) return w.lv == a.d - 1 ? w.sigma : 0.0f;
  return w.lv == a.d + 1 ? srow / w.sigma * (1 + w.delta) : 0.0f;
}

// The state of neighbour nb, or an off-level placeholder for a skipped entry.
__device__ inline VState load_state(const BcArgs& a, bool need, uint32_t nb) {
  if (need) return a.st[nb];
  VState v;
  v.lv = -1;
  v.sigma = 0.0f;
  v.delta = 0.0f;
  v.pad = 0.0f;
  return v;
}

template <bool kFwd>
__device__ inline void finish(const BcArgs& a, uint32_t r, float S) {
  if (kFwd) {
    a.st[r].sigma = S;
  } else {
    a.st[r].delta = S;
    if (r != a.source) a.bc[r] = a.bc[r] + S;
  }
}

// One work item of a pass (all threads of the workgroup call it).
template <bool kFwd>
__device__ inline void bc_item(const BcArgs& a, uint64_t item) {
  __shared__ uint64_t rofs[kRowsPerItem + 1];
  __shared__ __align__(16) float vals[2][kMaxEdges];
  __shared__ uint8_t rowof[kMaxEdges];
  __shared__ float srow[kRowsPerItem];
  __shared__ uint8_t onf[kRowsPerItem];
  __shared__ float red[kBlock / 64];
  const RowItem it = a.items[item];
  if (it.npieces == 0) {
    const uint32_t nrows = it.row_end - it.row_begin;
    const uint32_t r = it.row_begin + threadIdx.x;
    const bool on = threadIdx.x < nrows && a.st[r].lv == a.d;
    if (!__syncthreads_or(on)) return;
    for (uint32_t i = threadIdx.x; i <= nrows; i += kBlock) rofs[i] = a.ofs[it.row_begin + i];
    __syncthreads();
    const uint64_t e0 = rofs[0];
    const uint32_t ne = uint32_t(rofs[nrows] - e0);
    const uint32_t lo = threadIdx.x < nrows ? uint32_t(rofs[threadIdx.x] - e0) : 0u;
    const uint32_t hi = threadIdx.x < nrows ? uint32_t(rofs[threadIdx.x + 1] - e0) : 0u;
    if (threadIdx.x < nrows) {
      onf[threadIdx.x] = on ? 1 : 0;
      srow[threadIdx.x] = on && !kFwd ? a.st[r].sigma : 0.0f;
      if (on)
        for (uint32_t j = lo; j < hi; j++) rowof[j] = uint8_t(threadIdx.x);
    }
    __syncthreads();
    // entries of rows off the level are neither loaded nor staged (rowof is
    // written for rows on the level only, so the row's range is checked too);
    // all neighbour IDs first, then their states, so the loads overlap
    uint32_t nb[kPer], rr[kPer];
    VState w[kPer];
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
    for (uint32_t k = 0; k < kPer; k++) w[k] = load_state(a, need[k], nb[k]);
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++)
      if (need[k]) vals[0][k * kBlock + threadIdx.x] = term<kFwd>(a, w[k], srow[rr[k]]);
    __syncthreads();
    if (on) {
      // the row's terms in edge order, as the reference's loop adds them
      float S = 0.0f;
      for (uint32_t j = lo; j < hi; j++) S = S + vals[0][j];
      finish<kFwd>(a, r, S);
    }
  } else if (!kFwd) {
    // Reverse pass, long row: k_bc_hub_terms has written the row's terms
    // (all pieces in parallel); the first piece's workgroup adds them in edge
    // order.  Waves 1-3 copy chunk c into LDS while lane 0 of wave 0 adds
    // chunk c-1, so the dependency is the reference's sequential sum.
    const uint32_t r = it.row_begin;
    if (it.piece != 0 || a.st[r].lv != a.d) return;  // uniform over the workgroup
    const uint64_t rb = a.ofs[r], re = a.ofs[r + 1];
    const uint32_t nch = uint32_t((re - rb + kMaxEdges - 1) / kMaxEdges);
    constexpr uint32_t kGather = kBlock - 64;
    constexpr uint32_t kPerG = kMaxEdges / kGather;
    static_assert(kMaxEdges % kGather == 0 && kMaxEdges % 32 == 0, "chunk layout");
    float S = 0.0f;
    for (uint32_t c = 0; c <= nch; c++) {
      if (threadIdx.x >= 64) {
        if (c < nch) {
          // entries past the row add +0: the sum is >= +0, so they change nothing
          const uint64_t c0 = rb + uint64_t(c) * kMaxEdges;
          const uint32_t t = threadIdx.x - 64;
          float* v = vals[c & 1];
          float x[kPerG];
#pragma unroll
          for (uint32_t k = 0; k < kPerG; k++) {
            const uint64_t e = c0 + k * kGather + t;
            x[k] = e < re ? a.hub[e] : 0.0f;
          }
#pragma unroll
          for (uint32_t k = 0; k < kPerG; k++) v[k * kGather + t] = x[k];
        }
      } else if (c > 0 && threadIdx.x == 0) {
        // the chunk in edge order; LDS reads of the next 32 terms are issued
        // before the additions of the current 32
        const float4* v4 = reinterpret_cast<const float4*>(vals[(c - 1) & 1]);
        constexpr uint32_t kG = 8, kGroups = kMaxEdges / 4 / kG;
        float4 cur[kG];
#pragma unroll
        for (uint32_t k = 0; k < kG; k++) cur[k] = v4[k];
        for (uint32_t g = 0; g < kGroups; g++) {
          float4 nxt[kG];
          if (g + 1 < kGroups) {
#pragma unroll
            for (uint32_t k = 0; k < kG; k++) nxt[k] = v4[(g + 1) * kG + k];
          }
#pragma unroll
          for (uint32_t k = 0; k < kG; k++) {
            S = S + cur[k].x;
            S = S + cur[k].y;
            S = S + cur[k].z;
            S = S + cur[k].w;
          }
#pragma unroll
          for (uint32_t k = 0; k < kG; k++) cur[k] = nxt[k];
        }
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) finish<false>(a, r, S);
  } else {
    // Forward pass, long row: path counts are whole numbers, exact in any
    // order below 2^24; pieces reduce in parallel, k_bc_fixup joins them.
    const uint32_t r = it.row_begin;
    if (a.st[r].lv != a.d) return;  // uniform over the workgroup
    const uint64_t lo = a.ofs[r] + uint64_t(it.piece) * kPiece;
    const uint64_t hi = min(a.ofs[r + 1], lo + kPiece);
    float S = 0.0f;
    for (uint64_t c0 = lo; c0 < hi; c0 += kMaxEdges) {
      uint32_t nb[kPer];
      VState w[kPer];
#pragma unroll
      for (uint32_t k = 0; k < kPer; k++) {
        const uint64_t e = c0 + k * kBlock + threadIdx.x;
        nb[k] = e < hi ? a.adj[e] : 0u;
      }
#pragma unroll
      for (uint32_t k = 0; k < kPer; k++)
        w[k] = load_state(a, c0 + k * kBlock + threadIdx.x < hi, nb[k]);
#pragma unroll
      for (uint32_t k = 0; k < kPer; k++) S = S + term<true>(a, w[k], 0.0f);
    }
    for (int o = 32; o > 0; o >>= 1) S = S + __shfl_down(S, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = S;
    __syncthreads();
    if (threadIdx.x == 0) {
      float t = red[0];
      for (int i = 1; i < kBlock / 64; i++) t = t + red[i];
      if (it.npieces == 1) finish<true>(a, r, t);
      else a.part[item] = t;
    }
  }
}

// A pass: one workgroup per item of the level's list.
template <bool kFwd>
__global__ __launch_bounds__(kBlock) void k_bc_pass(BcArgs a, const uint32_t* list) {
  bc_item<kFwd>(a, list[blockIdx.x]);
}

// Reverse pass, long rows on the level: each piece writes its edges' terms
// to a.hub at their out-edge positions (the list's first items are the long
// rows' pieces).
__global__ __launch_bounds__(kBlock) void k_bc_hub_terms(BcArgs a, const uint32_t* list) {
  const RowItem it = a.items[list[blockIdx.x]];
  const uint32_t r = it.row_begin;
  if (a.st[r].lv != a.d) return;
  const uint64_t lo = a.ofs[r] + uint64_t(it.piece) * kPiece;
  const uint64_t hi = min(a.ofs[r + 1], lo + kPiece);
  const float sr = a.st[r].sigma;
  for (uint64_t c0 = lo; c0 < hi; c0 += kMaxEdges) {
    uint32_t nb[kPer];
    VState w[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      const uint64_t e = c0 + k * kBlock + threadIdx.x;
      nb[k] = e < hi ? a.adj[e] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) w[k] = load_state(a, c0 + k * kBlock + threadIdx.x < hi, nb[k]);
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      const uint64_t e = c0 + k * kBlock + threadIdx.x;
      if (e < hi) a.hub[e] = term<false>(a, w[k], sr);
    }
  }
}

// Level lists: list[l * n + j] = the j-th item with a row on level l (levels
// >= 31 share list 31), counts in cnt[l].  Called for the long rows' pieces
// first and then for the other items, so each list starts with its long rows.
// Positions come from one global atomic per workgroup and level (the waves
// take their slots in the workgroup's share from an LDS counter).
constexpr unsigned kListBlock = 1024;  // workgroup of k_level_lists

__global__ __launch_bounds__(kListBlock) void k_level_lists(const RowItem* items, uint64_t n,
                                                            const uint32_t* mask, bool long_rows,
                                                            uint32_t* list, uint32_t* cnt) {
  constexpr uint32_t kWaves = kListBlock / 64;
  __shared__ uint32_t bcount[32], bbase[32], woff[kWaves][32];
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t m = 0;
  if (i < n && (items[i].npieces != 0) == long_rows) m = mask[i];
  if (threadIdx.x < 32) bcount[threadIdx.x] = 0;
  __syncthreads();
  uint32_t wm = m;
  for (int o = 32; o > 0; o >>= 1) wm |= __shfl_xor(wm, o, 64);
  // the wave's slot in the workgroup's share of each level's list
  for (uint32_t x = wm; x;) {
    const uint32_t l = uint32_t(__builtin_ctz(x));
    x &= x - 1;
    const uint64_t b = __ballot((m >> l) & 1u);
    if (lane == uint32_t(__builtin_ctzll(b))) woff[wave][l] = atomicAdd(&bcount[l], uint32_t(__popcll(b)));
  }
  __syncthreads();
  // one global atomic per workgroup and level
  if (threadIdx.x < 32 && bcount[threadIdx.x])
    bbase[threadIdx.x] = atomicAdd(&cnt[threadIdx.x], bcount[threadIdx.x]);
  __syncthreads();
  for (uint32_t x = wm; x;) {
    const uint32_t l = uint32_t(__builtin_ctz(x));
    x &= x - 1;
    const uint64_t b = __ballot((m >> l) & 1u);
    if ((m >> l) & 1u) {
      const uint64_t below = lane ? (b & (~0ull >> (64 - lane))) : 0ull;
      list[uint64_t(l) * n + bbase[l] + woff[wave][l] + uint32_t(__popcll(below))] = uint32_t(i);
    }
  }
}

__global__ void k_copy_u32(const uint32_t* src, uint32_t* dst, uint32_t n) {
  if (threadIdx.x < n) dst[threadIdx.x] = src[threadIdx.x];
}

// Per item, the levels its rows are on (bit min(level, 31)), one wave per item.
__global__ void k_item_levels(const RowItem* items, uint64_t n, const int32_t* lv,
                              uint32_t* mask) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nw = (uint64_t(gridDim.x) * blockDim.x) >> 6;
  for (uint64_t i = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6; i < n; i += nw) {
    const RowItem it = items[i];
    const uint32_t re = it.npieces == 0 ? it.row_end : it.row_begin + 1;
    uint32_t m = 0;
    for (uint32_t r = it.row_begin + lane; r < re; r += 64) {
      const int32_t l = lv[r];
      if (l != kUnreachedLevel) m |= 1u << min(l, 31);
    }
    for (int o = 32; o > 0; o >>= 1) m |= __shfl_xor(m, o, 64);
    if (lane == 0) mask[i] = m;
  }
}

// Forward pass, rows split over several pieces: combine the partials in
// piece order.
__global__ void k_bc_fixup(BcArgs a, uint64_t nitems) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nitems) return;
  const RowItem it = a.items[i];
  if (it.npieces > 1 && it.piece == 0 && a.st[it.row_begin].lv == a.d) {
    float t = a.part[i];
    for (uint32_t p = 1; p < it.npieces; p++) t = t + a.part[i + p];
    finish<true>(a, it.row_begin, t);
  }
}

// A source's state: the BFS levels, one path to the source, no dependency.
__global__ void k_bc_source_init(const int32_t* lv, VState* st, uint64_t V, uint32_t s) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V) {
    VState x;
    x.lv = lv[v];
    x.sigma = v == s ? 1.0f : 0.0f;
    x.delta = 0.0f;
    x.pad = 0.0f;
    st[v] = x;
  }
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
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
    const uint64_t nin = std::max<uint64_t>(in_rb.n, 1), nout = std::max<uint64_t>(out_rb.n, 1);
    DevBuf lv, visited, st, bc, part, in_mask, out_mask, in_list, out_list, cnt, hub;
    lv.alloc(V * 4, s);
    st.alloc(V * sizeof(VState), s);
    bc.alloc(V * 4, s);
    part.alloc(std::max(nin, nout) * 4, s);
    in_mask.alloc(nin * 4, s);
    out_mask.alloc(nout * 4, s);
    in_list.alloc(32 * nin * 4, s);
    out_list.alloc(32 * nout * 4, s);
    cnt.alloc(3 * 32 * 4, s);  // in, out, out long rows
    const unsigned mask_grid = 4u * unsigned(ctx->num_cus);
    k_fill_f32<<<grid_for(V, 256), 256, 0, s>>>(bc.as<float>(), V, 0.0f);
    uint32_t* c_in = cnt.as<uint32_t>();
    uint32_t* c_out = c_in + 32;
    uint32_t* c_long = c_in + 64;
    const uint64_t* dead = ensure_dead_rows(g, in);
    for (int64_t i = 0; i < ns; i++) {
      const uint32_t src = uint32_t(sources ? sources[i] : i);
      // the search's step count bounds the deepest level (a last step may
      // find nothing): a level without rows has an empty list
      const int32_t D =
          bfs_core(g, out, in, outdeg, src, dead, visited, lv.as<int32_t>()).levels;
      k_bc_source_init<<<grid_for(V, 256), 256, 0, s>>>(lv.as<int32_t>(), st.as<VState>(), V,
                                                        src);
      k_fill_u64<<<1, 64, 0, s>>>(reinterpret_cast<unsigned long long*>(c_in), 48);
      k_item_levels<<<mask_grid, 256, 0, s>>>(in_rb.items.as<RowItem>(), in_rb.n,
                                              lv.as<int32_t>(), in_mask.as<uint32_t>());
      k_item_levels<<<mask_grid, 256, 0, s>>>(out_rb.items.as<RowItem>(), out_rb.n,
                                              lv.as<int32_t>(), out_mask.as<uint32_t>());
      for (bool long_rows : {true, false})
        k_level_lists<<<grid_for(in_rb.n, kListBlock), kListBlock, 0, s>>>(
            in_rb.items.as<RowItem>(), in_rb.n, in_mask.as<uint32_t>(), long_rows,
            in_list.as<uint32_t>(), c_in);
      k_level_lists<<<grid_for(out_rb.n, kListBlock), kListBlock, 0, s>>>(
          out_rb.items.as<RowItem>(), out_rb.n, out_mask.as<uint32_t>(), true,
          out_list.as<uint32_t>(), c_out);
      k_copy_u32<<<1, 32, 0, s>>>(c_out, c_long, 32);
      k_level_lists<<<grid_for(out_rb.n, kListBlock), kListBlock, 0, s>>>(
          out_rb.items.as<RowItem>(), out_rb.n, out_mask.as<uint32_t>(), false,
          out_list.as<uint32_t>(), c_out);
      LLG_HIP(hipGetLastError());
      LLG_HIP(hipMemcpyAsync(ctx->readback, c_in, 96 * 4, hipMemcpyDeviceToHost, s));
      LLG_HIP(hipStreamSynchronize(s));
      uint32_t h[96];
      std::memcpy(h, ctx->readback, sizeof(h));
      // long rows' terms by out-edge position, allocated when a source first
      // reaches one
      for (int l = 0; l < 32 && !hub.p; l++)
        if (h[64 + l]) hub.alloc(out.E * 4, s);
      BcArgs a;
      a.st = st.as<VState>();
      a.bc = bc.as<float>();
      a.part = part.as<float>();
      a.hub = hub.as<float>();
      a.source = src;
      a.items = in_rb.items.as<RowItem>();
      a.ofs = in.ofs.as<uint64_t>();
      a.adj = in.adj.as<uint32_t>();
      for (int32_t d = 1; d <= D; d++) {
        const uint32_t l = uint32_t(std::min(d, 31));
        if (!h[l]) continue;
        a.d = d;
        {
          TimedLaunch t(ctx, "k_bc_pass");
          k_bc_pass<true><<<h[l], kBlock, 0, s>>>(a, in_list.as<uint32_t>() + l * nin);
        }
        if (in_rb.nsplit) k_bc_fixup<<<grid_for(in_rb.n, 256), 256, 0, s>>>(a, in_rb.n);
      }
      a.items = out_rb.items.as<RowItem>();
      a.ofs = out.ofs.as<uint64_t>();
      a.adj = out.adj.as<uint32_t>();
      // level 0 is the source alone, whose dependency is never added
      for (int32_t d = D - 1; d >= 1; d--) {
        const uint32_t l = uint32_t(std::min(d, 31));
        if (!h[32 + l]) continue;
        a.d = d;
        const uint32_t* lst = out_list.as<uint32_t>() + l * nout;
        if (h[64 + l]) {
          TimedLaunch t(ctx, "k_bc_hub_terms");
          k_bc_hub_terms<<<h[64 + l], kBlock, 0, s>>>(a, lst);
        }
        TimedLaunch t(ctx, "k_bc_pass");
        k_bc_pass<false><<<h[32 + l], kBlock, 0, s>>>(a, lst);
      }
      LLG_HIP(hipGetLastError());
    }
    if (bc_out) LLG_HIP(hipMemcpyAsync(bc_out, bc.p, V * 4, hipMemcpyDeviceToHost, s));
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
