// Pull PageRank over the snapshot in-CSR (task pagerank / pagerank_double,
// benchmark/benchmarks/pagerank.h:98-182).
//
// Per iteration one launch of k_pr_pull reduces, for every vertex t,
//   S[t] = sum_{w in in(t)} contrib[w],   contrib[w] = rank[w] / outdeg(w)
// and writes rank'[t] = (1-d)/N + d*S[t] and contrib'[t] = rank'[t]/outdeg(t)
// (0 for outdeg 0; such a vertex is never an in-neighbour).  The reference
// divides per edge (pagerank.h:147); dividing once per vertex moves the
// division but keeps the same terms.  Sums are FP32 with Neumaier
// compensation; rows split over several workgroups are combined in piece order
// by k_pr_fixup.  Bytes per iteration: 4 B in-neighbour ID + 4 B contribution
// gather per edge, 8 B offset + 4 B out-degree + 8 B rank/contrib writes per
// vertex.
//
// On large graphs the pull runs over column segments of the in-CSR (built by
// pr_segments.hip): the same kernel reduces every (row, segment) pair to a
// compensated partial, and k_pr_merge adds each row's partials in segment
// order before the finalisation above.
#include "ll_gpu_internal.h"
#include "rowblocks.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;

struct PrArgs {
  const RowItem* items;
  const uint64_t* ofs;
  const uint32_t* adj;
  const float* contrib_in;
  const float* rank_in;   // only read when diff != nullptr
  const uint32_t* outdeg;
  float* rank_out;
  float* contrib_out;
  const uint32_t* inv;    // optional: contribution slot of each vertex (locality order)
  Neu* part;              // per item partials of split rows
  double* diff;           // L1 change accumulator (last iteration only)
  Neu* pair_part;         // segmented pull: compensated partial per pair (row of the CSR)
  const uint16_t* perm;   // sorted items: original in-item position of each entry
  float base, d;
};

__device__ inline float finalize(const PrArgs& a, uint32_t r, float S) {
  float v = a.base + a.d * S;
  a.rank_out[r] = v;
  uint32_t od = a.outdeg[r];
  a.contrib_out[a.inv ? a.inv[r] : r] = od ? v / float(od) : 0.f;
  return a.diff ? fabsf(v - a.rank_in[r]) : 0.f;
}

// What a pull does with a finished row sum:
//   kFinal    -- finalise the vertex (rank, contribution);
//   kPartOut  -- store the compensated partial in pair_part[row] (a (row,
//                segment) pair, or a vertex in the hot pass of a hot/cold split);
//   kPartIn   -- start from pair_part[row] (the hot partial), then finalise.
enum PullMode : int { kFinal = 0, kPartOut = 1, kPartIn = 2 };

template <int kMode>
__device__ inline float finish(const PrArgs& a, uint32_t r, Neu s) {
  if (kMode == kPartOut) {
    a.pair_part[r] = s;
    return 0.f;
  }
  if (kMode == kPartIn) {
    Neu t = a.pair_part[r];
    t.merge(s);
    s = t;
  }
  return finalize(a, r, s.value());
}

__device__ inline void block_add_diff(const PrArgs& a, float x) {
  if (!a.diff) return;
  __shared__ float red[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    float t = 0;
    for (int i = 0; i < kBlock / 64; i++) t += red[i];
    atomicAdd(a.diff, double(t));
  }
}

// kSorted: every item's entries are sorted by slot (a wave's gathers then
// share cache lines) and a.perm restores the row order in LDS, so the row
// sums are the same as without sorting; pieces of long rows are summed in
// slot order.
template <uint32_t CAP, int kMode, bool kSorted>
__global__ __launch_bounds__(kBlock) void k_pr_pull(PrArgs a) {
  __shared__ float vals[CAP + CAP / 2];
  __shared__ uint64_t rofs[kRowsPerItem + 1];
  __shared__ Neu red[kBlock / 64];
  const RowItem it = a.items[blockIdx.x];
  float dsum = 0.f;
  if (it.npieces == 0) {
    const uint32_t nrows = it.row_end - it.row_begin;
    for (uint32_t i = threadIdx.x; i <= nrows; i += kBlock) rofs[i] = a.ofs[it.row_begin + i];
    __syncthreads();
    const uint64_t e0 = rofs[0];
    const uint32_t ne = uint32_t(rofs[nrows] - e0);
    // coalesced neighbour IDs first, then all gathers in flight together
    constexpr uint32_t kPer = (CAP + CAP / 2) / kBlock;
    uint32_t idx[kPer];
    float v[kPer];
    uint32_t pos[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      idx[k] = i < ne ? a.adj[e0 + i] : 0u;
      pos[k] = i;
      if (kSorted && i < ne) pos[k] = a.perm[e0 + i];
    }
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      v[k] = i < ne ? a.contrib_in[idx[k]] : 0.f;
    }
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      if (i < ne) vals[pos[k]] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < nrows) {
      const uint32_t lo = uint32_t(rofs[threadIdx.x] - e0), hi = uint32_t(rofs[threadIdx.x + 1] - e0);
      Neu s = Neu::zero();
      for (uint32_t j = lo; j < hi; j++) s.add(vals[j]);
      dsum = finish<kMode>(a, it.row_begin + threadIdx.x, s);
    }
  } else {
    const uint32_t r = it.row_begin;
    const uint64_t rb = a.ofs[r], re = a.ofs[r + 1];
    const uint64_t lo = rb + uint64_t(it.piece) * kPiece;
    const uint64_t hi = min(re, lo + kPiece);
    Neu s = Neu::zero();
    for (uint64_t e = lo + threadIdx.x; e < hi; e += kBlock) s.add(a.contrib_in[a.adj[e]]);
    s = wave_reduce(s);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      Neu t = red[0];
      for (int i = 1; i < kBlock / 64; i++) t.merge(red[i]);
      if (it.npieces == 1) dsum = finish<kMode>(a, r, t);
      else a.part[blockIdx.x] = t;
    }
  }
  block_add_diff(a, dsum);
}

// Rows split into several pieces: combine the partials in piece order.
template <int kMode>
__global__ void k_pr_fixup(PrArgs a, uint64_t nitems) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  float dsum = 0.f;
  if (i < nitems) {
    const RowItem it = a.items[i];
    if (it.npieces > 1 && it.piece == 0) {
      Neu t = a.part[i];
      for (uint32_t p = 1; p < it.npieces; p++) t.merge(a.part[i + p]);
      dsum = finish<kMode>(a, it.row_begin, t);
    }
  }
  block_add_diff(a, dsum);
}

// Segmented pull: block b finalises rows [256 b, 256 b + 256), adding their
// pair partials segment by segment (a row has at most one pair per segment,
// so the threads of one segment never share an accumulator).
__global__ __launch_bounds__(kBlock) void k_pr_merge(PrArgs a, const uint32_t* pair_row,
                                                     const uint32_t* mt, uint32_t S, uint64_t nb,
                                                     uint64_t V) {
  static_assert(kBlock == kRowsPerItem, "one thread per row of the block");
  __shared__ Neu acc[kBlock];
  const uint64_t b = blockIdx.x;
  const uint32_t r0 = uint32_t(b * kRowsPerItem);
  acc[threadIdx.x] = Neu::zero();
  __syncthreads();
  for (uint32_t s = 0; s < S; s++) {
    const uint32_t lo = mt[uint64_t(s) * (nb + 1) + b], hi = mt[uint64_t(s) * (nb + 1) + b + 1];
    if (lo == hi) continue;  // uniform over the block
    const uint32_t p = lo + threadIdx.x;
    if (p < hi) acc[pair_row[p] - r0].merge(a.pair_part[p]);
    __syncthreads();
  }
  float dsum = 0.f;
  const uint64_t r = uint64_t(r0) + threadIdx.x;
  if (r < V) dsum = finalize(a, uint32_t(r), acc[threadIdx.x].value());
  block_add_diff(a, dsum);
}

__global__ void k_pr_init(const uint32_t* outdeg, const uint32_t* inv, float* rank,
                          float* contrib, uint64_t V, float r0) {
  uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= V) return;
  rank[v] = r0;
  uint32_t od = outdeg[v];
  contrib[inv ? inv[v] : v] = od ? r0 / float(od) : 0.f;
}

// Locality order for the contribution vector: slot = rank of the vertex by
// descending out-degree (ties by vertex ID), so the entries most edges gather
// share few cache lines (the hottest ones stay in each XCD's L2).  A stable
// radix sort of (maxdeg - outdeg) with the vertex IDs as values needs only the
// bits of maxdeg.
__global__ void k_max_u32(const uint32_t* x, uint64_t n, uint32_t* out) {
  uint32_t m = 0;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    m = max(m, x[i]);
  for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
  __shared__ uint32_t part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < int(blockDim.x / 64); i++) m = max(m, part[i]);
    atomicMax(out, m);
  }
}
__global__ void k_deg_keys(const uint32_t* outdeg, uint32_t maxdeg, uint32_t* keys, uint32_t* ids,
                           uint64_t V) {
  uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V) {
    keys[v] = maxdeg - outdeg[v];
    ids[v] = uint32_t(v);
  }
}
__global__ void k_inv_from_sorted(const uint32_t* sorted_ids, uint32_t* inv, uint64_t V) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < V) inv[sorted_ids[i]] = uint32_t(i);
}
__global__ void k_relabel(const uint32_t* adj, const uint32_t* inv, uint32_t* out, uint64_t E) {
  uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < E) out[e] = inv[adj[e]];
}

// Per-item slot order (option pr_item_sort): the entries of every row-block
// item (a run of short rows, or one piece of a long row; at most 8192) sorted
// ascending by contribution slot with a bitonic network in LDS, keeping each
// entry's original in-item position.
constexpr uint32_t kSortMax = 8192;
static_assert(kSortMax >= kPiece && kSortMax >= kMaxCap + kMaxCap / 2, "items fit the sort");
__global__ __launch_bounds__(kBlock) void k_item_sort(const RowItem* items, const uint64_t* ofs,
                                                      const uint32_t* adj, uint32_t* sadj,
                                                      uint16_t* sperm) {
  __shared__ uint64_t key[kSortMax];
  const RowItem it = items[blockIdx.x];
  uint64_t lo, hi;
  if (it.npieces == 0) {
    lo = ofs[it.row_begin];
    hi = ofs[it.row_end];
  } else {
    lo = ofs[it.row_begin] + uint64_t(it.piece) * kPiece;
    hi = min(ofs[it.row_begin + 1], lo + kPiece);
  }
  const uint32_t n = uint32_t(hi - lo);
  uint32_t P = 1;
  while (P < n) P <<= 1;
  for (uint32_t i = threadIdx.x; i < P; i += kBlock)
    key[i] = i < n ? (uint64_t(adj[lo + i]) << 16) | i : ~uint64_t(0);
  __syncthreads();
  for (uint32_t k = 2; k <= P; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = threadIdx.x; i < P; i += kBlock) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const uint64_t x = key[i], y = key[l];
          if ((x > y) == ((i & k) == 0)) {
            key[i] = y;
            key[l] = x;
          }
        }
      }
      __syncthreads();
    }
  for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
    sadj[lo + i] = uint32_t(key[i] >> 16);
    sperm[lo + i] = uint16_t(key[i] & 0xffffu);
  }
}

__global__ void k_widen(const float* in, double* out, uint64_t n) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = double(in[i]);
}

__global__ void k_zero_f64(double* p) { *p = 0.0; }
__global__ void k_fill_u32(uint32_t* p, uint32_t v) { *p = v; }

}  // namespace

template <int kMode, bool kSorted>
void launch_pr_pull_t(uint32_t cap, unsigned nitems, const PrArgs& a, hipStream_t s) {
  switch (cap) {
    case 512: k_pr_pull<512, kMode, kSorted><<<nitems, kBlock, 0, s>>>(a); break;
    case 1024: k_pr_pull<1024, kMode, kSorted><<<nitems, kBlock, 0, s>>>(a); break;
    case 2048: k_pr_pull<2048, kMode, kSorted><<<nitems, kBlock, 0, s>>>(a); break;
    case 4096: k_pr_pull<4096, kMode, kSorted><<<nitems, kBlock, 0, s>>>(a); break;
    default: throw Error{LL_GPU_EINVAL, "unsupported PageRank row-block capacity"};
  }
}

template <int kMode>
void launch_pr_pull(uint32_t cap, bool sorted, unsigned nitems, const PrArgs& a, hipStream_t s) {
  if (sorted && kMode == kFinal) launch_pr_pull_t<kFinal, true>(cap, nitems, a, s);
  else launch_pr_pull_t<kMode, false>(cap, nitems, a, s);
}

// Run `iters` pull iterations; leaves FP32 ranks in rank_final (device).
void pagerank_device(ll_gpu_graph* g, int level, int iters, double damping, uint32_t cap,
                     DevBuf& rank_final, double* last_diff) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V;
  // Binned pull over the out-CSR (pr_bins.hip) unless turned off: bins of
  // 2^pr_bin_bits vertices.  Automatic (-1), while E < 2^32: bins of 2^13
  // while a chunk (the slots from one bin into another) holds 32 slots on
  // average, else 2^14 (R-MAT s24: 69 slots per chunk, 1.10 ms per iteration
  // against 1.49 ms with 2^14; s27: 8.6 slots per chunk, 17.1 ms against
  // 9.2 ms with 2^14), as long as the bins of 2^14 number at most 16384.
  {
    const long bin_bits = ctx->options.count("pr_bin_bits") ? ctx->options["pr_bin_bits"] : -1;
    if (PrBins* pb = ensure_pr_bins(g, level, bin_bits)) {
      const Csr& out = ensure_csr(g, LL_GPU_DIR_OUT, level);
      const uint32_t* od = ensure_degrees(g, LL_GPU_DIR_OUT, level);
      pagerank_bins(g, out, od, *pb, iters, damping, rank_final, last_diff);
      return;
    }
  }
  const Csr& in = ensure_csr(g, LL_GPU_DIR_IN, level);
  const uint32_t* outdeg = ensure_degrees(g, LL_GPU_DIR_OUT, level);
  // contributions in descending-out-degree slot order unless turned off
  const bool reorder = !ctx->options.count("pr_reorder") || ctx->options["pr_reorder"] != 0;
  // column segments of 2^pr_seg_bits slots (0 = off).  Automatic (-1) = off:
  // at R-MAT scale 24 the degree-ordered plain pull is faster (2.24 ms vs
  // 2.44 ms per iteration; the merge costs more than the L2 misses it saves).
  long seg_bits = ctx->options.count("pr_seg_bits") ? ctx->options["pr_seg_bits"] : -1;
  if (seg_bits < 0) seg_bits = 0;
  LLG_CHECK(seg_bits <= 31, LL_GPU_EINVAL, "pr_seg_bits must be <= 31");
  const bool segmented = seg_bits > 0;
  const uint32_t* adj = in.adj.as<uint32_t>();
  const uint32_t* inv = nullptr;
  if (reorder) {
    if (g->pr_level != in.level) {
      TimedLaunch t(ctx, "pr_reorder");
      DevBuf mx, keys, ids, skeys, sids;
      mx.alloc(4, s);
      k_fill_u32<<<1, 1, 0, s>>>(mx.as<uint32_t>(), 0u);
      k_max_u32<<<std::min(grid_for(V, 256), 4u * unsigned(ctx->num_cus)), 256, 0, s>>>(
          outdeg, V, mx.as<uint32_t>());
      LLG_HIP(hipMemcpyAsync(ctx->readback, mx.p, 4, hipMemcpyDeviceToHost, s));
      LLG_HIP(hipStreamSynchronize(s));
      const uint32_t maxdeg = *static_cast<const uint32_t*>(ctx->readback);
      int bits = 1;
      while (bits < 32 && (maxdeg >> bits) != 0) bits++;
      keys.alloc(V * 4, s);
      ids.alloc(V * 4, s);
      skeys.alloc(V * 4, s);
      sids.alloc(V * 4, s);
      k_deg_keys<<<grid_for(V, 256), 256, 0, s>>>(outdeg, maxdeg, keys.as<uint32_t>(),
                                                  ids.as<uint32_t>(), V);
      sort_pairs_u32_u32(keys.as<uint32_t>(), skeys.as<uint32_t>(), ids.as<uint32_t>(),
                         sids.as<uint32_t>(), V, bits, s);
      g->pr_inv.alloc(V * 4, s);
      k_inv_from_sorted<<<grid_for(V, 256), 256, 0, s>>>(sids.as<uint32_t>(),
                                                         g->pr_inv.as<uint32_t>(), V);
      LLG_HIP(hipGetLastError());
      g->pr_adj.release();
      g->pr_sadj.release();
      g->pr_sperm.release();
      g->pr_level = in.level;
    }
    inv = g->pr_inv.as<uint32_t>();
    if (!segmented) {
      if (!g->pr_adj.p) {
        TimedLaunch t(ctx, "pr_reorder");
        g->pr_adj.alloc(std::max<uint64_t>(in.E, 1) * 4, s);
        k_relabel<<<grid_for(in.E, 256), 256, 0, s>>>(in.adj.as<uint32_t>(), inv,
                                                      g->pr_adj.as<uint32_t>(), in.E);
        LLG_HIP(hipGetLastError());
      }
      adj = g->pr_adj.as<uint32_t>();
    }
  }
  PrSegments& ps = g->pr_seg;
  if (segmented && !(ps.valid && ps.level == in.level && ps.req_zb == uint32_t(seg_bits) &&
                     ps.slots == reorder)) {
    TimedLaunch t(ctx, "pr_segments");
    build_pr_segments(g, in, inv, uint32_t(seg_bits), ps);
    ps.req_zb = uint32_t(seg_bits);
    ps.slots = reorder;
  }
  const RowBlocks& rb = segmented ? pr_segment_rowblocks(ctx, ps, cap)
                                  : ensure_rowblocks(g, LL_GPU_DIR_IN, level, cap, kPiece);
  // hot/cold split at slot 2^pr_hot_bits (0 = off, -1 = automatic = off)
  long hot_bits = ctx->options.count("pr_hot_bits") ? ctx->options["pr_hot_bits"] : -1;
  if (hot_bits < 0) hot_bits = 0;
  LLG_CHECK(hot_bits <= 31, LL_GPU_EINVAL, "pr_hot_bits must be <= 31");
  const bool hotcold = reorder && !segmented && hot_bits > 0;
  PrHotCold& hc = g->pr_hc;
  if (hotcold) {
    if (!(hc.valid && hc.level == in.level && hc.hb == uint32_t(hot_bits))) {
      TimedLaunch t(ctx, "pr_hotcold");
      build_pr_hotcold(g, in, adj, uint32_t(hot_bits), hc);
    }
    if (!hc.hot_rb.valid || hc.hot_rb.cap != cap) build_rowblocks(ctx, hc.hot, cap, kPiece, hc.hot_rb);
    if (!hc.cold_rb.valid || hc.cold_rb.cap != cap)
      build_rowblocks(ctx, hc.cold, cap, kPiece, hc.cold_rb);
  }
  // entries of every item sorted by slot (plain pull over slots only)
  const long sort_opt = ctx->options.count("pr_item_sort") ? ctx->options["pr_item_sort"] : -1;
  const bool item_sort = reorder && !segmented && !hotcold && sort_opt > 0;
  const uint16_t* perm = nullptr;
  if (item_sort) {
    if (!g->pr_sadj.p || g->pr_sort_cap != cap) {
      TimedLaunch t(ctx, "pr_item_sort");
      g->pr_sadj.alloc(std::max<uint64_t>(in.E, 1) * 4, s);
      g->pr_sperm.alloc(std::max<uint64_t>(in.E, 1) * 2, s);
      k_item_sort<<<unsigned(rb.n), kBlock, 0, s>>>(rb.items.as<RowItem>(), in.ofs.as<uint64_t>(),
                                                    adj, g->pr_sadj.as<uint32_t>(),
                                                    g->pr_sperm.as<uint16_t>());
      LLG_HIP(hipGetLastError());
      g->pr_sort_cap = cap;
    }
    adj = g->pr_sadj.as<uint32_t>();
    perm = g->pr_sperm.as<uint16_t>();
  }
  DevBuf rank[2], contrib[2], part, diff, pair_part;
  for (int i = 0; i < 2; i++) {
    rank[i].alloc(V * 4, s);
    contrib[i].alloc(V * 4, s);
  }
  uint64_t nparts = std::max<uint64_t>(rb.n, 1);
  if (hotcold) nparts = std::max(nparts, std::max(hc.hot_rb.n, hc.cold_rb.n));
  part.alloc(nparts * sizeof(Neu), s);
  if (segmented) pair_part.alloc(std::max<uint64_t>(ps.P, 1) * sizeof(Neu), s);
  if (hotcold) pair_part.alloc(V * sizeof(Neu), s);  // hot partial per vertex
  diff.alloc(8, s);
  const float N = float(V);
  k_pr_init<<<grid_for(V, 256), 256, 0, s>>>(outdeg, inv, rank[0].as<float>(),
                                             contrib[0].as<float>(), V, float(1.0 / double(V)));
  k_zero_f64<<<1, 1, 0, s>>>(diff.as<double>());
  LLG_HIP(hipGetLastError());
  int cur = 0;
  for (int it = 0; it < iters; it++) {
    PrArgs a;
    a.items = rb.items.as<RowItem>();
    a.ofs = segmented ? ps.pairs.ofs.as<uint64_t>() : in.ofs.as<uint64_t>();
    a.adj = segmented ? ps.pairs.adj.as<uint32_t>() : adj;
    a.inv = inv;
    a.contrib_in = contrib[cur].as<float>();
    a.rank_in = rank[cur].as<float>();
    a.outdeg = outdeg;
    a.rank_out = rank[cur ^ 1].as<float>();
    a.contrib_out = contrib[cur ^ 1].as<float>();
    a.part = part.as<Neu>();
    a.pair_part = pair_part.as<Neu>();
    a.perm = perm;
    a.diff = (it == iters - 1 && last_diff) ? diff.as<double>() : nullptr;
    a.base = float((1.0 - damping) / double(N));
    a.d = float(damping);
    if (segmented) {
      PrArgs pa = a;
      pa.diff = nullptr;
      if (rb.n) {
        {
          TimedLaunch t(ctx, "k_pr_pull");
          launch_pr_pull<kPartOut>(cap, false, unsigned(rb.n), pa, s);
        }
        TimedLaunch t(ctx, "k_pr_fixup");
        k_pr_fixup<kPartOut><<<grid_for(rb.n, 256), 256, 0, s>>>(pa, rb.n);
      }
      TimedLaunch t(ctx, "k_pr_merge");
      k_pr_merge<<<unsigned(ps.nb), kBlock, 0, s>>>(a, ps.pair_row.as<uint32_t>(),
                                                    ps.mt.as<uint32_t>(), ps.S, ps.nb, V);
    } else if (hotcold) {
      // hot pass: partial of every vertex over its hot in-edges (L2-resident
      // gathers); cold pass: continue from it over the cold in-edges, finalise
      PrArgs h = a;
      h.items = hc.hot_rb.items.as<RowItem>();
      h.ofs = hc.hot.ofs.as<uint64_t>();
      h.adj = hc.hot.adj.as<uint32_t>();
      h.diff = nullptr;
      {
        TimedLaunch t(ctx, "k_pr_pull");
        launch_pr_pull<kPartOut>(cap, false, unsigned(hc.hot_rb.n), h, s);
      }
      {
        TimedLaunch t(ctx, "k_pr_fixup");
        k_pr_fixup<kPartOut><<<grid_for(hc.hot_rb.n, 256), 256, 0, s>>>(h, hc.hot_rb.n);
      }
      PrArgs c = a;
      c.items = hc.cold_rb.items.as<RowItem>();
      c.ofs = hc.cold.ofs.as<uint64_t>();
      c.adj = hc.cold.adj.as<uint32_t>();
      {
        TimedLaunch t(ctx, "k_pr_pull_cold");
        launch_pr_pull<kPartIn>(cap, false, unsigned(hc.cold_rb.n), c, s);
      }
      TimedLaunch t(ctx, "k_pr_fixup");
      k_pr_fixup<kPartIn><<<grid_for(hc.cold_rb.n, 256), 256, 0, s>>>(c, hc.cold_rb.n);
    } else {
      {
        TimedLaunch t(ctx, "k_pr_pull");
        launch_pr_pull<kFinal>(cap, item_sort, unsigned(rb.n), a, s);
      }
      TimedLaunch t(ctx, "k_pr_fixup");
      k_pr_fixup<kFinal><<<grid_for(rb.n, 256), 256, 0, s>>>(a, rb.n);
    }
    LLG_HIP(hipGetLastError());
    cur ^= 1;
  }
  rank_final = std::move(rank[cur]);
  if (last_diff) {
    double h = 0;
    LLG_HIP(hipMemcpyAsync(&h, diff.p, 8, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    *last_diff = iters > 0 ? h : 0.0;
  }
}

void widen_to_host(ll_gpu_graph* g, const float* dev, double* host, uint64_t n) {
  hipStream_t s = g->ctx->stream;
  DevBuf w;
  w.alloc(n * 8, s);
  k_widen<<<grid_for(n, 256), 256, 0, s>>>(dev, w.as<double>(), n);
  LLG_HIP(hipGetLastError());
  LLG_HIP(hipMemcpyAsync(host, w.p, n * 8, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
}

}  // namespace llgpu

extern "C" {

int ll_gpu_pagerank_ex(ll_gpu_graph* g, int level, int iters, double damping, uint32_t cap,
                       double* ranks_out, double* last_diff) {
  if (!g) return LL_GPU_EINVAL;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    LLG_CHECK(iters >= 1, LL_GPU_EINVAL, "iters must be >= 1 (the reference runs do/while)");
    LLG_CHECK(damping >= 0 && damping <= 1, LL_GPU_EINVAL, "damping must be in [0, 1]");
    if (cap == 0) {
      auto it = g->ctx->options.find("pr_cap");
      cap = it != g->ctx->options.end() ? uint32_t(it->second) : 2048u;
    }
    llgpu::DevBuf r;
    llgpu::pagerank_device(g, level, iters, damping, cap, r, last_diff);
    if (ranks_out) llgpu::widen_to_host(g, r.as<float>(), ranks_out, g->V);
    else LLG_HIP(hipStreamSynchronize(g->ctx->stream));
    return LL_GPU_OK;
  } catch (const llgpu::Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

int ll_gpu_pagerank(ll_gpu_graph* g, int level, int iters, double damping, double* ranks_out,
                    double* last_diff) {
  return ll_gpu_pagerank_ex(g, level, iters, damping, 0, ranks_out, last_diff);
}

}  // extern "C"
