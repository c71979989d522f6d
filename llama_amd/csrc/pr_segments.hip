// Column segments of the PageRank in-CSR.
//
// The pull gathers one 4 B contribution per in-edge from a vector of V
// entries; at R-MAT scale 24 that vector is 64 MiB, far beyond an XCD's 4 MiB
// L2, so most gathers cost a full cache line from the fabric.  The
// contribution slots (vertices in descending out-degree order) are therefore
// cut into segments of 2^zb slots, small enough for one segment to stay
// resident in every XCD's L2 next to the streamed neighbour IDs, and the
// in-edges are stably partitioned by the segment of their source slot:
//
//   seg_adj   = in-neighbour slots, segment-major; inside a segment rows
//               ascend and each row keeps its snapshot edge order;
//   pair      = (row, segment) cell with at least one edge: a row of the
//               "pair CSR" (pair_ofs into seg_adj, pair_row = its row);
//   mt[s][b]  = first pair of segment s whose row is >= 256*b, so the merge
//               of row block b finds its pairs of every segment directly.
//
// An iteration reduces every pair to a compensated partial (the workgroups in
// flight at any moment gather from one segment) and merges each row's
// partials in segment order: the summation order is fixed, so results do not
// depend on scheduling.  Built once per snapshot level and cached.
#include "ll_gpu_internal.h"
#include "rowblocks.h"

namespace llgpu {

namespace {

constexpr int kB = 256;
constexpr uint32_t kTile = kB * 8;  // positions per block of the pair compaction

__device__ inline void item_range(const RowItem& it, const uint64_t* ofs, uint64_t& lo,
                                  uint64_t& hi) {
  if (it.npieces == 0) {
    lo = ofs[it.row_begin];
    hi = ofs[it.row_end];
  } else {
    const uint64_t rb = ofs[it.row_begin], re = ofs[it.row_begin + 1];
    lo = rb + uint64_t(it.piece) * kPiece;
    hi = min(re, lo + kPiece);
  }
}

// Lanes of the wave whose value `sg` (nbits wide) equals this lane's, among
// the lanes with `valid` set.
__device__ inline uint64_t wave_match(uint32_t sg, bool valid, uint32_t nbits) {
  uint64_t m = __ballot(valid);
  for (uint32_t b = 0; b < nbits; b++) {
    const bool bit = (sg >> b) & 1u;
    const uint64_t w = __ballot(bit);
    m &= bit ? w : ~w;
  }
  return m;
}

__global__ void k_slot_of(const uint32_t* adj, const uint32_t* inv, uint32_t* out, uint64_t E) {
  uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e < E) out[e] = inv[adj[e]];
}

// Per item of the in-CSR row blocks: edges per segment -> cnt[s * n + item].
__global__ __launch_bounds__(kB) void k_seg_hist(const RowItem* items, uint64_t n,
                                                 const uint64_t* ofs, const uint32_t* slot,
                                                 uint32_t zb, uint32_t S, uint32_t nbits,
                                                 uint32_t* cnt) {
  __shared__ uint32_t h[kMaxSegments];
  for (uint32_t i = threadIdx.x; i < S; i += kB) h[i] = 0;
  __syncthreads();
  uint64_t lo, hi;
  item_range(items[blockIdx.x], ofs, lo, hi);
  const uint32_t lane = threadIdx.x & 63;
  for (uint64_t r0 = lo; r0 < hi; r0 += kB) {
    const uint64_t e = r0 + threadIdx.x;
    const bool valid = e < hi;
    const uint32_t sg = valid ? min(slot[e] >> zb, S - 1) : 0u;
    const uint64_t m = wave_match(sg, valid, nbits);
    // the lowest lane of each group adds the group's size
    if (valid && (m & ((uint64_t(1) << lane) - 1)) == 0) atomicAdd(&h[sg], uint32_t(__popcll(m)));
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < S; i += kB) cnt[uint64_t(i) * n + blockIdx.x] = h[i];
}

// Stable scatter of every item's edges to their segment-major positions:
// base[s * n + item] (exclusive scan of cnt) + rank among the item's earlier
// edges of the same segment.  Positions from `split` on go to out_hi (at
// pos - split), the others to out_lo.  The segment of a slot is
// min(slot >> zb, S - 1), so the last segment may be wider.
__global__ __launch_bounds__(kB) void k_seg_scatter(const RowItem* items, uint64_t n,
                                                    const uint64_t* ofs, const uint32_t* slot,
                                                    uint32_t zb, uint32_t S, uint32_t nbits,
                                                    const uint64_t* base, uint64_t split,
                                                    uint32_t* out_lo, uint32_t* out_hi,
                                                    uint32_t* out_row) {
  __shared__ uint64_t rofs[kRowsPerItem + 1];
  __shared__ uint64_t sbase[kMaxSegments];
  __shared__ uint32_t wc[kB / 64][kMaxSegments];
  const RowItem it = items[blockIdx.x];
  const bool stream = it.npieces == 0;
  const uint32_t nrows = stream ? it.row_end - it.row_begin : 1u;
  if (stream)
    for (uint32_t i = threadIdx.x; i <= nrows; i += kB) rofs[i] = ofs[it.row_begin + i];
  for (uint32_t i = threadIdx.x; i < S; i += kB) sbase[i] = base[uint64_t(i) * n + blockIdx.x];
  __syncthreads();
  uint64_t lo, hi;
  if (stream) {
    lo = rofs[0];
    hi = rofs[nrows];
  } else {
    item_range(it, ofs, lo, hi);
  }
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint64_t r0 = lo; r0 < hi; r0 += kB) {
    for (uint32_t i = threadIdx.x; i < (kB / 64) * S; i += kB) wc[i / S][i % S] = 0;
    __syncthreads();
    const uint64_t e = r0 + threadIdx.x;
    const bool valid = e < hi;
    const uint32_t sl = valid ? slot[e] : 0u;
    const uint32_t sg = min(sl >> zb, S - 1);
    const uint64_t m = wave_match(sg, valid, nbits);
    const uint32_t rank = __popcll(m & ((uint64_t(1) << lane) - 1));
    if (valid && (m >> lane) == 1u) wc[w][sg] = rank + 1;  // highest lane of its group
    __syncthreads();
    if (valid) {
      uint64_t pos = sbase[sg] + rank;
      for (uint32_t v = 0; v < w; v++) pos += wc[v][sg];
      uint32_t row = it.row_begin;
      if (stream) {  // last row whose start is <= e
        uint32_t a = 0, b = nrows - 1;
        while (a < b) {
          const uint32_t mid = (a + b + 1) >> 1;
          if (rofs[mid] <= e) a = mid;
          else b = mid - 1;
        }
        row += a;
      }
      if (pos < split) out_lo[pos] = sl;
      else out_hi[pos - split] = sl;
      out_row[pos] = row;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < S; i += kB) {
      uint32_t t = 0;
      for (uint32_t v = 0; v < kB / 64; v++) t += wc[v][i];
      sbase[i] += t;
    }
    __syncthreads();
  }
}

__device__ inline bool pair_start(const uint32_t* sl, const uint32_t* row, uint64_t p,
                                  uint32_t zb) {
  return p == 0 || row[p] != row[p - 1] || (sl[p] >> zb) != (sl[p - 1] >> zb);
}

__global__ __launch_bounds__(kB) void k_pair_count(const uint32_t* sl, const uint32_t* row,
                                                   uint64_t E, uint32_t zb, uint32_t* tcnt) {
  const uint64_t t0 = uint64_t(blockIdx.x) * kTile;
  uint32_t c = 0;
  for (uint32_t k = 0; k < kTile / kB; k++) {
    const uint64_t p = t0 + k * kB + threadIdx.x;
    c += (p < E && pair_start(sl, row, p, zb)) ? 1u : 0u;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  __shared__ uint32_t red[kB / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < kB / 64; i++) t += red[i];
    tcnt[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(kB) void k_pair_fill(const uint32_t* sl, const uint32_t* row,
                                                  uint64_t E, uint32_t zb, const uint64_t* tpos,
                                                  uint64_t* pair_ofs, uint32_t* pair_row) {
  __shared__ uint32_t wsum[kB / 64];
  __shared__ uint64_t run;
  const uint64_t t0 = uint64_t(blockIdx.x) * kTile;
  if (threadIdx.x == 0) run = tpos[blockIdx.x];
  const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint32_t k = 0; k < kTile / kB; k++) {
    const uint64_t p = t0 + k * kB + threadIdx.x;
    const bool f = p < E && pair_start(sl, row, p, zb);
    const uint64_t m = __ballot(f);
    if (lane == 0) wsum[w] = __popcll(m);
    __syncthreads();
    uint32_t tot = 0;
    for (uint32_t v = 0; v < kB / 64; v++) tot += wsum[v];
    if (f) {
      uint64_t idx = run + __popcll(m & ((uint64_t(1) << lane) - 1));
      for (uint32_t v = 0; v < w; v++) idx += wsum[v];
      pair_ofs[idx] = p;
      pair_row[idx] = row[p];
    }
    __syncthreads();  // wsum and run are read before they change
    if (threadIdx.x == 0) run += tot;
  }
}

__global__ void k_set_u64(uint64_t* p, uint64_t v) { *p = v; }

// pseg[s] = first pair whose edges start at or after segment s's first edge.
__global__ void k_seg_first_pair(const uint64_t* pair_ofs, uint64_t P, const uint64_t* base,
                                 uint64_t n, uint32_t S, uint64_t E, uint32_t* pseg) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > S) return;
  const uint64_t target = s < S ? base[uint64_t(s) * n] : E;
  uint64_t a = 0, b = P;  // lower bound in pair_ofs[0, P)
  while (a < b) {
    const uint64_t mid = (a + b) >> 1;
    if (pair_ofs[mid] < target) a = mid + 1;
    else b = mid;
  }
  pseg[s] = uint32_t(a);
}

// mt[s * (nb + 1) + b] = first pair of segment s with row >= b * kRowsPerItem.
__global__ void k_merge_table(const uint32_t* pair_row, const uint32_t* pseg, uint32_t S,
                              uint64_t nb, uint32_t* mt) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= uint64_t(S) * (nb + 1)) return;
  const uint32_t s = uint32_t(i / (nb + 1));
  const uint64_t b = i % (nb + 1);
  uint32_t a = pseg[s], e = pseg[s + 1];
  if (b < nb) {
    const uint64_t target = b * kRowsPerItem;
    while (a < e) {
      const uint32_t mid = a + ((e - a) >> 1);
      if (pair_row[mid] < target) a = mid + 1;
      else e = mid;
    }
  } else {
    a = e;
  }
  mt[i] = a;
}

}  // namespace

void build_pr_segments(ll_gpu_graph* g, const Csr& in, const uint32_t* inv, uint32_t zb,
                       PrSegments& ps) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = in.V, E = in.E;
  LLG_CHECK(E < (uint64_t(1) << 32), LL_GPU_ERANGE, "segmented PageRank needs < 2^32 edges");
  uint32_t S = uint32_t((V + (uint64_t(1) << zb) - 1) >> zb);
  while (S > kMaxSegments) {
    zb++;
    S = uint32_t((V + (uint64_t(1) << zb) - 1) >> zb);
  }
  S = std::max(S, 1u);
  uint32_t nbits = 0;
  while ((1u << nbits) < S) nbits++;
  ps = PrSegments();
  ps.level = in.level;
  ps.zb = zb;
  ps.S = S;
  ps.nb = (V + kRowsPerItem - 1) / kRowsPerItem;
  // the partition walks the items of the in-CSR row blocks (edge order)
  const RowBlocks& irb = ensure_rowblocks(g, LL_GPU_DIR_IN, in.level, 2048, kPiece);
  const uint64_t n = irb.n;
  DevBuf slot, cnt, base, rowtmp;
  const uint32_t* sl = in.adj.as<uint32_t>();
  if (inv) {
    slot.alloc(std::max<uint64_t>(E, 1) * 4, s);
    k_slot_of<<<grid_for(E, 256), 256, 0, s>>>(in.adj.as<uint32_t>(), inv, slot.as<uint32_t>(), E);
    sl = slot.as<uint32_t>();
  }
  cnt.alloc(std::max<uint64_t>(uint64_t(S) * n, 1) * 4, s);
  base.alloc((uint64_t(S) * n + 1) * 8, s);
  if (n) {
    k_seg_hist<<<unsigned(n), kB, 0, s>>>(irb.items.as<RowItem>(), n, in.ofs.as<uint64_t>(), sl,
                                          zb, S, nbits, cnt.as<uint32_t>());
    LLG_HIP(hipGetLastError());
  }
  scan_u32_to_u64(cnt.as<uint32_t>(), base.as<uint64_t>(), uint64_t(S) * n, s);
  ps.pairs.adj.alloc(std::max<uint64_t>(E, 1) * 4, s);
  rowtmp.alloc(std::max<uint64_t>(E, 1) * 4, s);
  if (n) {
    k_seg_scatter<<<unsigned(n), kB, 0, s>>>(irb.items.as<RowItem>(), n, in.ofs.as<uint64_t>(),
                                             sl, zb, S, nbits, base.as<uint64_t>(), E,
                                             ps.pairs.adj.as<uint32_t>(), nullptr,
                                             rowtmp.as<uint32_t>());
    LLG_HIP(hipGetLastError());
  }
  slot.release();
  // pairs = starts of (row, segment) runs
  const uint64_t ntiles = (E + kTile - 1) / kTile;
  DevBuf tcnt, tpos;
  tcnt.alloc(std::max<uint64_t>(ntiles, 1) * 4, s);
  tpos.alloc((ntiles + 1) * 8, s);
  if (ntiles) {
    k_pair_count<<<unsigned(ntiles), kB, 0, s>>>(ps.pairs.adj.as<uint32_t>(), rowtmp.as<uint32_t>(), E,
                                                 zb, tcnt.as<uint32_t>());
    LLG_HIP(hipGetLastError());
  }
  scan_u32_to_u64(tcnt.as<uint32_t>(), tpos.as<uint64_t>(), ntiles, s);
  uint64_t P = 0;
  LLG_HIP(hipMemcpyAsync(&P, tpos.as<uint64_t>() + ntiles, 8, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
  LLG_CHECK(P < (uint64_t(1) << 32) - 1, LL_GPU_ERANGE, "too many (row, segment) pairs");
  ps.P = P;
  ps.pairs.ofs.alloc((P + 1) * 8, s);
  ps.pair_row.alloc(std::max<uint64_t>(P, 1) * 4, s);
  if (ntiles) {
    k_pair_fill<<<unsigned(ntiles), kB, 0, s>>>(ps.pairs.adj.as<uint32_t>(), rowtmp.as<uint32_t>(), E,
                                                zb, tpos.as<uint64_t>(), ps.pairs.ofs.as<uint64_t>(),
                                                ps.pair_row.as<uint32_t>());
  }
  k_set_u64<<<1, 1, 0, s>>>(ps.pairs.ofs.as<uint64_t>() + P, E);
  LLG_HIP(hipGetLastError());
  rowtmp.release();
  ps.pseg.alloc((S + 1) * 4, s);
  k_seg_first_pair<<<grid_for(S + 1, 64), 64, 0, s>>>(ps.pairs.ofs.as<uint64_t>(), P,
                                                      base.as<uint64_t>(), n, S, E,
                                                      ps.pseg.as<uint32_t>());
  const uint64_t nmt = uint64_t(S) * (ps.nb + 1);
  ps.mt.alloc(nmt * 4, s);
  k_merge_table<<<grid_for(nmt, 256), 256, 0, s>>>(ps.pair_row.as<uint32_t>(),
                                                   ps.pseg.as<uint32_t>(), S, ps.nb,
                                                   ps.mt.as<uint32_t>());
  LLG_HIP(hipGetLastError());
  ps.pairs.valid = true;
  ps.pairs.level = in.level;
  ps.pairs.V = P;
  ps.pairs.E = E;
  ps.valid = true;
}

namespace {

// hot_ofs[r] = first hot position of row r: rows ascend over the hot
// positions, so each run start fills the offsets of the rows up to its own.
__global__ void k_hot_ofs(const uint32_t* row, uint64_t e_hot, uint64_t V, uint64_t* hot_ofs) {
  const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p > e_hot) return;
  const uint64_t prev = p == 0 ? 0 : uint64_t(row[p - 1]) + 1;  // first row not yet covered
  const uint64_t cur = p == e_hot ? V : uint64_t(row[p]);
  for (uint64_t r = prev; r <= cur && r <= V; r++) hot_ofs[r] = p;
}

__global__ void k_cold_ofs(const uint64_t* ofs, const uint64_t* hot_ofs, uint64_t V,
                           uint64_t* cold_ofs) {
  const uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r <= V) cold_ofs[r] = ofs[r] - hot_ofs[r];
}

}  // namespace

void build_pr_hotcold(ll_gpu_graph* g, const Csr& in, const uint32_t* slot_adj, uint32_t hb,
                      PrHotCold& hc) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = in.V, E = in.E;
  hc = PrHotCold();
  hc.level = in.level;
  hc.hb = hb;
  const RowBlocks& irb = ensure_rowblocks(g, LL_GPU_DIR_IN, in.level, 2048, kPiece);
  const uint64_t n = irb.n;
  constexpr uint32_t S = 2, nbits = 1;
  DevBuf cnt, base, rowtmp;
  cnt.alloc(S * n * 4, s);
  base.alloc((S * n + 1) * 8, s);
  k_seg_hist<<<unsigned(n), kB, 0, s>>>(irb.items.as<RowItem>(), n, in.ofs.as<uint64_t>(),
                                        slot_adj, hb, S, nbits, cnt.as<uint32_t>());
  LLG_HIP(hipGetLastError());
  scan_u32_to_u64(cnt.as<uint32_t>(), base.as<uint64_t>(), S * n, s);
  uint64_t e_hot = 0;
  LLG_HIP(hipMemcpyAsync(&e_hot, base.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
  LLG_CHECK(e_hot <= E, LL_GPU_ESTATE, "hot/cold split: inconsistent counts");
  Csr& hot = hc.hot;
  Csr& cold = hc.cold;
  hot.V = cold.V = V;
  hot.E = e_hot;
  cold.E = E - e_hot;
  hot.level = cold.level = in.level;
  hot.adj.alloc(std::max<uint64_t>(hot.E, 1) * 4, s);
  cold.adj.alloc(std::max<uint64_t>(cold.E, 1) * 4, s);
  rowtmp.alloc(std::max<uint64_t>(E, 1) * 4, s);
  k_seg_scatter<<<unsigned(n), kB, 0, s>>>(irb.items.as<RowItem>(), n, in.ofs.as<uint64_t>(),
                                           slot_adj, hb, S, nbits, base.as<uint64_t>(), e_hot,
                                           hot.adj.as<uint32_t>(), cold.adj.as<uint32_t>(),
                                           rowtmp.as<uint32_t>());
  LLG_HIP(hipGetLastError());
  hot.ofs.alloc((V + 1) * 8, s);
  cold.ofs.alloc((V + 1) * 8, s);
  k_hot_ofs<<<grid_for(e_hot + 1, 256), 256, 0, s>>>(rowtmp.as<uint32_t>(), e_hot, V,
                                                     hot.ofs.as<uint64_t>());
  k_cold_ofs<<<grid_for(V + 1, 256), 256, 0, s>>>(in.ofs.as<uint64_t>(), hot.ofs.as<uint64_t>(), V,
                                                  cold.ofs.as<uint64_t>());
  LLG_HIP(hipGetLastError());
  hot.valid = cold.valid = true;
  hc.valid = true;
}

const RowBlocks& pr_segment_rowblocks(ll_gpu_ctx* ctx, PrSegments& ps, uint32_t cap) {
  if (ps.rb.valid && ps.rb.cap == cap) return ps.rb;
  if (ps.P == 0) {  // no edges: nothing to pull, the merge alone finalises
    ps.rb = RowBlocks();
    ps.rb.valid = true;
    ps.rb.cap = cap;
    ps.rb.piece = kPiece;
    return ps.rb;
  }
  build_rowblocks(ctx, ps.pairs, cap, kPiece, ps.rb);
  return ps.rb;
}

}  // namespace llgpu
