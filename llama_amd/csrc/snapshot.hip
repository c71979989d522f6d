// Snapshot materialisation: the multi-level iterator, run once per snapshot.
//
// For snapshot level L the adjacency of v is the chain of segments reached
// from VT_L[v] by following continuation records (iter_begin / iter_next /
// iter_descend, ll_mlcsr_sp.h:1994-2196), minus the words whose max-visible
// level is <= L (is_edge_deleted, :1290-1310).  Concatenated over v in ID
// order, chain order and stored order, the stored words form one "virtual"
// word sequence; the live words of that sequence, in order, are exactly the
// compact CSR adjacency in iterator order.  So:
//   1. thread per vertex: walk the chain through the continuation records only
//      (one dependent 16-byte load per segment), emit the segment count and
//      stored-word count;                                    k_seg_count
//   2. scans give each segment its slot and virtual word offset;
//   3. thread per vertex: write segment records (edge ID of the run start and
//      virtual offset);                                      k_seg_fill
//   4. the row offsets are the scan of the precomputed VT_L degrees (inside
//      an LL_MIN_LEVEL window, of the window's degrees), which already count
//      only live words, so compaction blocks that start at a vertex know
//      where their output starts: block b takes the vertices whose first
//      virtual word lies in the b-th window of 2048 words;
//   5. per block, each word's segment by binary search in an LDS copy of the
//      block's segment range, the word read (coalesced within segments) and
//      tested, the live ones written in order with a block-wide ballot
//      compaction from the first vertex's row offset;        k_compact
// Every block checks its live words against its rows' degrees (a mismatch
// means a missing deletion patch).
#include "ll_gpu_internal.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr int kWordsPerBlock = 2048;
// Segment records a compaction block stages in LDS: a block overlaps about
// 200 segments at R-MAT s24, so room for 512 keeps the LDS per workgroup at
// 8 KiB (eight workgroups per CU); a block overlapping more searches them in
// global memory.
constexpr int kSegLds = 512;
constexpr int kWaves = kBlock / 64;
constexpr int kBatch = kWordsPerBlock / kBlock + 1;  // rounds per compaction batch

__global__ void k_zero_u32(uint32_t* p) { *p = 0; }

__global__ void k_ofs_degrees(const uint64_t* ofs, uint32_t* deg, uint64_t V) {
  uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V) deg[v] = uint32_t(ofs[v + 1] - ofs[v]);
}

__global__ void k_degrees(VtPages vt, uint32_t* deg, uint64_t V) {
  uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V) deg[v] = vt[v].degree;
}

// Chain walk shared by the two per-vertex passes.  Levels strictly decrease
// along a chain and every run (plus its continuation record on levels > 0)
// lies inside its level's edge table; a violation flags the mirror as corrupt
// and stops the walk instead of reading out of bounds.
template <typename F>
__device__ inline void walk_chain(VtPages vt, const uint64_t* const* et,
                                  const uint64_t* et_words, uint32_t level, uint32_t min_level,
                                  uint64_t v, uint32_t* bad, F&& f) {
  VtEntry e = vt[v];
  int64_t start = e.start;
  uint32_t len = e.len;
  uint32_t kprev = level + 1;
  while (start != -1 && len > 0) {
    uint32_t k = edge_level(uint64_t(start));
    if (k < min_level) return;  // LL_MIN_LEVEL window (ll_mlcsr_sp.h:2021-2027, 2107-2111)
    uint64_t idx = edge_index(uint64_t(start));
    if (k >= kprev || idx + len + (k > 0 ? 2 : 0) > et_words[k]) {
      atomicOr(bad, 1u);
      return;
    }
    f(uint64_t(start), len);
    if (k == 0) return;
    const uint64_t* c = et[k] + idx + len;
    start = int64_t(ld_global(c));
    len = uint32_t(ld_global(c + 1));
    kprev = k;
  }
}

// Degrees inside an LL_MIN_LEVEL window: a vertex's live words in the
// segments of levels >= min_level (without deletions, the reference's
// degree(n) = VT_L.degree - VT_(min-1).degree, ll_mlcsr_sp.h:1913-1925; with
// deletions the reference leaves it unimplemented and the live words are
// counted).
__global__ void k_window_degrees(VtPages vt, const uint64_t* const* et, const uint64_t* et_words,
                                 uint32_t level, uint32_t min_level, bool deletions, uint64_t V,
                                 uint32_t* deg, uint32_t* bad) {
  uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= V) return;
  uint32_t d = 0;
  walk_chain(vt, et, et_words, level, min_level, v, bad, [&](uint64_t start, uint32_t len) {
    if (!deletions) {
      d += len;
      return;
    }
    const uint64_t* w = et[edge_level(start)] + edge_index(start);
    for (uint32_t j = 0; j < len; j++) d += word_maxlev(w[j]) > level ? 1u : 0u;
  });
  deg[v] = d;
}

__global__ void k_seg_count(VtPages vt, const uint64_t* const* et,
                            const uint64_t* et_words, uint32_t level, uint32_t min_level,
                            uint64_t V,
                            uint32_t* nseg, uint64_t* nwords, uint32_t* bad) {
  uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= V) return;
  uint32_t cnt = 0;
  uint64_t words = 0;
  walk_chain(vt, et, et_words, level, min_level, v, bad, [&](uint64_t, uint32_t len) {
    cnt++;
    words += len;
  });
  nseg[v] = cnt;
  nwords[v] = words;
}

__global__ void k_seg_fill(VtPages vt, const uint64_t* const* et,
                           const uint64_t* et_words, uint32_t level, uint32_t min_level,
                           uint64_t V,
                           const uint64_t* seg_base, const uint64_t* word_base,
                           uint64_t* seg_addr, uint64_t* seg_vofs, uint32_t* bad) {
  uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= V) return;
  uint64_t s = seg_base[v], w = word_base[v];
  walk_chain(vt, et, et_words, level, min_level, v, bad, [&](uint64_t start, uint32_t len) {
    seg_addr[s] = start;
    seg_vofs[s] = w;
    s++;
    w += len;
  });
}

__global__ void k_seg_total(uint64_t* seg_vofs, uint64_t nseg, uint64_t total) {
  seg_vofs[nseg] = total;
}

// Largest s in [lo, hi) with vofs[s] <= x (vofs ascending, vofs[lo] <= x).
__device__ inline uint64_t seg_search_global(const uint64_t* vofs, uint64_t lo, uint64_t hi,
                                             uint64_t x) {
  while (hi - lo > 1) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (vofs[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// Compaction blocks start at a vertex: block b takes the vertices whose first
// virtual word lies in [b * kWordsPerBlock, (b + 1) * kWordsPerBlock), bv[b]
// being the first of them (bv[nblk] = V), so its live words start at that
// vertex's row offset and no count pass is needed.  A vertex longer than the
// window makes its block longer; the windows it covers hold no vertex start
// and their blocks are empty.
__global__ void k_block_verts(const uint64_t* word_base, uint64_t V, uint64_t nblk,
                              uint64_t* bv) {
  const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b > nblk) return;
  if (b == nblk) {
    bv[b] = V;
    return;
  }
  const uint64_t x = b * kWordsPerBlock;
  uint64_t lo = 0, hi = V;  // first v with word_base[v] >= x
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (word_base[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  bv[b] = lo;
}

// Everything a compaction block needs before it starts, in one record: its
// virtual words [x0, x1), its rows' output range [out0, out1) and the
// segments overlapping it (s_lo, n; n = 0 for an empty block).  One thread per
// block: the lookups and binary searches run in parallel instead of at the
// head of each workgroup.
struct BlockPlan {
  uint64_t x0, x1, out0, out1, s_lo, n;
};

__global__ void k_block_plan(const uint64_t* seg_vofs, uint64_t nseg, const uint64_t* word_base,
                             const uint64_t* bv, const uint64_t* ofs, uint64_t nblk,
                             BlockPlan* plan) {
  const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const uint64_t va = bv[b], vb = bv[b + 1];
  BlockPlan p{word_base[va], word_base[vb], ofs[va], ofs[vb], 0, 0};
  if (p.x0 < p.x1) {
    p.s_lo = seg_search_global(seg_vofs, 0, nseg, p.x0);
    p.n = seg_search_global(seg_vofs, p.s_lo, nseg, p.x1 - 1) - p.s_lo + 1;
  }
  plan[b] = p;
}

__device__ inline uint32_t seg_search_lds(const uint64_t* l_vofs, uint32_t n, uint64_t x) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (l_vofs[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

// kWeights: also write each live word's weight (the out direction's
// per-level weight arrays, same level and index as the word) to wout.
// Block b: the vertices [bv[b], bv[b + 1]) (its plan record), their virtual
// words in batches of kBatch rounds of 256 (each thread requests its words of
// a batch before any is tested), live words written in virtual order from the
// first vertex's row offset.  The block's segments are looked up in LDS (in global
// memory when more than kSegLds overlap it).  Writes stay inside the
// block's rows whatever the mirror holds; a block whose live words differ in
// number from its rows' degrees counts itself in *bad.
// The words of one block, live ones written in order.  kLds: the block's
// segment records are in LDS (l_vofs / l_addr, n entries), else they are
// searched in global memory; the two forms are separate code, so neither
// pays for the other's address space.
template <bool kDeletions, bool kWeights, bool kLds>
__device__ inline void compact_block(const uint64_t* const* et, const uint64_t* seg_vofs,
                                     const uint64_t* seg_addr, const BlockPlan& pl,
                                     const uint64_t* l_vofs, const uint64_t* l_addr,
                                     uint32_t* wcnt, uint64_t E, uint32_t level, uint32_t* adj,
                                     const float* const* wt, float* wout, uint32_t* bad) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t x0 = pl.x0, x1 = pl.x1, s_lo = pl.s_lo, n = pl.n;
  const uint64_t out1 = min(pl.out1, E);
  uint64_t out = pl.out0;
  // batches of kBatch rounds: a block holds kWordsPerBlock words plus the
  // tail of its last vertex, so most blocks take one batch
  for (uint64_t xb = x0; xb < x1; xb += uint64_t(kBatch) * kBlock) {
    uint64_t wd[kBatch];
    float wv[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; j++) {
      const uint64_t x = xb + uint64_t(j) * kBlock + threadIdx.x;
      wd[j] = 0;
      wv[j] = 0.f;
      if (x < x1) {
        uint64_t a, v;
        if (kLds) {
          const uint32_t sg = seg_search_lds(l_vofs, uint32_t(n), x);
          a = l_addr[sg];
          v = l_vofs[sg];
        } else {
          const uint64_t sg = seg_search_global(seg_vofs, s_lo, s_lo + n, x);
          a = seg_addr[sg];
          v = seg_vofs[sg];
        }
        const uint32_t lvl = edge_level(a);
        const uint64_t idx = edge_index(a) + (x - v);
        wd[j] = ld_global(et[lvl] + idx);
        if (kWeights) wv[j] = wt[lvl][idx];
      }
    }
    // the batch's words in rounds of 256 consecutive words: every wave counts
    // its live words per round, then one barrier, and each word's output
    // position is the live words before it in virtual order (rounds before,
    // waves before in its round, lanes before in its wave)
    bool alive[kBatch];
    uint32_t rank[kBatch];
#pragma unroll
    for (int j = 0; j < kBatch; j++) {
      const uint64_t x = xb + uint64_t(j) * kBlock + threadIdx.x;
      alive[j] = x < x1 && (!kDeletions || word_maxlev(wd[j]) > level);
      const uint64_t mask = __ballot(alive[j]);
      rank[j] = __popcll(mask & ((uint64_t(1) << lane) - 1));
      if (lane == 0) wcnt[j * kWaves + wave] = __popcll(mask);
    }
    __syncthreads();
    uint64_t before = out;
#pragma unroll
    for (int j = 0; j < kBatch; j++) {
      uint32_t mine = 0, round = 0;
      for (int i = 0; i < kWaves; i++) {
        const uint32_t c = wcnt[j * kWaves + i];
        if (i < wave) mine += c;
        round += c;
      }
      const uint64_t pos = before + mine + rank[j];
      if (alive[j] && pos < out1) {
        adj[pos] = uint32_t(wd[j] & kIndexMask);
        if (kWeights) wout[pos] = wv[j];
      }
      before += round;
    }
    out = before;
    __syncthreads();  // wcnt is rewritten by the next batch
  }
  if (threadIdx.x == 0 && out != pl.out1) atomicAdd(bad, 1u);
}

template <bool kDeletions, bool kWeights>
__global__ __launch_bounds__(kBlock) void k_compact(const uint64_t* const* et,
                                                    const uint64_t* seg_vofs,
                                                    const uint64_t* seg_addr,
                                                    const BlockPlan* plan, uint64_t E,
                                                    uint32_t level, uint32_t* adj,
                                                    const float* const* wt, float* wout,
                                                    uint32_t* bad) {
  __shared__ uint64_t l_vofs[kSegLds + 1];
  __shared__ uint64_t l_addr[kSegLds];
  __shared__ uint32_t wcnt[kBatch * kWaves];
  const BlockPlan pl = plan[blockIdx.x];
  if (pl.x0 == pl.x1) return;  // uniform: a window inside a long vertex
  if (pl.n <= kSegLds) {  // uniform
    for (uint32_t i = threadIdx.x; i <= uint32_t(pl.n); i += blockDim.x) {
      l_vofs[i] = seg_vofs[pl.s_lo + i];
      if (i < pl.n) l_addr[i] = seg_addr[pl.s_lo + i];
    }
    __syncthreads();
    compact_block<kDeletions, kWeights, true>(et, seg_vofs, seg_addr, pl, l_vofs, l_addr, wcnt,
                                              E, level, adj, wt, wout, bad);
  } else {
    compact_block<kDeletions, kWeights, false>(et, seg_vofs, seg_addr, pl, l_vofs, l_addr, wcnt,
                                               E, level, adj, wt, wout, bad);
  }
}

}  // namespace

const uint32_t* ensure_degrees(ll_gpu_graph* g, uint32_t dir, int level) {
  level = resolve_level(g, level);
  if (g->deg_level[dir] == level) return g->deg[dir].as<uint32_t>();
  hipStream_t s = g->ctx->stream;
  g->deg[dir].alloc(g->V * sizeof(uint32_t), s);
  if (g->slcsr) {  // LL_SLCSR degree(n) = begin[n+1] - begin[n] (ll_slcsr.h:275-278)
    const Csr& c = ensure_csr(g, dir, level);
    k_ofs_degrees<<<grid_for(g->V, 256), 256, 0, s>>>(c.ofs.as<uint64_t>(),
                                                      g->deg[dir].as<uint32_t>(), g->V);
    LLG_HIP(hipGetLastError());
    g->deg_level[dir] = level;
    return g->deg[dir].as<uint32_t>();
  }
  {
    TimedLaunch t(g->ctx, "k_degrees");
    const VtPages vt{g->dir[dir].vt_ptab[level].as<const VtEntry* const>()};
    if (g->min_level == 0) {
      k_degrees<<<grid_for(g->V, 256), 256, 0, s>>>(vt, g->deg[dir].as<uint32_t>(), g->V);
    } else {
      DevBuf bad;
      bad.alloc(4, s);
      k_zero_u32<<<1, 1, 0, s>>>(bad.as<uint32_t>());
      k_window_degrees<<<grid_for(g->V, 256), 256, 0, s>>>(
          vt, g->dir[dir].et_table.as<const uint64_t* const>(),
          g->dir[dir].et_words_dev.as<const uint64_t>(), uint32_t(level), g->min_level,
          (g->flags & LL_GPU_HAS_DELETIONS) != 0, g->V, g->deg[dir].as<uint32_t>(),
          bad.as<uint32_t>());
      uint32_t hbad = 0;
      LLG_HIP(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, s));
      LLG_HIP(hipStreamSynchronize(s));
      LLG_CHECK(hbad == 0, LL_GPU_ESTATE, "corrupt mirror: a segment chain leaves its bounds");
    }
  }
  LLG_HIP(hipGetLastError());
  g->deg_level[dir] = level;
  return g->deg[dir].as<uint32_t>();
}

const Csr& ensure_csr(ll_gpu_graph* g, uint32_t dir, int level) {
  level = resolve_level(g, level);
  if (g->slcsr) {
    LLG_CHECK(g->sl[dir].valid, LL_GPU_ESTATE,
              "this direction of the single-level CSR is not pushed");
    return g->sl[dir];
  }
  Csr& c = g->csr[dir];
  if (c.valid && c.level == level) return c;
  c = Csr();
  g->rb[dir].clear();  // row blocks describe the previous level's rows
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V;
  const VtPages vt{g->dir[dir].vt_ptab[level].as<const VtEntry* const>()};
  const uint64_t* const* et = g->dir[dir].et_table.as<const uint64_t* const>();
  llgpu::EventPair ev;
  hipEvent_t e0 = ev.a, e1 = ev.b;
  LLG_HIP(hipEventRecord(e0, s));

  // 1-3: segments
  const uint64_t* et_words = g->dir[dir].et_words_dev.as<const uint64_t>();
  DevBuf nseg, nwords, seg_base, word_base, bad;
  nseg.alloc(V * 4, s);
  nwords.alloc(V * 8, s);
  bad.alloc(4, s);
  k_zero_u32<<<1, 1, 0, s>>>(bad.as<uint32_t>());
  {
    TimedLaunch t(ctx, "k_seg_count");
    k_seg_count<<<grid_for(V, 256), 256, 0, s>>>(vt, et, et_words, uint32_t(level),
                                                 g->min_level, V,
                                                 nseg.as<uint32_t>(), nwords.as<uint64_t>(),
                                                 bad.as<uint32_t>());
  }
  LLG_HIP(hipGetLastError());
  seg_base.alloc((V + 1) * 8, s);
  word_base.alloc((V + 1) * 8, s);
  scan_u32_to_u64(nseg.as<uint32_t>(), seg_base.as<uint64_t>(), V, s);
  scan_u64(nwords.as<uint64_t>(), word_base.as<uint64_t>(), V, s);
  uint64_t tot[2];
  uint32_t hbad = 0;
  LLG_HIP(hipMemcpyAsync(&tot[0], seg_base.as<uint64_t>() + V, 8, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipMemcpyAsync(&tot[1], word_base.as<uint64_t>() + V, 8, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
  LLG_CHECK(hbad == 0, LL_GPU_ESTATE,
            "corrupt mirror: a continuation chain leaves its edge table or does not descend");
  const uint64_t S = tot[0], W = tot[1];
  nseg.release();
  nwords.release();
  DevBuf seg_addr, seg_vofs;
  seg_addr.alloc(std::max<uint64_t>(S, 1) * 8, s);
  seg_vofs.alloc((S + 1) * 8, s);
  {
    TimedLaunch t(ctx, "k_seg_fill");
    k_seg_fill<<<grid_for(V, 256), 256, 0, s>>>(vt, et, et_words, uint32_t(level),
                                                g->min_level, V,
                                                seg_base.as<uint64_t>(), word_base.as<uint64_t>(),
                                                seg_addr.as<uint64_t>(), seg_vofs.as<uint64_t>(),
                                                bad.as<uint32_t>());
  }
  k_seg_total<<<1, 1, 0, s>>>(seg_vofs.as<uint64_t>(), S, W);
  LLG_HIP(hipGetLastError());
  seg_base.release();

  // Row offsets from the precomputed degrees.
  const uint32_t* deg = ensure_degrees(g, dir, level);
  c.ofs.alloc((V + 1) * 8, s);
  scan_u32_to_u64(deg, c.ofs.as<uint64_t>(), V, s);
  uint64_t E = 0;
  LLG_HIP(hipMemcpyAsync(&E, c.ofs.as<uint64_t>() + V, 8, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));

  // 4-5: compaction of live words, blocks aligned to vertices
  const uint64_t nblk = (W + kWordsPerBlock - 1) / kWordsPerBlock;
  const bool dels = (g->flags & LL_GPU_HAS_DELETIONS) != 0;
  c.adj.alloc(std::max<uint64_t>(E, 1) * 4, s);
  // out-edge weights too when every level up to this one carries them
  bool weighted = dir == LL_GPU_DIR_OUT && int(g->dir[dir].wt.size()) > level;
  for (int k = 0; weighted && k <= level; k++) weighted = g->dir[dir].wt[k].p != nullptr;
  if (weighted) c.w.alloc(std::max<uint64_t>(E, 1) * 4, s);
  if (nblk > 0) {
    LLG_CHECK(dels || W == E, LL_GPU_ESTATE,
              "snapshot has " + std::to_string(W) + " stored words but its degrees sum to " +
                  std::to_string(E) + " and the graph has no deletions");
    DevBuf bv, plan, nbad;
    bv.alloc((nblk + 1) * 8, s);
    plan.alloc(nblk * sizeof(BlockPlan), s);
    nbad.alloc(4, s);
    k_zero_u32<<<1, 1, 0, s>>>(nbad.as<uint32_t>());
    k_block_verts<<<grid_for(nblk + 1, 256), 256, 0, s>>>(word_base.as<uint64_t>(), V, nblk,
                                                         bv.as<uint64_t>());
    k_block_plan<<<grid_for(nblk, 256), 256, 0, s>>>(seg_vofs.as<uint64_t>(), S,
                                                     word_base.as<uint64_t>(), bv.as<uint64_t>(),
                                                     c.ofs.as<uint64_t>(), nblk,
                                                     plan.as<BlockPlan>());
    LLG_HIP(hipGetLastError());
    {
      TimedLaunch t(ctx, "k_compact");
      const float* const* wt = weighted ? g->dir[dir].wt_table.as<const float* const>() : nullptr;
      float* wout = weighted ? c.w.as<float>() : nullptr;
      const unsigned nb = unsigned(nblk);
      const uint64_t* sv = seg_vofs.as<uint64_t>();
      const uint64_t* sa = seg_addr.as<uint64_t>();
      const BlockPlan* bp = plan.as<BlockPlan>();
      uint32_t* lv = nbad.as<uint32_t>();
      uint32_t* adj = c.adj.as<uint32_t>();
      const uint32_t L = uint32_t(level);
      if (dels && weighted)
        k_compact<true, true><<<nb, kBlock, 0, s>>>(et, sv, sa, bp, E, L, adj, wt, wout, lv);
      else if (dels)
        k_compact<true, false><<<nb, kBlock, 0, s>>>(et, sv, sa, bp, E, L, adj, wt, wout, lv);
      else if (weighted)
        k_compact<false, true><<<nb, kBlock, 0, s>>>(et, sv, sa, bp, E, L, adj, wt, wout, lv);
      else
        k_compact<false, false><<<nb, kBlock, 0, s>>>(et, sv, sa, bp, E, L, adj, wt, wout, lv);
    }
    LLG_HIP(hipGetLastError());
    uint32_t hbad = 0;
    LLG_HIP(hipMemcpyAsync(&hbad, nbad.p, 4, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    LLG_CHECK(hbad == 0, LL_GPU_ESTATE,
              "snapshot: " + std::to_string(hbad) + " blocks of vertices hold more or fewer "
              "live words than their degrees count (missing deletion patches?)");
  } else {
    LLG_CHECK(E == 0, LL_GPU_ESTATE, "degrees are non-zero but no segments are reachable");
  }
  word_base.release();
  LLG_HIP(hipEventRecord(e1, s));
  LLG_HIP(hipEventSynchronize(e1));
  float ms = 0;
  LLG_HIP(hipEventElapsedTime(&ms, e0, e1));
  c.valid = true;
  c.level = level;
  c.V = V;
  c.E = E;
  c.stored = W;
  c.segments = S;
  c.ms = ms;
  return c;
}

}  // namespace llgpu

extern "C" {

int ll_gpu_snapshot_stats(ll_gpu_graph* g, int level, ll_gpu_snapshot_info* info) {
  if (!g || !info) return LL_GPU_EINVAL;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    const llgpu::Csr& c = llgpu::ensure_csr(g, LL_GPU_DIR_OUT, level);
    info->num_nodes = c.V;
    info->live_edges = c.E;
    info->stored_words = c.stored;
    info->segments = c.segments;
    // vertices with >= 1 segment contribute (segments - 1) continuations
    uint64_t with_edges = 0;
    {
      const uint32_t* deg = llgpu::ensure_degrees(g, LL_GPU_DIR_OUT, level);
      std::vector<uint32_t> h(g->V);
      LLG_HIP(hipMemcpyAsync(h.data(), deg, g->V * 4, hipMemcpyDeviceToHost, g->ctx->stream));
      LLG_HIP(hipStreamSynchronize(g->ctx->stream));
      for (uint32_t d : h) with_edges += d > 0;
    }
    info->continuations = c.segments >= with_edges ? c.segments - with_edges : 0;
    info->materialize_ms = c.ms;
    return LL_GPU_OK;
  } catch (const llgpu::Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

int ll_gpu_snapshot_adjacency(ll_gpu_graph* g, int level, uint32_t dir, uint64_t* offsets,
                              uint32_t* neighbours) {
  if (!g || dir > 1) return LL_GPU_EINVAL;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    const llgpu::Csr& c = llgpu::ensure_csr(g, dir, level);
    hipStream_t s = g->ctx->stream;
    if (offsets)
      LLG_HIP(hipMemcpyAsync(offsets, c.ofs.p, (c.V + 1) * 8, hipMemcpyDeviceToHost, s));
    if (neighbours && c.E)
      LLG_HIP(hipMemcpyAsync(neighbours, c.adj.p, c.E * 4, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    return LL_GPU_OK;
  } catch (const llgpu::Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

int ll_gpu_snapshot_degrees(ll_gpu_graph* g, int level, uint32_t dir, uint32_t* degrees) {
  if (!g || dir > 1 || !degrees) return LL_GPU_EINVAL;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    const uint32_t* d = llgpu::ensure_degrees(g, dir, level);
    LLG_HIP(hipMemcpyAsync(degrees, d, g->V * 4, hipMemcpyDeviceToHost, g->ctx->stream));
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
