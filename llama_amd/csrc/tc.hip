// Triangle counting on the snapshot's simple undirected graph (tasks tc_u /
// tc_od, benchmark/benchmarks/triangle_counting.h:343-470, 682-831).
//
// The reference needs a single sorted, deduplicated level; here the snapshot's
// live out-edges are symmetrised and oriented on the device:
//   1. every live edge {u,v}, u != v, becomes the arc a -> b with
//      (deg(a), a) < (deg(b), b), deg = out-degree + in-degree (any strict
//      total order counts each triangle once; degree order bounds the arc
//      lists by O(sqrt(E))), encoded as the 64-bit key a<<32|b;
//   2. keys are radix-sorted and deduplicated: the oriented DAG as a CSR with
//      ascending neighbour lists;
//   3. for each arc (a,b) count |N+(a) & N+(b)|: one wave per row a (a whole
//      workgroup for rows of 257..4096 arcs) stages N+(a), a bucket index of
//      it (bucket = ID >> shift, about one entry per bucket), and for the arcs
//      with d+(b) > 0 the prefix sums of d+(b) and the list starts ofs[b] in
//      LDS.  The concatenated lists N+(b) are then streamed 256 entries per
//      wave step (coalesced within each list): every entry finds its arc from
//      a ballot of arc-end flags (no search), the four 64-entry loads of a
//      step are in flight together, and each loaded ID is searched in its
//      bucket's short run of N+(a).  Rows longer than 4096 arcs search N+(a)
//      in global memory.
// Every triangle a<b<c (in that order) is counted once, at arc (a,b).
//
// Task tc_i (triangle_counting.h:154-336) visits the pair (u, v), u < v by
// ID, once per directed edge between them (out-list, then in-list) and
// counts the common neighbours w > v: every triangle weighs the number of
// directed edges joining its two smallest IDs.  The DAG keeps that number per
// arc (the run length of its key among the sorted, not yet deduplicated,
// keys), and the weighted count adds it for the arc joining the two smallest
// of {a, b, w}.
#include "ll_gpu_internal.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr int kRowBlock = 1024;        // workgroup of the kWaveRow..kBlockRow rows
constexpr uint32_t kWaveRow = 256;     // row length handled by one wave (LDS)
constexpr uint32_t kBlockRow = 4096;   // row length handled by one workgroup (LDS)
constexpr uint32_t kU = 4;             // 64-element chunks a wave resolves together
constexpr uint32_t kPosShift = 48;     // staged list start | position of b in N+(a) << 48
constexpr uint64_t kOfsMask = (1ull << kPosShift) - 1;

// One wave per vertex row (grid-stride), lanes over the row's edges: a hub's
// row is spread over 64 lanes instead of serialising one thread.  Arcs are
// written at the row's own edge positions.
__global__ __launch_bounds__(256) void k_orient(const uint64_t* ofs, const uint32_t* adj,
                                                const uint32_t* dout, const uint32_t* din,
                                                uint64_t V, uint64_t* keys) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nw = uint64_t(gridDim.x) * (blockDim.x / 64);
  for (uint64_t u = uint64_t(blockIdx.x) * (blockDim.x / 64) + (threadIdx.x >> 6); u < V;
       u += nw) {
    const uint64_t du = uint64_t(dout[u]) + din[u];
    for (uint64_t e = ofs[u] + lane, end = ofs[u + 1]; e < end; e += 64) {
      const uint64_t v = adj[e];
      uint64_t key = ~0ull;  // self-loops sort last and are cut
      if (v != u) {
        const uint64_t dv = uint64_t(dout[v]) + din[v];
        const bool uv = du < dv || (du == dv && u < v);
        key = uv ? (u << 32) | v : (v << 32) | u;
      }
      keys[e] = key;
    }
  }
}

// Row offsets of the sorted arc keys: ofs[x] = first arc with tail >= x.
__global__ void k_dag_offsets(const uint64_t* keys, uint64_t m, uint64_t V, uint64_t* ofs,
                              uint32_t* nbr) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i > m) return;
  const uint64_t a = i < m ? keys[i] >> 32 : V;
  const uint64_t p = i == 0 ? 0 : (keys[i - 1] >> 32) + 1;
  for (uint64_t x = p; x <= a && x <= V; x++) ofs[x] = i;
  if (i < m) nbr[i] = uint32_t(keys[i]);
}

// Run starts of the sorted arc keys (before deduplication).
__global__ void k_run_starts(const uint64_t* sorted, uint64_t n, uint32_t* flag) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) flag[i] = (i == 0 || sorted[i] != sorted[i - 1]) ? 1u : 0u;
}

// Arc multiplicity = length of its key's run: the run of unique key r starts
// at sorted position start[r]; start[r + 1] (or n) ends it.
__global__ void k_run_lengths(const uint32_t* flag, const uint64_t* rank, uint64_t n,
                              uint64_t m, uint8_t* mult) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const uint64_t r = rank[i];
  if (r >= m) return;  // the self-loop sentinel's run
  uint64_t j = i + 1;
  while (j < n && !flag[j]) j++;
  mult[r] = uint8_t(min<uint64_t>(j - i, 255));
}

// Make a wave's LDS writes visible to its other lanes before they read them.
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Position of x in an ascending list of n entries, or n when absent.
__device__ inline uint32_t find_in(const uint32_t* list, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (list[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < n && list[lo] == x ? lo : n;
}

// Count common out-neighbours for all arcs of row a, with `lanes` threads of a
// workgroup; `na` = d+(a) > kBlockRow, so N+(a) is searched in global memory
// and `pre` (inclusive prefix sums of d+(b) over the row's b's) lies in a
// global buffer at the row's own arc range.  kWeighted: each triangle
// {a, b, w} adds mult of the arc joining its two smallest IDs.
template <bool kWeighted>
__device__ inline unsigned long long count_row_global(const uint64_t* ofs, const uint32_t* nbr,
                                                      const uint8_t* mult, uint64_t a,
                                                      uint32_t na, const uint32_t* pre,
                                                      uint32_t tid, uint32_t lanes) {
  const uint64_t a0 = ofs[a];
  const uint32_t work = na ? pre[na - 1] : 0;
  unsigned long long c = 0;
  for (uint32_t x = tid; x < work; x += lanes) {
    // arc index i with pre[i-1] <= x < pre[i]
    uint32_t lo = 0, hi = na;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (pre[mid] <= x) lo = mid + 1; else hi = mid;
    }
    const uint32_t b = nbr[a0 + lo];
    const uint32_t j = x - (lo ? pre[lo - 1] : 0);
    const uint64_t bw = ofs[b] + j;
    const uint32_t w = nbr[bw];
    const uint32_t p = find_in(nbr + a0, na, w);
    if (p == na) continue;
    if (!kWeighted) {
      c++;
    } else {
      // the arc between the two smallest IDs: (a,b) unless w is not the largest
      const uint64_t amax = uint64_t(a) > b && uint64_t(a) > w;
      const uint64_t arc = amax ? bw : (b > w ? a0 + p : a0 + lo);
      c += mult[arc];
    }
  }
  return c;
}

// Bucket index of a staged row: N+(a) is sorted, so the IDs of bucket
// b = id >> shift (nb = 2^ceil(log2 na) buckets over the ID range) form the
// run la[bs[b] .. bs[b+1]).  With IDs spread over [0, V) a bucket holds about
// one entry: a lookup reads two bucket starts and searches the short run,
// and an absent ID usually finds an empty run at once.  Clustered IDs only
// lengthen the runs (the search stays exact).
__device__ inline uint32_t bucket_shift(uint64_t V, uint32_t na) {
  const uint32_t vbits = V > 1 ? uint32_t(64 - __clzll((unsigned long long)(V - 1))) : 1u;
  const uint32_t nbits = na > 1 ? uint32_t(32 - __clz(int(na - 1))) : 0u;
  return vbits > nbits ? vbits - nbits : 0u;
}

// bs[b] = first i with (la[i] >> shift) >= b for b in [0, 2^nbits]; `lanes`
// threads from `tid` (la[0..na) staged and visible).
__device__ inline void build_buckets(const uint32_t* la, uint32_t na, uint32_t shift,
                                     uint32_t nb, uint16_t* bs, uint32_t tid, uint32_t lanes) {
  for (uint32_t i = tid; i < na; i += lanes) {
    const uint32_t bi = la[i] >> shift;
    for (uint32_t b = i ? (la[i - 1] >> shift) + 1u : 0u; b <= bi; b++) bs[b] = uint16_t(i);
  }
  for (uint32_t b = (la[na - 1] >> shift) + 1u + tid; b <= nb; b += lanes) bs[b] = uint16_t(na);
}

__device__ inline uint64_t lanes_below(uint32_t lane) {
  return lane ? (~0ull >> (64u - lane)) : 0ull;
}

// The staged row a (LDS): la[0..na) = N+(a) ascending, bs = its bucket index,
// lp[0..cnt) = inclusive prefix sums of d+(b) over the arcs with d+(b) > 0
// ("compacted arcs", in any order), lofs[k] = ofs[b] of compacted arc k, with
// the position of its b in la in the bits from kPosShift up (weighted count
// only; list starts stay below 2^kPosShift).  Element x in
// [0, lp[cnt-1]) is entry x - lp[k-1] of N+(b) for the compacted arc k with
// lp[k-1] <= x < lp[k].
//
// One wave counts the elements x in [base0, base0 + 64kU) below `end`, lane
// `lane` taking x = base0 + 64u + lane (u < kU).  *lo = the first compacted
// arc with lp > base0 on entry, the first with lp > base0 + 64kU on return.
// Every compacted arc holds at least one element, so at most 64kU arcs end
// inside the range, and they are arcs lo .. lo + 64kU - 1: the lanes load
// those, flag the range position after each one's last element, and a ballot
// of the flags per 64 positions plus popcounts give every element its arc
// without a search.  The kU neighbour loads are then issued together and
// each loaded ID is looked up in its bucket's run of la (the kU searches
// interleaved).
template <bool kWeighted>
__device__ inline unsigned long long count_group(const uint32_t* nbr, const uint8_t* mult,
                                                 uint64_t a, uint64_t a0, const uint32_t* la,
                                                 uint32_t na, const uint32_t* lp,
                                                 const uint64_t* lofs, uint32_t cnt,
                                                 const uint16_t* bs, uint32_t shift,
                                                 uint8_t* fl, uint32_t base0,
                                                 uint32_t end, uint32_t* lo, uint32_t lane) {
  const uint32_t lo0 = *lo;
#pragma unroll
  for (uint32_t u = 0; u < kU; u++) fl[u * 64 + lane] = 0;
  wave_sync();
#pragma unroll
  for (uint32_t k = 0; k < kU; k++) {
    const uint32_t t = lo0 + k * 64 + lane;
    if (t < cnt) {
      const uint32_t q = lp[t] - base0 - 1u;  // lp[t] > base0 for t >= lo0
      if (q < 64u * kU) fl[q] = 1;
    }
  }
  wave_sync();
  uint64_t M[kU];
#pragma unroll
  for (uint32_t u = 0; u < kU; u++) M[u] = __ballot(fl[u * 64 + lane] != 0);
  wave_sync();  // every flag read before the next group clears them
  uint64_t bw[kU];
  uint32_t arc[kU];
  bool act[kU];
  uint32_t before = 0;
#pragma unroll
  for (uint32_t u = 0; u < kU; u++) {
    const uint32_t r = lo0 + before + uint32_t(__popcll(M[u] & lanes_below(lane)));
    const uint32_t x = base0 + 64u * u + lane;
    act[u] = x < end;
    arc[u] = r;
    bw[u] = 0;
    if (act[u]) bw[u] = (lofs[r] & kOfsMask) + (x - (r ? lp[r - 1] : 0u));
    before += uint32_t(__popcll(M[u]));
  }
  *lo = lo0 + before;
  uint32_t w[kU];
#pragma unroll
  for (uint32_t u = 0; u < kU; u++) w[u] = act[u] ? nbr[bw[u]] : 0u;
  // lower bound of w in its bucket's run [bl, be)
  uint32_t bl[kU], bh[kU], be[kU];
#pragma unroll
  for (uint32_t u = 0; u < kU; u++) {
    const uint32_t b = w[u] >> shift;
    bl[u] = act[u] ? uint32_t(bs[b]) : 0u;
    be[u] = act[u] ? uint32_t(bs[b + 1]) : 0u;
    bh[u] = be[u];
  }
  for (;;) {
    bool open = false;
#pragma unroll
    for (uint32_t u = 0; u < kU; u++) open |= bl[u] < bh[u];
    if (!__ballot(open)) break;
#pragma unroll
    for (uint32_t u = 0; u < kU; u++)
      if (bl[u] < bh[u]) {
        const uint32_t mid = (bl[u] + bh[u]) >> 1;
        if (la[mid] < w[u]) bl[u] = mid + 1; else bh[u] = mid;
      }
  }
  unsigned long long c = 0;
#pragma unroll
  for (uint32_t u = 0; u < kU; u++) {
    const uint32_t pw = bl[u];
    if (pw >= be[u] || la[pw] != w[u]) continue;
    if (!kWeighted) {
      c++;
    } else {
      const uint32_t ik = uint32_t(lofs[arc[u]] >> kPosShift);
      const uint32_t b = la[ik];
      const bool amax = a > b && a > w[u];
      const uint64_t e = amax ? bw[u] : (b > w[u] ? a0 + pw : a0 + ik);
      c += mult[e];
    }
  }
  return c;
}

// Rows with 2 <= d+(a) <= kWaveRow, one wave each (a single arc closes no
// triangle).
template <bool kWeighted>
__global__ __launch_bounds__(kBlock) void k_tc_wave(const uint64_t* ofs, const uint32_t* nbr,
                                                    const uint8_t* mult, uint64_t V,
                                                    unsigned long long* partial) {
  constexpr uint32_t kWaves = kBlock / 64;
  __shared__ uint32_t l_a[kWaves][kWaveRow];
  __shared__ uint32_t l_p[kWaves][kWaveRow];
  __shared__ uint64_t l_o[kWaves][kWaveRow];
  __shared__ uint16_t l_bs[kWaves][kWaveRow + 1];
  __shared__ uint8_t l_fl[kWaves][kU * 64];
  __shared__ unsigned long long red[kWaves];
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  uint32_t* la = l_a[wave];
  uint32_t* lp = l_p[wave];
  uint64_t* lofs = l_o[wave];
  uint16_t* bs = l_bs[wave];
  const uint64_t nwaves = uint64_t(gridDim.x) * kWaves;
  unsigned long long c = 0;
  for (uint64_t a = uint64_t(blockIdx.x) * kWaves + wave; a < V; a += nwaves) {
    const uint64_t a0 = ofs[a];
    const uint32_t na = uint32_t(ofs[a + 1] - a0);
    if (na < 2 || na > kWaveRow) continue;
    uint32_t cnt = 0;
    for (uint32_t i0 = 0; i0 < na; i0 += 64) {
      const uint32_t i = i0 + lane;
      uint32_t d = 0;
      uint64_t ob = 0;
      if (i < na) {
        const uint32_t b = nbr[a0 + i];
        la[i] = b;
        ob = ofs[b];
        d = uint32_t(ofs[b + 1] - ob);
      }
      const uint64_t m = __ballot(d != 0);
      if (d) {
        const uint32_t pos = cnt + uint32_t(__popcll(m & lanes_below(lane)));
        lp[pos] = d;
        lofs[pos] = kWeighted ? ob | (uint64_t(i) << kPosShift) : ob;
      }
      cnt += uint32_t(__popcll(m));
    }
    wave_sync();
    const uint32_t shift = bucket_shift(V, na);
    build_buckets(la, na, shift, 1u << (32 - __clz(int(na - 1))), bs, lane, 64);
    wave_sync();
    // inclusive scan of lp[0..cnt) in chunks of 64
    uint32_t carry = 0;
    for (uint32_t base = 0; base < cnt; base += 64) {
      const uint32_t i = base + lane;
      uint32_t v = i < cnt ? lp[i] : 0;
      for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= uint32_t(o)) v += t;
      }
      if (i < cnt) lp[i] = v + carry;
      carry += __shfl(v, 63, 64);
    }
    wave_sync();
    const uint32_t total = carry;
    uint32_t lo = 0;
    for (uint32_t base0 = 0; base0 < total; base0 += 64u * kU)
      c += count_group<kWeighted>(nbr, mult, a, a0, la, na, lp, lofs, cnt, bs, shift,
                                  l_fl[wave], base0, total, &lo, lane);
    wave_sync();
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if (lane == 0) red[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (uint32_t i = 0; i < kWaves; i++) t += red[i];
    partial[blockIdx.x] = t;
  }
}

// Rows with kWaveRow < d+(a) <= kBlockRow, one workgroup each: the row is
// staged once and its groups of kU chunks are dealt to the waves in turn.
template <bool kWeighted>
__global__ __launch_bounds__(kRowBlock) void k_tc_block(const uint64_t* ofs, const uint32_t* nbr,
                                                        const uint8_t* mult, const uint32_t* rows,
                                                        uint64_t nrows, uint64_t V,
                                                        unsigned long long* partial) {
  constexpr uint32_t kWaves = kRowBlock / 64;
  __shared__ uint32_t la[kBlockRow];
  __shared__ uint32_t lp[kBlockRow];
  __shared__ uint64_t lofs[kBlockRow];
  __shared__ uint16_t bs[kBlockRow + 1];
  __shared__ uint8_t l_fl[kWaves][kU * 64];
  __shared__ uint32_t wtot[kWaves];
  __shared__ uint32_t ncomp;
  __shared__ unsigned long long red[kWaves];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  unsigned long long c = 0;
  for (uint64_t r = blockIdx.x; r < nrows; r += gridDim.x) {
    const uint64_t a = rows[r];
    const uint64_t a0 = ofs[a];
    const uint32_t na = uint32_t(ofs[a + 1] - a0);  // in (kWaveRow, kBlockRow]
    if (tid == 0) ncomp = 0;
    __syncthreads();
    for (uint32_t i0 = 0; i0 < na; i0 += kRowBlock) {
      const uint32_t i = i0 + tid;
      uint32_t d = 0;
      uint64_t ob = 0;
      if (i < na) {
        const uint32_t b = nbr[a0 + i];
        la[i] = b;
        ob = ofs[b];
        d = uint32_t(ofs[b + 1] - ob);
      }
      const uint64_t m = __ballot(d != 0);
      uint32_t wbase = 0;
      if (lane == 0 && m) wbase = atomicAdd(&ncomp, uint32_t(__popcll(m)));
      wbase = uint32_t(__shfl(int(wbase), 0, 64));
      if (d) {
        const uint32_t pos = wbase + uint32_t(__popcll(m & lanes_below(lane)));
        lp[pos] = d;
        lofs[pos] = kWeighted ? ob | (uint64_t(i) << kPosShift) : ob;
      }
    }
    __syncthreads();
    const uint32_t cnt = ncomp;
    const uint32_t shift = bucket_shift(V, na);
    build_buckets(la, na, shift, 1u << (32 - __clz(int(na - 1))), bs, tid, kRowBlock);
    __syncthreads();
    // block-wide inclusive scan of lp[0..cnt) in chunks of kRowBlock
    uint32_t carry = 0;
    for (uint32_t base = 0; base < cnt; base += kRowBlock) {
      const uint32_t i = base + tid;
      uint32_t v = i < cnt ? lp[i] : 0;
      for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= uint32_t(o)) v += t;
      }
      if (lane == 63) wtot[wave] = v;
      __syncthreads();
      uint32_t off = carry;
      for (uint32_t k = 0; k < wave; k++) off += wtot[k];
      uint32_t chunk = 0;
      for (uint32_t k = 0; k < kWaves; k++) chunk += wtot[k];
      if (i < cnt) lp[i] = v + off;
      carry += chunk;
      __syncthreads();
    }
    const uint32_t total = carry;
    // wave w takes the contiguous elements [w * per, (w + 1) * per)
    const uint32_t per = ((total + kWaves - 1) / kWaves + 63u) & ~63u;
    const uint32_t wb = min(total, wave * per), we = min(total, wb + per);
    uint32_t lo = 0, hi = cnt;  // first compacted arc with lp > wb
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (lp[mid] <= wb) lo = mid + 1; else hi = mid;
    }
    for (uint32_t base0 = wb; base0 < we; base0 += 64u * kU)
      c += count_group<kWeighted>(nbr, mult, a, a0, la, na, lp, lofs, cnt, bs, shift,
                                  l_fl[wave], base0, we, &lo, lane);
    __syncthreads();  // the row's LDS is read before the next row overwrites it
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if (lane == 0) red[wave] = c;
  __syncthreads();
  if (tid == 0) {
    unsigned long long t = 0;
    for (uint32_t i = 0; i < kWaves; i++) t += red[i];
    partial[blockIdx.x] = t;
  }
}

// Rows with d+(a) > kBlockRow, one workgroup each; the prefix sums of d+(b)
// live in gpre at the row's own arc range and N+(a) is searched in place.
template <bool kWeighted>
__global__ __launch_bounds__(kBlock) void k_tc_global(const uint64_t* ofs, const uint32_t* nbr,
                                                      const uint8_t* mult, const uint32_t* rows,
                                                      uint64_t nrows, uint32_t* gpre,
                                                      unsigned long long* partial) {
  __shared__ uint32_t wtot[kBlock / 64];
  __shared__ unsigned long long red[kBlock / 64];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long c = 0;
  for (uint64_t r = blockIdx.x; r < nrows; r += gridDim.x) {
    const uint64_t a = rows[r];
    const uint64_t a0 = ofs[a];
    const uint32_t na = uint32_t(ofs[a + 1] - a0);
    uint32_t* pre = gpre + a0;
    for (uint32_t i = threadIdx.x; i < na; i += kBlock) {
      const uint32_t b = nbr[a0 + i];
      pre[i] = uint32_t(ofs[b + 1] - ofs[b]);
    }
    __syncthreads();
    // block-wide inclusive scan over na entries in chunks of kBlock
    uint32_t carry = 0;
    for (uint32_t base = 0; base < na; base += kBlock) {
      const uint32_t i = base + threadIdx.x;
      uint32_t v = i < na ? pre[i] : 0;
      for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= uint32_t(o)) v += t;
      }
      if (lane == 63) wtot[wave] = v;
      __syncthreads();
      uint32_t off = carry;
      for (uint32_t k = 0; k < wave; k++) off += wtot[k];
      uint32_t chunk = 0;
      for (uint32_t k = 0; k < kBlock / 64; k++) chunk += wtot[k];
      if (i < na) pre[i] = v + off;
      carry += chunk;
      __syncthreads();
    }
    c += count_row_global<kWeighted>(ofs, nbr, mult, a, na, pre, threadIdx.x, kBlock);
    __syncthreads();
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_down(c, o, 64);
  if (lane == 0) red[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlock / 64; i++) t += red[i];
    partial[blockIdx.x] = t;
  }
}

// Rows by length: kWaveRow < d+ <= kBlockRow from the front of `rows`
// (count n[0]), d+ > kBlockRow from the back (count n[2]).
__global__ void k_classify_rows(const uint64_t* ofs, uint64_t V, uint32_t* rows,
                                unsigned long long* n) {
  const uint64_t a = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (a >= V) return;
  const uint64_t na = ofs[a + 1] - ofs[a];
  if (na > kBlockRow) rows[V - 1 - atomicAdd(n + 2, 1ull)] = uint32_t(a);
  else if (na > kWaveRow) rows[atomicAdd(n, 1ull)] = uint32_t(a);
}

// W = sum over arcs (a,b) of (d+(a) + d+(b)): intersection element reads,
// into *work; the largest per-row element count sum_b d+(b) into *max_row
// (the counting kernels keep element positions below 2^31).  One wave per row (grid-stride),
// one atomic of each kind per workgroup.
__global__ __launch_bounds__(256) void k_tc_work(const uint64_t* ofs, const uint32_t* nbr,
                                                 uint64_t V, unsigned long long* work,
                                                 unsigned long long* max_row) {
  __shared__ unsigned long long red[4], rmax[4];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t nw = uint64_t(gridDim.x) * 4;
  unsigned long long w = 0, mx = 0;
  for (uint64_t a = uint64_t(blockIdx.x) * 4 + wave; a < V; a += nw) {
    const uint64_t a0 = ofs[a], na = ofs[a + 1] - a0;
    unsigned long long el = 0;
    for (uint64_t i = lane; i < na; i += 64) {
      const uint32_t b = nbr[a0 + i];
      el += ofs[b + 1] - ofs[b];
    }
    w += na * ((na + 63 - lane) / 64) + el;  // na per arc handled by this lane
    for (int o = 32; o > 0; o >>= 1) el += __shfl_xor(el, o, 64);
    mx = el > mx ? el : mx;
  }
  for (int o = 32; o > 0; o >>= 1) w += __shfl_down(w, o, 64);
  if (lane == 0) {
    red[wave] = w;
    rmax[wave] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = red[0] + red[1] + red[2] + red[3];
    unsigned long long m = rmax[0];
    for (int i = 1; i < 4; i++) m = rmax[i] > m ? rmax[i] : m;
    if (t) atomicAdd(work, t);
    if (m) atomicMax(max_row, m);
  }
}

__global__ void k_sum_partials(const unsigned long long* p, uint64_t n, unsigned long long* out) {
  unsigned long long s = 0;
  for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) s += p[i];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  __shared__ unsigned long long red[16];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (unsigned i = 0; i < blockDim.x / 64; i++) t += red[i];
    *out = t;
  }
}

__global__ void k_zero_ull2(unsigned long long* p, int n) {
  for (int i = 0; i < n; i++) p[i] = 0;
}

}  // namespace

struct TcResult {
  uint64_t triangles = 0, arcs = 0, work = 0;
};

// Build (or reuse) the oriented DAG of the snapshot.
void ensure_dag(ll_gpu_graph* g, int level) {
  level = resolve_level(g, level);
  if (g->sym_level == level) return;
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V;
  const Csr& out = ensure_csr(g, LL_GPU_DIR_OUT, level);
  const uint32_t* dout = ensure_degrees(g, LL_GPU_DIR_OUT, level);
  const uint32_t* din = ensure_degrees(g, LL_GPU_DIR_IN, level);
  const uint64_t E = out.E;
  DevBuf keys, sorted;
  keys.alloc(std::max<uint64_t>(E, 1) * 8, s);
  {
    TimedLaunch t(ctx, "k_orient");
    const unsigned go = unsigned(std::min<uint64_t>(std::max<uint64_t>((V + 3) / 4, 1),
                                                    uint64_t(ctx->num_cus) * 32));
    k_orient<<<go, 256, 0, s>>>(out.ofs.as<uint64_t>(), out.adj.as<uint32_t>(), dout, din, V,
                                keys.as<uint64_t>());
  }
  LLG_HIP(hipGetLastError());
  sorted.alloc(std::max<uint64_t>(E, 1) * 8, s);
  sort_keys_u64(keys.as<uint64_t>(), sorted.as<uint64_t>(), E, 64, s);
  uint64_t m = unique_u64(sorted.as<uint64_t>(), keys.as<uint64_t>(), E, s);
  // drop the self-loop sentinel (all ones) if present: it sorts last
  if (m > 0) {
    uint64_t last = 0;
    LLG_HIP(hipMemcpyAsync(&last, keys.as<uint64_t>() + (m - 1), 8, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    if (last == ~0ull) m--;
  }
  // multiplicities (tc_i): runs of equal keys among the sorted keys
  {
    DevBuf flag, rank;
    flag.alloc(std::max<uint64_t>(E, 1) * 4, s);
    rank.alloc((E + 1) * 8, s);
    g->sym_mult.alloc(std::max<uint64_t>(m, 1), s);
    if (E) {
      k_run_starts<<<grid_for(E, 256), 256, 0, s>>>(sorted.as<uint64_t>(), E, flag.as<uint32_t>());
      scan_u32_to_u64(flag.as<uint32_t>(), rank.as<uint64_t>(), E, s);
      k_run_lengths<<<grid_for(E, 256), 256, 0, s>>>(flag.as<uint32_t>(), rank.as<uint64_t>(), E, m,
                                                     g->sym_mult.as<uint8_t>());
      LLG_HIP(hipGetLastError());
    }
  }
  sorted.release();
  g->sym_ofs.alloc((V + 1) * 8, s);
  g->sym_adj.alloc(std::max<uint64_t>(m, 1) * 4, s);
  k_dag_offsets<<<grid_for(m + 1, 256), 256, 0, s>>>(keys.as<uint64_t>(), m, V,
                                                     g->sym_ofs.as<uint64_t>(),
                                                     g->sym_adj.as<uint32_t>());
  LLG_HIP(hipGetLastError());
  g->sym_edges = m;
  // rows by length, W and the largest row element count (with the DAG)
  {
    const uint64_t* dofs = g->sym_ofs.as<uint64_t>();
    const uint32_t* dnbr = g->sym_adj.as<uint32_t>();
    DevBuf cnt;
    cnt.alloc(4 * 8, s);  // block rows, W, global rows, largest row element count
    k_zero_ull2<<<1, 1, 0, s>>>(cnt.as<unsigned long long>(), 4);
    g->sym_rows.alloc(std::max<uint64_t>(V, 1) * 4, s);
    k_classify_rows<<<grid_for(V, 256), 256, 0, s>>>(dofs, V, g->sym_rows.as<uint32_t>(),
                                                     cnt.as<unsigned long long>());
    k_tc_work<<<unsigned(std::min<uint64_t>(std::max<uint64_t>((V + 3) / 4, 1),
                                             uint64_t(ctx->num_cus) * 32)),
                256, 0, s>>>(dofs, dnbr, V, cnt.as<unsigned long long>() + 1,
                             cnt.as<unsigned long long>() + 3);
    LLG_HIP(hipGetLastError());
    unsigned long long h[4];
    LLG_HIP(hipMemcpyAsync(h, cnt.p, 32, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    g->sym_nblock = h[0];
    g->sym_work = h[1];
    g->sym_nglobal = h[2];
    g->sym_maxrow = h[3];
  }
  g->sym_level = level;
}

TcResult tc_device(ll_gpu_graph* g, int level, bool weighted) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V;
  ensure_dag(g, level);
  const uint64_t* ofs = g->sym_ofs.as<uint64_t>();
  const uint32_t* nbr = g->sym_adj.as<uint32_t>();
  const uint8_t* mult = g->sym_mult.as<uint8_t>();
  LLG_CHECK(g->sym_edges <= kOfsMask, LL_GPU_ERANGE, "oriented arcs beyond 2^48");
  TcResult r;
  r.arcs = g->sym_edges;
  DevBuf gpre, part_w, part_b, part_g, total;
  unsigned long long h[3];
  const uint64_t nblock = g->sym_nblock, nglobal = g->sym_nglobal;
  r.work = g->sym_work;
  // element positions of a row stay below 2^31, so every offset added to one
  // (a wave step, a per-wave range) still fits in 32 bits
  LLG_CHECK(g->sym_maxrow < (1ull << 31), LL_GPU_ERANGE,
            "a DAG row's neighbour lists hold 2^31 or more entries in total");
  LLG_CHECK(nblock + nglobal <= V, LL_GPU_ESTATE, "row classification overflow");
  const uint32_t* rows_b = g->sym_rows.as<uint32_t>();
  const uint32_t* rows_g = g->sym_rows.as<uint32_t>() + (V - nglobal);
  const unsigned gw = unsigned(std::min<uint64_t>(std::max<uint64_t>((V + 3) / 4, 1),
                                                  uint64_t(ctx->num_cus) * 32));
  part_w.alloc(gw * 8, s);
  {
    TimedLaunch t(ctx, "k_tc_wave");
    if (weighted)
      k_tc_wave<true><<<gw, kBlock, 0, s>>>(ofs, nbr, mult, V, part_w.as<unsigned long long>());
    else
      k_tc_wave<false><<<gw, kBlock, 0, s>>>(ofs, nbr, mult, V, part_w.as<unsigned long long>());
  }
  const unsigned gb = unsigned(std::min<uint64_t>(std::max<uint64_t>(nblock, 1),
                                                  uint64_t(ctx->num_cus) * 4));
  part_b.alloc(gb * 8, s);
  if (nblock) {
    TimedLaunch t(ctx, "k_tc_block");
    if (weighted)
      k_tc_block<true><<<gb, kRowBlock, 0, s>>>(ofs, nbr, mult, rows_b, nblock, V,
                                                part_b.as<unsigned long long>());
    else
      k_tc_block<false><<<gb, kRowBlock, 0, s>>>(ofs, nbr, mult, rows_b, nblock, V,
                                                 part_b.as<unsigned long long>());
  } else {
    k_zero_ull2<<<1, 1, 0, s>>>(part_b.as<unsigned long long>(), int(gb));
  }
  const unsigned gg = unsigned(std::min<uint64_t>(std::max<uint64_t>(nglobal, 1),
                                                  uint64_t(ctx->num_cus) * 8));
  part_g.alloc(gg * 8, s);
  if (nglobal) {
    gpre.alloc(std::max<uint64_t>(g->sym_edges, 1) * 4, s);
    TimedLaunch t(ctx, "k_tc_global");
    if (weighted)
      k_tc_global<true><<<gg, kBlock, 0, s>>>(ofs, nbr, mult, rows_g, nglobal,
                                              gpre.as<uint32_t>(), part_g.as<unsigned long long>());
    else
      k_tc_global<false><<<gg, kBlock, 0, s>>>(ofs, nbr, mult, rows_g, nglobal,
                                               gpre.as<uint32_t>(), part_g.as<unsigned long long>());
  } else {
    k_zero_ull2<<<1, 1, 0, s>>>(part_g.as<unsigned long long>(), int(gg));
  }
  LLG_HIP(hipGetLastError());
  total.alloc(24, s);
  k_sum_partials<<<1, 1024, 0, s>>>(part_w.as<unsigned long long>(), gw,
                                    total.as<unsigned long long>());
  k_sum_partials<<<1, 1024, 0, s>>>(part_b.as<unsigned long long>(), gb,
                                    total.as<unsigned long long>() + 1);
  k_sum_partials<<<1, 1024, 0, s>>>(part_g.as<unsigned long long>(), gg,
                                    total.as<unsigned long long>() + 2);
  LLG_HIP(hipGetLastError());
  LLG_HIP(hipMemcpyAsync(h, total.p, sizeof(h), hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
  r.triangles = h[0] + h[1] + h[2];
  return r;
}

}  // namespace llgpu

extern "C" {

int ll_gpu_tc_ex(ll_gpu_graph* g, int level, uint64_t* triangles, uint64_t* arcs,
                 uint64_t* work) {
  if (!g || !triangles) return LL_GPU_EINVAL;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    llgpu::TcResult r = llgpu::tc_device(g, level, false);
    *triangles = r.triangles;
    if (arcs) *arcs = r.arcs;
    if (work) *work = r.work;
    return LL_GPU_OK;
  } catch (const llgpu::Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

int ll_gpu_tc(ll_gpu_graph* g, int level, uint64_t* triangles) {
  return ll_gpu_tc_ex(g, level, triangles, nullptr, nullptr);
}

int ll_gpu_tc_i(ll_gpu_graph* g, int level, uint64_t* count) {
  if (!g || !count) return LL_GPU_EINVAL;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    *count = llgpu::tc_device(g, level, true).triangles;
    return LL_GPU_OK;
  } catch (const llgpu::Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

}  // extern "C"
