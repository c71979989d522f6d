// Multi-source BFS with sources packed into lanes of a per-vertex bit word
// (16, 32, 64 or 128 lanes).  Level l+1 of every source is one sweep:
//   next[v] = (OR_{u in in(v)} frontier[u]) & ~visited[v]
// taken one of three ways, chosen per level from the previous sweep's report:
//  - push, while the frontier's out-edges are few: a queue of the frontier
//    vertices, their out-edges OR-ed into next[] edge-parallel;
//  - binned, while the rows some search has not reached hold many in-edges:
//    over the binned out-edge layout, every slot streamed (no random reads);
//  - pull otherwise, on the row blocks of the in-CSR: a workgroup stages
//    frontier[u] for a run of short rows in LDS with coalesced neighbour loads
//    and ORs each row in one thread; long rows are split into pieces whose
//    partial ORs are combined in piece order.  Rows already visited by every
//    lane still searching are skipped.
// Levels are unique per source, so each lane's levels equal a single-source
// BFS from that lane's root (sssp_unweighted semantics, sssp.h:220-333).
#include <climits>
#include <string>

#include "bins.h"
#include "ll_gpu_internal.h"
#include "rowblocks.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr int32_t kUnreached = INT_MAX - 1;
constexpr uint32_t kCap = 1024;  // stream-item capacity of the in-CSR row blocks

struct W128 {
  uint64_t a, b;
};
__device__ __host__ inline W128 operator|(W128 x, W128 y) { return {x.a | y.a, x.b | y.b}; }
__device__ __host__ inline W128 operator&(W128 x, W128 y) { return {x.a & y.a, x.b & y.b}; }
__device__ __host__ inline W128 operator~(W128 x) { return {~x.a, ~x.b}; }
__device__ __host__ inline bool operator==(W128 x, W128 y) { return x.a == y.a && x.b == y.b; }

template <typename W> struct Lane;
template <> struct Lane<uint16_t> {
  static constexpr int kBits = 16;
  __device__ __host__ static uint16_t zero() { return 0; }
  __device__ __host__ static uint16_t low(int n) { return n >= 16 ? 0xffff : uint16_t((1u << n) - 1); }
  __device__ static bool test(uint16_t w, int s) { return (w >> s) & 1; }
  __device__ static int popc(uint16_t w) { return __popc(w); }
  __device__ __host__ static uint64_t lo64(uint16_t w) { return uint64_t(w); }
  __device__ __host__ static uint64_t hi64(uint16_t) { return 0; }
  __device__ __host__ static uint16_t from64(uint64_t lo, uint64_t) { return uint16_t(lo); }
};
template <> struct Lane<uint32_t> {
  static constexpr int kBits = 32;
  __device__ __host__ static uint32_t zero() { return 0; }
  __device__ __host__ static uint32_t low(int n) { return n >= 32 ? ~0u : (1u << n) - 1; }
  __device__ static bool test(uint32_t w, int s) { return (w >> s) & 1; }
  __device__ static int popc(uint32_t w) { return __popc(w); }
  __device__ __host__ static uint64_t lo64(uint32_t w) { return uint64_t(w); }
  __device__ __host__ static uint64_t hi64(uint32_t) { return 0; }
  __device__ __host__ static uint32_t from64(uint64_t lo, uint64_t) { return uint32_t(lo); }
};
template <> struct Lane<uint64_t> {
  static constexpr int kBits = 64;
  __device__ __host__ static uint64_t zero() { return 0; }
  __device__ __host__ static uint64_t low(int n) { return n >= 64 ? ~0ull : (1ull << n) - 1; }
  __device__ static bool test(uint64_t w, int s) { return (w >> s) & 1; }
  __device__ static int popc(uint64_t w) { return __popcll(w); }
  __device__ __host__ static uint64_t lo64(uint64_t w) { return w; }
  __device__ __host__ static uint64_t hi64(uint64_t) { return 0; }
  __device__ __host__ static uint64_t from64(uint64_t lo, uint64_t) { return lo; }
};
template <> struct Lane<W128> {
  static constexpr int kBits = 128;
  __device__ __host__ static W128 zero() { return {0, 0}; }
  __device__ __host__ static W128 low(int n) {
    return n >= 128 ? W128{~0ull, ~0ull}
                    : n >= 64 ? W128{~0ull, (n == 64 ? 0ull : (1ull << (n - 64)) - 1)}
                              : W128{(1ull << n) - 1, 0ull};
  }
  __device__ static bool test(W128 w, int s) { return s < 64 ? (w.a >> s) & 1 : (w.b >> (s - 64)) & 1; }
  __device__ static int popc(W128 w) { return __popcll(w.a) + __popcll(w.b); }
  __device__ __host__ static uint64_t lo64(W128 w) { return w.a; }
  __device__ __host__ static uint64_t hi64(W128 w) { return w.b; }
  __device__ __host__ static W128 from64(uint64_t lo, uint64_t hi) { return {lo, hi}; }
};

template <typename W>
struct MsArgs {
  const RowItem* items;
  const uint64_t* ofs;
  const uint32_t* adj;
  const W* fr;
  W* vis;
  W* nx;
  W* part;
  int32_t* lv;            // n x V, or nullptr
  uint64_t V;
  int32_t next_lv;
  int src_base, n_src;    // lanes [0, n_src) map to sources src_base + lane
  W full;
  W done;                 // lanes whose search has ended (empty frontier): count as reached
  const uint32_t* outdeg; // out-degrees (frontier work estimate)
  unsigned long long* found;  // striped OR of the lanes that discovered vertices
};

// Row r discovered by lanes nw (not zero) over visited lanes vr: visited |=
// nw, levels; adds the out-degree of r to `work` (next frontier's edges) and
// returns it.
template <typename W>
__device__ inline uint32_t discover(const MsArgs<W>& a, uint32_t r, W vr, W nw,
                                    unsigned long long& work) {
  a.vis[r] = vr | nw;
  const uint32_t od = a.outdeg[r];
  work += od;
  if (a.lv) {
    for (int s = 0; s < a.n_src; s++)
      if (Lane<W>::test(nw, s)) a.lv[uint64_t(a.src_base + s) * a.V + r] = a.next_lv;
  }
  return od;
}

// Row r's new lanes: discover(), next frontier = the new lanes of a vertex
// with out-edges (the frontier words of a vertex without any stay zero: no
// sweep propagates from it); adds the row's in-degree `deg` to `open` while
// a lane still searching has not reached it.  Returns the new lanes.
template <typename W>
__device__ inline W finalize(const MsArgs<W>& a, uint32_t r, W acc, unsigned long long& work,
                             unsigned long long& open, uint64_t deg) {
  const W vr = a.vis[r];
  const W nw = acc & ~vr;
  if (!((vr | nw | a.done) == a.full)) open += deg;
  W keep = Lane<W>::zero();
  if (!(nw == Lane<W>::zero()) && discover(a, r, vr, nw, work) > 0) keep = nw;
  a.nx[r] = keep;
  return nw;
}

// Sum over the workgroup, returned to every thread.
__device__ inline uint32_t block_sum_u32(uint32_t x) {
  __shared__ uint32_t part[kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
  __syncthreads();
  uint32_t t = 0;
  for (int i = 0; i < kBlock / 64; i++) t += part[i];
  __syncthreads();  // part may be reused by the next call
  return t;
}

// Sweep report: the lanes that discovered vertices (OR), the out-degree sum
// of the discovered vertices and the in-degree sum of the rows a lane still
// searching has not reached (pull and binned sweeps).  A workgroup reduces
// its threads' values and adds them to its stripe (blockIdx % kFlagStripes,
// one 128-byte line each: word 0/1 = lanes 0-63/64-127, word 2 = out-degree
// sum, word 3 = open in-degree sum), so hundreds of thousands of workgroups
// spread over many L2 lines; the host combines the stripes.
constexpr int kFlagStripes = 64;
constexpr int kFlagStride = 16;  // u64 per 128-byte line

template <typename W, int kThreads = kBlock>
__device__ inline void block_report(unsigned long long* flags, W x, unsigned long long work,
                                    unsigned long long open) {
  __shared__ unsigned long long part[4][kThreads / 64];
  unsigned long long lo = Lane<W>::lo64(x), hi = Lane<W>::hi64(x);
  for (int o = 32; o > 0; o >>= 1) {
    lo |= __shfl_xor(lo, o, 64);
    hi |= __shfl_xor(hi, o, 64);
    work += __shfl_xor(work, o, 64);
    open += __shfl_xor(open, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    part[0][threadIdx.x >> 6] = lo;
    part[1][threadIdx.x >> 6] = hi;
    part[2][threadIdx.x >> 6] = work;
    part[3][threadIdx.x >> 6] = open;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    unsigned long long t = 0;
    for (int i = 0; i < kThreads / 64; i++)
      t = threadIdx.x < 2 ? (t | part[threadIdx.x][i]) : t + part[threadIdx.x][i];
    unsigned long long* dst = &flags[(blockIdx.x % kFlagStripes) * kFlagStride + threadIdx.x];
    if (t) {
      if (threadIdx.x < 2) atomicOr(dst, t);
      else atomicAdd(dst, t);
    }
  }
}

// Stripe 0's word 4 counts the vertices a push sweep reaches (kTouchCount),
// word 5 those of them with out-edges, listed by the settle (kListCount);
// word 6 is set when a push workgroup's edges span more queue entries than
// it stages (kOverflow: an entry without out-edges, which no list holds).
constexpr int kTouchCount = 4;
constexpr int kListCount = 5;
constexpr int kOverflow = 6;

__global__ void k_zero_flags(unsigned long long* flags) {  // <<<1, kFlagStripes>>>
  for (int w = 0; w <= kOverflow; w++) flags[threadIdx.x * kFlagStride + w] = 0ull;
}

template <typename W>
__global__ __launch_bounds__(kBlock) void k_ms_pull(MsArgs<W> a) {
  __shared__ W vals[kCap + kCap / 2];
  __shared__ uint64_t rofs[kRowsPerItem + 1];
  __shared__ W red[kBlock / 64];
  const RowItem it = a.items[blockIdx.x];
  W cnt = Lane<W>::zero();  // lanes this thread's row was discovered by
  unsigned long long work = 0;  // out-edges of the vertices it discovered
  unsigned long long open = 0;  // in-edges of its row while a lane has not reached it
  if (it.npieces == 0) {
    const uint32_t nrows = it.row_end - it.row_begin;
    // rows already reached from every lane still searching need no gathers
    const bool live =
        threadIdx.x < nrows && !((a.vis[it.row_begin + threadIdx.x] | a.done) == a.full);
    if (!__syncthreads_or(live)) {
      if (threadIdx.x < nrows) a.nx[it.row_begin + threadIdx.x] = Lane<W>::zero();
      return;
    }
    for (uint32_t i = threadIdx.x; i <= nrows; i += kBlock) rofs[i] = a.ofs[it.row_begin + i];
    __syncthreads();
    const uint64_t e0 = rofs[0];
    const uint32_t ne = uint32_t(rofs[nrows] - e0);
    // few open edges (late sweeps): each open row walks its own edges and stops
    // once every lane has reached it
    const uint32_t lo0 = threadIdx.x < nrows ? uint32_t(rofs[threadIdx.x] - e0) : 0u;
    const uint32_t hi0 = threadIdx.x < nrows ? uint32_t(rofs[threadIdx.x + 1] - e0) : 0u;
    const uint32_t open_edges = block_sum_u32(live ? hi0 - lo0 : 0u);
    if (open_edges * 4 < ne) {
      if (threadIdx.x < nrows) {
        const uint32_t r = it.row_begin + threadIdx.x;
        if (live) {
          const W vr = a.vis[r];
          W acc = Lane<W>::zero();
          for (uint32_t j = lo0; j < hi0; j++) {
            acc = acc | a.fr[a.adj[e0 + j]];
            if ((acc | vr | a.done) == a.full) break;
          }
          cnt = finalize(a, r, acc, work, open, hi0 - lo0);
        } else {
          a.nx[r] = Lane<W>::zero();
        }
      }
      block_report(a.found, cnt, work, open);
      return;
    }
    // entries of rows every lane has reached are neither loaded nor gathered:
    // rowof maps an entry to its row in the item
    __shared__ uint8_t rowof[kCap + kCap / 2];
    __shared__ uint8_t openf[kRowsPerItem];
    if (threadIdx.x < nrows) {
      openf[threadIdx.x] = live ? 1 : 0;
      if (live)
        for (uint32_t j = lo0; j < hi0; j++) rowof[j] = uint8_t(threadIdx.x);
    }
    __syncthreads();
    constexpr uint32_t kPer = (kCap + kCap / 2) / kBlock;
    uint32_t idx[kPer];
    bool need[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      // rowof is only written for open rows: test the row range, not just the byte
      need[k] = false;
      if (i < ne) {
        const uint32_t rr = rowof[i];
        need[k] = rr < nrows && openf[rr] && i >= uint32_t(rofs[rr] - e0) &&
                  i < uint32_t(rofs[rr + 1] - e0);
      }
      idx[k] = need[k] ? a.adj[e0 + i] : 0u;
    }
#pragma unroll
    for (uint32_t k = 0; k < kPer; k++) {
      uint32_t i = threadIdx.x + k * kBlock;
      if (need[k]) vals[i] = a.fr[idx[k]];
    }
    __syncthreads();
    if (threadIdx.x < nrows) {
      const uint32_t r = it.row_begin + threadIdx.x;
      if (!((a.vis[r] | a.done) == a.full)) {
        W acc = Lane<W>::zero();
        const uint32_t lo = uint32_t(rofs[threadIdx.x] - e0), hi = uint32_t(rofs[threadIdx.x + 1] - e0);
        for (uint32_t j = lo; j < hi; j++) acc = acc | vals[j];
        cnt = finalize(a, r, acc, work, open, hi - lo);
      } else {
        a.nx[r] = Lane<W>::zero();
      }
    }
  } else {
    const uint32_t r = it.row_begin;
    const uint64_t lo = a.ofs[r] + uint64_t(it.piece) * kPiece;
    const uint64_t hi = min(a.ofs[r + 1], lo + kPiece);
    W acc = Lane<W>::zero();
    if (!((a.vis[r] | a.done) == a.full))
      for (uint64_t e = lo + threadIdx.x; e < hi; e += kBlock) acc = acc | a.fr[a.adj[e]];
    // OR-reduce the workgroup through LDS
    __shared__ W wred[kBlock];
    wred[threadIdx.x] = acc;
    __syncthreads();
    for (unsigned st = kBlock / 2; st > 0; st >>= 1) {
      if (threadIdx.x < st) wred[threadIdx.x] = wred[threadIdx.x] | wred[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      if (it.npieces == 1) cnt = finalize(a, r, wred[0], work, open, hi - lo);
      else a.part[blockIdx.x] = wred[0];
    }
    (void)red;
  }
  block_report(a.found, cnt, work, open);
}

template <typename W>
__global__ void k_ms_fixup(MsArgs<W> a, uint64_t nitems) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  W cnt = Lane<W>::zero();
  unsigned long long work = 0, open = 0;
  if (i < nitems) {
    const RowItem it = a.items[i];
    if (it.npieces > 1 && it.piece == 0) {
      W acc = a.part[i];
      for (uint32_t p = 1; p < it.npieces; p++) acc = acc | a.part[i + p];
      const uint32_t r = it.row_begin;
      cnt = finalize(a, r, acc, work, open, a.ofs[r + 1] - a.ofs[r]);
    }
  }
  block_report(a.found, cnt, work, open);
}

template <typename W>
__global__ void k_ms_init(W* vis, W* fr, W* nx, uint64_t V, W init) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V) {
    vis[v] = init;
    fr[v] = Lane<W>::zero();
    nx[v] = Lane<W>::zero();
  }
}

__global__ void k_fill_i32(int32_t* p, uint64_t n, int32_t x) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = x;
}

// The roots' lanes; the first frontier: the roots with out-edges, listed
// once each, their count into stripe 0's word kListCount and their out-edges
// into word 2.
template <typename W>
__global__ void k_ms_roots(W* vis, W* fr, const int64_t* roots, int n, int32_t* lv, uint64_t V,
                           int src_base, const uint32_t* outdeg, uint32_t* lst,
                           unsigned long long* flags) {
  // one thread: roots may repeat across lanes
  unsigned long long nl = 0, work = 0;
  for (int s = 0; s < n; s++) {
    const uint64_t r = uint64_t(roots[s]);
    W bit = Lane<W>::low(s + 1) & ~Lane<W>::low(s);
    vis[r] = vis[r] | bit;
    if (lv) lv[uint64_t(src_base + s) * V + r] = 0;
    const uint32_t od = outdeg[r];
    if (od == 0) continue;
    if (fr[r] == Lane<W>::zero()) {
      lst[nl++] = uint32_t(r);
      work += od;
    }
    fr[r] = fr[r] | bit;
  }
  flags[kListCount] = nl;
  flags[2] = work;
}

__global__ void k_zero_ull(unsigned long long* p) { *p = 0; }

// ---- top-down (push) sweeps for small frontiers ---------------------------
// A push sweep works from the list of the frontier's vertices (every vertex
// with frontier lanes; all have out-edges) and lists the vertices it
// discovers that have out-edges: the next sweep's list.  After a pull or
// binned sweep the frontier is listed by a scan of every vertex: fixed grid;
// each workgroup owns a contiguous vertex range, counts it, takes one list
// allocation and writes its vertices in order.
template <typename W>
__global__ __launch_bounds__(kBlock) void k_ms_queue(const W* fr, uint64_t V, uint32_t* q,
                                                     unsigned long long* qn) {
  __shared__ unsigned long long base;
  __shared__ uint32_t wcnt[kBlock / 64];
  const uint64_t per = (V + gridDim.x - 1) / gridDim.x;
  const uint64_t v0 = uint64_t(blockIdx.x) * per, v1 = min(V, v0 + per);
  uint32_t c = 0;
  for (uint64_t v = v0 + threadIdx.x; v < v1; v += kBlock)
    c += !(fr[v] == Lane<W>::zero()) ? 1u : 0u;
  const uint32_t tot = block_sum_u32(c);
  if (threadIdx.x == 0) base = tot ? atomicAdd(qn, (unsigned long long)tot) : 0ull;
  __syncthreads();
  unsigned long long off = base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint64_t r0 = v0; r0 < v1; r0 += kBlock) {
    const uint64_t v = r0 + threadIdx.x;
    const bool in = v < v1 && !(fr[v] == Lane<W>::zero());
    const uint64_t m = __ballot(in);
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    uint32_t before = 0, round = 0;
    for (int i = 0; i < kBlock / 64; i++) {
      if (i < wave) before += wcnt[i];
      round += wcnt[i];
    }
    if (in) q[off + before + __popcll(m & ((uint64_t(1) << lane) - 1))] = uint32_t(v);
    off += round;
    __syncthreads();
  }
}

__global__ void k_queue_deg(const uint32_t* q, uint64_t n, const uint32_t* outdeg, uint32_t* d) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) d[i] = outdeg[q[i]];
}

constexpr uint32_t kPushEdges = 1024;  // queue out-edges per push workgroup

// First and last queue entry of each push workgroup (parallel binary searches).
__global__ void k_push_blocks(const uint64_t* qofs, uint64_t nq, uint64_t m_f, uint64_t nblk,
                              uint64_t* blk) {
  const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
  const uint64_t x0 = b * kPushEdges, x1 = min(m_f, x0 + kPushEdges);
  auto search = [&](uint64_t lo, uint64_t x) {  // largest i with qofs[i] <= x
    uint64_t hi = nq;
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) / 2;
      if (qofs[mid] <= x) lo = mid;
      else hi = mid;
    }
    return lo;
  };
  const uint64_t lo = search(0, x0);
  blk[2 * b] = lo;
  blk[2 * b + 1] = search(lo, x1 - 1);
}

// nx[v] |= m (m not zero); returns true for exactly one of the ORs into a
// word that started at zero.  Up to 64 lanes: the OR that found the word
// zero.  128 lanes take two atomics, so v's bit in `touched` decides.
__device__ inline bool lanes_or(uint16_t* p, uint16_t m, uint32_t*, uint32_t) {
  // 16-bit lanes: OR into the containing aligned 32-bit word
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  const int sh = (a & 2) ? 16 : 0;
  const uint32_t old =
      atomicOr(reinterpret_cast<unsigned int*>(a & ~uintptr_t(3)), uint32_t(m) << sh);
  return ((old >> sh) & 0xffffu) == 0;
}
__device__ inline bool lanes_or(uint32_t* p, uint32_t m, uint32_t*, uint32_t) {
  return atomicOr(p, m) == 0;
}
__device__ inline bool lanes_or(uint64_t* p, uint64_t m, uint32_t*, uint32_t) {
  return atomicOr(reinterpret_cast<unsigned long long*>(p), (unsigned long long)m) == 0;
}
__device__ inline bool lanes_or(W128* p, W128 m, uint32_t* touched, uint32_t v) {
  if (m.a) atomicOr(reinterpret_cast<unsigned long long*>(&p->a), (unsigned long long)m.a);
  if (m.b) atomicOr(reinterpret_cast<unsigned long long*>(&p->b), (unsigned long long)m.b);
  const uint32_t bit = 1u << (v & 31);
  return (atomicOr(&touched[v >> 5], bit) & bit) == 0;
}
// Lane words whose first touch is read off the OR itself.
template <typename W> constexpr bool kOrFindsFirst = true;
template <> constexpr bool kOrFindsFirst<W128> = false;

// Edge-parallel push over the queue's out-edges: nx[v] |= fr[u] & ~vis[v]
// (nx is all zero before the sweep).  The OR that first gives v lanes
// (lanes_or) lists v in `tl`: the workgroup gathers its vertices in LDS and
// takes their list positions with one atomic, so every vertex the sweep
// reaches is listed once.  Every queue entry has out-edges, so a workgroup's
// kPushEdges edges span at most kPushEdges + 1 entries.
template <typename W>
__global__ __launch_bounds__(kBlock) void k_ms_push(const uint32_t* q, const uint64_t* qofs,
                                                    const uint64_t* blk, uint64_t m_f,
                                                    const uint64_t* ofs, const uint32_t* adj,
                                                    const W* fr, const W* vis, W* nx,
                                                    uint32_t* touched, uint32_t* tl,
                                                    unsigned long long* flags) {
  __shared__ uint64_t l_qofs[kPushEdges + 1];
  __shared__ uint64_t l_base[kPushEdges];
  __shared__ W l_fr[kPushEdges];
  __shared__ uint32_t l_new[kPushEdges];
  __shared__ uint32_t l_nnew;
  __shared__ unsigned long long l_pos;
  if (threadIdx.x == 0) l_nnew = 0;
  const uint64_t x0 = uint64_t(blockIdx.x) * kPushEdges;
  const uint64_t x1 = min(m_f, x0 + kPushEdges);
  const uint64_t lo = blk[2 * blockIdx.x];
  const uint64_t span = blk[2 * blockIdx.x + 1] - lo + 1;
  // at most kPushEdges entries are staged; more is reported (the edges past
  // the staged entries are skipped below)
  if (span > kPushEdges && threadIdx.x == 0) flags[kOverflow] = 1ull;
  const uint32_t n = uint32_t(min(span, uint64_t(kPushEdges)));
  for (uint32_t i = threadIdx.x; i <= n; i += kBlock) {
    l_qofs[i] = qofs[lo + i];
    if (i < n) {
      const uint32_t u = q[lo + i];
      l_base[i] = ofs[u];
      l_fr[i] = fr[u];
    }
  }
  __syncthreads();
  // x < l_qofs[n]: the edge belongs to one of this workgroup's queue entries
  for (uint64_t x = x0 + threadIdx.x; x < x1 && x < l_qofs[n]; x += kBlock) {
    uint32_t a = 0, b = n;
    while (b - a > 1) {
      const uint32_t mid = (a + b) >> 1;
      if (l_qofs[mid] <= x) a = mid;
      else b = mid;
    }
    const uint32_t v = adj[l_base[a] + (x - l_qofs[a])];
    const W m = l_fr[a] & ~vis[v];
    if (!(m == Lane<W>::zero()) && lanes_or(&nx[v], m, touched, v))
      l_new[atomicAdd(&l_nnew, 1u)] = v;  // at most one per edge: kPushEdges
  }
  __syncthreads();
  const uint32_t nn = l_nnew;
  if (threadIdx.x == 0 && nn) l_pos = atomicAdd(&flags[kTouchCount], (unsigned long long)nn);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nn; i += kBlock) tl[l_pos + i] = l_new[i];
}

// After a push: each reached vertex (list `tl`, count in stripe 0's word
// kTouchCount) settles its pushed lanes like a pull row would (and clears its
// `touched` bit where lanes_or used it); those with out-edges are listed in
// `lst` (count in word kListCount; one atomic per workgroup round), the
// others' frontier words cleared.  The pushed lanes were unvisited, so nx[v]
// is v's new lanes as it stands.  Fixed grid.  The open in-degree sum is not
// kept.
template <typename W>
__global__ __launch_bounds__(kBlock) void k_ms_settle(MsArgs<W> a, const uint32_t* tl,
                                                      uint32_t* touched, uint32_t* lst) {
  __shared__ uint32_t l_keep[kBlock];
  __shared__ uint32_t l_nkeep;
  __shared__ unsigned long long l_pos;
  const uint64_t n = a.found[kTouchCount];
  W cnt = Lane<W>::zero();
  unsigned long long work = 0;
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  // rounds of kBlock vertices, the same for every thread of the workgroup
  for (uint64_t i0 = uint64_t(blockIdx.x) * kBlock; i0 < n; i0 += stride) {
    if (threadIdx.x == 0) l_nkeep = 0;
    __syncthreads();
    const uint64_t i = i0 + threadIdx.x;
    if (i < n) {
      const uint32_t v = tl[i];
      const W vr = a.vis[v];
      const W nw = a.nx[v];
      if (discover(a, v, vr, nw, work) > 0) l_keep[atomicAdd(&l_nkeep, 1u)] = v;
      else a.nx[v] = Lane<W>::zero();
      cnt = cnt | nw;
      if (!kOrFindsFirst<W>) atomicAnd(&touched[v >> 5], ~(1u << (v & 31)));
    }
    __syncthreads();
    const uint32_t nk = l_nkeep;
    if (threadIdx.x == 0 && nk) l_pos = atomicAdd(&a.found[kListCount], (unsigned long long)nk);
    __syncthreads();
    if (threadIdx.x < nk) lst[l_pos + threadIdx.x] = l_keep[threadIdx.x];
    __syncthreads();  // l_keep and l_nkeep are reused by the next round
  }
  block_report(a.found, cnt, work, 0ull);
}

// Clears the lane words of the listed vertices.
template <typename W>
__global__ void k_clear_listed(W* p, const uint32_t* lst, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[lst[i]] = Lane<W>::zero();
}

// ---- binned sweeps for levels with many open rows ---------------------------
// Over the binned out-edge layout (bins.h): a scatter launch writes each
// source's frontier lanes into its slots (one workgroup per source bin, the
// bin's lanes staged in LDS, the slots written in chunk runs), then a gather
// launch ORs each destination bin's slots into per-vertex lane words in LDS
// (streamed, no random reads) and settles the bin's vertices as a pull row
// would.  Every slot is read and written once per sweep, whatever the rows'
// state: taken while the rows still open hold a large share of the edges.
constexpr int kBinBlock = 1024;
constexpr int kBinUnroll = 8;  // gather: slots per thread in flight
constexpr int kBinGroup = 8;   // scatter: chunks per wave in flight

// LDS lane word of the gather (16-bit lanes OR into 32-bit words)
template <typename W> struct BinAcc { using T = W; };
template <> struct BinAcc<uint16_t> { using T = uint32_t; };

__device__ inline void lds_or(uint32_t* p, uint32_t x) { atomicOr(p, x); }
__device__ inline void lds_or(uint64_t* p, uint64_t x) {
  atomicOr(reinterpret_cast<unsigned long long*>(p), (unsigned long long)x);
}
__device__ inline void lds_or(W128* p, W128 x) {
  if (x.a) atomicOr(reinterpret_cast<unsigned long long*>(&p->a), (unsigned long long)x.a);
  if (x.b) atomicOr(reinterpret_cast<unsigned long long*>(&p->b), (unsigned long long)x.b);
}

template <uint32_t BB, typename W>
__global__ __launch_bounds__(kBinBlock) void k_msb_scatter(const W* fr, uint64_t V, uint32_t nb,
                                                           const uint2* ct, const uint16_t* sidx,
                                                           W* sv) {
  constexpr uint32_t R = 1u << BB;
  __shared__ W cl[R];
  // neighbouring source bins write neighbouring chunks of every destination
  // bin: keep them on one XCD
  const uint64_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t v0 = b << BB;
  const uint32_t nrows = uint32_t(min(uint64_t(R), V - v0));
  for (uint32_t i = threadIdx.x; i < R; i += kBinBlock)
    cl[i] = i < nrows ? fr[v0 + i] : Lane<W>::zero();
  __syncthreads();
  scatter_chunks<kBinBlock, kBinGroup>(ct + b * nb, nb, sidx, cl, sv);
}

template <uint32_t BB, typename W>
__global__ __launch_bounds__(kBinBlock) void k_msb_gather(MsArgs<W> a, const W* sv,
                                                          const uint16_t* ldst,
                                                          const uint64_t* bstart,
                                                          const uint32_t* indeg) {
  using A = typename BinAcc<W>::T;
  constexpr uint32_t R = 1u << BB;
  constexpr uint32_t kPer = R / kBinBlock;
  static_assert(kPer >= 1 && R % kBinBlock == 0, "bin rows must be a multiple of the block");
  __shared__ A acc[R];
  const uint64_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t v0 = b << BB;
  const uint32_t nrows = uint32_t(min(uint64_t(R), a.V - v0));
  for (uint32_t i = threadIdx.x; i < R; i += kBinBlock) acc[i] = A(Lane<W>::zero());
  __syncthreads();
  const uint64_t k0 = bstart[b], k1 = bstart[b + 1];
  for (uint64_t k = k0 + threadIdx.x; k < k1; k += kBinUnroll * uint64_t(kBinBlock)) {
    W x[kBinUnroll];
    uint16_t d[kBinUnroll];
#pragma unroll
    for (int u = 0; u < kBinUnroll; u++) {
      const uint64_t q = k + uint64_t(u) * kBinBlock;
      x[u] = q < k1 ? sv[q] : Lane<W>::zero();
      d[u] = q < k1 ? ldst[q] : uint16_t(0);
    }
#pragma unroll
    for (int u = 0; u < kBinUnroll; u++)
      if (!(x[u] == Lane<W>::zero())) lds_or(&acc[d[u]], A(x[u]));
  }
  __syncthreads();
  W cnt = Lane<W>::zero();
  unsigned long long work = 0, open = 0;
#pragma unroll
  for (uint32_t j = 0; j < kPer; j++) {
    const uint32_t i = threadIdx.x + j * kBinBlock;
    if (i < nrows) {
      const uint32_t r = uint32_t(v0 + i);
      const W vr = a.vis[r];
      const W nw = W(acc[i]) & ~vr;
      if (!((vr | nw | a.done) == a.full)) open += indeg[r];
      W keep = Lane<W>::zero();
      if (!(nw == Lane<W>::zero())) {
        if (discover(a, r, vr, nw, work) > 0) keep = nw;
        cnt = cnt | nw;
      }
      a.nx[r] = keep;
    }
  }
  block_report<W, kBinBlock>(a.found, cnt, work, open);
}

// The bin sizes a lane width can gather: a bin's lane words fill at most
// 128 KiB of LDS.
template <typename W>
constexpr bool bins_fit(uint32_t bb) {
  return (uint64_t(sizeof(typename BinAcc<W>::T)) << bb) <= 128 * 1024;
}

template <uint32_t BB, typename W>
void launch_binned(const PrBins& pb, const W* fr, W* sv, const MsArgs<W>& a,
                   const uint32_t* indeg, hipStream_t s) {
  if constexpr (bins_fit<W>(BB)) {
    k_msb_scatter<BB, W><<<pb.nbins, kBinBlock, 0, s>>>(fr, a.V, pb.nbins, pb.ct.as<uint2>(),
                                                        pb.sidx.as<uint16_t>(), sv);
    k_msb_gather<BB, W><<<pb.nbins, kBinBlock, 0, s>>>(a, sv, pb.ldst.as<uint16_t>(),
                                                       pb.bstart.as<uint64_t>(), indeg);
  } else {
    throw Error{LL_GPU_EINVAL, "binned sweep: bins too large for the lane width"};
  }
}

template <typename W>
__global__ void k_zero_lanes(W* p, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = Lane<W>::zero();
}

// Per-lane visited counts: for each bit s of each 64-bit half of the lane
// word, popcount(ballot(bit s)) over a wave's 64 vertices, kept by thread s
// (the shifts are constants: the loop over s is unrolled, no branches).  A
// workgroup sums its waves in LDS and adds once per lane.  Grid-stride over a
// fixed grid of kReachBlock-thread workgroups.
constexpr int kReachBlock = 1024;

template <typename W>
__global__ __launch_bounds__(kReachBlock) void k_ms_reached(const W* vis, uint64_t V, int n_src,
                                                            unsigned long long* out) {
  constexpr int kHalves = Lane<W>::kBits > 64 ? 2 : 1;
  constexpr int kBitsPerHalf = Lane<W>::kBits > 64 ? 64 : Lane<W>::kBits;
  __shared__ unsigned long long acc[128];
  const int lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < 128; i += kReachBlock) acc[i] = 0;
  __syncthreads();
  uint32_t c[kHalves] = {};
  const uint64_t stride = uint64_t(gridDim.x) * kReachBlock;
  for (uint64_t v = uint64_t(blockIdx.x) * kReachBlock + threadIdx.x; v - threadIdx.x < V;
       v += stride) {
    // every lane of a wave runs the same rounds (v - threadIdx.x is uniform)
    const W w = v < V ? vis[v] : Lane<W>::zero();
    uint64_t h[kHalves];
    h[0] = Lane<W>::lo64(w);
    if (kHalves > 1) h[kHalves - 1] = Lane<W>::hi64(w);
#pragma unroll
    for (int b = 0; b < kBitsPerHalf; b++) {
#pragma unroll
      for (int j = 0; j < kHalves; j++) {
        const uint32_t k = uint32_t(__popcll(__ballot((h[j] >> b) & 1)));
        c[j] += lane == b ? k : 0u;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kHalves; j++) {
    const int sl = j * 64 + lane;
    if (sl < n_src && lane < kBitsPerHalf && c[j]) atomicAdd(&acc[sl], (unsigned long long)c[j]);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_src; i += kReachBlock)
    if (acc[i]) atomicAdd(&out[i], acc[i]);
}

}  // namespace

template <typename W>
void msbfs_batch(ll_gpu_graph* g, int level, const int64_t* d_roots, int n, int src_base,
                 int32_t* d_lv, unsigned long long* d_reached, uint64_t* levels_run) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V;
  const Csr& in = ensure_csr(g, LL_GPU_DIR_IN, level);
  const RowBlocks& rb = ensure_rowblocks(g, LL_GPU_DIR_IN, level, kCap, kPiece);
  const uint32_t* outdeg = ensure_degrees(g, LL_GPU_DIR_OUT, level);
  DevBuf vis, fr, nx, part, found;
  vis.alloc(V * sizeof(W), s);
  fr.alloc(V * sizeof(W) + 16, s);  // padded: 16-bit lanes are OR-ed through 32-bit words
  nx.alloc(V * sizeof(W) + 16, s);
  part.alloc(std::max<uint64_t>(rb.n, 1) * sizeof(W), s);
  found.alloc(kFlagStripes * kFlagStride * 8, s);
  // vertex lists of the frontiers (two, alternating), the vertices a push
  // reaches and its touched bits (all clear between sweeps)
  DevBuf lst[2], tl, touched;
  lst[0].alloc(V * 4, s);
  lst[1].alloc(V * 4, s);
  tl.alloc(V * 4, s);
  const uint64_t tw = (V + 31) / 32;
  touched.alloc(tw * 4, s);
  k_fill_i32<<<grid_for(tw, 256), 256, 0, s>>>(touched.as<int32_t>(), tw, 0);
  const W full = Lane<W>::low(Lane<W>::kBits);
  const W used = Lane<W>::low(n);
  k_ms_init<W><<<grid_for(V, 256), 256, 0, s>>>(vis.as<W>(), fr.as<W>(), nx.as<W>(), V,
                                                full & ~used);
  k_zero_flags<<<1, kFlagStripes, 0, s>>>(found.as<unsigned long long>());
  k_ms_roots<W><<<1, 1, 0, s>>>(vis.as<W>(), fr.as<W>(), d_roots + src_base, n, d_lv, V, src_base,
                                outdeg, lst[0].as<uint32_t>(), found.as<unsigned long long>());
  LLG_HIP(hipGetLastError());
  static_assert(kFlagStripes * kFlagStride * 8 <= kReadbackBytes, "flags fit the read-back buffer");
  const uint64_t* fl = static_cast<const uint64_t*>(ctx->readback);
  auto read_report = [&](uint64_t& lo, uint64_t& hi, uint64_t& work, uint64_t& open) {
    LLG_HIP(hipMemcpyAsync(ctx->readback, found.p, kFlagStripes * kFlagStride * 8,
                           hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    lo = hi = work = open = 0;
    for (int i = 0; i < kFlagStripes; i++) {
      lo |= fl[i * kFlagStride];
      hi |= fl[i * kFlagStride + 1];
      work += fl[i * kFlagStride + 2];
      open += fl[i * kFlagStride + 3];
    }
  };
  uint64_t lo, hi, m_f, open;
  read_report(lo, hi, m_f, open);  // m_f: out-edges of the frontier (the roots)
  // Which vertices hold lanes in each lane-word buffer: the list buffer
  // (0/1) naming them, kUnlisted (a pull or binned sweep wrote every vertex),
  // or kAllZero.  A push sweep needs its output buffer all zero and its
  // frontier listed.
  constexpr int kUnlisted = -1, kAllZero = -2;
  int fr_list = 0, nx_list = kAllZero;
  uint64_t fr_n = fl[kListCount], nx_n = 0;
  // Push while the frontier's out-edges are under 1/kPushRatio of the edges
  // (the first sweeps from a handful of roots).  Otherwise a binned sweep
  // while the rows some searching lane has not reached hold at least
  // kBinOpen of the in-edges (a pull skips the other rows; the binned sweep
  // streams every edge at about half a full pull's cost), else a pull.
  constexpr uint64_t kPushRatio = 16;
  const long bins_opt = ctx->options.count("bfs_batch_bins") ? ctx->options["bfs_batch_bins"] : -1;
  const long open_opt =
      ctx->options.count("bfs_batch_bin_open") ? ctx->options["bfs_batch_bin_open"] : -1;
  const double kBinOpen = open_opt >= 0 ? double(open_opt) / 100.0 : 0.4;
  const PrBins* pb = nullptr;
  const uint32_t* indeg = nullptr;
  if (bins_opt != 0) {
    const long bb = ctx->options.count("pr_bin_bits") ? ctx->options["pr_bin_bits"] : -1;
    PrBins* p = ensure_pr_bins(g, level, bb);
    if (!p && bins_opt == 1) p = ensure_pr_bins(g, level, 12);
    if (p && bins_fit<W>(p->bb)) {
      pb = p;
      indeg = ensure_degrees(g, LL_GPU_DIR_IN, level);
    }
  }
  uint64_t open_in = in.E;  // in-edges of the rows a searching lane has not reached
  DevBuf sv;                // binned sweeps: the slots' lane words
  const Csr* out = nullptr;
  DevBuf qdeg, qofs, qn, blk;
  uint64_t lv_count = 0;
  W done = full & ~used;  // lanes not searching (unused); later also finished ones
  for (int32_t l = 1;; l++) {
    MsArgs<W> a;
    a.items = rb.items.as<RowItem>();
    a.ofs = in.ofs.as<uint64_t>();
    a.adj = in.adj.as<uint32_t>();
    a.fr = fr.as<W>();
    a.vis = vis.as<W>();
    a.nx = nx.as<W>();
    a.part = part.as<W>();
    a.lv = d_lv;
    a.V = V;
    a.next_lv = l;
    a.src_base = src_base;
    a.n_src = n;
    a.full = full;
    a.done = done;
    a.outdeg = outdeg;
    a.found = found.as<unsigned long long>();
    k_zero_flags<<<1, kFlagStripes, 0, s>>>(a.found);
    const bool push = m_f * kPushRatio < in.E;
    const bool binned =
        !push && pb && (bins_opt == 1 || double(open_in) >= kBinOpen * double(in.E));
    int out_list = kUnlisted;
    if (push) {
      if (!out) {
        out = &ensure_csr(g, LL_GPU_DIR_OUT, level);
        qdeg.alloc(V * 4, s);
        qofs.alloc((V + 1) * 8, s);
        qn.alloc(8, s);
      }
      // the output words start at zero: clear the listed vertices, or all
      if (nx_list == kUnlisted || (nx_list >= 0 && nx_n > V / 16)) {
        k_zero_lanes<W><<<grid_for(V, 256), 256, 0, s>>>(nx.as<W>(), V);
      } else if (nx_list >= 0 && nx_n > 0) {
        k_clear_listed<W><<<grid_for(nx_n, 256), 256, 0, s>>>(nx.as<W>(),
                                                              lst[nx_list].as<uint32_t>(), nx_n);
      }
      if (fr_list == kUnlisted) {
        // the list buffers are free: the output's list (if any) was read above
        k_zero_ull<<<1, 1, 0, s>>>(qn.as<unsigned long long>());
        k_ms_queue<W><<<std::min(grid_for(V, kBlock), 4u * unsigned(ctx->num_cus)), kBlock, 0,
                        s>>>(fr.as<W>(), V, lst[0].as<uint32_t>(), qn.as<unsigned long long>());
        LLG_HIP(hipMemcpyAsync(ctx->readback, qn.p, 8, hipMemcpyDeviceToHost, s));
        LLG_HIP(hipStreamSynchronize(s));
        fr_list = 0;
        fr_n = *static_cast<const uint64_t*>(ctx->readback);
      }
      out_list = 1 - fr_list;
      const uint32_t* q = lst[fr_list].as<uint32_t>();
      // m_f = the out-degree sum over the listed frontier (each sweep adds the
      // out-degree of every vertex it discovers, once)
      const uint64_t nblk = (m_f + kPushEdges - 1) / kPushEdges;
      if (fr_n > 0 && nblk > 0) {
        k_queue_deg<<<grid_for(fr_n, 256), 256, 0, s>>>(q, fr_n, outdeg, qdeg.as<uint32_t>());
        scan_u32_to_u64(qdeg.as<uint32_t>(), qofs.as<uint64_t>(), fr_n, s);
        blk.alloc(nblk * 16, s);
        k_push_blocks<<<grid_for(nblk, 256), 256, 0, s>>>(qofs.as<uint64_t>(), fr_n, m_f, nblk,
                                                          blk.as<uint64_t>());
        TimedLaunch t(ctx, "k_ms_push");
        k_ms_push<W><<<unsigned(nblk), kBlock, 0, s>>>(
            q, qofs.as<uint64_t>(), blk.as<uint64_t>(), m_f, out->ofs.as<uint64_t>(),
            out->adj.as<uint32_t>(), fr.as<W>(), vis.as<W>(), nx.as<W>(), touched.as<uint32_t>(),
            tl.as<uint32_t>(), a.found);
      }
      k_ms_settle<W><<<std::min(grid_for(V, kBlock), 4u * unsigned(ctx->num_cus)), kBlock, 0, s>>>(
          a, tl.as<uint32_t>(), touched.as<uint32_t>(), lst[out_list].as<uint32_t>());
    } else if (binned) {
      if (!sv.p) sv.alloc(std::max<uint64_t>(in.E, 1) * sizeof(W), s);
      TimedLaunch t(ctx, "k_msb_sweep");
      switch (pb->bb) {
        case 12: launch_binned<12, W>(*pb, fr.as<W>(), sv.as<W>(), a, indeg, s); break;
        case 13: launch_binned<13, W>(*pb, fr.as<W>(), sv.as<W>(), a, indeg, s); break;
        default: launch_binned<14, W>(*pb, fr.as<W>(), sv.as<W>(), a, indeg, s); break;
      }
    } else {
      {
        TimedLaunch t(ctx, "k_ms_pull");
        k_ms_pull<W><<<unsigned(rb.n), kBlock, 0, s>>>(a);
      }
      k_ms_fixup<W><<<grid_for(rb.n, 256), 256, 0, s>>>(a, rb.n);
    }
    LLG_HIP(hipGetLastError());
    // lanes that discovered vertices = the next frontier's lanes
    read_report(lo, hi, m_f, open);
    LLG_CHECK(fl[kOverflow] == 0, LL_GPU_ESTATE,
              "multi-source BFS: a push queue entry has no out-edges");
    if (!push) open_in = open;
    lv_count++;
    if ((lo | hi) == 0) break;
    done = full & ~Lane<W>::from64(lo, hi);
    std::swap(fr, nx);
    // the input words keep their list (no sweep writes them)
    nx_list = fr_list;
    nx_n = fr_n;
    fr_list = out_list;
    fr_n = push ? fl[kListCount] : 0;
  }
  if (d_reached)
    k_ms_reached<W><<<std::min(grid_for(V, kReachBlock), unsigned(ctx->num_cus)), kReachBlock, 0,
                      s>>>(vis.as<W>(), V, n, d_reached + src_base);
  LLG_HIP(hipGetLastError());
  if (levels_run) *levels_run += lv_count;
}

// Time one batch per lane width on the leading roots; keep the best time per
// source for this snapshot shape.
int autotune_lanes(ll_gpu_graph* g, int level, const int64_t* roots, int n) {
  const Csr& in = ensure_csr(g, LL_GPU_DIR_IN, level);
  // decisions per request size: the smallest power of two >= min(n, 128)
  int bucket = 16;
  while (bucket < std::min(n, 128)) bucket <<= 1;
  const std::string key = "bfs_lanes:" + std::to_string(g->V) + ":" + std::to_string(in.E) + ":" +
                          std::to_string(resolve_level(g, level)) + ":" + std::to_string(bucket);
  auto it = g->ctx->tuned.find(key);
  if (it != g->ctx->tuned.end()) return it->second;
  hipStream_t s = g->ctx->stream;
  DevBuf d_roots, d_reached;
  const int m = std::min(n, 128);
  d_roots.alloc(size_t(m) * 8, s);
  d_reached.alloc(size_t(m) * 8, s);
  LLG_HIP(hipMemcpyAsync(d_roots.p, roots, size_t(m) * 8, hipMemcpyHostToDevice, s));
  llgpu::EventPair ev;
  hipEvent_t e0 = ev.a, e1 = ev.b;
  int best = 64;
  double best_t = 1e300;
  for (int w : {16, 32, 64, 128}) {
    const int k = std::min(m, w);
    LLG_HIP(hipEventRecord(e0, s));
    switch (w) {
      case 16: msbfs_batch<uint16_t>(g, level, d_roots.as<int64_t>(), k, 0, nullptr, nullptr, nullptr); break;
      case 32: msbfs_batch<uint32_t>(g, level, d_roots.as<int64_t>(), k, 0, nullptr, nullptr, nullptr); break;
      case 64: msbfs_batch<uint64_t>(g, level, d_roots.as<int64_t>(), k, 0, nullptr, nullptr, nullptr); break;
      default: msbfs_batch<W128>(g, level, d_roots.as<int64_t>(), k, 0, nullptr, nullptr, nullptr); break;
    }
    LLG_HIP(hipEventRecord(e1, s));
    LLG_HIP(hipEventSynchronize(e1));
    float ms = 0;
    LLG_HIP(hipEventElapsedTime(&ms, e0, e1));
    // a sweep costs about the same for k or w lanes: the request of n
    // sources takes ceil(n / w) batches of this width
    const double est = double(ms) * double((n + w - 1) / w);
    if (est < best_t) { best_t = est; best = w; }
  }
  g->ctx->tuned[key] = best;
  return best;
}

}  // namespace llgpu

extern "C" {

int ll_gpu_bfs_batch_ex(ll_gpu_graph* g, int level, const int64_t* roots, int n, int lane_width,
                        int32_t* levels_out, uint64_t* reached_out, uint64_t* sweeps) {
  if (!g) return LL_GPU_EINVAL;
  using namespace llgpu;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    LLG_CHECK(n >= 0 && (n == 0 || roots), LL_GPU_EINVAL, "bad root list");
    LLG_CHECK(lane_width == 0 || lane_width == 16 || lane_width == 32 || lane_width == 64 ||
                  lane_width == 128,
              LL_GPU_EINVAL, "lane_width must be 0 (autotune), 16, 32, 64 or 128");
    for (int i = 0; i < n; i++)
      LLG_CHECK(roots[i] >= 0 && uint64_t(roots[i]) < g->V, LL_GPU_EINVAL,
                "root outside [0, max_nodes)");
    if (sweeps) *sweeps = 0;
    if (n == 0) return LL_GPU_OK;
    hipStream_t s = g->ctx->stream;
    if (lane_width == 0) lane_width = autotune_lanes(g, level, roots, n);
    const uint64_t V = g->V;
    DevBuf d_roots, d_lv, d_reached;
    d_roots.alloc(size_t(n) * 8, s);
    LLG_HIP(hipMemcpyAsync(d_roots.p, roots, size_t(n) * 8, hipMemcpyHostToDevice, s));
    if (levels_out) {
      d_lv.alloc(size_t(n) * V * 4, s);
      k_fill_i32<<<grid_for(uint64_t(n) * V, 256), 256, 0, s>>>(d_lv.as<int32_t>(),
                                                                uint64_t(n) * V, kUnreached);
    }
    d_reached.alloc(size_t(n) * 8, s);
    k_fill_i32<<<grid_for(uint64_t(n) * 2, 256), 256, 0, s>>>(d_reached.as<int32_t>(),
                                                            uint64_t(n) * 2, 0);
    for (int b = 0; b < n; b += lane_width) {
      const int m = std::min(lane_width, n - b);
      int32_t* lvp = levels_out ? d_lv.as<int32_t>() : nullptr;
      unsigned long long* rp = d_reached.as<unsigned long long>();
      switch (lane_width) {
        case 16: msbfs_batch<uint16_t>(g, level, d_roots.as<int64_t>(), m, b, lvp, rp, sweeps); break;
        case 32: msbfs_batch<uint32_t>(g, level, d_roots.as<int64_t>(), m, b, lvp, rp, sweeps); break;
        case 64: msbfs_batch<uint64_t>(g, level, d_roots.as<int64_t>(), m, b, lvp, rp, sweeps); break;
        default: msbfs_batch<W128>(g, level, d_roots.as<int64_t>(), m, b, lvp, rp, sweeps); break;
      }
    }
    if (levels_out)
      LLG_HIP(hipMemcpyAsync(levels_out, d_lv.p, size_t(n) * V * 4, hipMemcpyDeviceToHost, s));
    if (reached_out)
      LLG_HIP(hipMemcpyAsync(reached_out, d_reached.p, size_t(n) * 8, hipMemcpyDeviceToHost, s));
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

extern "C" int ll_gpu_bfs_batch(ll_gpu_graph* g, int level, const int64_t* roots, int n,
                                int lane_width, int32_t* levels_out, uint64_t* reached_out) {
  return ll_gpu_bfs_batch_ex(g, level, roots, n, lane_width, levels_out, reached_out, nullptr);
}
