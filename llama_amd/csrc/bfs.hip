// Direction-optimising BFS over the snapshot (task sssp_unweighted / bfs,
// benchmark/benchmarks/sssp.h:220-333, bfs.h:85-128).
//
// The reference traverses top-down only (ll_bfs_template.h:152-293).  Levels
// of a BFS are unique, so any correct traversal order yields the same
// dist[] = level array; here each level runs either
//   * top-down over a queue of frontier vertices with out-edges: the
//     frontier's out-edge lists are concatenated by an exclusive scan of their
//     degrees and processed edge-parallel (2048 edges per workgroup; each edge
//     finds its frontier vertex by binary search in an LDS copy of the
//     workgroup's slice of the scan), claiming vertices with an atomic OR on
//     the visited bitmap; each workgroup stages its new vertices in LDS and
//     reserves their queue slots with one atomic;
//   * bottom-up over lists of the open rows (unvisited and not excluded),
//     listed in ascending order from the visited bitmap before each level:
//     rows of at most 32 in-edges a thread each, four neighbour IDs and then
//     their four frontier words at a time, stopping at the first neighbour in
//     the frontier (the V/8-byte bitmap stays in L2); longer rows eight
//     lanes each, eight rows of a wave side by side, eight neighbours per
//     row and step, a row stopping once its lanes find one; a wave's found
//     rows OR their next-frontier bits per bitmap word before one atomic per
//     word.
// Rows no search from the root can enter start as visited: rows without a
// live in-edge always, and, when the root's successors all lie in the
// snapshot's reach set (the vertices reachable from its highest-degree
// vertex, closed under successors), every row outside that set.
// The switch follows Beamer et al. (SC'12): top-down -> bottom-up when the
// frontier's out-edges exceed the unexplored edges / alpha, back when the
// frontier shrinks below V / beta (options "bfs_alpha", "bfs_beta"; default
// alpha 10).  Default beta: 24 when only dead rows are excluded; when the
// rows outside the reach set are excluded too, a search that has gone
// bottom-up stays bottom-up -- the open rows it still lists are the rows of
// the reach set it has not found yet, so each further level walks only those
// and the lists shrink, while a return to top-down pays for the queue and the
// frontier's out-edges (R-MAT s24, 64 sources, when the choice was measured:
// 115 against 101 GTEPS with beta 24).  Such a search launches its bottom-up levels kBuBatch at a time,
// each with its own counters, and reads them back once per batch: the only
// decision left is where the search ended, and the levels launched past that
// find nothing (their frontier is empty or has no out-edges).
#include <climits>
#include <optional>

#include "ll_gpu_internal.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr int kEdgesPerBlock = 2048;
constexpr int32_t kUnreached = INT_MAX - 1;

// Level counters, striped over kStripes cache lines: workgroups add their
// totals to stripe blockIdx % kStripes, so hundreds of thousands of
// workgroups do not serialise on one L2 address.  Queue slots are still
// allocated from stripe 0's n_next (fetch-add needs one counter); the host
// sums the stripes.
constexpr int kStripes = 64;
// bottom-up levels launched per host read-back once a search stays
// bottom-up; each has its own set of striped counters
constexpr int kBuBatch = 4;
struct alignas(128) Counters {
  unsigned long long n_next;     // vertices discovered this level
  unsigned long long m_next;     // their out-edges
  unsigned long long examined;   // edges inspected this level
  unsigned long long aux;        // stripe 0: 1 when the search excludes the reach complement
  unsigned long long found;      // vertices a top-down level claimed (queued or not)
  unsigned long long n_short;    // stripe 0: open rows walked by a thread each
  unsigned long long n_long;     // stripe 0: open rows walked by a lane group each
};
static_assert(sizeof(Counters) == 128, "one stripe per 128-byte line");

__device__ inline Counters* stripe(Counters* c) { return c + (blockIdx.x & (kStripes - 1)); }

// `excluded` (may be null) marks rows no search from this root enters; when
// `alt` is given and the root's check set use_alt (Counters::aux of stripe
// 0), `alt` is used instead.
__global__ void k_bfs_init(int32_t* lv, uint64_t* visited, const uint64_t* excluded,
                           const uint64_t* alt, const Counters* c, uint64_t V, uint64_t words) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (alt && c->aux) excluded = alt;
  if (lv && i < V) lv[i] = kUnreached;
  if (i < words) {
    // bits beyond V (and excluded vertices) count as visited, so no step claims them
    uint64_t lo = i * 64;
    uint64_t w = 0;
    if (lo + 64 > V) w = V > lo ? ~((uint64_t(1) << (V - lo)) - 1) : ~uint64_t(0);
    if (excluded) w |= excluded[i];
    visited[i] = w;
  }
}

__global__ void k_bfs_root(int32_t* lv, uint64_t* visited, uint32_t* queue, uint32_t root,
                           const uint32_t* outdeg, Counters* c) {
  if (lv) lv[root] = 0;
  visited[root >> 6] |= uint64_t(1) << (root & 63);
  queue[0] = root;
  c->n_next = outdeg[root] > 0 ? 1 : 0;
  c->m_next = outdeg[root];
  c->examined = 0;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__device__ inline void zero_stripe(Counters* c) {  // aux is left as it is
  c->n_next = 0;
  c->m_next = 0;
  c->examined = 0;
  c->found = 0;
  c->n_short = 0;
  c->n_long = 0;
}

__global__ void k_zero_counters(Counters* c) {  // <<<sets, kStripes>>>
  zero_stripe(c + uint64_t(blockIdx.x) * kStripes + threadIdx.x);
}

// Counters::aux of stripe 0 = 1 when every vertex the search from `root` can
// reach lies in the reach set R (the complement bitmap `notr` clear): the
// root is in R, or all its out-neighbours are (R is closed under successors).
// One workgroup.
__global__ void k_reach_check(uint32_t root, const uint64_t* notr, const uint64_t* ofs,
                              const uint32_t* adj, Counters* c) {
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const bool in_r = !((notr[root >> 6] >> (root & 63)) & 1);
  if (!in_r)
    for (uint64_t e = ofs[root] + threadIdx.x; e < ofs[root + 1]; e += blockDim.x) {
      const uint32_t w = adj[e];
      if ((notr[w >> 6] >> (w & 63)) & 1) bad = 1;
    }
  __syncthreads();
  if (threadIdx.x == 0) c->aux = (in_r || !bad) ? 1ull : 0ull;
}

// Stage the calling lane's vertex (when it has out-edges) in the workgroup's
// LDS queue lq (count *lqn) with one LDS atomic per wave; the frontier
// out-edge count goes to the striped counters.  The workgroup later reserves
// its whole share of the next queue with one atomic: the queue counter is a
// single address every workgroup adds to.
__device__ inline void wave_stage(bool take, uint32_t v, uint32_t outdeg, uint32_t* lq,
                                  uint32_t* lqn, Counters* c) {
  const uint64_t m = __ballot(take);
  const uint64_t mq = __ballot(take && outdeg > 0);
  const int lane = threadIdx.x & 63;
  const int leader = m ? __ffsll((long long)m) - 1 : 0;
  uint32_t qbase = 0;
  uint64_t md = take ? outdeg : 0;
  for (int o = 32; o > 0; o >>= 1) md += __shfl_xor(md, o, 64);
  if (lane == leader && m) {
    atomicAdd(&stripe(c)->m_next, (unsigned long long)md);
    atomicAdd(&stripe(c)->found, (unsigned long long)__popcll(m));
    if (mq) qbase = atomicAdd(lqn, uint32_t(__popcll(mq)));
  }
  qbase = uint32_t(__shfl(int(qbase), leader, 64));
  if (take && outdeg > 0) lq[qbase + uint32_t(__popcll(mq & ((uint64_t(1) << lane) - 1)))] = v;
}

__global__ __launch_bounds__(kBlock) void k_td_expand(
    const uint32_t* q, const uint64_t* qofs, uint64_t nq, uint64_t m_f, const uint64_t* ofs,
    const uint32_t* adj, const uint32_t* outdeg, uint64_t* visited, int32_t* lv, int32_t next_lv,
    uint32_t* q_next, Counters* c, Counters* zero_next) {
  __shared__ uint64_t l_qofs[kEdgesPerBlock + 1];
  __shared__ uint64_t l_base[kEdgesPerBlock];
  __shared__ uint64_t sh_lo, sh_hi;
  __shared__ uint32_t l_q[kEdgesPerBlock];  // at most one append per edge
  __shared__ uint32_t l_qn;
  __shared__ unsigned long long l_qbase;
  const uint64_t x0 = uint64_t(blockIdx.x) * kEdgesPerBlock;
  const uint64_t x1 = min(m_f, x0 + kEdgesPerBlock);
  // the counter set of the next top-down level (read by the host before this
  // launch) starts at zero
  if (blockIdx.x == 0 && threadIdx.x < kStripes) zero_stripe(zero_next + threadIdx.x);
  if (threadIdx.x == 0) {
    uint64_t lo = 0, hi = nq;
    while (hi - lo > 1) { uint64_t mid = (lo + hi) / 2; if (qofs[mid] <= x0) lo = mid; else hi = mid; }
    sh_lo = lo;
    uint64_t lo2 = lo, hi2 = nq;
    while (hi2 - lo2 > 1) { uint64_t mid = (lo2 + hi2) / 2; if (qofs[mid] <= x1 - 1) lo2 = mid; else hi2 = mid; }
    sh_hi = lo2;
    l_qn = 0;
  }
  __syncthreads();
  const uint64_t lo = sh_lo;
  const uint32_t n = uint32_t(sh_hi - lo + 1);
  for (uint32_t i = threadIdx.x; i <= n; i += kBlock) {
    l_qofs[i] = qofs[lo + i];
    if (i < n) l_base[i] = ofs[q[lo + i]];
  }
  __syncthreads();
  for (uint64_t base = x0; base < x1; base += kBlock) {
    const uint64_t x = base + threadIdx.x;
    bool take = false;
    uint32_t v = 0, od = 0;
    if (x < x1) {
      uint32_t a = 0, b = n;
      while (b - a > 1) { uint32_t mid = (a + b) >> 1; if (l_qofs[mid] <= x) a = mid; else b = mid; }
      v = adj[l_base[a] + (x - l_qofs[a])];
      const uint64_t bit = uint64_t(1) << (v & 63);
      if (!(visited[v >> 6] & bit)) {
        const uint64_t old = atomicOr((unsigned long long*)&visited[v >> 6], (unsigned long long)bit);
        if (!(old & bit)) {
          take = true;
          if (lv) lv[v] = next_lv;
          od = outdeg[v];
        }
      }
    }
    wave_stage(take, v, od, l_q, &l_qn, c);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&stripe(c)->examined, (unsigned long long)(x1 - x0));
    if (l_qn) l_qbase = atomicAdd(&c->n_next, (unsigned long long)l_qn);
  }
  __syncthreads();
  for (uint32_t k = threadIdx.x; k < l_qn; k += kBlock) q_next[l_qbase + k] = l_q[k];
}

__global__ void k_queue_to_bitmap(const uint32_t* q, uint64_t n, uint64_t* bm) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    uint32_t v = q[i];
    atomicOr((unsigned long long*)&bm[v >> 6], (unsigned long long)(uint64_t(1) << (v & 63)));
  }
}

// Bottom-up level over the open rows of the reverse CSR (unvisited and not
// excluded), each walking its in-neighbours until one is in the frontier
// (bitmap of V/8 bytes, L2-resident) -- most rows that are found stop after a
// few edges.  Rows of in-degree <= kLongRow: a thread per row, four
// neighbours loaded before their frontier words; longer rows: kSubLanes lanes
// per row, stopping once the row's lanes find one.  Found rows take
// the next level and their bit of the next frontier.  The row lists are
// rebuilt from the visited bitmap before each bottom-up level, their lengths
// stay on the device (stripe 0 of the counters), and both kernels run a
// fixed grid over them.
constexpr uint32_t kLongRow = 32;

// Lists of the open rows, short and long, from the visited bitmap: a thread
// per kOpenWords bitmap words, one list reservation (atomic) per workgroup and
// list, rows in ascending order.  `merge` (may be null): a bitmap of rows
// found by the level before, OR-ed into visited first (and written back);
// `clear` (may be null): a bitmap zeroed over the same words (the level's
// next frontier).  A row is long by the snapshot's long-row bitmap when one
// is given, else by its offsets.
constexpr int kOpenWords = 4;

__global__ __launch_bounds__(kBlock) void k_open_rows(uint64_t* visited, const uint64_t* merge,
                                                      uint64_t* clear, uint64_t words,
                                                      const uint64_t* ofs,
                                                      const uint64_t* longrows,
                                                      uint32_t* rows_short, uint32_t* rows_long,
                                                      Counters* c) {
  __shared__ uint32_t wsum[2][kBlock / 64];
  __shared__ unsigned long long base[2];
  const uint64_t w0 = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) * kOpenWords;
  uint64_t open[kOpenWords], longm[kOpenWords];
  uint32_t nl = 0, ns = 0;
#pragma unroll
  for (int k = 0; k < kOpenWords; k++) {
    const uint64_t w = w0 + uint64_t(k);
    open[k] = 0;
    longm[k] = 0;
    if (w < words) {
      uint64_t v = visited[w];
      if (merge) {
        v |= merge[w];
        visited[w] = v;
      }
      if (clear) clear[w] = 0;
      open[k] = ~v;  // padding bits are set in visited
      if (longrows) {
        longm[k] = open[k] & longrows[w];
      } else {
        for (uint64_t b = open[k]; b; b &= b - 1) {
          const int i = __ffsll((long long)b) - 1;
          const uint64_t r = w * 64 + uint64_t(i);
          if (ofs[r + 1] - ofs[r] > kLongRow) longm[k] |= uint64_t(1) << i;
        }
      }
      nl += uint32_t(__popcll(longm[k]));
      ns += uint32_t(__popcll(open[k] & ~longm[k]));
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t xs = ns, xl = nl;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t ys = __shfl_up(xs, o, 64), yl = __shfl_up(xl, o, 64);
    if (lane >= o) {
      xs += ys;
      xl += yl;
    }
  }
  if (lane == 63) {
    wsum[0][wave] = xs;
    wsum[1][wave] = xl;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint32_t t = 0;
    for (int k = 0; k < kBlock / 64; k++) t += wsum[threadIdx.x][k];
    unsigned long long* dst = threadIdx.x == 0 ? &c->n_short : &c->n_long;
    base[threadIdx.x] = t ? atomicAdd(dst, (unsigned long long)t) : 0ull;
  }
  __syncthreads();
  unsigned long long os = base[0] + xs - ns, ol = base[1] + xl - nl;
  for (int k = 0; k < wave; k++) {
    os += wsum[0][k];
    ol += wsum[1][k];
  }
#pragma unroll
  for (int k = 0; k < kOpenWords; k++)
    for (uint64_t b = open[k]; b; b &= b - 1) {
      const int i = __ffsll((long long)b) - 1;
      const uint32_t r = uint32_t((w0 + uint64_t(k)) * 64 + uint64_t(i));
      if ((longm[k] >> i) & 1) rows_long[ol++] = r;
      else rows_short[os++] = r;
    }
}

__device__ inline void found_row(uint32_t r, int32_t* lv, int32_t next_lv, uint64_t* nbm) {
  if (lv) lv[r] = next_lv;
  atomicOr((unsigned long long*)&nbm[r >> 6], (unsigned long long)(uint64_t(1) << (r & 63)));
}

// Add the workgroup's level totals to its stripe (one atomic per total).
__device__ inline void block_counters(unsigned long long n_new, unsigned long long m_new,
                                      unsigned long long exam, Counters* c) {
  __shared__ unsigned long long red[3][kBlock / 64];
  for (int o = 32; o > 0; o >>= 1) {
    n_new += __shfl_down(n_new, o, 64);
    m_new += __shfl_down(m_new, o, 64);
    exam += __shfl_down(exam, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = n_new;
    red[1][threadIdx.x >> 6] = m_new;
    red[2][threadIdx.x >> 6] = exam;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlock / 64; i++) t += red[threadIdx.x][i];
    if (t) {
      Counters* cs = stripe(c);
      unsigned long long* dst = threadIdx.x == 0 ? &cs->n_next : threadIdx.x == 1 ? &cs->m_next
                                                                                  : &cs->examined;
      atomicAdd(dst, t);
    }
  }
}

// A wave's found rows set their bits of the next frontier: lanes holding the
// same bitmap word (the row lists ascend, so they are neighbours) OR their
// bits together first, and the first lane of each such run issues the
// atomic.  Every lane of the wave calls it.
__device__ inline void wave_set_bits(uint64_t* bm, bool found, uint32_t r, bool valid) {
  const int lane = threadIdx.x & 63;
  const uint64_t word = valid ? uint64_t(r >> 6) : ~uint64_t(0);
  uint64_t bits = found ? uint64_t(1) << (r & 63) : 0ull;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t ob = __shfl_down(bits, o, 64);
    const uint64_t ow = __shfl_down(word, o, 64);
    if (lane + o < 64 && ow == word) bits |= ob;
  }
  const uint64_t prev = __shfl_up(word, 1, 64);
  if (bits && (lane == 0 || prev != word))
    atomicOr((unsigned long long*)&bm[word], (unsigned long long)bits);
}

__global__ __launch_bounds__(kBlock) void k_bu_short(const uint32_t* rows, const uint64_t* ofs,
                                                     const uint32_t* adj, const uint32_t* outdeg,
                                                     const uint64_t* fbm, uint64_t* nbm,
                                                     int32_t* lv, int32_t next_lv, Counters* c) {
  const uint64_t n = c->n_short;  // stripe 0
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  unsigned long long n_new = 0, m_new = 0, exam = 0;
  // Rows are walked one after another per thread; the list entry two rows
  // ahead and the offsets of the next row are loaded before the current row
  // is walked, so they arrive while its neighbours are tested.
  const uint64_t i0 = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  uint32_t r_next = 0, r_after = 0;
  uint64_t lo_next = 0, hi_next = 0;
  if (i0 < n) {
    r_next = rows[i0];
    lo_next = ofs[r_next];
    hi_next = ofs[r_next + 1];
  }
  if (i0 + stride < n) r_after = rows[i0 + stride];
  // rounds the same for every lane of a wave (wave_set_bits)
  for (uint64_t base = uint64_t(blockIdx.x) * kBlock; base < n; base += stride) {
    const uint64_t i = base + threadIdx.x;
    const bool valid = i < n;
    const uint32_t r = r_next;
    uint64_t e = lo_next;
    const uint64_t hi = hi_next;
    if (i + stride < n) {
      r_next = r_after;
      lo_next = ofs[r_next];
      hi_next = ofs[r_next + 1];
    }
    if (i + 2 * stride < n) r_after = rows[i + 2 * stride];
    bool found = false;
    if (valid) {
      while (e < hi && !found) {
        const uint32_t m = uint32_t(min(uint64_t(4), hi - e));
        uint32_t u[4];
        uint64_t fw[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) u[k] = adj[e + (k < m ? k : 0)];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) fw[k] = fbm[u[k] >> 6];
        uint32_t hit = 0;  // 1 + position of the first neighbour in the frontier
#pragma unroll
        for (int k = 3; k >= 0; k--)
          if (uint32_t(k) < m && ((fw[k] >> (u[k] & 63)) & 1)) hit = uint32_t(k) + 1;
        found = hit != 0;
        exam += found ? hit : m;
        e += m;
      }
      if (found) {
        if (lv) lv[r] = next_lv;
        n_new++;
        m_new += outdeg[r];
      }
    }
    wave_set_bits(nbm, found, r, valid);
  }
  block_counters(n_new, m_new, exam, c);
}

// Rows of more than kLongRow in-edges: kSubLanes lanes per row, 64 /
// kSubLanes rows of a wave walked side by side (each sub-group kSubLanes
// neighbours per step), so a wave's rows are not one chain of dependent loads
// after another; the wave steps until every one of its rows has found a
// parent or run out of in-edges. Widths 4 to 32 give identical counts run
// after run; the whole-wave form (64) gave run-to-run different counts in
// measurement and its cause is not found, so it is not allowed.
constexpr int kSubLanes = 8;
static_assert(kSubLanes >= 4 && kSubLanes <= 32 && (kSubLanes & (kSubLanes - 1)) == 0,
              "sub-group width must be a power of two in [4, 32]");

__global__ __launch_bounds__(kBlock) void k_bu_long(const uint32_t* rows, const uint64_t* ofs,
                                                    const uint32_t* adj, const uint32_t* outdeg,
                                                    const uint64_t* fbm, uint64_t* nbm,
                                                    int32_t* lv, int32_t next_lv, Counters* c) {
  constexpr int kPerWave = 64 / kSubLanes;
  constexpr uint64_t kSubMask = (uint64_t(1) << kSubLanes) - 1;
  const uint64_t n = c->n_long;  // stripe 0
  const int lane = threadIdx.x & 63;
  const int sub = lane / kSubLanes, sl = lane % kSubLanes;
  const uint64_t nw = uint64_t(gridDim.x) * (kBlock / 64);
  const uint64_t wid = uint64_t(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6);
  unsigned long long n_new = 0, m_new = 0, exam = 0;
  // rounds the same for every lane of a wave
  for (uint64_t i0 = wid * kPerWave; i0 < n; i0 += nw * kPerWave) {
    const uint64_t i = i0 + uint64_t(sub);
    uint32_t r = 0;
    uint64_t e0 = 0, hi = 0;
    if (i < n) {
      r = rows[i];
      e0 = ofs[r];
      hi = ofs[r + 1];
    }
    bool live = e0 < hi, found = false;
    unsigned long long ex = 0;  // the same in every lane of the sub-group
    while (__ballot(live)) {
      bool f = false;
      if (live) {
        const uint64_t e = e0 + uint64_t(sl);
        if (e < hi) {
          const uint32_t u = adj[e];
          f = (fbm[u >> 6] >> (u & 63)) & 1;
        }
      }
      const uint64_t mine = (__ballot(f) >> (sub * kSubLanes)) & kSubMask;
      if (live) {
        if (mine) {
          found = true;
          ex += uint64_t(__ffsll((long long)mine));
          live = false;
        } else {
          ex += min(uint64_t(kSubLanes), hi - e0);
          e0 += kSubLanes;
          live = e0 < hi;
        }
      }
    }
    if (sl == 0) {
      exam += ex;
      if (found) {
        found_row(r, lv, next_lv, nbm);
        n_new++;
        m_new += outdeg[r];
      }
    }
  }
  block_counters(n_new, m_new, exam, c);
}

__global__ void k_visited_or(uint64_t* visited, const uint64_t* nbm, uint64_t words) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < words) visited[i] |= nbm[i];
}

// Highest (out + in) degree, smallest ID among equals: max of deg << 32 | ~v.
__global__ void k_hub(const uint32_t* dout, const uint32_t* din, uint64_t V,
                      unsigned long long* best) {
  unsigned long long m = 0;
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
       v += uint64_t(gridDim.x) * blockDim.x) {
    const unsigned long long k = (uint64_t(dout[v]) + din[v]) << 32 | (0xffffffffull - v);
    m = k > m ? k : m;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long y = __shfl_xor(m, o, 64);
    m = y > m ? y : m;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(best, m);
}

// Complement of the reach set from the searched `visited` bitmap (which also
// holds the excluded dead rows and the padding bits beyond V): notr = rows
// outside R, padding included.
__global__ void k_reach_complement(const uint64_t* visited, const uint64_t* dead, uint64_t V,
                                   uint64_t words, uint64_t* notr) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= words) return;
  uint64_t pad = 0;
  const uint64_t lo = i * 64;
  if (lo + 64 > V) pad = V > lo ? ~((uint64_t(1) << (V - lo)) - 1) : ~uint64_t(0);
  notr[i] = ~visited[i] | dead[i] | pad;
}

// Out-edges of the vertices in R (grid-stride, one atomic per workgroup).
__global__ void k_reach_edges(const uint64_t* notr, const uint32_t* dout, uint64_t V,
                              unsigned long long* out) {
  unsigned long long n = 0;
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
       v += uint64_t(gridDim.x) * blockDim.x)
    if (!((notr[v >> 6] >> (v & 63)) & 1)) n += dout[v];
  block_atomic_add(out, n);
}

// Frontier bitmap -> queue of its vertices with out-edges.  One thread per
// 64-bit bitmap word, a workgroup scan of the per-word counts and one queue
// allocation (atomic) per workgroup of 256 words.
__global__ __launch_bounds__(kBlock) void k_bitmap_to_queue(const uint64_t* bm, uint64_t V,
                                                            const uint32_t* outdeg, uint32_t* q,
                                                            Counters* c) {
  __shared__ uint32_t wsum[kBlock / 64];
  __shared__ unsigned long long base;
  const uint64_t words = (V + 63) / 64;
  const uint64_t w = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
  uint64_t bits = w < words ? bm[w] : 0ull;
  if (w == words - 1 && (V & 63)) bits &= (uint64_t(1) << (V & 63)) - 1;  // padding bits
  uint64_t keep = 0;  // frontier vertices of this word with out-edges
  for (uint64_t b = bits; b; b &= b - 1) {
    const int i = __ffsll((long long)b) - 1;
    if (outdeg[w * 64 + i] > 0) keep |= uint64_t(1) << i;
  }
  const uint32_t cnt = __popcll(keep);
  // inclusive scan over the wave, then wave offsets through LDS
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t x = cnt;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < kBlock / 64; i++) t += wsum[i];
    base = t ? atomicAdd(&c->n_next, (unsigned long long)t) : 0ull;
  }
  __syncthreads();
  unsigned long long off = base + x - cnt;
  for (int i = 0; i < wave; i++) off += wsum[i];
  for (uint64_t b = keep; b; b &= b - 1) q[off++] = uint32_t(w * 64 + (__ffsll((long long)b) - 1));
}

// Word w of the dead-row bitmap: bit i set when row 64w+i of the CSR has no
// entry; word w of the long-row bitmap (at out + words): bit i set when the
// row has more than kLongRow entries.  A wave covers 64 consecutive rows
// (blocks of 256 start at multiples of 64), so one ballot forms each word.
__global__ void k_dead_rows(const uint64_t* ofs, uint64_t V, uint64_t words, uint64_t* out) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t n = v < V ? ofs[v + 1] - ofs[v] : 0;
  const uint64_t m = __ballot(v < V && n == 0);
  const uint64_t l = __ballot(n > kLongRow);
  if ((threadIdx.x & 63) == 0 && v < V) {
    out[v >> 6] = m;
    out[words + (v >> 6)] = l;
  }
}

// Summed stripes (aux from stripe 0); with `extra`, one more device word is
// read back in the same round trip into *extra_out.
// The summed stripes of n consecutive counter sets, one read-back.
void read_counter_sets(ll_gpu_ctx* ctx, const Counters* c, int n, Counters* out) {
  static_assert(kBuBatch * kStripes * sizeof(Counters) <= kReadbackBytes,
                "counter sets fit the read-back buffer");
  const Counters* st = static_cast<const Counters*>(ctx->readback);
  LLG_HIP(hipMemcpyAsync(ctx->readback, c, size_t(n) * kStripes * sizeof(Counters),
                         hipMemcpyDeviceToHost, ctx->stream));
  LLG_HIP(hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < n; k++) {
    const Counters* sk = st + k * kStripes;
    Counters h = sk[0];
    for (int i = 1; i < kStripes; i++) {
      h.n_next += sk[i].n_next;
      h.m_next += sk[i].m_next;
      h.examined += sk[i].examined;
      h.found += sk[i].found;
    }
    out[k] = h;
  }
}

Counters read_counters(ll_gpu_ctx* ctx, const Counters* c, const uint64_t* extra = nullptr,
                       uint64_t* extra_out = nullptr) {
  static_assert(kStripes * sizeof(Counters) + 8 <= kReadbackBytes,
                "stripes fit the read-back buffer");
  const Counters* st = static_cast<const Counters*>(ctx->readback);
  uint64_t* ex = reinterpret_cast<uint64_t*>(static_cast<char*>(ctx->readback) +
                                             kStripes * sizeof(Counters));
  LLG_HIP(hipMemcpyAsync(ctx->readback, c, kStripes * sizeof(Counters), hipMemcpyDeviceToHost,
                         ctx->stream));
  if (extra) LLG_HIP(hipMemcpyAsync(ex, extra, 8, hipMemcpyDeviceToHost, ctx->stream));
  LLG_HIP(hipStreamSynchronize(ctx->stream));
  if (extra) *extra_out = *ex;
  Counters h = st[0];
  for (int i = 1; i < kStripes; i++) {
    h.n_next += st[i].n_next;
    h.m_next += st[i].m_next;
    h.examined += st[i].examined;
    h.found += st[i].found;
  }
  return h;
}

}  // namespace

const uint64_t* ensure_dead_rows(ll_gpu_graph* g, const Csr& in) {
  if (g->dead_level != in.level || !g->dead_in.p) {
    hipStream_t s = g->ctx->stream;
    const uint64_t V = g->V;
    const uint64_t words = (V + 63) / 64;
    g->dead_in.alloc(2 * words * 8, s);  // dead rows, then long rows
    k_dead_rows<<<grid_for(V, 256), 256, 0, s>>>(in.ofs.as<uint64_t>(), V, words,
                                                 g->dead_in.as<uint64_t>());
    LLG_HIP(hipGetLastError());
    g->dead_level = in.level;
  }
  return g->dead_in.as<uint64_t>();
}

namespace {

// Device memory of one search: bitmaps of ceil(V/64) words, queues of V
// entries, queue offsets of V+1, the striped counters and the scan's
// temporary memory.  `timed` records per-kernel times when profiling is on.
struct BfsWork {
  uint64_t* visited;
  uint64_t* fbm;
  uint64_t* nbm;
  uint32_t* q[2];
  uint64_t* qofs;
  Counters* cnt;
  void* scan_tmp;       // scan_tmp_bytes(V)
  uint32_t* rows_short; // open rows of a bottom-up level (V entries each)
  uint32_t* rows_long;
  bool timed;
  const uint64_t* longrows = nullptr;  // long rows of `rev`, when known
};

// Launch timer that records only for searches that may allocate.
struct MaybeTimed {
  std::optional<TimedLaunch> t;
  MaybeTimed(ll_gpu_ctx* ctx, const char* name, bool on) {
    if (on) t.emplace(ctx, name);
  }
};

// bfs_core: see ll_gpu_internal.h; this form runs in caller-provided memory.
// `notr` (may be null) is the complement of a reach set R (closed under
// successors) with `e_reach` out-edges inside R: when the root's successors
// all lie in R, rows outside R are excluded instead of `excluded`.
BfsResult bfs_core_ws(ll_gpu_graph* g, const Csr& fwd, const Csr& rev, const uint32_t* fdeg,
                      uint32_t root, const uint64_t* excluded, const uint64_t* notr,
                      uint64_t e_reach, const BfsWork& w, int32_t* lv) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V;
  const uint64_t words = (V + 63) / 64;
  uint64_t* fbm = w.fbm;  // frontier / next-frontier bitmaps, swapped per bottom-up level
  uint64_t* nbm = w.nbm;
  // set 0: the root and bottom-up levels; top-down levels alternate sets 1
  // and 0, each expansion zeroing the other
  k_zero_counters<<<2, kStripes, 0, s>>>(w.cnt);
  if (notr)
    k_reach_check<<<1, 256, 0, s>>>(root, notr, fwd.ofs.as<uint64_t>(), fwd.adj.as<uint32_t>(),
                                    w.cnt);
  else
    k_fill_u64<<<1, 1, 0, s>>>(reinterpret_cast<uint64_t*>(&w.cnt->aux), 1, 0);
  k_bfs_init<<<grid_for(std::max(V, words), 256), 256, 0, s>>>(lv, w.visited, excluded, notr,
                                                              w.cnt, V, words);
  k_bfs_root<<<1, 1, 0, s>>>(lv, w.visited, w.q[0], root, fdeg,
                             w.cnt);
  LLG_HIP(hipGetLastError());
  Counters h = read_counters(ctx, w.cnt);
  BfsResult r;
  const unsigned bu_grid = 8u * unsigned(ctx->num_cus);
  uint64_t nq = h.n_next;       // queue entries (vertices with out-edges)
  uint64_t n_f = 1;             // frontier vertices
  uint64_t m_f = h.m_next;      // frontier out-edges
  uint64_t m_visited = h.m_next;
  r.reached = 1;
  bool bottom_up = false;
  int td = 1;  // counter set of the next top-down level (zeroed)
  int cur = 0;
  int32_t lvl = 0;
  // out-edges a search can still reach: those inside R when R excludes the rest
  const uint64_t E = h.aux ? e_reach : fwd.E;
  const long a_opt = ctx->options.count("bfs_alpha") ? ctx->options["bfs_alpha"] : -1;
  const long b_opt = ctx->options.count("bfs_beta") ? ctx->options["bfs_beta"] : -1;
  const uint64_t alpha = a_opt > 0 ? uint64_t(a_opt) : 10;
  // beta 0: no return to top-down
  const uint64_t beta = b_opt > 0 ? uint64_t(b_opt) : (h.aux ? 0 : 24);
  while (n_f > 0 && m_f > 0) {
    const uint64_t m_u = E > m_visited ? E - m_visited : 0;
    if (!bottom_up && m_f > m_u / alpha) {
      bottom_up = true;
      k_fill_u64<<<grid_for(words, 256), 256, 0, s>>>(fbm, words, 0);
      k_queue_to_bitmap<<<grid_for(nq, 256), 256, 0, s>>>(w.q[cur], nq,
                                                          fbm);
    }
    if (bottom_up && beta == 0) {
      // no return to top-down: kBuBatch levels per read-back; the levels
      // after the one that finds nothing find nothing either
      k_zero_counters<<<kBuBatch, kStripes, 0, s>>>(w.cnt);  // the batch's counter sets
      for (int k = 0; k < kBuBatch; k++) {
        Counters* ck = w.cnt + k * kStripes;
        {
          MaybeTimed t(ctx, "k_bu_step", w.timed);
          // the rows the level before found join visited here (the frontier
          // bitmap); the last level's are never merged: no later step reads
          // visited in a search that stays bottom-up.  The next frontier's
          // bitmap is cleared in the same pass.
          k_open_rows<<<grid_for(words, kBlock * kOpenWords), kBlock, 0, s>>>(
              w.visited, fbm, nbm, words, rev.ofs.as<uint64_t>(), w.longrows, w.rows_short,
              w.rows_long, ck);
          k_bu_short<<<bu_grid, kBlock, 0, s>>>(w.rows_short, rev.ofs.as<uint64_t>(),
                                                rev.adj.as<uint32_t>(), fdeg, fbm, nbm, lv,
                                                lvl + 1 + k, ck);
          k_bu_long<<<bu_grid, kBlock, 0, s>>>(w.rows_long, rev.ofs.as<uint64_t>(),
                                               rev.adj.as<uint32_t>(), fdeg, fbm, nbm, lv,
                                               lvl + 1 + k, ck);
        }
        std::swap(fbm, nbm);
      }
      LLG_HIP(hipGetLastError());
      Counters hs[kBuBatch];
      read_counter_sets(ctx, w.cnt, kBuBatch, hs);
      bool ended = false;
      for (int k = 0; k < kBuBatch && !ended; k++) {
        r.examined += hs[k].examined;
        r.reached += hs[k].n_next;
        m_visited += hs[k].m_next;
        lvl++;
        n_f = hs[k].n_next;
        m_f = hs[k].m_next;
        ended = n_f == 0 || m_f == 0;
      }
      if (ended) break;
      continue;
    }
    Counters* const cl = bottom_up ? w.cnt : w.cnt + td * kStripes;  // this level's counters
    if (bottom_up) {
      k_zero_counters<<<1, kStripes, 0, s>>>(w.cnt);
      {
        MaybeTimed t(ctx, "k_bu_step", w.timed);
        k_open_rows<<<grid_for(words, kBlock * kOpenWords), kBlock, 0, s>>>(
            w.visited, nullptr, nbm, words, rev.ofs.as<uint64_t>(), w.longrows, w.rows_short,
            w.rows_long, w.cnt);
        k_bu_short<<<bu_grid, kBlock, 0, s>>>(w.rows_short, rev.ofs.as<uint64_t>(),
                                              rev.adj.as<uint32_t>(), fdeg, fbm, nbm, lv,
                                              lvl + 1, w.cnt);
        k_bu_long<<<bu_grid, kBlock, 0, s>>>(w.rows_long, rev.ofs.as<uint64_t>(),
                                             rev.adj.as<uint32_t>(), fdeg, fbm, nbm, lv,
                                             lvl + 1, w.cnt);
      }
      k_visited_or<<<grid_for(words, 256), 256, 0, s>>>(w.visited, nbm, words);
    } else {
      // queue offsets: the prefix of the queued vertices' out-degrees
      scan_gather_u32_to_u64_ws(w.q[cur], fdeg, w.qofs, nq, w.scan_tmp, s);
      MaybeTimed t(ctx, "k_td_expand", w.timed);
      k_td_expand<<<grid_for(m_f, kEdgesPerBlock), kBlock, 0, s>>>(
          w.q[cur], w.qofs, nq, m_f, fwd.ofs.as<uint64_t>(),
          fwd.adj.as<uint32_t>(), fdeg, w.visited, lv, lvl + 1,
          w.q[cur ^ 1], cl, w.cnt + (1 - td) * kStripes);
    }
    LLG_HIP(hipGetLastError());
    h = read_counters(ctx, cl);
    r.examined += h.examined;
    const uint64_t n_next = h.n_next;
    lvl++;
    m_visited += h.m_next;
    if (bottom_up) {
      r.reached += n_next;
      const bool shrinking = n_next < n_f;
      n_f = n_next;
      m_f = h.m_next;
      std::swap(fbm, nbm);
      if (n_f == 0) break;
      if (beta > 0 && shrinking && n_f < V / beta) {
        bottom_up = false;
        // set 0 counts the queue; set 1 serves the next top-down level
        k_zero_counters<<<2, kStripes, 0, s>>>(w.cnt);
        k_bitmap_to_queue<<<grid_for(words, kBlock), kBlock, 0, s>>>(
            fbm, V, fdeg, w.q[cur], w.cnt);
        LLG_HIP(hipGetLastError());
        nq = read_counters(ctx, w.cnt).n_next;
        td = 1;
      }
    } else {
      // top-down n_next counts the queued vertices (those with out-edges),
      // found every claimed vertex
      r.reached += h.found;
      nq = n_next;
      n_f = n_next;
      m_f = h.m_next;
      cur ^= 1;
      td ^= 1;
    }
  }
  r.levels = lvl;
  r.traversed = m_visited;
  return r;
}

}  // namespace

BfsResult bfs_core(ll_gpu_graph* g, const Csr& fwd, const Csr& rev, const uint32_t* fdeg,
                   uint32_t root, const uint64_t* excluded, DevBuf& visited, int32_t* lv) {
  hipStream_t s = g->ctx->stream;
  const uint64_t V = g->V;
  const uint64_t words = (V + 63) / 64;
  DevBuf fbm, nbm, q[2], qofs, cnt, stmp;
  visited.alloc(words * 8, s);
  fbm.alloc(words * 8, s);
  nbm.alloc(words * 8, s);
  q[0].alloc(V * 4, s);
  q[1].alloc(V * 4, s);
  qofs.alloc((V + 1) * 8, s);
  cnt.alloc(kBuBatch * kStripes * sizeof(Counters), s);
  const size_t sb = scan_tmp_bytes(V);
  if (sb) stmp.alloc(sb, s);
  DevBuf rs, rl;
  rs.alloc(std::max<uint64_t>(V, 1) * 4, s);
  rl.alloc(std::max<uint64_t>(V, 1) * 4, s);
  BfsWork w{visited.as<uint64_t>(), fbm.as<uint64_t>(), nbm.as<uint64_t>(),
            {q[0].as<uint32_t>(), q[1].as<uint32_t>()}, qofs.as<uint64_t>(),
            cnt.as<Counters>(), stmp.p, rs.as<uint32_t>(), rl.as<uint32_t>(), true};
  return bfs_core_ws(g, fwd, rev, fdeg, root, excluded, nullptr, 0, w, lv);
}

// Working memory of one single-source search, carved into 256-byte aligned
// regions: three bitmaps, two queues, the queue offsets, the striped
// counters and the scan's temporary memory, and the two lists of open rows.
namespace {
struct WsLayout {
  size_t visited, fbm, nbm, q0, q1, qofs, cnt, stmp, rs, rl, total;
};
WsLayout ws_layout(uint64_t V) {
  const uint64_t words = (V + 63) / 64;
  auto up = [](size_t x) { return (x + 255) & ~size_t(255); };
  WsLayout L{};
  size_t o = 0;
  auto take = [&](size_t bytes) { const size_t at = o; o += up(bytes); return at; };
  L.visited = take(words * 8);
  L.fbm = take(words * 8);
  L.nbm = take(words * 8);
  L.q0 = take(V * 4);
  L.q1 = take(V * 4);
  L.qofs = take((V + 1) * 8);
  L.cnt = take(kBuBatch * kStripes * sizeof(Counters));
  L.stmp = take(scan_tmp_bytes(V));
  L.rs = take(std::max<uint64_t>(V, 1) * 4);
  L.rl = take(std::max<uint64_t>(V, 1) * 4);
  L.total = o;
  return L;
}
}  // namespace

size_t bfs_work_bytes(uint64_t V) { return ws_layout(V).total; }

namespace {

// The reach set of the snapshot: R = the vertices reachable from its hub, the
// vertex of highest out + in degree (smallest ID among equals).  R is closed
// under successors, so a search from any vertex of R -- or from a vertex all
// of whose out-neighbours are in R -- never leaves R, and its bottom-up steps
// can treat rows outside R as visited from the start (exactly as the dead
// rows, which R never holds).  Built with one search per snapshot level.
void build_reach(ll_gpu_graph* g, int level, const BfsPrepared& p) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V, words = (V + 63) / 64;
  LLG_CHECK(V > 0, LL_GPU_EINVAL, "a graph without vertices has no search");
  const uint32_t* din = ensure_degrees(g, LL_GPU_DIR_IN, level);
  DevBuf best;
  best.alloc(16, s);
  k_fill_u64<<<1, 2, 0, s>>>(best.as<uint64_t>(), 2, 0);
  const unsigned gr = std::min(grid_for(V, 256), 4u * unsigned(ctx->num_cus));
  k_hub<<<gr, 256, 0, s>>>(p.outdeg, din, V, best.as<unsigned long long>());
  LLG_HIP(hipGetLastError());
  LLG_HIP(hipMemcpyAsync(ctx->readback, best.p, 8, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
  const uint64_t key = *static_cast<const uint64_t*>(ctx->readback);
  const uint32_t hub = 0xffffffffu - uint32_t(key & 0xffffffffu);
  DevBuf visited;
  (void)bfs_core(g, *p.out, *p.in, p.outdeg, hub, p.dead, visited, nullptr);
  g->reach_notr.alloc(words * 8, s);
  k_reach_complement<<<grid_for(words, 256), 256, 0, s>>>(visited.as<uint64_t>(), p.dead, V,
                                                          words, g->reach_notr.as<uint64_t>());
  k_reach_edges<<<gr, 256, 0, s>>>(g->reach_notr.as<uint64_t>(), p.outdeg, V,
                                   best.as<unsigned long long>() + 1);
  LLG_HIP(hipGetLastError());
  LLG_HIP(hipMemcpyAsync(ctx->readback, best.as<uint64_t>() + 1, 8, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
  g->reach_e = *static_cast<const uint64_t*>(ctx->readback);
  g->reach_hub = hub;
  g->reach_level = level;
}

}  // namespace

bool bfs_peek(ll_gpu_graph* g, int level, BfsPrepared* p) {
  level = resolve_level(g, level);
  const Csr* c[2];
  for (uint32_t d = 0; d < 2; d++) {
    c[d] = g->slcsr ? &g->sl[d] : &g->csr[d];
    if (!c[d]->valid || (!g->slcsr && c[d]->level != level)) return false;
  }
  if (g->deg_level[LL_GPU_DIR_OUT] != level || !g->deg[LL_GPU_DIR_OUT].p) return false;
  if (g->dead_level != c[LL_GPU_DIR_IN]->level || !g->dead_in.p) return false;
  if (g->reach_level != level || !g->reach_notr.p) return false;
  *p = BfsPrepared{c[LL_GPU_DIR_OUT], c[LL_GPU_DIR_IN], g->deg[LL_GPU_DIR_OUT].as<uint32_t>(),
                   g->dead_in.as<uint64_t>(), g->reach_notr.as<uint64_t>(), g->reach_e,
                   g->dead_in.as<uint64_t>() + (g->V + 63) / 64};
  return true;
}

BfsPrepared bfs_prepare(ll_gpu_graph* g, int level) {
  level = resolve_level(g, level);
  BfsPrepared p;
  p.out = &ensure_csr(g, LL_GPU_DIR_OUT, level);
  p.in = &ensure_csr(g, LL_GPU_DIR_IN, level);
  p.outdeg = ensure_degrees(g, LL_GPU_DIR_OUT, level);
  p.dead = ensure_dead_rows(g, *p.in);
  p.longrows = p.dead + (g->V + 63) / 64;
  if (g->reach_level != level || !g->reach_notr.p) {
    TimedLaunch t(g->ctx, "bfs_reach_index");
    build_reach(g, level, p);
  }
  p.notr = g->reach_notr.as<uint64_t>();
  p.e_reach = g->reach_e;
  return p;
}

BfsResult bfs_single_ws(ll_gpu_graph* g, const BfsPrepared& p, uint32_t root, void* ws,
                        int32_t* lv, bool timed) {
  const WsLayout L = ws_layout(g->V);
  char* b = static_cast<char*>(ws);
  BfsWork w{reinterpret_cast<uint64_t*>(b + L.visited), reinterpret_cast<uint64_t*>(b + L.fbm),
            reinterpret_cast<uint64_t*>(b + L.nbm),
            {reinterpret_cast<uint32_t*>(b + L.q0), reinterpret_cast<uint32_t*>(b + L.q1)},
            reinterpret_cast<uint64_t*>(b + L.qofs),
            reinterpret_cast<Counters*>(b + L.cnt), b + L.stmp,
            reinterpret_cast<uint32_t*>(b + L.rs), reinterpret_cast<uint32_t*>(b + L.rl), timed,
            p.longrows};
  // option bfs_reach 0: dead rows only (for comparison)
  const auto& opts = g->ctx->options;
  const bool use_reach = !opts.count("bfs_reach") || opts.at("bfs_reach") != 0;
  return bfs_core_ws(g, *p.out, *p.in, p.outdeg, root, p.dead, use_reach ? p.notr : nullptr,
                     p.e_reach, w, lv);
}

// Single-source BFS; per-vertex levels stay in `lv` (device, int32[V]).
BfsResult bfs_device(ll_gpu_graph* g, int level, uint32_t root, DevBuf& lv) {
  hipStream_t s = g->ctx->stream;
  const BfsPrepared p = bfs_prepare(g, level);
  lv.alloc(g->V * 4, s);
  DevBuf ws;
  ws.alloc(bfs_work_bytes(g->V), s);
  return bfs_single_ws(g, p, root, ws.p, lv.as<int32_t>(), true);
}

}  // namespace llgpu

extern "C" {

int ll_gpu_bfs_ex(ll_gpu_graph* g, int level, int64_t root, int32_t* levels_out,
                  uint64_t* reached, uint64_t* edges_examined, uint64_t* edges_traversed) {
  if (!g) return LL_GPU_EINVAL;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    LLG_CHECK(root >= 0 && uint64_t(root) < g->V, LL_GPU_EINVAL, "root outside [0, max_nodes)");
    llgpu::DevBuf lv;
    llgpu::BfsResult r = llgpu::bfs_device(g, level, uint32_t(root), lv);
    if (levels_out) {
      LLG_HIP(hipMemcpyAsync(levels_out, lv.p, g->V * 4, hipMemcpyDeviceToHost, g->ctx->stream));
      LLG_HIP(hipStreamSynchronize(g->ctx->stream));
    }
    if (reached) *reached = r.reached;
    if (edges_examined) *edges_examined = r.examined;
    if (edges_traversed) *edges_traversed = r.traversed;
    return LL_GPU_OK;
  } catch (const llgpu::Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

int ll_gpu_bfs(ll_gpu_graph* g, int level, int64_t root, int32_t* levels_out, uint64_t* reached,
               uint64_t* edges_examined) {
  return ll_gpu_bfs_ex(g, level, root, levels_out, reached, edges_examined, nullptr);
}

}  // extern "C"
