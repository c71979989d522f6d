// Weighted single-source shortest paths (task sssp_weighted,
// benchmark/benchmarks/sssp.h:59-172) over the snapshot out-CSR and its
// per-edge FP32 weights (the levels' `weight` property, compacted with the
// snapshot).
//
// The reference runs synchronous Bellman-Ford rounds over the vertices
// updated in the previous round.  Here the rounds follow the near/far
// schedule (Davidson, Baxter, Garland, Owens, IPDPS 2014): a round relaxes
// the out-edges of a queue of vertices edge-parallel (LDS binary search over
// the queue's degree prefix, as the top-down BFS does) with atomicMin on the
// distances' bit patterns -- non-negative FP32 values order like their
// unsigned bit patterns.  A vertex whose distance dropped below the threshold
// T is queued for the next round once (round stamp); one whose new distance is
// >= T is flagged far.  When the queue runs dry, T grows by delta and the far
// vertices now below T whose distance changed since their edges were last
// relaxed form the next queue.  Every order of relaxations ends at the same
// fixed point, dist[v] = min over paths of the FP32 sums dist[u] + w taken in
// path order (each sum is monotone in dist[u] and never below it), so
// distances match the reference bit for bit; the schedule only cuts repeated
// relaxations.  Unreached vertices keep INT_MAX-1 converted to float, as the
// reference stores them.
#include <algorithm>
#include <climits>
#include <cstring>

#include "ll_gpu_internal.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kEdgesPerBlock = 2048;
constexpr float kUnreachedDist = float(INT_MAX - 1);
constexpr double kAutoDeltaMilli = 5;      // threshold step: 0.5% of the mean edge weight
constexpr uint64_t kAutoMinQueueDiv = 32;  // smallest tail phase queue: V / 32 vertices

__global__ void k_sssp_init(uint32_t* dist, int32_t* stamp, uint8_t* far, uint32_t* relaxed,
                            uint64_t V, uint32_t root, uint32_t* queue) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V) {
    dist[v] = v == root ? __float_as_uint(0.0f) : __float_as_uint(kUnreachedDist);
    stamp[v] = v == root ? 0 : -1;
    far[v] = 0;
    relaxed[v] = ~0u;  // a NaN pattern: equal to no distance
  }
  if (v == 0) queue[0] = root;  // a root without out-edges gives an empty first round
}

// Queue entry i: out-degree, and the distance its edges are relaxed with
// (recorded in relaxed[v]; a later drop queues v again).
__global__ void k_queue_prep(const uint32_t* q, uint64_t n, const uint64_t* ofs,
                             const uint32_t* dist, uint32_t* d, uint32_t* qd,
                             uint32_t* relaxed, unsigned long long* n_next) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i == 0) *n_next = 0;
  if (i < n) {
    const uint32_t v = q[i];
    d[i] = uint32_t(ofs[v + 1] - ofs[v]);
    const uint32_t x = dist[v];
    qd[i] = x;
    relaxed[v] = x;
  }
}

// First and last queue entry of each relaxation block of kEdgesPerBlock
// edges; the edge total m = qofs[nq] is read on the device (fixed grid).
__global__ void k_relax_blocks(const uint64_t* qofs, uint64_t nq, uint64_t* blk) {
  const uint64_t m = qofs[nq];
  const uint64_t nblk = (m + kEdgesPerBlock - 1) / kEdgesPerBlock;
  for (uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; b < nblk;
       b += uint64_t(gridDim.x) * blockDim.x) {
  const uint64_t x0 = b * kEdgesPerBlock, x1 = min(m, x0 + kEdgesPerBlock);
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
}

// Relax the queue's out-edges.  An improved vertex with out-edges is
// appended to the next queue once per round (stamp) when its distance is
// below the threshold (bits t_bits), else flagged far.  An edge block's
// appends are staged in LDS and reserved in the next queue with one atomic,
// not one per wave: the queue counter is a single address every workgroup
// adds to.
__global__ __launch_bounds__(kBlock) void k_relax(const uint32_t* q, const uint32_t* qd,
                                                  const uint64_t* qofs, uint64_t nq,
                                                  const uint64_t* blk, const uint64_t* ofs,
                                                  const uint32_t* adj, const float* w,
                                                  uint32_t* dist, int32_t* stamp, uint8_t* far,
                                                  uint32_t t_bits, int32_t round,
                                                  uint32_t* q_next, unsigned long long* n_next) {
  __shared__ uint64_t l_qofs[kEdgesPerBlock + 1];
  __shared__ uint64_t l_base[kEdgesPerBlock];
  __shared__ float l_dist[kEdgesPerBlock];
  __shared__ uint32_t l_q[kEdgesPerBlock];  // at most one append per edge
  __shared__ uint32_t l_qn;
  __shared__ unsigned long long l_qbase;
  const uint64_t m = qofs[nq];
  const uint64_t nblk = (m + kEdgesPerBlock - 1) / kEdgesPerBlock;
  const int lane = threadIdx.x & 63;
  if (blockIdx.x == 0 && threadIdx.x == 0) n_next[1] = m;  // the round's edges, read back with n_next[0]
  // equal-sized edge blocks: a fixed grid takes them in turn
  for (uint64_t bi = blockIdx.x; bi < nblk; bi += gridDim.x) {
    const uint64_t x0 = bi * kEdgesPerBlock;
    const uint64_t x1 = min(m, x0 + kEdgesPerBlock);
    const uint64_t lo = blk[2 * bi];
    const uint32_t n = uint32_t(blk[2 * bi + 1] - lo + 1);
    for (uint32_t i = threadIdx.x; i <= n; i += kBlock) {
      l_qofs[i] = qofs[lo + i];
      if (i < n) {
        l_dist[i] = __uint_as_float(qd[lo + i]);
        l_base[i] = ofs[q[lo + i]];
      }
    }
    if (threadIdx.x == 0) l_qn = 0;
    __syncthreads();
    for (uint64_t base = x0; base < x1; base += kBlock) {
      const uint64_t x = base + threadIdx.x;
      bool take = false;
      uint32_t v = 0;
      if (x < x1) {
        uint32_t a = 0, b = n;
        while (b - a > 1) {
          const uint32_t mid = (a + b) >> 1;
          if (l_qofs[mid] <= x) a = mid;
          else b = mid;
        }
        const uint64_t e = l_base[a] + (x - l_qofs[a]);
        v = adj[e];
        const float cand = l_dist[a] + w[e];
        const uint32_t cb = __float_as_uint(cand);
        if (cb < dist[v]) {
          const uint32_t old = atomicMin(&dist[v], cb);
          // only vertices with out-edges are queued: every queue entry then
          // owns at least one edge, so an edge block spans at most
          // kEdgesPerBlock entries (the LDS arrays above)
          if (cb < old && ofs[v + 1] > ofs[v]) {
            if (cb < t_bits) take = atomicExch(&stamp[v], round + 1) != round + 1;
            else far[v] = 1;
          }
        }
      }
      const uint64_t mk = __ballot(take);
      if (mk) {
        const int leader = __ffsll((long long)mk) - 1;
        uint32_t qb = 0;
        if (lane == leader) qb = atomicAdd(&l_qn, uint32_t(__popcll(mk)));
        qb = uint32_t(__shfl(int(qb), leader, 64));
        if (take) l_q[qb + uint32_t(__popcll(mk & ((uint64_t(1) << lane) - 1)))] = v;
      }
    }
    __syncthreads();
    if (threadIdx.x == 0 && l_qn) l_qbase = atomicAdd(n_next, (unsigned long long)l_qn);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < l_qn; k += kBlock) q_next[l_qbase + k] = l_q[k];
    __syncthreads();  // the LDS slice and queue are reloaded for the next block
  }
}

// The far pile.  far[v] = 1 marks a far vertex; a split bins every far
// vertex by distance into steps of delta above the current threshold `base`
// (bin 0 also holds anything below base + delta, the last bin everything past
// the others) and stores the bin in far[v] as 2 + bin, so the scatter takes
// exactly the vertices the histogram counted.  Far vertices whose edges were
// relaxed with their current distance are dropped.  The kernels run over
// fixed vertex ranges, one per workgroup, without shared counters.
constexpr unsigned kSplitBlocks = 2048;
constexpr unsigned kBins = 128;

__device__ inline uint32_t split_range(uint64_t V) {
  return uint32_t(((V + kSplitBlocks - 1) / kSplitBlocks + kBlock - 1) / kBlock * kBlock);
}

__global__ __launch_bounds__(kBlock) void k_far_hist(uint8_t* far, const uint32_t* dist,
                                                     const uint32_t* relaxed, uint64_t V,
                                                     float base, float delta, uint32_t* hist,
                                                     uint32_t* mins) {
  __shared__ uint32_t h[kBins];
  __shared__ uint32_t rm[kBlock / 64];
  for (uint32_t i = threadIdx.x; i < kBins; i += kBlock) h[i] = 0;
  __syncthreads();
  const uint32_t chunk = split_range(V);
  const uint64_t v0 = uint64_t(blockIdx.x) * chunk, v1 = min(V, v0 + chunk);
  uint32_t mn = ~0u;
  for (uint64_t v = v0 + threadIdx.x; v < v1; v += kBlock) {
    if (!far[v]) continue;
    const uint32_t d = dist[v];
    if (d == relaxed[v]) {
      far[v] = 0;
      continue;
    }
    const float x = (__uint_as_float(d) - base) / delta;
    const uint32_t bin = x >= 1.0f ? uint32_t(min(x, float(kBins - 1))) : 0u;
    far[v] = uint8_t(2 + bin);
    atomicAdd(&h[bin], 1u);
    mn = min(mn, d);
  }
  for (int o = 32; o > 0; o >>= 1) mn = min(mn, uint32_t(__shfl_xor(int(mn), o, 64)));
  if ((threadIdx.x & 63) == 0) rm[threadIdx.x >> 6] = mn;
  __syncthreads();
  // stored as inclusive prefix counts: hist[range][b] = far vertices in bins <= b
  if (threadIdx.x == 0) {
    uint32_t run = 0;
    for (uint32_t i = 0; i < kBins; i++) {
      run += h[i];
      h[i] = run;
    }
    for (int i = 1; i < kBlock / 64; i++) mn = min(mn, rm[i]);
    mins[blockIdx.x] = mn;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < kBins; i += kBlock) hist[uint64_t(blockIdx.x) * kBins + i] = h[i];
}

// Column sums of the per-range prefix counts, one workgroup per bin, and the
// smallest far distance (workgroup kBins): tot[b] = far vertices in bins <= b,
// tot[kBins] = the smallest distance's bits.
__global__ __launch_bounds__(kBlock) void k_hist_total(const uint32_t* hist, const uint32_t* mins,
                                                       unsigned long long* tot) {
  __shared__ unsigned long long rs[kBlock / 64];
  __shared__ uint32_t rm[kBlock / 64];
  const uint32_t bin = blockIdx.x;
  unsigned long long t = 0;
  uint32_t mn = ~0u;
  for (uint32_t r = threadIdx.x; r < kSplitBlocks; r += kBlock) {
    if (bin < kBins) t += hist[uint64_t(r) * kBins + bin];
    else mn = min(mn, mins[r]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    t += __shfl_xor(t, o, 64);
    mn = min(mn, uint32_t(__shfl_xor(int(mn), o, 64)));
  }
  if ((threadIdx.x & 63) == 0) {
    rs[threadIdx.x >> 6] = t;
    rm[threadIdx.x >> 6] = mn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kBlock / 64; i++) {
      t += rs[i];
      mn = min(mn, rm[i]);
    }
    tot[bin] = bin < kBins ? t : (unsigned long long)mn;
  }
}

// Offsets of each range's taken vertices (bins 0..k), exclusive scan over the
// ranges in one workgroup, kBlock ranges at a time.
__global__ __launch_bounds__(kBlock) void k_far_offsets(const uint32_t* hist, uint32_t k,
                                                        uint32_t* offs) {
  __shared__ uint32_t part[kBlock];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  for (uint32_t c0 = 0; c0 < kSplitBlocks; c0 += kBlock) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t t = hist[uint64_t(i) * kBins + k];
    part[threadIdx.x] = t;
    __syncthreads();
    for (uint32_t o = 1; o < kBlock; o <<= 1) {  // inclusive Hillis-Steele scan
      const uint32_t x = threadIdx.x >= o ? part[threadIdx.x - o] : 0u;
      __syncthreads();
      part[threadIdx.x] += x;
      __syncthreads();
    }
    offs[i] = carry + part[threadIdx.x] - t;
    __syncthreads();
    if (threadIdx.x == kBlock - 1) carry += part[kBlock - 1];
    __syncthreads();
  }
}

// Take the far vertices of bins 0..k into the queue, in vertex order.
__global__ __launch_bounds__(kBlock) void k_far_scatter(uint8_t* far, uint64_t V, uint32_t k,
                                                        const uint32_t* offs, uint32_t* q) {
  const uint32_t chunk = split_range(V);
  const uint64_t v0 = uint64_t(blockIdx.x) * chunk, v1 = min(V, v0 + chunk);
  __shared__ uint32_t wcnt[kBlock / 64];
  uint32_t pos = offs[blockIdx.x];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint64_t b = v0; b < v1; b += kBlock) {
    const uint64_t v = b + threadIdx.x;
    const uint32_t f = v < v1 ? far[v] : 0u;
    const bool take = f >= 2 && f - 2 <= k;
    if (take) far[v] = 0;
    const uint64_t mk = __ballot(take);
    if (lane == 0) wcnt[wave] = uint32_t(__popcll(mk));
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int i = 0; i < kBlock / 64; i++) {
      before += i < wave ? wcnt[i] : 0u;
      total += wcnt[i];
    }
    if (take) q[pos + before + __popcll(mk & ((uint64_t(1) << lane) - 1))] = uint32_t(v);
    pos += total;
    __syncthreads();
  }
}

// A queue too small for a round of its own rejoins the far pile.
__global__ void k_queue_to_far(const uint32_t* q, uint64_t n, uint8_t* far) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) far[q[i]] = 1;
}

// Sum of the edge weights (for the threshold step), FP64, fixed grid.
__global__ void k_weight_sum(const float* w, uint64_t n, double* out) {
  double t = 0;
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    t += w[i];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_down(t, o, 64);
  if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(out, t);
}

__global__ void k_zero_f64(double* p) { *p = 0; }

// Reached vertices and the largest finite distance (fixed grid).
__global__ void k_sssp_summary(const uint32_t* dist, uint64_t V, unsigned long long* reached,
                               uint32_t* max_bits) {
  unsigned long long n = 0;
  uint32_t mx = 0;
  const uint32_t unreached = __float_as_uint(kUnreachedDist);
  for (uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; v < V;
       v += uint64_t(gridDim.x) * blockDim.x) {
    const uint32_t d = dist[v];
    if (d < unreached) {
      n++;
      mx = max(mx, d);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    n += __shfl_down(n, o, 64);
    mx = max(mx, __shfl_down(mx, o, 64));
  }
  __shared__ unsigned long long pn[4];
  __shared__ uint32_t pm[4];
  if ((threadIdx.x & 63) == 0) {
    pn[threadIdx.x >> 6] = n;
    pm[threadIdx.x >> 6] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < int(blockDim.x / 64); i++) {
      n += pn[i];
      mx = max(mx, pm[i]);
    }
    if (n) atomicAdd(reached, n);
    atomicMax(max_bits, mx);
  }
}

__global__ void k_sssp_summary_init(unsigned long long* reached, uint32_t* max_bits) {
  *reached = 0;
  *max_bits = 0;
}

}  // namespace

}  // namespace llgpu

extern "C" {

int ll_gpu_sssp(ll_gpu_graph* g, int level, int64_t root, float* dist_out, uint64_t* reached,
                double* max_dist, uint64_t* rounds) {
  if (!g) return LL_GPU_EINVAL;
  using namespace llgpu;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    ll_gpu_ctx* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    const uint64_t V = g->V;
    LLG_CHECK(root >= 0 && uint64_t(root) < V, LL_GPU_EINVAL, "root outside [0, max_nodes)");
    const Csr& out = ensure_csr(g, LL_GPU_DIR_OUT, level);
    LLG_CHECK(out.w.p != nullptr, LL_GPU_ESTATE,
              "no edge weights for this snapshot (ll_gpu_graph_push_weights for every level)");
    // near/far threshold step in thousandths of the mean edge weight (0 = no
    // far pile: every improved vertex is queued)
    const long dm = ctx->options.count("sssp_delta_milli") ? ctx->options["sssp_delta_milli"] : -1;
    const double delta_milli = dm < 0 ? kAutoDeltaMilli : double(dm);
    const auto* rb = static_cast<const unsigned long long*>(ctx->readback);
    if (delta_milli > 0 && out.w_mean < 0) {
      DevBuf sum;
      sum.alloc(8, s);
      k_zero_f64<<<1, 1, 0, s>>>(sum.as<double>());
      k_weight_sum<<<std::min(grid_for(out.E, 256), 4u * unsigned(ctx->num_cus)), 256, 0, s>>>(
          out.w.as<float>(), out.E, sum.as<double>());
      LLG_HIP(hipMemcpyAsync(ctx->readback, sum.p, 8, hipMemcpyDeviceToHost, s));
      LLG_HIP(hipStreamSynchronize(s));
      double tot;
      std::memcpy(&tot, ctx->readback, 8);
      out.w_mean = out.E ? tot / double(out.E) : 0.0;
    }
    const double delta = delta_milli > 0 ? out.w_mean * delta_milli / 1000.0 : 0.0;
    const bool near_far = delta > 0;
    auto bits_of = [](double x) {
      const float f = float(std::min(x, double(kUnreachedDist)));
      uint32_t b;
      std::memcpy(&b, &f, 4);
      return b;
    };
    // a phase starts with at least min_queue vertices when the far pile has
    // them (smaller queues rejoin it); 0 = no merging
    const long mq_opt = ctx->options.count("sssp_min_queue") ? ctx->options["sssp_min_queue"] : -1;
    const uint64_t min_queue = mq_opt < 0 ? std::max<uint64_t>(V / kAutoMinQueueDiv, 1)
                                          : uint64_t(mq_opt);
    DevBuf dist, stamp, far, relaxed, q[2], qdeg, qd, qofs, cnt, blk, hist, tot;
    dist.alloc(V * 4, s);
    stamp.alloc(V * 4, s);
    far.alloc(V, s);
    relaxed.alloc(V * 4, s);
    q[0].alloc(V * 4, s);
    q[1].alloc(V * 4, s);
    qdeg.alloc(V * 4, s);
    qd.alloc(V * 4, s);
    qofs.alloc((V + 1) * 8, s);
    cnt.alloc(16, s);
    hist.alloc((uint64_t(kSplitBlocks) * kBins + 2 * kSplitBlocks) * 4, s);
    tot.alloc((kBins + 1) * 8, s);
    uint32_t* h_bins = hist.as<uint32_t>();
    uint32_t* h_mins = h_bins + uint64_t(kSplitBlocks) * kBins;
    uint32_t* h_offs = h_mins + kSplitBlocks;
    // a round queues each vertex once, so its edges are at most E
    blk.alloc((out.E / kEdgesPerBlock + 1) * 16, s);
    k_sssp_init<<<grid_for(V, 256), 256, 0, s>>>(dist.as<uint32_t>(), stamp.as<int32_t>(),
                                                 far.as<uint8_t>(), relaxed.as<uint32_t>(), V,
                                                 uint32_t(root), q[0].as<uint32_t>());
    LLG_HIP(hipGetLastError());
    const unsigned relax_grid = 3u * unsigned(ctx->num_cus);  // 48 KB LDS per workgroup
    uint64_t nq = 1, nrounds = 0;
    int cur = 0;
    double T = near_far ? delta : double(kUnreachedDist);
    uint32_t t_bits = near_far ? bits_of(T) : ~0u;
    bool force = true;  // the root's queue, and a queue just taken from the far pile
    uint64_t work_done = 0;  // edges relaxed so far
    for (;;) {
      // small queues rejoin the far pile only in the tail, once edges worth
      // half the snapshot have been relaxed: early phases keep the step delta
      const bool tail = work_done * 2 >= out.E;
      if (nq > 0 && (!near_far || force || !tail || nq >= min_queue)) {
        // one round; one host round trip: the edge total stays on the device
        k_queue_prep<<<grid_for(nq, 256), 256, 0, s>>>(
            q[cur].as<uint32_t>(), nq, out.ofs.as<uint64_t>(), dist.as<uint32_t>(),
            qdeg.as<uint32_t>(), qd.as<uint32_t>(), relaxed.as<uint32_t>(),
            cnt.as<unsigned long long>());
        scan_u32_to_u64(qdeg.as<uint32_t>(), qofs.as<uint64_t>(), nq, s);
        k_relax_blocks<<<256, 256, 0, s>>>(qofs.as<uint64_t>(), nq, blk.as<uint64_t>());
        {
          TimedLaunch t(ctx, "k_relax");
          k_relax<<<relax_grid, kBlock, 0, s>>>(
              q[cur].as<uint32_t>(), qd.as<uint32_t>(), qofs.as<uint64_t>(), nq,
              blk.as<uint64_t>(), out.ofs.as<uint64_t>(), out.adj.as<uint32_t>(),
              out.w.as<float>(), dist.as<uint32_t>(), stamp.as<int32_t>(), far.as<uint8_t>(),
              t_bits, int32_t(nrounds), q[cur ^ 1].as<uint32_t>(), cnt.as<unsigned long long>());
        }
        LLG_HIP(hipGetLastError());
        LLG_HIP(hipMemcpyAsync(ctx->readback, cnt.p, 16, hipMemcpyDeviceToHost, s));
        LLG_HIP(hipStreamSynchronize(s));
        nq = rb[0];
        work_done += rb[1];
        nrounds++;
        cur ^= 1;
        force = false;
        continue;
      }
      if (!near_far) break;
      if (nq > 0)
        k_queue_to_far<<<grid_for(nq, 256), 256, 0, s>>>(q[cur].as<uint32_t>(), nq,
                                                         far.as<uint8_t>());
      // split: histogram of the far pile in steps of delta above T
      uint64_t c[kBins + 1];
      for (;;) {
        k_far_hist<<<kSplitBlocks, kBlock, 0, s>>>(far.as<uint8_t>(), dist.as<uint32_t>(),
                                                   relaxed.as<uint32_t>(), V, float(T),
                                                   float(delta), h_bins, h_mins);
        k_hist_total<<<kBins + 1, kBlock, 0, s>>>(h_bins, h_mins, tot.as<unsigned long long>());
        LLG_HIP(hipGetLastError());
        LLG_HIP(hipMemcpyAsync(ctx->readback, tot.p, sizeof(c), hipMemcpyDeviceToHost, s));
        LLG_HIP(hipStreamSynchronize(s));
        std::memcpy(c, ctx->readback, sizeof(c));
        for (unsigned i = kBins - 1; i > 0; i--) c[i] -= c[i - 1];  // prefix -> per bin
        uint64_t below_last = 0;
        for (unsigned i = 0; i + 1 < kBins; i++) below_last += c[i];
        if (below_last > 0 || c[kBins - 1] == 0) break;
        // everything lies past the histogram: restart it at the smallest distance
        float mf;
        const uint32_t mb = uint32_t(c[kBins]);
        std::memcpy(&mf, &mb, 4);
        T = std::max(T, double(mf));
      }
      // from the first non-empty bin on, enough bins to hold min_queue
      // vertices (in the tail only), never past the last non-empty one
      int first = -1, last = -1;
      for (unsigned i = 0; i < kBins; i++)
        if (c[i]) {
          if (first < 0) first = int(i);
          last = int(i);
        }
      if (first < 0) break;  // the far pile is empty: done
      const uint64_t want = work_done * 2 >= out.E ? min_queue : 0;
      uint32_t k = uint32_t(first);
      uint64_t take = c[k];
      while (int(k) < last && take < want) take += c[++k];
      k_far_offsets<<<1, kBlock, 0, s>>>(h_bins, k, h_offs);
      k_far_scatter<<<kSplitBlocks, kBlock, 0, s>>>(far.as<uint8_t>(), V, k, h_offs,
                                                    q[cur].as<uint32_t>());
      LLG_HIP(hipGetLastError());
      T = k + 1 < kBins ? T + double(k + 1) * delta : double(kUnreachedDist);
      t_bits = bits_of(T);
      nq = take;
      force = true;
    }
    k_sssp_summary_init<<<1, 1, 0, s>>>(cnt.as<unsigned long long>(),
                                        reinterpret_cast<uint32_t*>(cnt.as<unsigned long long>() + 1));
    k_sssp_summary<<<std::min(grid_for(V, 256), 4u * unsigned(ctx->num_cus)), 256, 0, s>>>(
        dist.as<uint32_t>(), V, cnt.as<unsigned long long>(),
        reinterpret_cast<uint32_t*>(cnt.as<unsigned long long>() + 1));
    LLG_HIP(hipGetLastError());
    if (dist_out) LLG_HIP(hipMemcpyAsync(dist_out, dist.p, V * 4, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipMemcpyAsync(ctx->readback, cnt.p, 16, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    if (reached) *reached = rb[0];
    if (max_dist) {
      const uint32_t bits = uint32_t(rb[1]);
      float f;
      std::memcpy(&f, &bits, 4);
      *max_dist = double(f);
    }
    if (rounds) *rounds = nrounds;
    return LL_GPU_OK;
  } catch (const Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

}  // extern "C"
