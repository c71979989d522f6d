// Weighted single-source shortest paths (task sssp_weighted,
// benchmark/benchmarks/sssp.h:59-172) over the snapshot out-CSR and its
// per-edge FP32 weights (the levels' `weight` property, compacted with the
// snapshot).
//
// The reference runs synchronous Bellman-Ford rounds over the vertices
// updated in the previous round.  Here each round is a frontier queue: the
// out-edges of the queued vertices are relaxed edge-parallel (LDS binary
// search over the queue's degree prefix, as the top-down BFS does) with
// atomicMin on the distances' bit patterns -- non-negative FP32 values order
// like their unsigned bit patterns -- and a vertex whose distance dropped is
// queued for the next round once (round stamp).  Relaxations may see
// distances lowered earlier in the same round; the result is the same fixed
// point, dist[v] = min over paths of the FP32 sums dist[u] + w taken in path
// order, so distances match the reference bit for bit.  Unreached vertices
// keep INT_MAX-1 converted to float, as the reference stores them.
#include <climits>
#include <cstring>

#include "ll_gpu_internal.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kEdgesPerBlock = 2048;
constexpr float kUnreachedDist = float(INT_MAX - 1);

__global__ void k_sssp_init(uint32_t* dist, int32_t* stamp, uint64_t V, uint32_t root,
                            uint32_t* queue) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V) {
    dist[v] = v == root ? __float_as_uint(0.0f) : __float_as_uint(kUnreachedDist);
    stamp[v] = v == root ? 0 : -1;
  }
  if (v == 0) queue[0] = root;  // a root without out-edges gives an empty first round
}

__global__ void k_queue_deg(const uint32_t* q, uint64_t n, const uint64_t* ofs, uint32_t* d) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) d[i] = uint32_t(ofs[q[i] + 1] - ofs[q[i]]);
}

// First and last queue entry of each relaxation workgroup.
__global__ void k_relax_blocks(const uint64_t* qofs, uint64_t nq, uint64_t m, uint64_t nblk,
                               uint64_t* blk) {
  const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b >= nblk) return;
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

// Relax the queue's out-edges; newly improved vertices are appended to the
// next queue once per round (stamp) with one atomic per wave.
__global__ __launch_bounds__(kBlock) void k_relax(const uint32_t* q, const uint64_t* qofs,
                                                  const uint64_t* blk, uint64_t m,
                                                  const uint64_t* ofs, const uint32_t* adj,
                                                  const float* w, uint32_t* dist, int32_t* stamp,
                                                  int32_t round, uint32_t* q_next,
                                                  unsigned long long* n_next) {
  __shared__ uint64_t l_qofs[kEdgesPerBlock + 1];
  __shared__ uint64_t l_base[kEdgesPerBlock];
  __shared__ uint32_t l_src[kEdgesPerBlock];
  const uint64_t x0 = uint64_t(blockIdx.x) * kEdgesPerBlock;
  const uint64_t x1 = min(m, x0 + kEdgesPerBlock);
  const uint64_t lo = blk[2 * blockIdx.x];
  const uint32_t n = uint32_t(blk[2 * blockIdx.x + 1] - lo + 1);
  for (uint32_t i = threadIdx.x; i <= n; i += kBlock) {
    l_qofs[i] = qofs[lo + i];
    if (i < n) {
      const uint32_t u = q[lo + i];
      l_src[i] = u;
      l_base[i] = ofs[u];
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
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
      const float cand = __uint_as_float(dist[l_src[a]]) + w[e];
      const uint32_t cb = __float_as_uint(cand);
      if (cb < dist[v]) {
        const uint32_t old = atomicMin(&dist[v], cb);
        // only vertices with out-edges are queued: every queue entry then owns
        // at least one edge, so a workgroup's edge range spans at most
        // kEdgesPerBlock entries (the LDS arrays above)
        if (cb < old && ofs[v + 1] > ofs[v])
          take = atomicExch(&stamp[v], round + 1) != round + 1;
      }
    }
    const uint64_t mk = __ballot(take);
    if (mk) {
      const int leader = __ffsll((long long)mk) - 1;
      unsigned long long qb = 0;
      if (lane == leader) qb = atomicAdd(n_next, (unsigned long long)__popcll(mk));
      qb = __shfl(qb, leader, 64);
      if (take) q_next[qb + __popcll(mk & ((uint64_t(1) << lane) - 1))] = v;
    }
  }
}

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

__global__ void k_zero_ull(unsigned long long* p) { *p = 0; }
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
    DevBuf dist, stamp, q[2], qdeg, qofs, cnt, blk;
    dist.alloc(V * 4, s);
    stamp.alloc(V * 4, s);
    q[0].alloc(V * 4, s);
    q[1].alloc(V * 4, s);
    qdeg.alloc(V * 4, s);
    qofs.alloc((V + 1) * 8, s);
    cnt.alloc(16, s);
    k_sssp_init<<<grid_for(V, 256), 256, 0, s>>>(dist.as<uint32_t>(), stamp.as<int32_t>(), V,
                                                 uint32_t(root), q[0].as<uint32_t>());
    LLG_HIP(hipGetLastError());
    uint64_t nq = 1, nrounds = 0;
    int cur = 0;
    const auto* rb = static_cast<const unsigned long long*>(ctx->readback);
    while (nq > 0) {
      k_queue_deg<<<grid_for(nq, 256), 256, 0, s>>>(q[cur].as<uint32_t>(), nq,
                                                    out.ofs.as<uint64_t>(), qdeg.as<uint32_t>());
      scan_u32_to_u64(qdeg.as<uint32_t>(), qofs.as<uint64_t>(), nq, s);
      k_zero_ull<<<1, 1, 0, s>>>(cnt.as<unsigned long long>());
      LLG_HIP(hipMemcpyAsync(ctx->readback, qofs.as<uint64_t>() + nq, 8, hipMemcpyDeviceToHost, s));
      LLG_HIP(hipStreamSynchronize(s));
      const uint64_t m = rb[0];
      if (m > 0) {
        const uint64_t nblk = (m + kEdgesPerBlock - 1) / kEdgesPerBlock;
        blk.alloc(nblk * 16, s);
        k_relax_blocks<<<grid_for(nblk, 256), 256, 0, s>>>(qofs.as<uint64_t>(), nq, m, nblk,
                                                           blk.as<uint64_t>());
        TimedLaunch t(ctx, "k_relax");
        k_relax<<<unsigned(nblk), kBlock, 0, s>>>(
            q[cur].as<uint32_t>(), qofs.as<uint64_t>(), blk.as<uint64_t>(), m,
            out.ofs.as<uint64_t>(), out.adj.as<uint32_t>(), out.w.as<float>(),
            dist.as<uint32_t>(), stamp.as<int32_t>(), int32_t(nrounds), q[cur ^ 1].as<uint32_t>(),
            cnt.as<unsigned long long>());
      }
      LLG_HIP(hipGetLastError());
      LLG_HIP(hipMemcpyAsync(ctx->readback, cnt.p, 8, hipMemcpyDeviceToHost, s));
      LLG_HIP(hipStreamSynchronize(s));
      nq = rb[0];
      nrounds++;
      cur ^= 1;
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
  }
}

}  // extern "C"
