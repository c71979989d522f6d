// Batched personalized PageRank: B single-vertex seeds per pass, ranks kept as
// vertex-major FP32 blocks R[V][B] (one vertex's B ranks are one contiguous
// 4B-byte row, so gathering an in-neighbour's row is a coalesced load).
//
//   R0 = e_s,  R'[t][j] = (1-d) [t == s_j] + d * sum_{w in in(t)} R[w][j] / outdeg(w)
//
// With a uniform teleport this is the reference pull recurrence
// (pagerank.h:98-182); no dangling redistribution.  One iteration is one
// launch of a persistent kernel over the in-CSR row blocks (grid-stride over
// items, so per-column L1 changes are reduced in registers and flushed with B
// atomics per workgroup), a fix-up for rows split into pieces, and a check
// that raises a device flag once every column's L1 change is <= tol; later
// iterations see the flag and return at once.  All iterations are captured in
// one hipGraph and replayed per batch of seeds.  Sums are FP32 Neumaier.
// B in {16, 32, 64, 128}; 0 = autotune on the first batch.
#include <chrono>
#include <string>

#include "ll_gpu_internal.h"
#include "rowblocks.h"

namespace llgpu {

namespace {

constexpr int kBlock = 256;
constexpr uint32_t kPprCap = 256;
constexpr uint32_t kPprPiece = 1024;
constexpr uint32_t kNoSeed = 0xffffffffu;

struct PprState {
  unsigned long long converged;  // 1 once every column met the tolerance
  unsigned long long iters_done;
};

struct PprArgs {
  const RowItem* items;
  const uint32_t* slots;
  uint64_t nitems;
  const uint64_t* ofs;
  const uint32_t* adj;
  const float* invdeg;
  const uint32_t* seeds;   // [B]
  const float* rin;
  float* rout;
  float* part_s;           // [nsplit][B]
  float* part_c;
  float* diff;             // [B]
  PprState* st;
  float d, one_minus_d;
};

template <int B>
struct Geo {
  static constexpr int EPL = B == 128 ? 2 : 1;     // floats per lane
  static constexpr int LPR = B / EPL;              // lanes per row
  static constexpr int RPW = 64 / LPR;             // rows per wave
  static constexpr int SLOTS = (kBlock / 64) * RPW;  // row slots per workgroup
};

template <int B>
__device__ inline void finalize_row(const PprArgs& a, uint32_t t, int li, const Neu* acc,
                                    float* dsum) {
  constexpr int EPL = Geo<B>::EPL;
#pragma unroll
  for (int k = 0; k < EPL; k++) {
    const int col = li * EPL + k;
    float v = a.d * acc[k].value();
    if (a.seeds[col] == t) v += a.one_minus_d;
    const uint64_t idx = uint64_t(t) * B + col;
    dsum[k] += fabsf(v - a.rin[idx]);
    a.rout[idx] = v;
  }
}

// acc += sum over edges e in [lo, hi) step `stride` of rin[adj[e]][cols] / outdeg:
// four rows are loaded before any is accumulated, so each lane group keeps
// four independent gathers in flight.
template <int B>
__device__ inline void accumulate_rows(const PprArgs& a, uint64_t lo, uint64_t hi,
                                       uint64_t stride, int li, Neu* acc) {
  constexpr int EPL = Geo<B>::EPL;
  constexpr int U = 4;
  uint64_t e = lo;
  for (; e + (U - 1) * stride < hi; e += U * stride) {
    uint32_t w[U];
    float f[U], x[U][EPL];
#pragma unroll
    for (int u = 0; u < U; u++) w[u] = a.adj[e + u * stride];
#pragma unroll
    for (int u = 0; u < U; u++) {
      f[u] = a.invdeg[w[u]];
      const float* row = a.rin + uint64_t(w[u]) * B + li * EPL;
#pragma unroll
      for (int k = 0; k < EPL; k++) x[u][k] = row[k];
    }
#pragma unroll
    for (int u = 0; u < U; u++)
#pragma unroll
      for (int k = 0; k < EPL; k++) acc[k].add(x[u][k] * f[u]);
  }
  for (; e < hi; e += stride) {
    const uint32_t w = a.adj[e];
    const float f = a.invdeg[w];
    const float* row = a.rin + uint64_t(w) * B + li * EPL;
#pragma unroll
    for (int k = 0; k < EPL; k++) acc[k].add(row[k] * f);
  }
}

template <int B>
__global__ __launch_bounds__(kBlock) void k_ppr_iter(PprArgs a) {
  if (a.st->converged) return;
  constexpr int EPL = Geo<B>::EPL, LPR = Geo<B>::LPR, RPW = Geo<B>::RPW, SLOTS = Geo<B>::SLOTS;
  __shared__ float ls[SLOTS][B];
  __shared__ float lc[SLOTS][B];
  __shared__ float ld[SLOTS][B];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane % LPR, sr = lane / LPR;
  const int slot = wave * RPW + sr;
  float dsum[EPL];
#pragma unroll
  for (int k = 0; k < EPL; k++) dsum[k] = 0.f;
  for (uint64_t i = blockIdx.x; i < a.nitems; i += gridDim.x) {
    const RowItem it = a.items[i];
    if (it.npieces == 0) {
      for (uint32_t t = it.row_begin + slot; t < it.row_end; t += SLOTS) {
        Neu acc[EPL];
#pragma unroll
        for (int k = 0; k < EPL; k++) acc[k] = Neu::zero();
        accumulate_rows<B>(a, a.ofs[t], a.ofs[t + 1], 1, li, acc);
        finalize_row<B>(a, t, li, acc, dsum);
      }
    } else {
      const uint32_t t = it.row_begin;
      const uint64_t lo = a.ofs[t] + uint64_t(it.piece) * kPprPiece;
      const uint64_t hi = min(a.ofs[t + 1], lo + kPprPiece);
      Neu acc[EPL];
#pragma unroll
      for (int k = 0; k < EPL; k++) acc[k] = Neu::zero();
      accumulate_rows<B>(a, lo + slot, hi, SLOTS, li, acc);
#pragma unroll
      for (int k = 0; k < EPL; k++) {
        ls[slot][li * EPL + k] = acc[k].s;
        lc[slot][li * EPL + k] = acc[k].c;
      }
      __syncthreads();
      if (slot == 0) {
        Neu tot[EPL];
#pragma unroll
        for (int k = 0; k < EPL; k++) {
          tot[k] = Neu{ls[0][li * EPL + k], lc[0][li * EPL + k]};
          for (int q = 1; q < SLOTS; q++) tot[k].merge(Neu{ls[q][li * EPL + k], lc[q][li * EPL + k]});
        }
        if (it.npieces == 1) {
          finalize_row<B>(a, t, li, tot, dsum);
        } else {
          const uint64_t base = uint64_t(a.slots[i]) * B;
#pragma unroll
          for (int k = 0; k < EPL; k++) {
            a.part_s[base + li * EPL + k] = tot[k].s;
            a.part_c[base + li * EPL + k] = tot[k].c;
          }
        }
      }
      __syncthreads();
    }
  }
  // per-column L1 change: reduce the workgroup's slots, one atomic per column
#pragma unroll
  for (int k = 0; k < EPL; k++) ld[slot][li * EPL + k] = dsum[k];
  __syncthreads();
  for (int col = threadIdx.x; col < B; col += kBlock) {
    float t = 0.f;
    for (int q = 0; q < SLOTS; q++) t += ld[q][col];
    if (t != 0.f) atomicAdd(&a.diff[col], t);
  }
}

template <int B>
__global__ void k_ppr_fixup(PprArgs a) {
  if (a.st->converged) return;
  constexpr int EPL = Geo<B>::EPL, LPR = Geo<B>::LPR;
  // one wave-slice of LPR lanes per split row head
  const uint64_t gid = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t i = gid / LPR;
  const int li = int(gid % LPR);
  float dsum[EPL];
#pragma unroll
  for (int k = 0; k < EPL; k++) dsum[k] = 0.f;
  if (i < a.nitems) {
    const RowItem it = a.items[i];
    if (it.npieces > 1 && it.piece == 0) {
      Neu tot[EPL];
#pragma unroll
      for (int k = 0; k < EPL; k++) tot[k] = Neu::zero();
      for (uint32_t p = 0; p < it.npieces; p++) {
        const uint64_t base = uint64_t(a.slots[i + p]) * B;
#pragma unroll
        for (int k = 0; k < EPL; k++)
          tot[k].merge(Neu{a.part_s[base + li * EPL + k], a.part_c[base + li * EPL + k]});
      }
      finalize_row<B>(a, it.row_begin, li, tot, dsum);
#pragma unroll
      for (int k = 0; k < EPL; k++)
        if (dsum[k] != 0.f) atomicAdd(&a.diff[li * EPL + k], dsum[k]);
    }
  }
}

template <int B>
__global__ void k_ppr_check(PprState* st, float* diff, float tol) {
  if (st->converged) return;
  // single workgroup of B threads
  __shared__ int ok;
  if (threadIdx.x == 0) ok = 1;
  __syncthreads();
  if (threadIdx.x < B) {
    if (!(diff[threadIdx.x] <= tol)) atomicAnd(&ok, 0);
  }
  __syncthreads();
  if (threadIdx.x < B) diff[threadIdx.x] = 0.f;
  if (threadIdx.x == 0) {
    st->iters_done += 1;
    if (ok) st->converged = 1;
  }
}

template <int B>
__global__ void k_ppr_init(float* r, const uint32_t* seeds, uint64_t V, float* diff,
                           PprState* st) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < V * B) {
    const uint32_t t = uint32_t(i / B);
    r[i] = seeds[i % B] == t ? 1.f : 0.f;
  }
  if (i < B) diff[i] = 0.f;
  if (i == 0) {
    st->converged = 0;
    st->iters_done = 0;
  }
}

__global__ void k_invdeg(const uint32_t* deg, float* inv, uint64_t V) {
  const uint64_t v = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < V) inv[v] = deg[v] ? 1.f / float(deg[v]) : 0.f;
}

struct PprPass {
  hipGraphExec_t exec = nullptr;
  ~PprPass() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

}  // namespace

struct PprBufs {
  DevBuf r[2], seeds, diff, st, part_s, part_c, invdeg;
};

template <int B>
void ppr_capture(ll_gpu_graph* g, const Csr& in, const RowBlocks& rb, PprBufs& b, int iters,
                 double d, double tol, PprPass& pass) {
  hipStream_t s = g->ctx->stream;
  const uint64_t V = g->V;
  PprArgs a;
  a.items = rb.items.as<RowItem>();
  a.slots = rb.slots.as<uint32_t>();
  a.nitems = rb.n;
  a.ofs = in.ofs.as<uint64_t>();
  a.adj = in.adj.as<uint32_t>();
  a.invdeg = b.invdeg.as<float>();
  a.seeds = b.seeds.as<uint32_t>();
  a.part_s = b.part_s.as<float>();
  a.part_c = b.part_c.as<float>();
  a.diff = b.diff.as<float>();
  a.st = b.st.as<PprState>();
  a.d = float(d);
  a.one_minus_d = float(1.0 - d);
  const unsigned grid = unsigned(std::min<uint64_t>(std::max<uint64_t>(rb.n, 1),
                                                    uint64_t(g->ctx->num_cus) * 8));
  const uint64_t fix_threads = rb.n * Geo<B>::LPR;
  hipGraph_t graph;
  LLG_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  k_ppr_init<B><<<grid_for(V * B, 256), 256, 0, s>>>(b.r[0].as<float>(), a.seeds, V, a.diff, a.st);
  for (int it = 0; it < iters; it++) {
    a.rin = b.r[it & 1].as<float>();
    a.rout = b.r[(it & 1) ^ 1].as<float>();
    k_ppr_iter<B><<<grid, kBlock, 0, s>>>(a);
    if (rb.nsplit) k_ppr_fixup<B><<<grid_for(fix_threads, 256), 256, 0, s>>>(a);
    k_ppr_check<B><<<1, B < 64 ? 64 : B, 0, s>>>(a.st, a.diff, float(tol));
  }
  hipError_t ce = hipStreamEndCapture(s, &graph);
  LLG_HIP(ce);
  hipError_t ie = hipGraphInstantiate(&pass.exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  LLG_HIP(ie);
}

template <int B>
double ppr_run(ll_gpu_graph* g, int level, const int64_t* seeds, int n, int iters, double d,
               double tol, float* ranks_out, int* iters_done, int max_batches) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V;
  const Csr& in = ensure_csr(g, LL_GPU_DIR_IN, level);
  const uint32_t* outdeg = ensure_degrees(g, LL_GPU_DIR_OUT, level);
  const RowBlocks& rb = ensure_rowblocks(g, LL_GPU_DIR_IN, level, kPprCap, kPprPiece);
  PprBufs b;
  b.r[0].alloc(V * B * 4, s);
  b.r[1].alloc(V * B * 4, s);
  b.seeds.alloc(B * 4, s);
  b.diff.alloc(B * 4, s);
  b.st.alloc(sizeof(PprState), s);
  b.part_s.alloc(std::max<uint64_t>(rb.nsplit, 1) * B * 4, s);
  b.part_c.alloc(std::max<uint64_t>(rb.nsplit, 1) * B * 4, s);
  b.invdeg.alloc(V * 4, s);
  k_invdeg<<<grid_for(V, 256), 256, 0, s>>>(outdeg, b.invdeg.as<float>(), V);
  LLG_HIP(hipGetLastError());
  LLG_HIP(hipStreamSynchronize(s));
  PprPass pass;
  ppr_capture<B>(g, in, rb, b, iters, d, tol, pass);
  int done_max = 0;
  // several batches gather into one device block [V x n] (device-side strided
  // copies), copied to the host in one piece: a strided copy straight into
  // host memory moves one V-row at a time
  DevBuf gathered;
  const bool direct = ranks_out && n <= B;
  if (ranks_out && !direct) gathered.alloc(V * uint64_t(n) * 4, s);
  llgpu::EventPair ev;
  hipEvent_t e0 = ev.a, e1 = ev.b;
  double ms_total = 0;
  int batches = 0;
  for (int c0 = 0; c0 < n && batches < max_batches; c0 += B, batches++) {
    const int m = std::min(B, n - c0);
    uint32_t hs[B];
    for (int j = 0; j < B; j++) hs[j] = j < m ? uint32_t(seeds[c0 + j]) : kNoSeed;
    LLG_HIP(hipMemcpyAsync(b.seeds.p, hs, sizeof hs, hipMemcpyHostToDevice, s));
    LLG_HIP(hipEventRecord(e0, s));
    LLG_HIP(hipGraphLaunch(pass.exec, s));
    LLG_HIP(hipEventRecord(e1, s));
    PprState hst;
    LLG_HIP(hipMemcpyAsync(&hst, b.st.p, sizeof hst, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    float ms = 0;
    LLG_HIP(hipEventElapsedTime(&ms, e0, e1));
    ms_total += ms;
    const int done = int(hst.iters_done);
    done_max = std::max(done_max, done);
    if (ranks_out) {
      const float* src = b.r[done & 1].as<float>();
      if (direct) {
        LLG_HIP(hipMemcpy2DAsync(ranks_out, size_t(n) * 4, src, size_t(B) * 4, size_t(m) * 4,
                                 V, hipMemcpyDeviceToHost, s));
        LLG_HIP(hipStreamSynchronize(s));
      } else {
        LLG_HIP(hipMemcpy2DAsync(gathered.as<float>() + c0, size_t(n) * 4, src, size_t(B) * 4,
                                 size_t(m) * 4, V, hipMemcpyDeviceToDevice, s));
      }
    }
  }
  if (ranks_out && !direct) {
    LLG_HIP(hipMemcpyAsync(ranks_out, gathered.p, V * uint64_t(n) * 4, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
  }
  if (ctx->timer.on) {
    auto& acc = ctx->timer.acc["ppr_graph"];
    acc.first += batches;
    acc.second += ms_total;
  }
  if (iters_done) *iters_done = done_max;
  return batches ? ms_total / batches : 0.0;
}

double ppr_dispatch(int B, ll_gpu_graph* g, int level, const int64_t* seeds, int n, int iters,
                    double d, double tol, float* out, int* done, int max_batches) {
  switch (B) {
    case 16: return ppr_run<16>(g, level, seeds, n, iters, d, tol, out, done, max_batches);
    case 32: return ppr_run<32>(g, level, seeds, n, iters, d, tol, out, done, max_batches);
    case 64: return ppr_run<64>(g, level, seeds, n, iters, d, tol, out, done, max_batches);
    case 128: return ppr_run<128>(g, level, seeds, n, iters, d, tol, out, done, max_batches);
  }
  throw Error{LL_GPU_EINVAL, "batch_width must be 0, 16, 32, 64 or 128"};
}

}  // namespace llgpu

extern "C" {

int ll_gpu_ppr_batch(ll_gpu_graph* g, int level, const int64_t* seeds, int n, int iters,
                     double damping, double tol, int batch_width, float* ranks_out,
                     int* iters_done) {
  if (!g) return LL_GPU_EINVAL;
  using namespace llgpu;
  try {
    LLG_HIP(hipSetDevice(g->ctx->device));
    LLG_CHECK(n >= 0 && (n == 0 || seeds), LL_GPU_EINVAL, "bad seed list");
    LLG_CHECK(iters >= 1 && iters <= 100000, LL_GPU_EINVAL, "iters must be in [1, 100000]");
    LLG_CHECK(damping >= 0 && damping <= 1 && tol >= 0, LL_GPU_EINVAL, "bad damping/tolerance");
    for (int i = 0; i < n; i++)
      LLG_CHECK(seeds[i] >= 0 && uint64_t(seeds[i]) < g->V, LL_GPU_EINVAL,
                "seed outside [0, max_nodes)");
    if (iters_done) *iters_done = 0;
    if (n == 0) return LL_GPU_OK;
    if (batch_width == 0) {
      // autotune: one batch per width on the leading seeds; keep the width
      // with the least estimated time for the whole request
      const Csr& in = ensure_csr(g, LL_GPU_DIR_IN, level);
      int bucket = 16;  // decisions per request size: smallest power of two >= min(n, 128)
      while (bucket < std::min(n, 128)) bucket <<= 1;
      const std::string key = "ppr:" + std::to_string(g->V) + ":" + std::to_string(in.E) +
                              ":" + std::to_string(resolve_level(g, level)) + ":" +
                              std::to_string(bucket);
      auto it = g->ctx->tuned.find(key);
      if (it == g->ctx->tuned.end()) {
        int best = 64;
        double best_t = 1e300;
        const int probe_iters = std::min(iters, 3);
        for (int w : {16, 32, 64, 128}) {
          const double ms = ppr_dispatch(w, g, level, seeds, std::min(n, w), probe_iters,
                                         damping, 0.0, nullptr, nullptr, 1);
          // the request of n seeds takes ceil(n / w) batches of this width
          const double est = ms * double((n + w - 1) / w);
          if (est < best_t) { best_t = est; best = w; }
        }
        g->ctx->tuned[key] = best;
        batch_width = best;
      } else {
        batch_width = it->second;
      }
    }
    ppr_dispatch(batch_width, g, level, seeds, n, iters, damping, tol, ranks_out, iters_done,
                 1 << 30);
    return LL_GPU_OK;
  } catch (const Error& e) {
    g->ctx->err = e.msg;
    return e.code;
  } catch (...) {
    return llgpu::api_exception(g->ctx);
  }
}

int ll_gpu_tuned(ll_gpu_ctx* ctx, const char* prefix, int* value) {
  if (!ctx || !prefix || !value) return LL_GPU_EINVAL;
  for (auto& kv : ctx->tuned)
    if (kv.first.rfind(prefix, 0) == 0) {
      *value = kv.second;
      return LL_GPU_OK;
    }
  return LL_GPU_EINVAL;
}

}  // extern "C"
