// Binned pull PageRank (propagation blocking) over the snapshot out-CSR
// (task pagerank / pagerank_double, benchmark/benchmarks/pagerank.h:98-182).
//
// The plain pull gathers one 4-byte contribution per in-edge from a V-entry
// vector: at R-MAT scale 24 that vector is 16x an XCD's L2, every gather costs
// a cache-line request and the sweep is bound by request latency, not bytes.
// Here no iteration gathers at random.  Destinations are cut into bins of
// 2^bb consecutive vertices; every out-edge (u, v) owns one slot of a binned
// value array, grouped by v's bin.  One launch per iteration, one workgroup per
// bin b:
//   A. streams the slots of bin b -- value (the source's contribution, FP32)
//      and the destination's offset in the bin (u16) -- and adds each value,
//      as a 2^-61 fixed-point integer, into a per-destination 64-bit
//      accumulator in LDS;
//   B. finalises the bin's vertices: rank = (1-d)/N + d*S, contribution =
//      rank/outdeg (FP32), kept in LDS;
//   C. as vertices of the same 2^bb range are sources too, scatters their
//      contributions along their out-edges into the slots of the next
//      iteration's binned array (positions precomputed: binpos, and the
//      source's offset in its range: srcl).
// Integer addition is associative, so a row sum does not depend on the order
// the slots are added in: every run gives the same bits, whatever the
// schedule.  Each term is the FP32 contribution truncated to a multiple of
// 2^-61 (at most 2^-61 absolute per term; the rank sum never exceeds 1, so the
// accumulators cannot overflow).
//
// Bytes per iteration: 4 B value read + 2 B destination offset + 4 B binpos +
// 2 B source offset + 4 B value written per edge, and 4 B out-degree + 4 B rank
// written per vertex.  The layout (binpos, destination offsets, bin starts,
// source offsets) is built once per snapshot: a per-tile histogram of the
// destination bins, a scan of the (bin, tile) counts, and a placement pass in
// which each tile hands out positions within its (bin, tile) ranges with LDS
// counters -- slots of one bin in any order, since the sums do not depend on
// it.
#include "ll_gpu_internal.h"

namespace llgpu {

namespace {

constexpr int kIterBlock = 1024;
constexpr int kBuildBlock = 256;

__global__ __launch_bounds__(kBuildBlock) void k_pb_hist(const uint32_t* adj, uint64_t E,
                                                         uint32_t bb, uint32_t nbins, uint64_t T,
                                                         uint32_t* th, uint64_t ntiles) {
  extern __shared__ uint32_t h[];
  for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) h[i] = 0;
  __syncthreads();
  const uint64_t e0 = uint64_t(blockIdx.x) * T;
  const uint64_t e1 = min(E, e0 + T);
  for (uint64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) atomicAdd(&h[adj[e] >> bb], 1u);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x)
    th[uint64_t(i) * ntiles + blockIdx.x] = h[i];
}

__global__ __launch_bounds__(kBuildBlock) void k_pb_place(const uint32_t* adj, uint64_t E,
                                                          uint32_t bb, uint32_t nbins, uint64_t T,
                                                          const uint64_t* toff, uint64_t ntiles,
                                                          uint32_t* binpos, uint16_t* ldst) {
  extern __shared__ uint32_t cur[];
  for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) cur[i] = 0;
  __syncthreads();
  const uint64_t e0 = uint64_t(blockIdx.x) * T;
  const uint64_t e1 = min(E, e0 + T);
  const uint32_t mask = (1u << bb) - 1;
  for (uint64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const uint32_t v = adj[e];
    const uint32_t b = v >> bb;
    const uint32_t loc = atomicAdd(&cur[b], 1u);
    const uint64_t pos = toff[uint64_t(b) * ntiles + blockIdx.x] + loc;
    binpos[e] = uint32_t(pos);
    ldst[pos] = uint16_t(v & mask);
  }
}

// bstart[b] = first slot of bin b (the offset of its first tile), bstart[nbins] = E.
__global__ void k_pb_bstart(const uint64_t* toff, uint64_t ntiles, uint32_t nbins, uint64_t E,
                            uint64_t* bstart) {
  const uint64_t b = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (b < nbins) bstart[b] = toff[b * ntiles];
  if (b == nbins) bstart[b] = E;
}

// srcl[e] = offset of out-edge e's source in its 2^bb range; one wave per 64
// rows, all lanes on one row at a time (coalesced for long rows).
__global__ void k_pb_srcl(const uint64_t* ofs, uint64_t V, uint32_t bb, uint16_t* srcl) {
  const uint64_t w = (uint64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t r0 = w * 64;
  if (r0 >= V) return;
  const uint32_t mask = (1u << bb) - 1;
  const uint64_t r1 = min(V, r0 + 64);
  for (uint64_t r = r0; r < r1; r++) {
    const uint64_t lo = ofs[r], hi = ofs[r + 1];
    for (uint64_t e = lo + lane; e < hi; e += 64) srcl[e] = uint16_t(r & mask);
  }
}

struct PbArgs {
  const float* val_in;
  float* val_out;
  const uint16_t* ldst;
  const uint64_t* bstart;
  const uint32_t* binpos;
  const uint16_t* srcl;
  const uint64_t* ofs_out;
  const uint32_t* outdeg;
  const float* rank_in;   // read when diff is accumulated
  float* rank_out;
  double* diff;
  uint64_t V;
  double base, d, r0;
};

// kInit: no accumulation, every rank = r0 (iteration 0's input); kScatter:
// write the next iteration's binned values; kDiff: add |rank' - rank| of the
// bin's vertices to *diff.
template <uint32_t BB, bool kInit, bool kScatter, bool kDiff>
__global__ __launch_bounds__(kIterBlock) void k_pb_iter(PbArgs a) {
  constexpr uint32_t R = 1u << BB;
  constexpr uint32_t kPer = R / kIterBlock;
  static_assert(kPer >= 1 && R % kIterBlock == 0, "bin rows must be a multiple of the block");
  __shared__ unsigned long long acc[R];
  const uint64_t b = blockIdx.x;
  const uint64_t v0 = b << BB;
  const uint32_t nrows = uint32_t(min(uint64_t(R), a.V - v0));
  if (!kInit) {
    for (uint32_t i = threadIdx.x; i < R; i += kIterBlock) acc[i] = 0;
    __syncthreads();
    const uint64_t k0 = a.bstart[b], k1 = a.bstart[b + 1];
    for (uint64_t k = k0 + threadIdx.x; k < k1; k += kIterBlock) {
      const float x = a.val_in[k];
      atomicAdd(&acc[a.ldst[k]], (unsigned long long)(double(x) * 0x1p61));
    }
    __syncthreads();
  }
  float c[kPer];
  float dsum = 0.f;
#pragma unroll
  for (uint32_t j = 0; j < kPer; j++) {
    const uint32_t i = threadIdx.x + j * kIterBlock;
    c[j] = 0.f;
    if (i < nrows) {
      float r;
      if (kInit) r = float(a.r0);
      else r = float(a.base + a.d * (double(acc[i]) * 0x1p-61));
      a.rank_out[v0 + i] = r;
      if (kDiff) dsum += fabsf(r - a.rank_in[v0 + i]);
      const uint32_t od = a.outdeg[v0 + i];
      c[j] = od ? r / float(od) : 0.f;
    }
  }
  float* cl = reinterpret_cast<float*>(acc);
  if (kScatter) {
    if (!kInit) __syncthreads();  // accumulators read by every thread
#pragma unroll
    for (uint32_t j = 0; j < kPer; j++) cl[threadIdx.x + j * kIterBlock] = c[j];
    __syncthreads();
    const uint64_t e0 = a.ofs_out[v0], e1 = a.ofs_out[v0 + nrows];
    for (uint64_t e = e0 + threadIdx.x; e < e1; e += kIterBlock)
      a.val_out[a.binpos[e]] = cl[a.srcl[e]];
  }
  if (kDiff) {
    for (int o = 32; o > 0; o >>= 1) dsum += __shfl_down(dsum, o, 64);
    __syncthreads();  // the LDS words below are free again
    float* red = reinterpret_cast<float*>(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dsum;
    __syncthreads();
    if (threadIdx.x == 0) {
      float t = 0.f;
      for (int i = 0; i < kIterBlock / 64; i++) t += red[i];
      atomicAdd(a.diff, double(t));
    }
  }
}

template <uint32_t BB>
void launch_iter(bool init, bool scatter, bool diff, unsigned nb, const PbArgs& a,
                 hipStream_t s) {
  if (init) k_pb_iter<BB, true, true, false><<<nb, kIterBlock, 0, s>>>(a);
  else if (scatter) k_pb_iter<BB, false, true, false><<<nb, kIterBlock, 0, s>>>(a);
  else if (diff) k_pb_iter<BB, false, false, true><<<nb, kIterBlock, 0, s>>>(a);
  else k_pb_iter<BB, false, false, false><<<nb, kIterBlock, 0, s>>>(a);
}

__global__ void k_zero_f64(double* p) { *p = 0.0; }

__global__ void k_zero_u32(uint32_t* p, uint64_t n) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

}  // namespace

void build_pr_bins(ll_gpu_graph* g, const Csr& out, uint32_t bb, PrBins& pb) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V, E = out.E;
  LLG_CHECK(E < (uint64_t(1) << 32), LL_GPU_ERANGE, "binned PageRank needs E < 2^32");
  LLG_CHECK(bb >= 10 && bb <= 16, LL_GPU_EINVAL, "bin bits must be in [10, 16]");
  pb = PrBins();
  const uint32_t nbins = uint32_t((V + (uint64_t(1) << bb) - 1) >> bb);
  // tiles large enough that the (bin, tile) table stays small
  const uint64_t T = std::max<uint64_t>(16384, uint64_t(nbins) * 64);
  const uint64_t ntiles = std::max<uint64_t>((E + T - 1) / T, 1);
  const size_t lds = size_t(nbins) * 4;
  LLG_CHECK(lds <= 160 * 1024, LL_GPU_ERANGE, "too many PageRank bins for one workgroup's LDS");
  DevBuf th, toff;
  th.alloc(uint64_t(nbins) * ntiles * 4, s);
  toff.alloc((uint64_t(nbins) * ntiles + 1) * 8, s);
  pb.binpos.alloc(std::max<uint64_t>(E, 1) * 4, s);
  pb.ldst.alloc(std::max<uint64_t>(E, 1) * 2, s);
  pb.srcl.alloc(std::max<uint64_t>(E, 1) * 2, s);
  pb.bstart.alloc((uint64_t(nbins) + 1) * 8, s);
  if (lds > 64 * 1024) {
    LLG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pb_hist),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    LLG_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_pb_place),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
  }
  const uint32_t* adj = out.adj.as<uint32_t>();
  {
    TimedLaunch t(ctx, "pr_bins_build");
    if (E) {
      k_pb_hist<<<unsigned(ntiles), kBuildBlock, lds, s>>>(adj, E, bb, nbins, T,
                                                           th.as<uint32_t>(), ntiles);
      LLG_HIP(hipGetLastError());
    } else {
      k_zero_u32<<<grid_for(uint64_t(nbins) * ntiles, 256), 256, 0, s>>>(
          th.as<uint32_t>(), uint64_t(nbins) * ntiles);
    }
    scan_u32_to_u64(th.as<uint32_t>(), toff.as<uint64_t>(), uint64_t(nbins) * ntiles, s);
    if (E) {
      k_pb_place<<<unsigned(ntiles), kBuildBlock, lds, s>>>(adj, E, bb, nbins, T,
                                                            toff.as<uint64_t>(), ntiles,
                                                            pb.binpos.as<uint32_t>(),
                                                            pb.ldst.as<uint16_t>());
      LLG_HIP(hipGetLastError());
    }
    k_pb_bstart<<<grid_for(uint64_t(nbins) + 1, 256), 256, 0, s>>>(toff.as<uint64_t>(), ntiles,
                                                                   nbins, E,
                                                                   pb.bstart.as<uint64_t>());
    k_pb_srcl<<<grid_for(((V + 63) / 64) * 64, 256), 256, 0, s>>>(out.ofs.as<uint64_t>(), V, bb,
                                                                  pb.srcl.as<uint16_t>());
    LLG_HIP(hipGetLastError());
  }
  pb.valid = true;
  pb.level = out.level;
  pb.bb = bb;
  pb.nbins = nbins;
}

void pagerank_bins(ll_gpu_graph* g, const Csr& out, const uint32_t* outdeg, PrBins& pb,
                   int iters, double damping, DevBuf& rank_final, double* last_diff) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V, E = out.E;
  DevBuf val[2], rank[2], diff;
  for (int i = 0; i < 2; i++) {
    val[i].alloc(std::max<uint64_t>(E, 1) * 4, s);
    rank[i].alloc(std::max<uint64_t>(V, 1) * 4, s);
  }
  diff.alloc(8, s);
  k_zero_f64<<<1, 1, 0, s>>>(diff.as<double>());
  PbArgs a{};
  a.ldst = pb.ldst.as<uint16_t>();
  a.bstart = pb.bstart.as<uint64_t>();
  a.binpos = pb.binpos.as<uint32_t>();
  a.srcl = pb.srcl.as<uint16_t>();
  a.ofs_out = out.ofs.as<uint64_t>();
  a.outdeg = outdeg;
  a.V = V;
  a.base = (1.0 - damping) / double(V);
  a.d = damping;
  a.r0 = 1.0 / double(V);
  auto launch = [&](bool init, bool scatter, bool dflag) {
    switch (pb.bb) {
      case 12: launch_iter<12>(init, scatter, dflag, pb.nbins, a, s); break;
      case 13: launch_iter<13>(init, scatter, dflag, pb.nbins, a, s); break;
      case 14: launch_iter<14>(init, scatter, dflag, pb.nbins, a, s); break;
      default: throw Error{LL_GPU_EINVAL, "binned PageRank supports 12, 13 or 14 bin bits"};
    }
    LLG_HIP(hipGetLastError());
  };
  {
    TimedLaunch t(ctx, "k_pb_init");
    a.val_out = val[0].as<float>();
    a.rank_out = rank[0].as<float>();
    launch(true, true, false);
  }
  for (int it = 0; it < iters; it++) {
    const int cur = it & 1;
    a.val_in = val[cur].as<float>();
    a.val_out = val[cur ^ 1].as<float>();
    a.rank_in = rank[cur].as<float>();
    a.rank_out = rank[cur ^ 1].as<float>();
    const bool last = it == iters - 1;
    a.diff = (last && last_diff) ? diff.as<double>() : nullptr;
    TimedLaunch t(ctx, "k_pb_iter");
    launch(false, !last, a.diff != nullptr);
  }
  rank_final = std::move(rank[iters & 1]);
  if (last_diff) {
    double h = 0;
    LLG_HIP(hipMemcpyAsync(ctx->readback, diff.p, 8, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    h = *static_cast<const double*>(ctx->readback);
    *last_diff = iters > 0 ? h : 0.0;
  }
}

}  // namespace llgpu
