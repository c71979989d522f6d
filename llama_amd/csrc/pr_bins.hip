// Binned pull PageRank (propagation blocking) over the snapshot out-CSR
// (task pagerank / pagerank_double, benchmark/benchmarks/pagerank.h:98-182).
//
// The plain pull gathers one 4-byte contribution per in-edge from a V-entry
// vector: at R-MAT scale 24 that vector is 16x an XCD's L2, every gather costs
// a cache-line request and the sweep is bound by request latency, not bytes.
// Here no iteration gathers at random.  Vertices are cut into bins of 2^bb
// consecutive IDs.  Every out-edge (u, v) owns one slot of a binned value
// array; the slots of destination bin b' hold the edges into b', grouped by
// the source's bin b (a "chunk" (b', b), contiguous), in out-edge order.  One
// launch per iteration, one workgroup per bin b:
//   A. streams the slots of bin b -- value (the source's contribution, FP32)
//      and the destination's offset in the bin (u16) -- and adds each value,
//      as a 2^-61 fixed-point integer, into a per-destination 64-bit
//      accumulator in LDS;
//   B. finalises the bin's vertices: rank = (1-d)/N + d*S, contribution =
//      rank/outdeg (FP32), kept in LDS;
//   C. as the same vertices are the sources of chunks (b', b) for every b',
//      writes the next iteration's values into those chunks (each slot names
//      its source's offset in the bin, u16): a wave takes the chunks kGroupC
//      at a time, their first 128 slots loaded together, so its writes are
//      contiguous runs and several loads are in flight.  Neighbouring bins
//      run on one XCD, whose L2 merges the partial lines at chunk ends.
// Integer addition is associative, so a row sum does not depend on the order
// its slots are added in: every run gives the same bits, whatever the
// schedule or bin size.  Each term is the FP32 contribution truncated to a
// multiple of 2^-61 (at most 2^-61 absolute per term; the rank sum never
// exceeds 1, so the accumulators cannot overflow).
//
// Bytes per iteration: 4 B value read + 2 B destination offset + 2 B source
// offset + 4 B value written per edge; 4 B out-degree + 4 B rank written per
// vertex; 8 B per chunk of the chunk table.  The layout is built once per
// snapshot: the out-edges stably sorted by destination bin (the slots), and
// per-(piece of a source bin, destination bin) counts scanned
// destination-major into the chunk table.
#include <algorithm>
#include <vector>

#include "bins.h"
#include "ll_gpu_internal.h"

namespace llgpu {

namespace {

constexpr int kIterBlock = 1024;
constexpr int kUnrollA = 8;  // phase A: slots per thread in flight
constexpr int kGroupC = 8;   // phase C: chunks per wave in flight
constexpr int kBuildBlock = 512;

// Build items: pieces of one source bin's out-edges, a bin holding more than
// max(2^18, E / 2048) cut into equal pieces, so that no workgroup of the build
// walks a hub bin alone.
constexpr uint64_t kMinPiece = uint64_t(1) << 18;

struct PbItem {
  uint32_t bin, e0, e1, pad;  // out-edges [e0, e1) of source bin `bin`
};

// bo[b] = first out-edge of source bin b, b = 0..nb (bo[nb] = E).
__global__ void k_pb_bounds(const uint64_t* ofs, uint64_t V, uint32_t bb, uint32_t nb,
                            uint64_t* bo) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i <= nb) bo[i] = ofs[min(V, i << bb)];
}

// Per out-edge of an item: the destination bin as the sort key and (source
// offset << 16 | destination offset) as the value.  The item is walked in
// tiles of kKeyTile edges.  A tile's source rows: r0, the row holding its
// first edge (a binary search over the bin's row offsets, staged in LDS, for
// the item's first tile; then the row of the previous tile's last edge), and
// every later row starting inside the tile marks its first edge in LDS; a
// running maximum over the tile then gives each edge its row.  Reads and
// writes follow the edge order (coalesced); a tile's destinations are loaded
// while the tile before it is processed.  The same pass counts the item's
// out-edges per destination bin (LDS), into th[b' * ni + i].
constexpr int kKeyPer = 8;
constexpr uint32_t kKeyTile = kBuildBlock * kKeyPer;

template <uint32_t BB>
__global__ __launch_bounds__(kBuildBlock) void k_pb_keys(const uint64_t* ofs, const uint32_t* adj,
                                                         const PbItem* items, uint32_t ni,
                                                         uint64_t V, uint32_t nb, uint32_t* key,
                                                         uint32_t* val, uint32_t* th) {
  constexpr uint32_t R = 1u << BB;
  __shared__ uint32_t ro[R + 1];
  __shared__ uint32_t mark[kKeyTile];
  __shared__ uint32_t wmax[kBuildBlock / 64];
  extern __shared__ uint32_t h[];
  // neighbouring items write neighbouring counts: keep them on one XCD
  const uint32_t item = xcd_block(blockIdx.x, gridDim.x);
  const PbItem it = items[item];
  const uint64_t v0 = uint64_t(it.bin) << BB;
  const uint32_t nrows = uint32_t(min(uint64_t(R), V - v0));
  for (uint32_t i = threadIdx.x; i <= nrows; i += blockDim.x) ro[i] = uint32_t(ofs[v0 + i]);
  for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) h[i] = 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  // r0: the last row starting at or before the tile's first edge; found by a
  // binary search for the item's first tile (ro[0] <= e0 < ro[nrows]), then
  // carried: the next tile's r0 is the row of this tile's last edge
  uint32_t r0;
  {
    uint32_t lo = 0, hi = nrows;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (ro[mid] <= it.e0) lo = mid;
      else hi = mid;
    }
    r0 = lo;
  }
  // each tile's destinations are requested while the tile before is ranked
  uint32_t cur[kKeyPer];
#pragma unroll
  for (int j = 0; j < kKeyPer; j++) {
    const uint64_t e = it.e0 + uint64_t(j) * kBuildBlock + threadIdx.x;
    cur[j] = e < it.e1 ? adj[e] : 0u;
  }
  for (uint64_t t0 = it.e0; t0 < it.e1; t0 += kKeyTile) {
    const uint64_t t1 = min(t0 + kKeyTile, uint64_t(it.e1));
    uint32_t nxt[kKeyPer];
#pragma unroll
    for (int j = 0; j < kKeyPer; j++) {
      const uint64_t e = t1 + uint64_t(j) * kBuildBlock + threadIdx.x;
      nxt[j] = e < it.e1 ? adj[e] : 0u;
    }
#pragma unroll
    for (int j = 0; j < kKeyPer; j++) mark[threadIdx.x + j * kBuildBlock] = 0;
    __syncthreads();
    // rows after r0 that start inside the tile (ro ascending: a thread stops
    // at its first row starting past the tile)
    for (uint32_t r = r0 + 1 + threadIdx.x; r < nrows && ro[r] < t1; r += kBuildBlock)
      atomicMax(&mark[ro[r] - t0], r);
    __syncthreads();
    // running maximum over the tile, thread t owning [kKeyPer t, kKeyPer (t+1))
    uint32_t m[kKeyPer];
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < kKeyPer; j++) {
      run = max(run, mark[threadIdx.x * kKeyPer + j]);
      m[j] = run;
    }
    uint32_t inc = run;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(inc, o, 64);
      if (lane >= o) inc = max(inc, y);
    }
    if (lane == 63) wmax[wave] = inc;
    __syncthreads();
    uint32_t before = __shfl_up(inc, 1, 64);
    if (lane == 0) before = 0;
    uint32_t last = r0;  // the row of the tile's last edge
    for (int w = 0; w < kBuildBlock / 64; w++) {
      if (w < wave) before = max(before, wmax[w]);
      last = max(last, wmax[w]);
    }
#pragma unroll
    for (int j = 0; j < kKeyPer; j++) mark[threadIdx.x * kKeyPer + j] = max(before, m[j]);
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kKeyPer; j++) {
      const uint64_t e = t0 + uint64_t(j) * kBuildBlock + threadIdx.x;
      if (e < t1) {
        const uint32_t v = cur[j];
        const uint32_t row = max(r0, mark[uint32_t(e - t0)]);
        key[e] = v >> BB;
        val[e] = row << 16 | (v & (R - 1));
        atomicAdd(&h[v >> BB], 1u);
      }
    }
    r0 = last;
#pragma unroll
    for (int j = 0; j < kKeyPer; j++) cur[j] = nxt[j];
    __syncthreads();  // mark and wmax are rewritten by the next tile
  }
  for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) th[uint64_t(i) * ni + item] = h[i];
}

// Chunk table by source bin: ct[b * nb + b'] = {first slot, slots} of chunk
// (b', b), from X, the exclusive prefix of the item counts in (destination
// bin, item) order: the slots are the out-edges stably sorted by destination
// bin, so item i's out-edges into b' start at slot X[b' * ni + i], and the
// items of source bin b (first[b] .. first[b + 1] - 1, consecutive) fill
// chunk (b', b); bstart[b'] = first slot of destination bin b', bstart[nb] = E.
__global__ void k_pb_table(const uint64_t* X, const uint32_t* first, uint32_t ni, uint32_t nb,
                           uint64_t E, uint2* ct, uint64_t* bstart) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < uint64_t(nb) * nb) {
    const uint64_t b = i / nb, bp = i % nb;
    const uint64_t a = X[bp * ni + first[b]], z = X[bp * ni + first[b + 1]];
    ct[i] = make_uint2(uint32_t(a), uint32_t(z - a));
  }
  if (i < nb) bstart[i] = X[i * ni];
  if (i == nb) bstart[nb] = E;
}

struct PbArgs {
  const float* val_in;
  float* val_out;
  const uint16_t* ldst;
  const uint16_t* sidx;
  const uint64_t* bstart;
  const uint2* ct;
  const uint32_t* outdeg;
  const float* rank_in;  // read when diff is accumulated
  float* rank_out;
  double* diff;
  uint64_t V;
  uint32_t nb;
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
  // neighbouring source bins write neighbouring chunks of every destination
  // bin: keep them on one XCD
  const uint64_t b = xcd_block(blockIdx.x, gridDim.x);
  const uint64_t v0 = b << BB;
  const uint32_t nrows = uint32_t(min(uint64_t(R), a.V - v0));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (!kInit) {
    for (uint32_t i = threadIdx.x; i < R; i += kIterBlock) acc[i] = 0;
    __syncthreads();
    const uint64_t k0 = a.bstart[b], k1 = a.bstart[b + 1];
    // kUnrollA slots per thread requested before any is added
    for (uint64_t k = k0 + threadIdx.x; k < k1; k += kUnrollA * uint64_t(kIterBlock)) {
      float x[kUnrollA];
      uint16_t d[kUnrollA];
#pragma unroll
      for (int u = 0; u < kUnrollA; u++) {
        const uint64_t q = k + uint64_t(u) * kIterBlock;
        x[u] = q < k1 ? a.val_in[q] : 0.f;
        d[u] = q < k1 ? a.ldst[q] : uint16_t(0);
      }
#pragma unroll
      for (int u = 0; u < kUnrollA; u++)
        if (k + uint64_t(u) * kIterBlock < k1)
          atomicAdd(&acc[d[u]], (unsigned long long)(double(x[u]) * 0x1p61));
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
    // chunks (b', b) of every destination bin b'
    scatter_chunks<kIterBlock, kGroupC>(a.ct + b * a.nb, a.nb, a.sidx, cl, a.val_out);
  }
  if (kDiff) {
    for (int o = 32; o > 0; o >>= 1) dsum += __shfl_down(dsum, o, 64);
    __syncthreads();  // the LDS words below are free again
    float* red = reinterpret_cast<float*>(acc);
    if (lane == 0) red[wave] = dsum;
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

}  // namespace

void build_pr_bins(ll_gpu_graph* g, const Csr& out, uint32_t bb, PrBins& pb) {
  ll_gpu_ctx* ctx = g->ctx;
  hipStream_t s = ctx->stream;
  const uint64_t V = g->V, E = out.E;
  LLG_CHECK(E < (uint64_t(1) << 32), LL_GPU_ERANGE, "binned PageRank needs E < 2^32");
  LLG_CHECK(bb >= 12 && bb <= 14, LL_GPU_EINVAL, "bin bits must be 12, 13 or 14");
  pb = PrBins();
  const uint32_t nb = uint32_t((V + (uint64_t(1) << bb) - 1) >> bb);
  const uint64_t nc = uint64_t(nb) * nb;
  const size_t lds = size_t(nb) * 4;
  LLG_CHECK(lds <= 64 * 1024, LL_GPU_ERANGE, "too many PageRank bins for one workgroup's LDS");
  {
    // the key pass holds the bin's row offsets, a tile of row marks and the
    // bin counters in LDS
    int dev = 0, max_lds = 0;
    LLG_HIP(hipGetDevice(&dev));
    LLG_HIP(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    LLG_CHECK(lds + ((size_t(1) << bb) + 1) * 4 + kKeyTile * 4 + 64 <= size_t(max_lds),
              LL_GPU_ERANGE,
              "binned PageRank: " + std::to_string(nb) + " bins of 2^" + std::to_string(bb) +
                  " vertices need more LDS per workgroup than the device has");
  }
  const uint32_t* adj = out.adj.as<uint32_t>();
  const uint64_t* ofs = out.ofs.as<uint64_t>();
  pb.ldst.alloc(std::max<uint64_t>(E, 1) * 2, s);
  pb.sidx.alloc(std::max<uint64_t>(E, 1) * 2, s);
  pb.bstart.alloc((uint64_t(nb) + 1) * 8, s);
  pb.ct.alloc(nc * 8, s);
  TimedLaunch t(ctx, "pr_bins_build");
  // items: each source bin cut into pieces of at most `piece` out-edges
  std::vector<uint64_t> hb(size_t(nb) + 1);
  {
    DevBuf bo;
    bo.alloc((uint64_t(nb) + 1) * 8, s);
    k_pb_bounds<<<grid_for(uint64_t(nb) + 1, 256), 256, 0, s>>>(ofs, V, bb, nb, bo.as<uint64_t>());
    LLG_HIP(hipGetLastError());
    LLG_HIP(hipMemcpyAsync(hb.data(), bo.p, hb.size() * 8, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
  }
  const uint64_t piece = std::max(kMinPiece, (E + 2047) / 2048);
  std::vector<PbItem> items;
  std::vector<uint32_t> first(size_t(nb) + 1);
  for (uint32_t b = 0; b < nb; b++) {
    first[b] = uint32_t(items.size());
    const uint64_t n = hb[b + 1] - hb[b];
    const uint64_t k = std::max<uint64_t>(1, (n + piece - 1) / piece);
    for (uint64_t j = 0; j < k; j++)
      items.push_back({b, uint32_t(hb[b] + n * j / k), uint32_t(hb[b] + n * (j + 1) / k), 0u});
  }
  first[nb] = uint32_t(items.size());
  const uint32_t ni = uint32_t(items.size());
  const uint64_t nx = uint64_t(nb) * ni;
  DevBuf d_items, d_first, th, X;
  d_items.alloc(items.size() * sizeof(PbItem), s);
  d_first.alloc(first.size() * 4, s);
  LLG_HIP(hipMemcpyAsync(d_items.p, items.data(), d_items.bytes, hipMemcpyHostToDevice, s));
  LLG_HIP(hipMemcpyAsync(d_first.p, first.data(), d_first.bytes, hipMemcpyHostToDevice, s));
  th.alloc(nx * 4, s);
  X.alloc((nx + 1) * 8, s);
  {
    // keys and counts in one pass, then the slots: out-edges stably sorted by
    // destination bin, the last sort pass writing the two u16 arrays
    DevBuf key, val, kw, vw;
    key.alloc(std::max<uint64_t>(E, 1) * 4, s);
    val.alloc(std::max<uint64_t>(E, 1) * 4, s);
    kw.alloc(std::max<uint64_t>(E, 1) * 4, s);
    vw.alloc(std::max<uint64_t>(E, 1) * 4, s);
    const PbItem* di = d_items.as<PbItem>();
    uint32_t* tk = key.as<uint32_t>();
    uint32_t* tv = val.as<uint32_t>();
    uint32_t* tc = th.as<uint32_t>();
    switch (bb) {
      case 12: k_pb_keys<12><<<ni, kBuildBlock, lds, s>>>(ofs, adj, di, ni, V, nb, tk, tv, tc); break;
      case 13: k_pb_keys<13><<<ni, kBuildBlock, lds, s>>>(ofs, adj, di, ni, V, nb, tk, tv, tc); break;
      default: k_pb_keys<14><<<ni, kBuildBlock, lds, s>>>(ofs, adj, di, ni, V, nb, tk, tv, tc); break;
    }
    LLG_HIP(hipGetLastError());
    scan_u32_to_u64(th.as<uint32_t>(), X.as<uint64_t>(), nx, s);
    int kbits = 1;
    while (kbits < 32 && (uint64_t(nb - 1) >> kbits) != 0) kbits++;
    sort_pairs_u32_u32_split(tk, kw.as<uint32_t>(), tv, vw.as<uint32_t>(), E, kbits,
                             pb.ldst.as<uint16_t>(), pb.sidx.as<uint16_t>(), s);
  }
  k_pb_table<<<grid_for(std::max<uint64_t>(nc, nb + 1), 256), 256, 0, s>>>(
      X.as<uint64_t>(), d_first.as<uint32_t>(), ni, nb, E, pb.ct.as<uint2>(),
      pb.bstart.as<uint64_t>());
  LLG_HIP(hipGetLastError());
  // the host tables above are read by the copies
  LLG_HIP(hipStreamSynchronize(s));
  pb.valid = true;
  pb.level = out.level;
  pb.bb = bb;
  pb.nbins = nb;
}

PrBins* ensure_pr_bins(ll_gpu_graph* g, int level, long bin_bits) {
  const Csr& out = ensure_csr(g, LL_GPU_DIR_OUT, level);
  const uint64_t V = g->V;
  if (bin_bits < 0) {
    // bins of 2^13 while a chunk holds 32 slots on average, else 2^14, while
    // E < 2^32 and the bins of 2^14 number at most 16384; otherwise none
    const uint64_t nb13 = (V + (uint64_t(1) << 13) - 1) >> 13;
    const uint64_t nb14 = (V + (uint64_t(1) << 14) - 1) >> 14;
    if (out.E >= (uint64_t(1) << 32) || nb14 > 16384) bin_bits = 0;
    else bin_bits = out.E >= 32 * nb13 * nb13 ? 13 : 14;
  }
  if (bin_bits == 0) return nullptr;
  LLG_CHECK(bin_bits >= 12 && bin_bits <= 14, LL_GPU_EINVAL, "bin bits must be 0 or 12..14");
  PrBins& pb = g->pr_bins;
  if (!(pb.valid && pb.level == out.level && pb.bb == uint32_t(bin_bits)))
    build_pr_bins(g, out, uint32_t(bin_bits), pb);
  return &pb;
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
  a.sidx = pb.sidx.as<uint16_t>();
  a.bstart = pb.bstart.as<uint64_t>();
  a.ct = pb.ct.as<uint2>();
  a.outdeg = outdeg;
  a.V = V;
  a.nb = pb.nbins;
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
    LLG_HIP(hipMemcpyAsync(ctx->readback, diff.p, 8, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    *last_diff = iters > 0 ? *static_cast<const double*>(ctx->readback) : 0.0;
  }
}

}  // namespace llgpu
