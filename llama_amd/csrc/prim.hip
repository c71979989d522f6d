// Device-wide scans, stable radix sorts and run compaction, written for
// 64-wide waves.  They carry no graph logic: they turn per-vertex / per-block
// counts into offsets and order key arrays for the snapshot, PageRank and
// triangle-counting preprocessing.
//
// Scan: tiles of 4096 items (256 threads x 16 rounds); tile totals are
// scanned recursively, then every tile rescans its items from its prefix.
// Integer sums, so the result does not depend on the schedule.
//
// Radix sort: least-significant digit first, stable.  One pre-pass counts
// every 8-bit digit position at once; byte positions whose value is the same
// for all keys are dropped (the host reads the 8 x 256 counts back), and the
// remaining bits are cut into the fewest passes of at most 8 bits, as even
// as possible (11 bits: 6 + 5), so each tile's runs per digit stay long
// enough to fill cache lines.  Callers that know the key width skip the
// pre-pass.  A pass counts digits per tile, scans the counts digit-major
// (digit, tile) and scatters: the tile is ordered by digit in LDS (each round
// of 256 keys ranks a key among the equal digits of its wave with one ballot
// per digit bit, adds the counts of the waves before it and the digit's
// running count) and then written out run by run.
#include <algorithm>

#include "ll_gpu_internal.h"

namespace llgpu {

namespace {

constexpr int kB = 256;
constexpr int kRounds = 16;
constexpr uint64_t kTile = uint64_t(kB) * kRounds;

struct SrcU32 {
  const uint32_t* p;
  __device__ uint64_t operator()(uint64_t i) const { return p[i]; }
};
struct SrcU64 {
  const uint64_t* p;
  __device__ uint64_t operator()(uint64_t i) const { return p[i]; }
};
// val[idx[i]]
struct SrcGatherU32 {
  const uint32_t* idx;
  const uint32_t* val;
  __device__ uint64_t operator()(uint64_t i) const { return val[idx[i]]; }
};
// 1 where a sorted run of equal values starts
struct SrcRunStart {
  const uint64_t* p;
  __device__ uint64_t operator()(uint64_t i) const { return (i == 0 || p[i] != p[i - 1]) ? 1 : 0; }
};

// Exclusive prefix of x over the workgroup (all threads call); *total = sum.
__device__ inline uint64_t block_exclusive_scan(uint64_t x, uint64_t* total) {
  __shared__ uint64_t wsum[kB / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t inc = x;
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(inc, o, 64);
    if (lane >= o) inc += y;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint64_t wbase = 0, t = 0;
  for (int i = 0; i < kB / 64; i++) {
    const uint64_t w = wsum[i];
    if (i < wave) wbase += w;
    t += w;
  }
  __syncthreads();  // wsum is reused by the next call
  *total = t;
  return wbase + inc - x;
}

template <typename Src>
__global__ __launch_bounds__(kB) void k_tile_sums(Src src, uint64_t n, uint64_t* sums) {
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  uint64_t x = 0;
  for (int r = 0; r < kRounds; r++) {
    const uint64_t i = base + uint64_t(r) * kB + threadIdx.x;
    if (i < n) x += src(i);
  }
  uint64_t tot;
  (void)block_exclusive_scan(x, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// out[i] = prefix[tile] + exclusive prefix of src within the tile; the last
// tile also writes out[n] = total.
template <typename Src>
__global__ __launch_bounds__(kB) void k_tile_scan(Src src, uint64_t n, const uint64_t* prefix,
                                                  uint64_t* out) {
  uint64_t carry = prefix ? prefix[blockIdx.x] : 0;
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  for (int r = 0; r < kRounds; r++) {
    if (base + uint64_t(r) * kB >= n) break;  // uniform over the workgroup
    const uint64_t i = base + uint64_t(r) * kB + threadIdx.x;
    const uint64_t x = i < n ? src(i) : 0;
    uint64_t tot;
    const uint64_t ex = block_exclusive_scan(x, &tot);
    if (i < n) out[i] = carry + ex;
    carry += tot;
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = carry;
}

uint64_t ntiles_of(uint64_t n) { return (n + kTile - 1) / kTile; }

template <typename Src>
void scan_rec(Src src, uint64_t n, uint64_t* out, uint64_t* tmp, hipStream_t s) {
  if (n <= kTile) {
    k_tile_scan<<<1, kB, 0, s>>>(src, n, nullptr, out);
    LLG_HIP(hipGetLastError());
    return;
  }
  const uint64_t nt = ntiles_of(n);
  uint64_t* sums = tmp;
  uint64_t* pre = tmp + nt;  // nt + 1 entries
  k_tile_sums<<<unsigned(nt), kB, 0, s>>>(src, n, sums);
  LLG_HIP(hipGetLastError());
  scan_rec(SrcU64{sums}, nt, pre, tmp + 2 * nt + 1, s);
  k_tile_scan<<<unsigned(nt), kB, 0, s>>>(src, n, pre, out);
  LLG_HIP(hipGetLastError());
}

template <typename Src>
void scan_alloc(Src src, uint64_t n, uint64_t* out, hipStream_t s) {
  DevBuf t;
  const size_t b = scan_tmp_bytes(n);
  if (b) t.alloc(b, s);
  scan_rec(src, n, out, t.as<uint64_t>(), s);
}

// --- radix sort ----------------------------------------------------------------

// Counts of every digit position (P = sizeof(K) bytes) over all keys.
template <typename K>
__global__ __launch_bounds__(kB) void k_rs_all_hist(const K* keys, uint64_t n,
                                                    unsigned long long* ghist) {
  constexpr int P = sizeof(K);
  __shared__ uint32_t h[P][256];
  for (int p = 0; p < P; p++) h[p][threadIdx.x] = 0;
  __syncthreads();
  for (uint64_t i = uint64_t(blockIdx.x) * kB + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * kB) {
    const K k = keys[i];
#pragma unroll
    for (int p = 0; p < P; p++) atomicAdd(&h[p][uint32_t(k >> (8 * p)) & 255u], 1u);
  }
  __syncthreads();
  for (int p = 0; p < P; p++)
    if (h[p][threadIdx.x]) atomicAdd(&ghist[p * 256 + threadIdx.x], (unsigned long long)h[p][threadIdx.x]);
}

// Per-tile digit counts, digit-major: th[d * ntiles + tile] for the
// 2^bits digits of (key >> shift).
template <typename K>
__global__ __launch_bounds__(kB) void k_rs_tile_hist(const K* keys, uint64_t n, int shift,
                                                     int bits, uint32_t* th, uint64_t ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t dm = (1u << bits) - 1;
  const uint64_t base = uint64_t(blockIdx.x) * kTile;
  K k[kRounds];  // every key of the thread requested before any is counted
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const uint64_t i = base + uint64_t(r) * kB + threadIdx.x;
    k[r] = i < n ? keys[i] : K(0);
  }
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const uint64_t i = base + uint64_t(r) * kB + threadIdx.x;
    if (i < n) atomicAdd(&h[uint32_t(k[r] >> shift) & dm], 1u);
  }
  __syncthreads();
  if (threadIdx.x <= dm) th[uint64_t(threadIdx.x) * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Scatter of one tile.  The tile's keys are first ordered by digit in LDS,
// stably: each round of 256 keys ranks a key among the equal digits of its
// wave with one ballot per digit bit, then adds the counts of the waves
// before it and the digit's running count in the tile.  They are then written
// out in that order, so consecutive lanes write consecutive positions of one
// digit's run (a tile's run per digit is written as one contiguous piece).
// With lo16 given, the pass writes no keys and each value as its two 16-bit
// halves (lo16, hi16) instead.
template <typename K, bool kVals>
__global__ __launch_bounds__(kB) void k_rs_scatter(const K* kin, const uint32_t* vin, K* kout,
                                                   uint32_t* vout, uint64_t n, int shift, int bits,
                                                   const uint32_t* th, const uint64_t* tofs,
                                                   uint64_t ntiles, uint16_t* lo16,
                                                   uint16_t* hi16) {
  __shared__ K skey[kTile];
  __shared__ uint32_t sval[kVals ? kTile : 1];
  __shared__ uint64_t gbase[256];  // first output position of the tile's run per digit
  __shared__ uint32_t lbase[256];  // first LDS position of the digit in the tile
  __shared__ uint32_t wcnt[kB / 64][256];  // per wave: digit counts, then their offsets
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint32_t dm = (1u << bits) - 1;
  // neighbouring tiles write neighbouring pieces of every digit's run: keep
  // them on one XCD
  const uint64_t tile = xcd_block(blockIdx.x, gridDim.x);
  const bool mine = uint32_t(t) <= dm;
  const uint64_t cnt = mine ? th[uint64_t(t) * ntiles + tile] : 0;
  uint64_t tot;
  lbase[t] = uint32_t(block_exclusive_scan(cnt, &tot));
  gbase[t] = mine ? tofs[uint64_t(t) * ntiles + tile] : 0;
  for (int w = 0; w < kB / 64; w++) wcnt[w][t] = 0;
  __syncthreads();
  const uint64_t base = tile * kTile;
  const uint32_t nt = uint32_t(min(kTile, n - base));
  // wave w owns the tile's keys [w * 64 * kRounds, (w + 1) * 64 * kRounds) and
  // ranks them in order with its own running digit counts (no barrier); the
  // waves' counts then give each wave's offset per digit, which keeps the
  // tile's order within a digit.  Every key (and value) of the thread is
  // requested before any is ranked.
  const uint32_t w0 = uint32_t(wave) * 64u * kRounds;
  K kr[kRounds];
  uint32_t vr[kVals ? kRounds : 1];
  uint32_t lp[kRounds];  // rank among the wave's keys of the same digit
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const uint32_t j = w0 + uint32_t(r) * 64u + uint32_t(lane);
    kr[r] = j < nt ? kin[base + j] : K(0);
    if (kVals) vr[r] = j < nt ? vin[base + j] : 0u;
  }
  const uint64_t lt = (uint64_t(1) << lane) - 1;
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const uint32_t j = w0 + uint32_t(r) * 64u + uint32_t(lane);
    const bool valid = j < nt;
    const uint32_t d = uint32_t(kr[r] >> shift) & dm;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; b++) {
      const uint64_t m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    // a wave's LDS accesses take effect in program order: every lane reads
    // the count before the digit's first lane writes it back
    const uint32_t c = valid ? wcnt[wave][d] : 0u;
    const uint64_t below = peers & lt;
    lp[r] = c + uint32_t(__popcll(below));
    if (valid && below == 0) wcnt[wave][d] = c + uint32_t(__popcll(peers));
  }
  __syncthreads();
  if (mine) {
    uint32_t acc = 0;
    for (int w = 0; w < kB / 64; w++) {
      const uint32_t c = wcnt[w][t];
      wcnt[w][t] = lbase[t] + acc;
      acc += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kRounds; r++) {
    const uint32_t j = w0 + uint32_t(r) * 64u + uint32_t(lane);
    if (j < nt) {
      const uint32_t d = uint32_t(kr[r] >> shift) & dm;
      const uint32_t pos = wcnt[wave][d] + lp[r];
      skey[pos] = kr[r];
      if (kVals) sval[pos] = vr[r];
    }
  }
  __syncthreads();
  for (uint32_t j = uint32_t(t); j < nt; j += kB) {
    const K key = skey[j];
    const uint32_t d = uint32_t(key >> shift) & dm;
    const uint64_t pos = gbase[d] + (j - lbase[d]);
    if (kVals && lo16) {
      const uint32_t v = sval[j];
      lo16[pos] = uint16_t(v & 0xffffu);
      hi16[pos] = uint16_t(v >> 16);
    } else {
      kout[pos] = key;
      if (kVals) vout[pos] = sval[j];
    }
  }
}

// probe: count the digit positions first and skip the bits that never vary;
// without it, bits [0, end_bit) are sorted as given.  lo16 / hi16: the last
// pass writes the values' halves there instead of keys and values (kout and
// vout then only serve as working memory).
template <typename K, bool kVals>
void radix_sort(const K* kin, K* kout, const uint32_t* vin, uint32_t* vout, uint64_t n,
                int end_bit, hipStream_t s, bool probe = true, uint16_t* lo16 = nullptr,
                uint16_t* hi16 = nullptr) {
  if (n == 0) return;
  LLG_CHECK(!lo16 || (!probe && end_bit > 0), LL_GPU_EINVAL,
            "split output needs a sort of known, non-empty key width");
  constexpr int P = sizeof(K);
  const int npass_max = std::min(P, (end_bit + 7) / 8);
  struct Pass {
    int shift, bits;
  };
  std::vector<Pass> passes;
  // bits [lo, hi) in the fewest even passes of at most 8 bits
  auto cut = [&](int lo, int hi) {
    const int nbits = hi - lo;
    const int np = (nbits + 7) / 8;
    for (int k = 0, at = lo; k < np; k++) {
      const int w = (nbits - (at - lo) + (np - k) - 1) / (np - k);
      passes.push_back({at, w});
      at += w;
    }
  };
  if (!probe) {
    cut(0, std::min(end_bit, 8 * P));
  } else {
    // which passes move anything
    DevBuf gh;
    gh.alloc(P * 256 * sizeof(unsigned long long), s);
    LLG_HIP(hipMemsetAsync(gh.p, 0, gh.bytes, s));
    {
      int dev = 0, cus = 0;
      LLG_HIP(hipGetDevice(&dev));
      LLG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      const uint64_t g = std::min<uint64_t>((n + kB - 1) / kB, uint64_t(std::max(cus, 1)) * 8);
      k_rs_all_hist<K><<<unsigned(g), kB, 0, s>>>(kin, n, gh.as<unsigned long long>());
      LLG_HIP(hipGetLastError());
    }
    std::vector<unsigned long long> h(P * 256);
    LLG_HIP(hipMemcpyAsync(h.data(), gh.p, gh.bytes, hipMemcpyDeviceToHost, s));
    LLG_HIP(hipStreamSynchronize(s));
    // the bits to sort: bytes whose value differs between keys, cut into
    // the fewest even passes of at most 8 bits within each run of such bytes
    for (int p = 0; p < npass_max;) {
      bool trivial = false;
      for (int d = 0; d < 256; d++)
        if (h[p * 256 + d] == n) trivial = true;
      if (trivial) {
        p++;
        continue;
      }
      int q = p;  // bytes p..q-1 all vary
      while (q < npass_max) {
        bool t2 = false;
        for (int d = 0; d < 256; d++)
          if (h[q * 256 + d] == n) t2 = true;
        if (t2) break;
        q++;
      }
      int lo = 8 * p, hi = std::min(8 * q, end_bit);
      // the highest byte may carry fewer significant bits than 8
      if (q == npass_max && hi > lo) {
        const int top = 8 * (q - 1);
        int tb = 0;
        for (int d = 0; d < 256; d++)
          if (h[(q - 1) * 256 + d]) tb = std::max(tb, 32 - __builtin_clz(unsigned(d) | 1u));
        hi = std::min(hi, top + tb);
      }
      cut(lo, hi);
      p = q;
    }
  }
  const uint64_t nt = ntiles_of(n);
  if (passes.empty()) {
    LLG_HIP(hipMemcpyAsync(kout, kin, n * sizeof(K), hipMemcpyDeviceToDevice, s));
    if (kVals) LLG_HIP(hipMemcpyAsync(vout, vin, n * 4, hipMemcpyDeviceToDevice, s));
    return;
  }
  DevBuf th, tofs, kt, vt, stmp;
  th.alloc(256 * nt * 4, s);
  tofs.alloc((256 * nt + 1) * 8, s);
  const size_t sb = scan_tmp_bytes(256 * nt);
  if (sb) stmp.alloc(sb, s);
  if (passes.size() > 1) {
    kt.alloc(n * sizeof(K), s);
    if (kVals) vt.alloc(n * 4, s);
  }
  const int np = int(passes.size());
  const K* ksrc = kin;
  const uint32_t* vsrc = vin;
  for (int j = 0; j < np; j++) {
    // the last pass lands in the output; earlier ones alternate so that it does
    const bool to_out = ((np - 1 - j) % 2) == 0;
    K* kdst = to_out ? kout : kt.as<K>();
    uint32_t* vdst = to_out ? vout : vt.as<uint32_t>();
    const int shift = passes[j].shift, bits = passes[j].bits;
    const uint64_t nd = uint64_t(1) << bits;
    k_rs_tile_hist<K><<<unsigned(nt), kB, 0, s>>>(ksrc, n, shift, bits, th.as<uint32_t>(), nt);
    LLG_HIP(hipGetLastError());
    scan_rec(SrcU32{th.as<uint32_t>()}, nd * nt, tofs.as<uint64_t>(), stmp.as<uint64_t>(), s);
    const bool split = lo16 && j == np - 1;
    k_rs_scatter<K, kVals><<<unsigned(nt), kB, 0, s>>>(ksrc, vsrc, kdst, vdst, n, shift, bits,
                                                       th.as<uint32_t>(), tofs.as<uint64_t>(), nt,
                                                       split ? lo16 : nullptr,
                                                       split ? hi16 : nullptr);
    LLG_HIP(hipGetLastError());
    ksrc = kdst;
    vsrc = vdst;
  }
}

__global__ void k_unique_scatter(const uint64_t* in, const uint64_t* pos, uint64_t n,
                                 uint64_t* out) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && (i == 0 || in[i] != in[i - 1])) out[pos[i]] = in[i];
}

}  // namespace

size_t scan_tmp_bytes(uint64_t n) {
  size_t w = 0;
  while (n > kTile) {
    const uint64_t nt = ntiles_of(n);
    w += 2 * nt + 1;
    n = nt;
  }
  return w * 8;
}

void scan_u32_to_u64_ws(const uint32_t* in, uint64_t* out, uint64_t n, void* tmp,
                        hipStream_t s) {
  scan_rec(SrcU32{in}, n, out, static_cast<uint64_t*>(tmp), s);
}

void scan_gather_u32_to_u64_ws(const uint32_t* idx, const uint32_t* val, uint64_t* out,
                               uint64_t n, void* tmp, hipStream_t s) {
  scan_rec(SrcGatherU32{idx, val}, n, out, static_cast<uint64_t*>(tmp), s);
}

void scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint64_t n, hipStream_t s) {
  scan_alloc(SrcU32{in}, n, out, s);
}

void scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t s) {
  scan_alloc(SrcU64{in}, n, out, s);
}

void sort_pairs_u64_u32(const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in,
                        uint32_t* vals_out, uint64_t n, int end_bit, hipStream_t s) {
  radix_sort<uint64_t, true>(keys_in, keys_out, vals_in, vals_out, n, end_bit, s);
}

void sort_pairs_u32_u32(const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in,
                        uint32_t* vals_out, uint64_t n, int end_bit, hipStream_t s) {
  radix_sort<uint32_t, true>(keys_in, keys_out, vals_in, vals_out, n, end_bit, s);
}

void sort_pairs_u32_u32_split(const uint32_t* keys_in, uint32_t* keys_work,
                              const uint32_t* vals_in, uint32_t* vals_work, uint64_t n,
                              int end_bit, uint16_t* lo16, uint16_t* hi16, hipStream_t s) {
  radix_sort<uint32_t, true>(keys_in, keys_work, vals_in, vals_work, n, end_bit, s, false, lo16,
                             hi16);
}

void sort_keys_u64(const uint64_t* keys_in, uint64_t* keys_out, uint64_t n, int end_bit,
                   hipStream_t s) {
  radix_sort<uint64_t, false>(keys_in, keys_out, nullptr, nullptr, n, end_bit, s);
}

uint64_t unique_u64(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t s) {
  if (n == 0) return 0;
  DevBuf pos;
  pos.alloc((n + 1) * 8, s);
  scan_alloc(SrcRunStart{in}, n, pos.as<uint64_t>(), s);
  k_unique_scatter<<<grid_for(n, 256), 256, 0, s>>>(in, pos.as<uint64_t>(), n, out);
  LLG_HIP(hipGetLastError());
  uint64_t h = 0;
  LLG_HIP(hipMemcpyAsync(&h, pos.as<uint64_t>() + n, sizeof h, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace llgpu
