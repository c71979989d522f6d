// Row blocks for load-balanced row reductions over a snapshot CSR.
//
// R-MAT degree distributions are heavy-tailed, so a fixed rows-per-block split
// leaves most lanes idle on short rows and serialises hubs.  Rows are grouped
// into work items instead (the CSR-adaptive scheme):
//   * a "stream" item is a run of short rows (each < cap/2 entries) whose
//     starts fall into one cap-sized bucket of the edge array, at most 256 rows:
//     its entries (< 1.5*cap) are gathered into LDS with coalesced loads and
//     each row is then reduced by one thread;
//   * a "long" row (>= cap/2 entries) gets ceil(deg / piece) items, each
//     reduced by a whole workgroup; rows with several pieces are combined in
//     piece order by a fix-up pass, so results do not depend on scheduling.
// Items are built with two flag/scan passes and cached with the snapshot.
#include "ll_gpu_internal.h"
#include "rowblocks.h"

namespace llgpu {

namespace {

__device__ inline bool is_long(const uint64_t* ofs, uint64_t r, uint32_t cap) {
  return ofs[r + 1] - ofs[r] >= cap / 2;
}

__device__ inline bool row_starts_item(const uint64_t* ofs, uint64_t r, uint32_t cap) {
  if (r == 0 || (r % kRowsPerItem) == 0) return true;
  if (is_long(ofs, r, cap) || is_long(ofs, r - 1, cap)) return true;
  return ofs[r] / cap != ofs[r - 1] / cap;
}

__global__ void k_item_count(const uint64_t* ofs, uint64_t V, uint32_t cap, uint32_t piece,
                             uint32_t* cnt) {
  uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= V) return;
  uint32_t c = 0;
  if (row_starts_item(ofs, r, cap)) {
    uint64_t deg = ofs[r + 1] - ofs[r];
    c = is_long(ofs, r, cap) ? uint32_t((deg + piece - 1) / piece) : 1u;
  }
  cnt[r] = c;
}

__global__ void k_item_fill(const uint64_t* ofs, uint64_t V, uint32_t cap, uint32_t piece,
                            const uint64_t* pos, RowItem* items) {
  uint64_t r = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= V || !row_starts_item(ofs, r, cap)) return;
  uint64_t p = pos[r];
  if (is_long(ofs, r, cap)) {
    uint64_t deg = ofs[r + 1] - ofs[r];
    uint32_t np = uint32_t((deg + piece - 1) / piece);
    for (uint32_t i = 0; i < np; i++) items[p + i] = RowItem{uint32_t(r), uint32_t(r + 1), i, np};
  } else {
    // row_end is patched to the next item's first row by k_item_ends
    items[p] = RowItem{uint32_t(r), 0u, 0u, 0u};
  }
}

__global__ void k_item_ends(RowItem* items, uint64_t n, uint64_t V) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n || items[i].npieces != 0) return;
  items[i].row_end = i + 1 < n ? items[i + 1].row_begin : uint32_t(V);
}

// slot of each item of a row split over several pieces (partials buffer index)
__global__ void k_split_flags(const RowItem* items, uint64_t n, uint32_t* f) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) f[i] = items[i].npieces > 1 ? 1u : 0u;
}

__global__ void k_split_slots(const uint32_t* f, const uint64_t* pos, uint64_t n, uint32_t* slot) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) slot[i] = f[i] ? uint32_t(pos[i]) : 0xffffffffu;
}

}  // namespace

void build_rowblocks(ll_gpu_ctx* ctx, const Csr& c, uint32_t cap, uint32_t piece,
                     RowBlocks& rb) {
  LLG_CHECK(cap >= 64 && cap <= kMaxCap && (cap & (cap - 1)) == 0, LL_GPU_EINVAL,
            "row-block capacity must be a power of two in [64, 4096]");
  LLG_CHECK(piece >= 64, LL_GPU_EINVAL, "row piece must be >= 64 entries");
  rb = RowBlocks();
  hipStream_t s = ctx->stream;
  const uint64_t V = c.V;
  DevBuf cnt, pos;
  cnt.alloc(V * 4, s);
  pos.alloc((V + 1) * 8, s);
  k_item_count<<<grid_for(V, 256), 256, 0, s>>>(c.ofs.as<uint64_t>(), V, cap, piece,
                                                cnt.as<uint32_t>());
  LLG_HIP(hipGetLastError());
  scan_u32_to_u64(cnt.as<uint32_t>(), pos.as<uint64_t>(), V, s);
  uint64_t n = 0;
  LLG_HIP(hipMemcpyAsync(&n, pos.as<uint64_t>() + V, 8, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
  LLG_CHECK(n < (uint64_t(1) << 31), LL_GPU_ERANGE, "too many row blocks");
  rb.items.alloc(std::max<uint64_t>(n, 1) * sizeof(RowItem), s);
  k_item_fill<<<grid_for(V, 256), 256, 0, s>>>(c.ofs.as<uint64_t>(), V, cap, piece,
                                               pos.as<uint64_t>(), rb.items.as<RowItem>());
  k_item_ends<<<grid_for(n, 256), 256, 0, s>>>(rb.items.as<RowItem>(), n, V);
  LLG_HIP(hipGetLastError());
  // partial-buffer slots for rows split into several pieces
  DevBuf f, fpos;
  f.alloc(std::max<uint64_t>(n, 1) * 4, s);
  fpos.alloc((n + 1) * 8, s);
  k_split_flags<<<grid_for(n, 256), 256, 0, s>>>(rb.items.as<RowItem>(), n, f.as<uint32_t>());
  scan_u32_to_u64(f.as<uint32_t>(), fpos.as<uint64_t>(), n, s);
  uint64_t nsplit = 0;
  LLG_HIP(hipMemcpyAsync(&nsplit, fpos.as<uint64_t>() + n, 8, hipMemcpyDeviceToHost, s));
  LLG_HIP(hipStreamSynchronize(s));
  rb.slots.alloc(std::max<uint64_t>(n, 1) * 4, s);
  k_split_slots<<<grid_for(n, 256), 256, 0, s>>>(f.as<uint32_t>(), fpos.as<uint64_t>(), n,
                                                 rb.slots.as<uint32_t>());
  LLG_HIP(hipGetLastError());
  rb.valid = true;
  rb.cap = cap;
  rb.piece = piece;
  rb.n = n;
  rb.nsplit = nsplit;
}

const RowBlocks& ensure_rowblocks(ll_gpu_graph* g, uint32_t dir, int level, uint32_t cap,
                                  uint32_t piece) {
  const Csr& c = ensure_csr(g, dir, level);
  for (RowBlocks& rb : g->rb[dir])
    if (rb.valid && rb.cap == cap && rb.piece == piece) return rb;
  g->rb[dir].emplace_back();
  build_rowblocks(g->ctx, c, cap, piece, g->rb[dir].back());
  return g->rb[dir].back();
}

}  // namespace llgpu
