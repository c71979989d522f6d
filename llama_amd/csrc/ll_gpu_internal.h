// Internal structures shared by the HIP translation units of libll_gpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <map>
#include <memory>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "ll_gpu.h"

namespace llgpu {

// LLAMA edge-ID / edge-table word encoding (ll_mlcsr_helpers.h:55-107).
constexpr int kBitsIndex = 38;
constexpr uint64_t kIndexMask = (uint64_t(1) << kBitsIndex) - 1;
constexpr uint32_t kLevelMask = (1u << 26) - 1;

struct VtEntry {
  int64_t start;
  uint32_t len;
  uint32_t degree;
};
static_assert(sizeof(VtEntry) == 16, "VT entry must be 16 bytes");

// Vertex tables are held in pages of kVtPageEntries entries, as the reference's
// copy-on-write vertex table (ll_mem_array_swcow, LL_ENTRIES_PER_PAGE): each
// level has a table of page addresses, and pages a level did not write are
// the previous level's.
constexpr uint32_t kVtPageBits = 9;
constexpr uint64_t kVtPageEntries = uint64_t(1) << kVtPageBits;
// A load through a pointer read from a table (page tables, per-level edge
// tables): the memory is device global memory, so the load is issued as a
// global access, which waits only on the vector memory counter (a generic
// access also waits on the LDS counter).
template <typename T>
__device__ inline T ld_global(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

struct VtPages {
  const VtEntry* const* pages;
  __device__ inline VtEntry operator[](uint64_t v) const {
    const uint64_t* p =
        reinterpret_cast<const uint64_t*>(pages[v >> kVtPageBits] + (v & (kVtPageEntries - 1)));
    const uint64_t a = ld_global(p), b = ld_global(p + 1);
    VtEntry e;
    e.start = int64_t(a);
    e.len = uint32_t(b);
    e.degree = uint32_t(b >> 32);
    return e;
  }
};

__host__ __device__ inline uint32_t edge_level(uint64_t e) { return uint32_t(e >> kBitsIndex); }
__host__ __device__ inline uint64_t edge_index(uint64_t e) { return e & kIndexMask; }
__host__ __device__ inline uint32_t word_maxlev(uint64_t w) {
  return uint32_t(w >> kBitsIndex) & kLevelMask;
}

// Error carried out of helpers; converted to a status at the C boundary.
struct Error {
  int code;
  std::string msg;
};

#define LLG_HIP(expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      throw ::llgpu::Error{LL_GPU_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define LLG_CHECK(cond, code, msg)                       \
  do {                                                   \
    if (!(cond)) throw ::llgpu::Error{(code), (msg)};    \
  } while (0)

// Stream-ordered device buffer (hipMallocAsync / hipFreeAsync on the context stream).
// Two timing events, destroyed on scope exit (also when a helper throws).
struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  EventPair() {
    LLG_HIP(hipEventCreate(&a));
    const hipError_t e = hipEventCreate(&b);
    if (e != hipSuccess) {
      (void)hipEventDestroy(a);
      a = nullptr;
      LLG_HIP(e);
    }
  }
  EventPair(const EventPair&) = delete;
  EventPair& operator=(const EventPair&) = delete;
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
};

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipStream_t s = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes), s(o.s) { o.p = nullptr; o.bytes = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; s = o.s; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n, hipStream_t stream) {
    release();
    s = stream;
    bytes = n;
    if (n == 0) return;
    hipError_t e = hipMallocAsync(&p, n, stream);
    if (e != hipSuccess) {
      // The device's stream-ordered pool keeps freed memory (the context sets
      // no release threshold), in blocks that may each be smaller than this
      // request: hand them back to the device and try once more.
      (void)hipGetLastError();
      int dev = 0;
      hipMemPool_t pool = nullptr;
      if (hipStreamSynchronize(stream) == hipSuccess && hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess && pool &&
          hipMemPoolTrimTo(pool, 0) == hipSuccess)
        e = hipMallocAsync(&p, n, stream);
    }
    if (e != hipSuccess) {
      p = nullptr;
      bytes = 0;
      (void)hipGetLastError();
      throw Error{LL_GPU_ENOMEM, "device allocation of " + std::to_string(n) + " bytes failed"};
    }
  }
  void release() {
    if (p) (void)hipFreeAsync(p, s);
    p = nullptr;
    bytes = 0;
  }
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Per-kernel device time accumulated from HIP events (profiling mode).
struct KernelTimer {
  struct Rec { std::string name; hipEvent_t a, b; };
  bool on = false;
  std::vector<Rec> pending;
  std::vector<hipEvent_t> pool;
  std::map<std::string, std::pair<uint64_t, double>> acc;
};

}  // namespace llgpu

constexpr size_t kReadbackBytes = size_t(64) << 10;

namespace llgpu {
// Size of the fixed device array ll_gpu_bfs_noalloc works in (one per GPU).
constexpr size_t kScratchBytes = size_t(64) << 20;
// Address of that array on `device` (the current device), looked up once.
void* scratch_setup(int device);
}  // namespace llgpu

struct ll_gpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int num_cus = 0;
  llgpu::KernelTimer timer;
  void* pinned = nullptr;      // staging for uploads
  size_t pinned_bytes = 0;
  hipEvent_t pinned_ev[2] = {nullptr, nullptr};
  void* readback = nullptr;    // small pinned buffer for per-level counter reads
  std::map<std::string, int> tuned;    // autotuner decisions
  std::map<std::string, long> options; // ll_gpu_set_option knobs
  void* scratch = nullptr;     // device address of the library's fixed array on `device`
};

namespace llgpu {
// Status of the exception in flight inside an entry point's catch (...):
// std::bad_alloc -> LL_GPU_ENOMEM, anything else -> LL_GPU_ESTATE, with the
// message kept in the context.  No exception crosses the C ABI.
inline int api_exception(ll_gpu_ctx* c) noexcept {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    if (c) c->err = "out of host memory";
    return LL_GPU_ENOMEM;
  } catch (const std::exception& e) {
    if (c) c->err = e.what();
    return LL_GPU_ESTATE;
  } catch (...) {
    if (c) c->err = "unknown exception";
    return LL_GPU_ESTATE;
  }
}
}  // namespace llgpu

namespace llgpu {

// Compact CSR of one direction at one snapshot level (iterator order).
struct Csr {
  bool valid = false;
  int level = -1;
  uint64_t V = 0, E = 0, stored = 0, segments = 0;
  DevBuf ofs;   // uint64_t[V+1]
  DevBuf adj;   // uint32_t[E]
  DevBuf w;     // float[E]: out-edge weights when every level up to `level` has them
  double ms = 0;
  mutable double w_mean = -1;  // mean of w, computed on first use (weighted SSSP)
};

// Row blocks of a CSR for the load-balanced row-reduction kernels.
struct RowItem {
  uint32_t row_begin, row_end, piece, npieces;
};
struct RowBlocks {
  bool valid = false;
  uint32_t cap = 0, piece = 0;
  uint64_t n = 0;       // items
  uint64_t nsplit = 0;  // items of rows split into several pieces
  DevBuf items;         // RowItem[n]
  DevBuf slots;         // uint32_t[n]: partials slot of split-row items, else ~0
};

// Column-segmented in-CSR for the PageRank pull (built by build_pr_segments):
// contribution slots cut into S segments of 2^zb; rows of `pairs` are the
// (row, segment) cells with edges, segment-major, rows ascending per segment.
constexpr uint32_t kMaxSegments = 64;
struct PrSegments {
  bool valid = false;
  int level = -1;
  uint32_t zb = 0, S = 0;
  uint32_t req_zb = 0; // segment bits asked for (zb may be larger)
  bool slots = false;  // sources are degree-ordered slots (else vertex IDs)
  uint64_t P = 0;      // pairs
  uint64_t nb = 0;     // merge row blocks of 256 rows
  Csr pairs;           // ofs: uint64_t[P+1] into adj = source slots (uint32_t[E])
  DevBuf pair_row;     // uint32_t[P]
  DevBuf pseg;         // uint32_t[S+1]: first pair of each segment
  DevBuf mt;           // uint32_t[S*(nb+1)]: first pair of segment s at row block b
  RowBlocks rb;        // row blocks of `pairs`
};

// Hot/cold split of the PageRank in-CSR (built by build_pr_hotcold): per row,
// the in-edges whose source slot is below 2^hb (hot) and the others (cold),
// each in snapshot order, as two CSRs over all V rows.
struct PrHotCold {
  bool valid = false;
  int level = -1;
  uint32_t hb = 0;
  Csr hot, cold;
  RowBlocks hot_rb, cold_rb;
};

// Binned PageRank layout of the out-CSR (built by build_pr_bins): vertices in
// bins of 2^bb; every out-edge owns one slot of the binned value array, the
// slots of a destination bin grouped by source bin (chunks), in out-edge
// order within a chunk.
struct PrBins {
  bool valid = false;
  int level = -1;
  uint32_t bb = 0, nbins = 0;
  DevBuf ldst;    // uint16_t[E]: slot -> destination offset in its bin
  DevBuf sidx;    // uint16_t[E]: slot -> source offset in its bin
  DevBuf bstart;  // uint64_t[nbins+1]: first slot of each destination bin
  DevBuf ct;      // uint2[nbins * nbins]: [source bin][destination bin] = {first slot, slots}
};

struct DirMirror {
  std::vector<DevBuf> vt;   // per level: the pages the level wrote (all of them for a full push)
  std::vector<std::vector<const VtEntry*>> vt_pages;  // per level: device address of every page
  std::vector<DevBuf> vt_ptab;  // per level: vt_pages on the device (VtPages)
  std::vector<DevBuf> et;   // uint64_t[words] per level
  std::vector<uint64_t> et_words;
  DevBuf et_table;          // const uint64_t*[levels] on device
  DevBuf et_words_dev;      // uint64_t[levels]: words per level (bounds of the walk)
  std::vector<DevBuf> wt;   // out direction: float weight per ET word per level (may be empty)
  DevBuf wt_table;          // const float*[levels] on device (null for levels without)
};

}  // namespace llgpu

struct ll_gpu_graph {
  ll_gpu_ctx* ctx = nullptr;
  uint64_t V = 0;
  uint32_t flags = 0;
  llgpu::DirMirror dir[2];
  uint64_t upload_bytes = 0;
  uint32_t min_level = 0;              // LL_MIN_LEVEL window (ll_gpu_graph_set_min_level)
  // LL_SLCSR graphs (ll_gpu_graph_push_slcsr): one level given as a plain CSR
  bool slcsr = false;
  llgpu::Csr sl[2];
  std::vector<char> level_deleted;     // per level: storage freed (ll_gpu_graph_delete_level)
  double last_upload_ms = 0;
  // snapshot caches
  int deg_level[2] = {-1, -1};
  llgpu::DevBuf deg[2];       // uint32_t[V]
  llgpu::Csr csr[2];
  std::list<llgpu::RowBlocks> rb[2];  // per direction, one per (cap, piece)
  // derived structures (built on demand per snapshot level)
  int sym_level = -1;
  llgpu::DevBuf sym_ofs, sym_adj;  // degree-oriented DAG for triangle counting
  llgpu::DevBuf sym_mult;          // uint8 per arc: directed live edges joining its ends (1 or 2)
  uint64_t sym_edges = 0;
  // DAG rows by length (workgroup rows from the front, global-memory rows from
  // the back), W and the largest row element count, built with the DAG
  llgpu::DevBuf sym_rows;
  uint64_t sym_nblock = 0, sym_nglobal = 0, sym_work = 0, sym_maxrow = 0;
  // rows of the in-CSR with no live in-edge (bitmap, excluded from BFS), then
  // the rows with more than 32 (bitmap, walked by lane groups bottom-up)
  int dead_level = -1;
  llgpu::DevBuf dead_in;
  // complement of the vertices reachable from the hub (bitmap), their out-edges
  int reach_level = -1;
  llgpu::DevBuf reach_notr;
  uint64_t reach_e = 0;
  uint32_t reach_hub = 0;
  // PageRank locality order of the in-CSR (option pr_reorder)
  int pr_level = -1;
  llgpu::DevBuf pr_inv, pr_adj;
  llgpu::DevBuf pr_sadj, pr_sperm;  // pr_adj with every row-block item sorted by slot
  uint32_t pr_sort_cap = 0;         // row-block capacity the sorted copy was built for
  llgpu::PrSegments pr_seg;   // column segments of the in-CSR in slot order
  llgpu::PrHotCold pr_hc;     // hot/cold split of the in-CSR in slot order
  llgpu::PrBins pr_bins;      // binned PageRank layout of the out-CSR
  // weighted in-CSR (rows = heads, sources ascending) + map to out-edge positions
  int w_level = -1;
  llgpu::Csr wcsr;
  llgpu::DevBuf wperm;        // uint32_t[E]: weighted in-edge -> out-edge position
  std::list<llgpu::RowBlocks> wrb;
};

namespace llgpu {

// Launch helpers that record per-kernel device time when profiling is on.
struct TimedLaunch {
  ll_gpu_ctx* ctx;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  TimedLaunch(ll_gpu_ctx* c, const char* n);
  ~TimedLaunch();
};

struct BfsResult {
  uint64_t reached = 0, examined = 0, traversed = 0;
  int levels = 0;
};
// Direction-optimising BFS from `root` over `fwd` (top-down) / `rev`
// (bottom-up); `fdeg` = degrees of `fwd`; `excluded` (bitmap or null) is never
// entered; `visited` receives the reached set (plus excluded/padding bits);
// `lv` (int32[V] or null) the levels.
BfsResult bfs_core(ll_gpu_graph* g, const Csr& fwd, const Csr& rev, const uint32_t* fdeg,
                   uint32_t root, const uint64_t* excluded, DevBuf& visited, int32_t* lv);
// Bitmap of the in-CSR rows (of `in`, the snapshot's in-edges) with no
// entry: no search from another vertex can reach them, so single-source
// searches pass it as `excluded` and the bottom-up steps never open them.
// Cached per snapshot level.
const uint64_t* ensure_dead_rows(ll_gpu_graph* g, const Csr& in);

// What a single-source search reads besides its working memory: both CSRs
// of the snapshot, the out-degrees, the dead-row bitmap and the reach set.
struct BfsPrepared {
  const Csr* out = nullptr;
  const Csr* in = nullptr;
  const uint32_t* outdeg = nullptr;
  const uint64_t* dead = nullptr;
  const uint64_t* notr = nullptr;  // complement of the snapshot's reach set
  uint64_t e_reach = 0;            // out-edges inside the reach set
  const uint64_t* longrows = nullptr;  // in-CSR rows of more than 32 entries
};
// Builds what is missing (allocates).
BfsPrepared bfs_prepare(ll_gpu_graph* g, int level);
// The cached structures of `level` only: false when any is missing.
bool bfs_peek(ll_gpu_graph* g, int level, BfsPrepared* p);
// Bytes of one search's working memory for V vertices.
size_t bfs_work_bytes(uint64_t V);
// Single-source search in `ws` (bfs_work_bytes(V) bytes, 256-byte aligned):
// levels into lv (device int32[V]); reached = vertices with a level.  Makes
// no device allocation; `timed` records per-kernel times when profiling.
BfsResult bfs_single_ws(ll_gpu_graph* g, const BfsPrepared& p, uint32_t root, void* ws,
                        int32_t* lv, bool timed);

int resolve_level(const ll_gpu_graph* g, int level);
const uint32_t* ensure_degrees(ll_gpu_graph* g, uint32_t dir, int level);
const Csr& ensure_csr(ll_gpu_graph* g, uint32_t dir, int level);
const RowBlocks& ensure_rowblocks(ll_gpu_graph* g, uint32_t dir, int level, uint32_t cap,
                                  uint32_t piece);
void build_rowblocks(ll_gpu_ctx* ctx, const Csr& c, uint32_t cap, uint32_t piece,
                     RowBlocks& rb);
void drop_snapshots(ll_gpu_graph* g);
// Column segments of `in` (source IDs mapped through `inv` when given) with
// segments of 2^zb slots (zb raised until at most kMaxSegments segments).
void build_pr_segments(ll_gpu_graph* g, const Csr& in, const uint32_t* inv, uint32_t zb,
                       PrSegments& ps);
const RowBlocks& pr_segment_rowblocks(ll_gpu_ctx* ctx, PrSegments& ps, uint32_t cap);
// Hot/cold split of `in` whose adjacency holds slots (`slot_adj`): hot = slot < 2^hb.
void build_pr_hotcold(ll_gpu_graph* g, const Csr& in, const uint32_t* slot_adj, uint32_t hb,
                      PrHotCold& hc);

// Binned PageRank (pr_bins.hip): layout of `out` with bins of 2^bb
// destinations, and `iters` iterations over it (FP32 ranks in rank_final).
void build_pr_bins(ll_gpu_graph* g, const Csr& out, uint32_t bb, PrBins& pb);
// The layout of level `level`'s out-CSR with bins of 2^bin_bits vertices
// (12..14), built on first use; -1 = automatic (bins of 2^13 while the edges
// between two such bins number 32 or more on average, else 2^14, while E <
// 2^32 and the bins of 2^14 number at most 16384; otherwise none); nullptr
// when there are none (0 or the automatic rule's "none").
PrBins* ensure_pr_bins(ll_gpu_graph* g, int level, long bin_bits);
void pagerank_bins(ll_gpu_graph* g, const Csr& out, const uint32_t* outdeg, PrBins& pb,
                   int iters, double damping, DevBuf& rank_final, double* last_diff);

// Device-wide exclusive scan of n values into out (n+1 entries, total last).
void scan_u32_to_u64(const uint32_t* in, uint64_t* out, uint64_t n, hipStream_t s);
// The same in caller memory: tmp holds scan_tmp_bytes(n) bytes (8-byte aligned).
size_t scan_tmp_bytes(uint64_t n);
void scan_u32_to_u64_ws(const uint32_t* in, uint64_t* out, uint64_t n, void* tmp, hipStream_t s);
// The same over val[idx[i]], i < n.
void scan_gather_u32_to_u64_ws(const uint32_t* idx, const uint32_t* val, uint64_t* out,
                               uint64_t n, void* tmp, hipStream_t s);
void scan_u64(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t s);
void sort_pairs_u64_u32(const uint64_t* keys_in, uint64_t* keys_out, const uint32_t* vals_in,
                        uint32_t* vals_out, uint64_t n, int end_bit, hipStream_t s);
void sort_keys_u64(const uint64_t* keys_in, uint64_t* keys_out, uint64_t n, int end_bit,
                   hipStream_t s);
void sort_pairs_u32_u32(const uint32_t* keys_in, uint32_t* keys_out, const uint32_t* vals_in,
                        uint32_t* vals_out, uint64_t n, int end_bit, hipStream_t s);
// The same with bits [0, end_bit) sorted as given (no pre-pass over the
// keys), the sorted values written as their low and high 16-bit halves
// (lo16, hi16); keys_work / vals_work (n entries each) are working memory.
void sort_pairs_u32_u32_split(const uint32_t* keys_in, uint32_t* keys_work,
                              const uint32_t* vals_in, uint32_t* vals_work, uint64_t n,
                              int end_bit, uint16_t* lo16, uint16_t* hi16, hipStream_t s);
uint64_t unique_u64(const uint64_t* in, uint64_t* out, uint64_t n, hipStream_t s);

// Sum x over the workgroup and add it to *p with a single atomic.  Every
// thread of the workgroup must call it (it synchronises the workgroup).
__device__ inline void block_atomic_add(unsigned long long* p, unsigned long long x) {
  __shared__ unsigned long long red[16];
  for (int o = 32; o > 0; o >>= 1) x += __shfl_down(x, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (unsigned i = 0; i < (blockDim.x + 63) / 64; i++) t += red[i];
    if (t) atomicAdd(p, t);
  }
}

// Workgroups are dealt to the chip's 8 XCDs round-robin by ID (workgroup i
// runs on XCD i mod 8).  xcd_block(i, n) renumbers workgroup i of n so that
// every XCD takes a contiguous range of work items, in order: items that are
// neighbours in memory then share one XCD's L2, where the partial cache lines
// that neighbouring items write merge before they go to HBM.
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcd_block(uint32_t i, uint32_t n) {
  const uint32_t x = i % kXcds, q = i / kXcds;
  const uint32_t per = n / kXcds, rem = n % kXcds;
  return x * per + (x < rem ? x : rem) + q;
}

inline unsigned grid_for(uint64_t n, unsigned block) {
  uint64_t g = (n + block - 1) / block;
  return unsigned(g == 0 ? 1 : g);
}

}  // namespace llgpu
