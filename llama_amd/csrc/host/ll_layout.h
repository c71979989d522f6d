// Host-side multi-level CSR in LLAMA's read-only layout.
//
// One direction (out- or in-edges) is a list of frozen levels.  Level k holds
//   * a vertex table VT_k[V+1] of 16-byte entries {adj_list_start, level_length,
//     degree}  (ll_mlcsr_core__begin_t, llama/include/llama/ll_mlcsr_helpers.h:141-149),
//     flattened to a contiguous array instead of LLAMA's copy-on-write page table
//     (ll_mem_array.h:918-1485);
//   * an edge table ET_k of 64-bit words: per vertex with new edges a run of
//     `level_length` neighbour words, followed (levels > 0) by a 2-word
//     continuation record that copies VT_{k-1}[v]  (ll_mlcsr_sp.h:1474-1497).
// With deletions on, a word keeps the neighbour in its low 38 bits and the
// exclusive max-visible level in its high 26 bits (ll_mlcsr_helpers.h:90-107).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace llhost {

// Allocator that leaves trivially constructible elements uninitialised, so
// large arrays are first touched by the parallel loops that fill them (serial
// zero-filling of multi-GB arrays otherwise dominates host build time).
template <typename T>
struct UninitAlloc : std::allocator<T> {
  template <typename U>
  struct rebind { using other = UninitAlloc<U>; };
  UninitAlloc() = default;
  template <typename U>
  UninitAlloc(const UninitAlloc<U>&) {}
  template <typename U, typename... Args>
  void construct(U* p, Args&&... args) {
    if constexpr (sizeof...(Args) == 0) ::new (static_cast<void*>(p)) U;
    else ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...);
  }
};
template <typename T>
using vec = std::vector<T, UninitAlloc<T>>;

constexpr int kBitsLevel = 26;                                   // ll_mlcsr_helpers.h:63
constexpr int kBitsIndex = 64 - kBitsLevel;                      // :68
constexpr uint64_t kIndexMask = (uint64_t(1) << kBitsIndex) - 1;
constexpr uint64_t kLevelMask = (uint64_t(1) << kBitsLevel) - 1;
constexpr uint32_t kMaxLevel = uint32_t(kLevelMask) - 2;         // LL_MAX_LEVEL  :75
constexpr uint32_t kNeverDeleted = kMaxLevel + 2;                // LL_VALUE_CREATE default  :97
constexpr int64_t kNilEdge = -1;                                 // LL_NIL_EDGE

inline uint64_t edge_create(uint32_t level, uint64_t index) {
  return index | (uint64_t(level) << kBitsIndex);
}
inline uint32_t edge_level(int64_t e) { return uint32_t(uint64_t(e) >> kBitsIndex); }
inline uint64_t edge_index(int64_t e) { return uint64_t(e) & kIndexMask; }
inline uint64_t word_create(uint64_t payload, uint32_t maxlev) {
  return payload | (uint64_t(maxlev) << kBitsIndex);
}
inline uint64_t word_payload(uint64_t w) { return w & kIndexMask; }
inline uint32_t word_maxlev(uint64_t w) { return uint32_t(w >> kBitsIndex) & uint32_t(kLevelMask); }

struct VtEntry {
  int64_t start;    // edge ID of the newest segment, or -1
  uint32_t len;     // words in that segment
  uint32_t degree;  // live degree at this level (LL_PRECOMPUTED_DEGREE)
};
static_assert(sizeof(VtEntry) == 16, "VT entry must be 16 bytes");

// Vertex-table page of the reference's copy-on-write array (LL_ENTRIES_PER_PAGE,
// ll_common.h:361-363).
constexpr uint64_t kVtPageEntries = 512;

struct Level {
  vec<VtEntry> vt;     // V+1 entries; VT[V] is the sentinel
  vec<uint64_t> et;    // edge-table words incl. continuation records
  vec<float> wt;       // out direction of a weighted graph: weight per ET word (0 for
                       // continuation words), the level's `weight` edge property
  bool sorted_segments = true; // every segment ascending by neighbour
  std::vector<char> page_written;  // per VT page of kVtPageEntries: written by this level
};

// Deletions applied while freezing a level: rewrites of older words.
struct PatchList {
  std::vector<uint64_t> edge_ids;  // level << 38 | index of each rewritten word
  std::vector<uint32_t> maxlev;    // new (exclusive) max-visible level
};

struct Direction {
  std::vector<Level> levels;
};

class Graph {
 public:
  Graph(uint64_t num_nodes, bool deletions) : V_(num_nodes), deletions_(deletions) {}

  uint64_t num_nodes() const { return V_; }
  bool deletions() const { return deletions_; }
  int num_levels() const { return int(dir_[0].levels.size()); }
  const Level& level(int dir, int k) const { return dir_[dir].levels[k]; }
  const PatchList& patches(int dir, int k) const { return patches_[dir][k]; }

  // Freeze one new level from a batch of inserted edges (in insertion order)
  // and deleted frozen edges (each live at the latest level).  Level 0 takes
  // no deletions.  Returns false with last_error() set on bad input.
  // ins_w (may be null): weight of each inserted edge; the first level decides
  // whether the graph is weighted (later unweighted levels get weight 1).
  bool add_level(const uint32_t* ins_src, const uint32_t* ins_dst, uint64_t n_ins,
                 const uint32_t* del_src, const uint32_t* del_dst, uint64_t n_del,
                 const float* ins_w = nullptr);
  bool weighted() const { return weighted_; }

  // Live degree at `level` (precomputed in the VT).
  uint32_t degree(int dir, int level, uint64_t v) const {
    return dir_[dir].levels[level].vt[v].degree;
  }

  // Flatten the snapshot at `level` into CSR (offsets V+1, neighbours) in
  // iterator order: newest segment first, stored order within a segment,
  // words deleted at <= level skipped.  This restates iter_begin/iter_next
  // (ll_mlcsr_sp.h:1994-2196) on the host for the update generator.
  void flatten(int dir, int level, vec<uint64_t>& ofs, vec<uint32_t>& adj) const;
  // Weights of the out-direction snapshot in the order of flatten(0, level).
  void flatten_weights(int level, vec<float>& w) const;

  const std::string& last_error() const { return err_; }

 private:
  bool build_direction(int dir, uint32_t k, const vec<uint32_t>& new_cnt,
                       const vec<uint32_t>& del_cnt, const uint32_t* key,
                       const uint32_t* val, uint64_t n, bool key_sorted,
                       const float* wts = nullptr);
  bool find_live_word(int dir, uint32_t u, uint32_t v, uint64_t* edge_id) const;

  uint64_t V_;
  bool deletions_;
  bool weighted_ = false;
  Direction dir_[2];
  std::vector<PatchList> patches_[2];
  std::string err_;
};

// Seeded counter-based generator (splitmix64 finaliser over (seed, stream, counter)).
uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t counter);

struct EdgeList {
  vec<uint32_t> src, dst;
};

// Graph500-style R-MAT, no per-bit noise: round(ef * 2^scale) edges, endpoints
// relabelled by a seeded random permutation, self-loops dropped, sorted by
// (src, dst) and deduplicated.
void gen_rmat(uint32_t scale, double edge_factor, double a, double b, double c,
              uint64_t seed, uint64_t perm_seed, EdgeList& out);

// Keeps the latest snapshot of a graph as a sorted CSR and draws one
// incremental level: floor(ins_frac * E_live) uniform new edges (u != v,
// absent at the latest level), floor(del_frac * E_live) distinct uniform
// live edges.  Inserts come out sorted by (src, dst); deletions too.
class Updater {
 public:
  explicit Updater(const Graph& g);
  uint64_t num_live_edges() const { return adj_.size(); }
  void next(double ins_frac, double del_frac, uint64_t seed, EdgeList& ins, EdgeList& del);
  // n_ins new edges with both ends in [lo, hi) and n_del live edges whose tail
  // is there (a level confined to a window of vertex IDs).
  void next_window(uint64_t n_ins, uint64_t n_del, uint64_t lo, uint64_t hi, uint64_t seed,
                   EdgeList& ins, EdgeList& del);
  // Also accepts explicit batches so the snapshot stays in step with the graph.
  void apply(const EdgeList& ins, const EdgeList& del);
  const vec<uint64_t>& offsets() const { return ofs_; }
  const vec<uint32_t>& neighbours() const { return adj_; }

 private:
  uint64_t V_;
  vec<uint64_t> ofs_;
  vec<uint32_t> adj_;  // sorted per vertex
};

}  // namespace llhost
