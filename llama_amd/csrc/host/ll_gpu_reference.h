// ll_gpu_reference.h - the reference side of the drop-in: what a maintainer of
// LLAMA includes in benchmark/benchmark.cc (after llama.h and the task
// headers) to run the GPU tasks of ll_gpu_tasks.h on ll_mlcsr_ro_graph.
//
// It gives ll_gpu_tasks.h the graph-access functions it asks of a graph type
// for the reference's read-only graph (llama/include/llama/ll_mlcsr_graph.h),
// the copy-on-write level push (ll_gpu_push_cow), and ll_benchmark<Graph>
// wrappers of the GPU tasks so run_benchmark (benchmark.cc:1140-1154) times
// their run() as it times the CPU tasks.  Registration is one row per task in
// g_runnable_things (benchmark.cc:878-946) and one LL_RT_COND_CREATE line in
// the task switch (:1466-1539); LL_GPU_REFERENCE_ROWS and
// LL_GPU_REFERENCE_CREATE below are those rows and lines.
//
// Build flags: -I<repo>/include -I<repo>/llama_amd/csrc/host, link
// -L<repo>/llama_amd -lll_gpu -Wl,-rpath,<repo>/llama_amd.
#pragma once

#include <cstdint>
#include <cstring>
#include <vector>

#include "ll_gpu.h"
#include "ll_gpu_tasks.h"

// --- graph access (ll_gpu_tasks.h) ------------------------------------------------

inline uint64_t ll_gpu_max_nodes(const ll_mlcsr_ro_graph& G) {
  return uint64_t(const_cast<ll_mlcsr_ro_graph&>(G).max_nodes());  // ll_mlcsr_graph.h:626
}

inline int ll_gpu_num_levels(const ll_mlcsr_ro_graph& G) {
  return int(const_cast<ll_mlcsr_ro_graph&>(G).num_levels());  // ll_mlcsr_graph.h:657
}

inline uint32_t ll_gpu_flags(const ll_mlcsr_ro_graph&) {
#ifdef LL_DELETIONS
  return LL_GPU_HAS_CONT | LL_GPU_HAS_DELETIONS;
#else
  return LL_GPU_HAS_CONT;
#endif
}

// The edge table of `level` (a flat ll_et_array) without its vertex table.
inline void ll_gpu_view_et_only(const ll_mlcsr_ro_graph& Gc, uint32_t dir, int level,
                                ll_gpu_level_view* v) {
  ll_mlcsr_ro_graph& G = const_cast<ll_mlcsr_ro_graph&>(Gc);
  auto& csr = dir == LL_GPU_DIR_OUT ? G.out() : G.in();  // ll_mlcsr_graph.h:456-488
  const uint64_t n = uint64_t(G.max_nodes());
  auto* et = csr.edge_table(level);                      // ll_mlcsr_sp.h:434
  const uint64_t w = uint64_t(csr.max_edges(level));     // ll_mlcsr_sp.h:609
  v->dir = dir;
  v->level = uint32_t(level);
  v->num_nodes = n;
  v->num_vt_entries = n + 1;
  v->num_et_words = w;
  v->vt = nullptr;
  v->et = w ? reinterpret_cast<const uint64_t*>(&(*et)[0]) : nullptr;
  v->flags = ll_gpu_flags(G);
  v->page_table = nullptr;
}

// The whole level: its vertex table (copy-on-write pages, or the flat table of
// LL_FLAT_VT / LL_ONE_VT -- as it stands now, i.e. right after the level's
// freeze when pushed then) staged contiguously with the sentinel entry.
inline void ll_gpu_view(const ll_mlcsr_ro_graph& Gc, uint32_t dir, int level,
                        ll_gpu_level_view* v) {
  static std::vector<ll_mlcsr_core__begin_t> staging;  // one push at a time
  ll_mlcsr_ro_graph& G = const_cast<ll_mlcsr_ro_graph&>(Gc);
  auto& csr = dir == LL_GPU_DIR_OUT ? G.out() : G.in();
  auto* vt = csr.vertex_table(level);  // ll_mlcsr_sp.h:362
  const size_t n = size_t(G.max_nodes());
  staging.resize(n + 1);
  for (size_t i = 0; i <= n; i++) staging[i] = (*vt)[node_t(i)];
  ll_gpu_view_et_only(Gc, dir, level, v);
  v->vt = staging.data();
}

// The `weight` property of level k by out-edge index, for sssp_weighted
// (ll_mlcsr_edge_property<float> stored as 32-bit words,
// ll_mlcsr_properties.h:729-1135; get by edge ID :985-1006).
inline bool ll_gpu_weights(const ll_mlcsr_ro_graph& Gc, int k, const float** w, uint64_t* n) {
  static std::vector<float> staging;
  ll_mlcsr_ro_graph& G = const_cast<ll_mlcsr_ro_graph&>(Gc);
  auto* prop = G.get_edge_property_32("weight");  // ll_mlcsr_graph.h:1247
  if (!prop) return false;
  const size_t m = size_t(G.out().max_edges(k));
  staging.assign(m, 0.0f);
  for (size_t i = 0; i < m; i++) {
    const uint32_t bits = prop->get(LL_EDGE_CREATE(k, i));
    std::memcpy(&staging[i], &bits, 4);
  }
  *w = staging.data();
  *n = m;
  return true;
}

// Freezing level k with deletions rewrote the max-visible level of older words
// to k (update_max_visible_level_lower_only, ll_mlcsr_sp.h:849-863): scan the
// older levels for them.  A continuation word that happens to match is
// harmless: the device only ever lowers a word's level.
inline void ll_gpu_patches(const ll_mlcsr_ro_graph& Gc, uint32_t dir, int k,
                           std::vector<uint64_t>& ids, std::vector<uint32_t>& ml) {
#ifdef LL_DELETIONS
  ll_mlcsr_ro_graph& G = const_cast<ll_mlcsr_ro_graph&>(Gc);
  auto& csr = dir == LL_GPU_DIR_OUT ? G.out() : G.in();
  for (int j = 0; j < k; j++) {
    auto* et = csr.edge_table(j);
    const size_t w = size_t(csr.max_edges(j));
    for (size_t i = 0; i < w; i++)
      if (LL_VALUE_MAX_LEVEL((*et)[i]) == uint64_t(k)) {
        ids.push_back(uint64_t(LL_EDGE_CREATE(j, i)));
        ml.push_back(uint32_t(k));
      }
  }
#else
  (void)Gc;
  (void)dir;
  (void)k;
  (void)ids;
  (void)ml;
#endif
}

// bc_random's sources (bc_random.h:177): the reference task's own draws.
// bc_random is registered without reverse edges (benchmark.cc:882-884), so its
// pick_random_node (ll_mlcsr_graph.h:677-682) accepts a draw of
// ll_rand64_positive (libc rand(), ll_utils.h:322-328) when the vertex has
// out-edges at the latest level; the GPU task needs the in-edges loaded, so it
// makes the same draws against the out-edges alone, which gives the reference
// task's sources.
inline int64_t ll_gpu_pick_source(const ll_mlcsr_ro_graph& Gc, uint64_t, int) {
  ll_mlcsr_ro_graph& G = const_cast<ll_mlcsr_ro_graph&>(Gc);
  for (;;) {
    const node_t n = ll_rand64_positive() % G.max_nodes();
    if (G.out().node_exists(n)) return int64_t(n);
  }
}

#if !defined(LL_FLAT_VT) && !defined(LL_SLCSR)
// Copy-on-write push of level k: only the vertex-table pages the level wrote
// (a page it did not write keeps the previous level's page ID,
// ll_mem_array.h:1326-1349) travel; the device points the others at its copy
// of the previous level's page.
inline int ll_gpu_push_cow(ll_gpu_graph* g, const ll_mlcsr_ro_graph& Gc, uint32_t dir, int k) {
  ll_mlcsr_ro_graph& G = const_cast<ll_mlcsr_ro_graph&>(Gc);
  auto& csr = dir == LL_GPU_DIR_OUT ? G.out() : G.in();
  auto* vt = csr.vertex_table(k);
  auto* prev = k > 0 ? csr.vertex_table(k - 1) : nullptr;
  const size_t np = (size_t(G.max_nodes()) + 1 + LL_ENTRIES_PER_PAGE - 1) / LL_ENTRIES_PER_PAGE;
  std::vector<const void*> pages(np);
  for (size_t p = 0; p < np; p++)
    pages[p] = prev && p < prev->pages() && vt->page_id(p) == prev->page_id(p)
                   ? nullptr
                   : static_cast<const void*>(vt->page(p));
  ll_gpu_level_view v;
  ll_gpu_view_et_only(Gc, dir, k, &v);
  return ll_gpu_graph_push_level_pages(g, &v, pages.data(), np);
}
#endif

// --- task wrappers ----------------------------------------------------------------

#define LL_GPU_WRAP(cls, inner, title)                                                \
  template <class Graph>                                                              \
  class cls : public ll_benchmark<Graph> {                                            \
    inner<Graph> t_;                                                                  \
                                                                                      \
   public:                                                                            \
    template <typename... A>                                                          \
    explicit cls(A... a) : ll_benchmark<Graph>(title), t_(a...) {}                    \
    virtual void initialize(Graph* g) {                                               \
      ll_benchmark<Graph>::initialize(g);                                             \
      t_.initialize(g);                                                               \
    }                                                                                 \
    virtual double run(void) { return t_.run(); }                                     \
    virtual double finalize(void) { return t_.finalize(); }                           \
    virtual void print_results(FILE* f) { t_.print_results(f); }                      \
  };

LL_GPU_WRAP(ll_b_gpu_pagerank, llgpu_tasks::gpu_pagerank, "GPU PageRank - pull")
LL_GPU_WRAP(ll_b_gpu_pagerank_push, llgpu_tasks::gpu_pagerank_push, "GPU PageRank - push")
LL_GPU_WRAP(ll_b_gpu_sssp_unweighted, llgpu_tasks::gpu_sssp_unweighted, "GPU BFS levels")
LL_GPU_WRAP(ll_b_gpu_bfs, llgpu_tasks::gpu_bfs, "GPU BFS - count vertices")
LL_GPU_WRAP(ll_b_gpu_sssp_weighted, llgpu_tasks::gpu_sssp_weighted, "GPU SSSP - weighted")
LL_GPU_WRAP(ll_b_gpu_bc_adj, llgpu_tasks::gpu_bc, "GPU Betweenness centrality - exact")
LL_GPU_WRAP(ll_b_gpu_bc_random, llgpu_tasks::gpu_bc, "GPU Betweenness centrality - randomized")
LL_GPU_WRAP(ll_b_gpu_tc, llgpu_tasks::gpu_triangle_counting, "GPU Triangle counting")
LL_GPU_WRAP(ll_b_gpu_tc_i, llgpu_tasks::gpu_tc_i, "GPU Triangle counting (tc_i)")
LL_GPU_WRAP(ll_b_gpu_scc, llgpu_tasks::gpu_scc, "GPU Strongly connected components")

// Rows for g_runnable_things ({class, identifier, name, needs_reverse_edges};
// reverse edges for the tasks that read in-edges).
#define LL_GPU_REFERENCE_ROWS                                                              \
  {"ll_b_gpu_pagerank", "gpu_pagerank", "PageRank on the GPU - pull, FP32", true},         \
  {"ll_b_gpu_pagerank_push", "gpu_pagerank_push", "PageRank on the GPU - push, FP64",      \
   true},                                                                                  \
  {"ll_b_gpu_sssp_unweighted", "gpu_sssp_unweighted", "BFS levels on the GPU", true},      \
  {"ll_b_gpu_bfs", "gpu_bfs_count", "BFS on the GPU - count vertices", true},              \
  {"ll_b_gpu_sssp_weighted", "gpu_sssp_weighted", "Weighted SSSP on the GPU", true},       \
  {"ll_b_gpu_bc_adj", "gpu_bc_adj", "Betweenness centrality on the GPU - exact", true},    \
  {"ll_b_gpu_bc_random", "gpu_bc_random", "Betweenness centrality on the GPU - randomized", \
   true},                                                                                  \
  {"ll_b_gpu_tc", "gpu_tc", "Triangle counting on the GPU (any level count)", true},       \
  {"ll_b_gpu_tc_i", "gpu_tc_i", "Triangle counting on the GPU, tc_i semantics", true},     \
  {"ll_b_gpu_scc", "gpu_scc", "Strongly connected components on the GPU", true},

// Lines for the task switch; `first` is the index of the first row above in
// g_runnable_things, iters and root the driver's -n and -R values.
#define LL_GPU_REFERENCE_CREATE(var, first, iters, root)                 \
  {                                                                      \
    LL_RT_COND_CREATE(var, (first) + 0, ll_b_gpu_pagerank, iters);       \
    LL_RT_COND_CREATE(var, (first) + 1, ll_b_gpu_pagerank_push, iters);  \
    LL_RT_COND_CREATE(var, (first) + 2, ll_b_gpu_sssp_unweighted, root); \
    LL_RT_COND_CREATE(var, (first) + 3, ll_b_gpu_bfs, root);             \
    LL_RT_COND_CREATE(var, (first) + 4, ll_b_gpu_sssp_weighted, root);   \
    LL_RT_COND_CREATE(var, (first) + 5, ll_b_gpu_bc_adj);                \
    LL_RT_COND_CREATE(var, (first) + 6, ll_b_gpu_bc_random, 100, 1ul);   \
    LL_RT_COND_CREATE(var, (first) + 7, ll_b_gpu_tc);                    \
    LL_RT_COND_CREATE(var, (first) + 8, ll_b_gpu_tc_i);                  \
    LL_RT_COND_CREATE(var, (first) + 9, ll_b_gpu_scc);                   \
  }
