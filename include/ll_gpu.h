/*
 * ll_gpu.h - C ABI of the MI355X (gfx950) read-only analytics path over LLAMA's
 * multi-level snapshot CSR.
 *
 * A graph handle mirrors, in HBM, the frozen levels of both directions of an
 * ll_mlcsr_ro_graph (llama/include/llama/ll_mlcsr_graph.h:83-2003): per-level
 * vertex tables (16-byte {adj_list_start, level_length, degree} entries,
 * ll_mlcsr_helpers.h:141-149, flattened from the copy-on-write pages of
 * ll_mem_array.h:918-1485), per-level edge tables of 64-bit words with their
 * continuation records (ll_mlcsr_sp.h:1474-1497), and the deletion vector kept
 * in the high 26 bits of each word (ll_mlcsr_helpers.h:90-107).  Each newly
 * frozen level is pushed with an asynchronous host-to-device copy; deletions
 * that rewrote words of older levels are replayed with
 * ll_gpu_graph_patch_deletions.
 *
 * Conventions (all entry points):
 *   - return 0 on success or a negative ll_gpu_status; the library never aborts;
 *     ll_gpu_last_error() gives a message for the last failure on that context;
 *   - host arrays are caller-owned and only borrowed for the call; device memory
 *     is owned by the context or graph; one host thread per context at a time;
 *   - `level` selects the snapshot: -1 = latest frozen level, otherwise 0..L;
 *     deleted words are those whose max-visible level <= level;
 *   - node IDs are 0..max_nodes-1 (uint32 on the device, int64 node_t at the
 *     boundary where the reference uses node_t).
 */
#ifndef LL_GPU_H
#define LL_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  LL_GPU_OK = 0,
  LL_GPU_EINVAL = -1,   /* bad argument                                       */
  LL_GPU_ENOMEM = -2,   /* device or pinned allocation failed                 */
  LL_GPU_EHIP = -3,     /* HIP runtime error                                  */
  LL_GPU_ESTATE = -4,   /* mirror inconsistent (level order, degree mismatch) */
  LL_GPU_ENOTSUP = -5,  /* unsupported flag or layout                         */
  LL_GPU_ERANGE = -6    /* size beyond a documented limit                     */
} ll_gpu_status;

#define LL_GPU_DIR_OUT 0u
#define LL_GPU_DIR_IN 1u

/* ll_gpu_level_view.flags / ll_gpu_graph_create flags */
#define LL_GPU_HAS_DELETIONS 0x1u   /* words carry max-visible level (LL_DELETIONS)   */
#define LL_GPU_HAS_CONT 0x2u        /* levels > 0 end runs with continuation records  */

typedef struct ll_gpu_ctx ll_gpu_ctx;
typedef struct ll_gpu_graph ll_gpu_graph;

/* One frozen level of one direction, as laid out by init_level_from_degrees +
 * init_node + finish_level_vertices (ll_mlcsr_sp.h:1375-1607, 786-808). */
typedef struct {
  uint32_t dir;              /* LL_GPU_DIR_OUT or LL_GPU_DIR_IN                  */
  uint32_t level;            /* must equal the number of levels already pushed  */
  uint64_t num_nodes;        /* max_nodes of this level                          */
  uint64_t num_vt_entries;   /* num_nodes + 1 (sentinel entry at num_nodes)      */
  uint64_t num_et_words;     /* edge-table words incl. continuation records      */
  const void* vt;            /* 16-byte entries {int64 start; uint32 len; uint32 degree} */
  const uint64_t* et;        /* edge-table words                                 */
  uint32_t flags;            /* LL_GPU_HAS_DELETIONS | LL_GPU_HAS_CONT           */
  const uint32_t* page_table;/* reserved; must be NULL (COW pages: see
                                ll_gpu_graph_push_level_pages)                 */
} ll_gpu_level_view;

/* Per-snapshot statistics (materialised on first use of a level). */
typedef struct {
  uint64_t num_nodes;
  uint64_t live_edges;       /* E_live: words visible at the snapshot level      */
  uint64_t stored_words;     /* words in all reachable segments (E_st)           */
  uint64_t segments;         /* reachable segments                               */
  uint64_t continuations;    /* C = sum over vertices of (segments - 1)          */
  double materialize_ms;     /* device time of the last materialisation          */
} ll_gpu_snapshot_info;

/* --- context and mirror ------------------------------------------------------- */
int ll_gpu_create(int device, ll_gpu_ctx** out);
void ll_gpu_destroy(ll_gpu_ctx* ctx);
const char* ll_gpu_last_error(const ll_gpu_ctx* ctx);
int ll_gpu_sync(ll_gpu_ctx* ctx);
/* Device stream the context launches on (a hipStream_t). */
void* ll_gpu_stream(ll_gpu_ctx* ctx);

/* Replaces: ll_database/ll_mlcsr_ro_graph ownership of the RO levels
 * (ll_database.h:61-80, ll_mlcsr_graph.h:180-309). */
int ll_gpu_graph_create(ll_gpu_ctx* ctx, uint64_t max_nodes, uint32_t flags, ll_gpu_graph** out);
void ll_gpu_graph_destroy(ll_gpu_graph* g);
/* Replaces: the level-freeze event ll_mlcsr_ro_graph::checkpoint
 * (ll_mlcsr_graph.h:1501-1831) / load_direct (loaders/ll_load_utils.h:289-880),
 * whose new VT_k / ET_k this call copies to HBM (async, pinned staging). */
int ll_gpu_graph_push_level(ll_gpu_graph* g, const ll_gpu_level_view* view);
/* Replaces: the same, for a level whose vertex table is the reference's
 * copy-on-write array (ll_mem_array_swcow, ll_mem_array.h:918-1140: an
 * indirection table of 512-entry pages; pages a level did not write are the
 * previous level's).  pages[p], p < num_pages = ceil((max_nodes + 1) / 512),
 * is the host address of page p (its _indirection entry) when the level wrote
 * it, NULL when it is shared with the previous level.  Only the written pages
 * are copied; the device keeps a page table per level that points at the
 * previous level's copy of every shared page.  view->vt is not read (may be
 * NULL); level 0 has no shared pages. */
int ll_gpu_graph_push_level_pages(ll_gpu_graph* g, const ll_gpu_level_view* view,
                                  const void* const* pages, uint64_t num_pages);
/* Replaces: the LL_SLCSR configuration's single-level CSR (ll_slcsr,
 * ll_slcsr.h:88-712; ll_common.h:280-297): one direction of the graph as
 * begin[max_nodes+1] (the vertex table's adj_list_start, degree(n) =
 * begin[n+1] - begin[n], :275-278) and values[num_edges] (the neighbour
 * IDs, iterated begin[n]..begin[n+1]-1 in stored order, :300-350).  The graph
 * must be created without LL_GPU_HAS_DELETIONS and hold no multi-level levels;
 * its only snapshot level is 0.  Push the in-direction too for the tasks that
 * read in-edges (the reference's needs_reverse_edges). */
int ll_gpu_graph_push_slcsr(ll_gpu_graph* g, uint32_t dir, const int64_t* begin,
                            const int64_t* values, uint64_t num_edges);
/* Replaces: ll_mlcsr_ro_graph::set_min_level (LL_MIN_LEVEL,
 * ll_mlcsr_graph.h:560-610 -> ll_mlcsr_sp.h:1801-1862): snapshots then visit
 * only segments of levels >= m (iter_begin / descend, ll_mlcsr_sp.h:2021-2027,
 * 2107-2117) and a vertex's degree is its live edges in the window (without
 * deletions the reference's degree(n) = VT_L.degree - VT_(m-1).degree,
 * ll_mlcsr_sp.h:1913-1925; the reference leaves the case with deletions
 * unimplemented).  m never decreases. */
int ll_gpu_graph_set_min_level(ll_gpu_graph* g, uint32_t m);
/* Replaces: ll_mlcsr_ro_graph::delete_level (ll_mlcsr_sp.h:962-1000, allowed
 * for level + 1 < min_level): frees the level's edge tables and weights in
 * both directions, and its vertex-table pages unless a later level shares
 * them.  The newest level stays.  A snapshot at a deleted level is refused;
 * later deletion patches of its words are dropped (no window reaches them). */
int ll_gpu_graph_delete_level(ll_gpu_graph* g, uint32_t level);
/* Replaces: update_max_visible_level_lower_only on old levels
 * (ll_mlcsr_sp.h:849-863; ll_mlcsr_graph.h:1091-1108).  edge_ids use the
 * level<<38|index encoding (ll_mlcsr_helpers.h:77-79). */
int ll_gpu_graph_patch_deletions(ll_gpu_graph* g, uint32_t dir, const uint64_t* edge_ids,
                                 const uint32_t* max_level, uint64_t n);
/* Replaces: ll_mlcsr_ro_graph::num_levels / max_nodes (ll_mlcsr_graph.h:626-659). */
int ll_gpu_graph_num_levels(const ll_gpu_graph* g, uint32_t dir);
uint64_t ll_gpu_graph_max_nodes(const ll_gpu_graph* g);
/* Total bytes copied host-to-device and device time of the last push (ms). */
int ll_gpu_graph_upload_stats(const ll_gpu_graph* g, uint64_t* bytes, double* last_ms);
/* Materialise (or reuse) the snapshot at `level` and report its statistics. */
int ll_gpu_snapshot_stats(ll_gpu_graph* g, int level, ll_gpu_snapshot_info* info);
/* Snapshot adjacency in iterator order (newest segment first, stored order
 * within a segment, deleted words skipped), copied to host arrays of
 * max_nodes+1 offsets and live_edges neighbours.  Replaces iterating
 * out_iter_begin/out_iter_next or inm_iter_* (ll_mlcsr_graph.h:751-915). */
int ll_gpu_snapshot_adjacency(ll_gpu_graph* g, int level, uint32_t dir, uint64_t* offsets,
                              uint32_t* neighbours);
/* Live degrees at the snapshot (precomputed VT degree, ll_mlcsr_sp.h:1900-1983). */
int ll_gpu_snapshot_degrees(ll_gpu_graph* g, int level, uint32_t dir, uint32_t* degrees);

/* --- algorithms -------------------------------------------------------------- */
/* Replaces: task `pagerank_double` / `pagerank` = ll_b_pagerank_pull_ext::run
 * (benchmark/benchmarks/pagerank.h:98-182): r0 = 1/N, N = max_nodes,
 * r' = (1-d)/N + d * sum_{w in in(t)} r[w]/outdeg(w), fixed `iters`, no dangling
 * redistribution.  Computed in FP32 with compensated sums; ranks_out[max_nodes]
 * receives the FP64-widened ranks; last_diff (may be NULL) the L1 change of
 * the final iteration (the reference's unused `diff`). */
int ll_gpu_pagerank(ll_gpu_graph* g, int level, int iters, double damping, double* ranks_out,
                    double* last_diff);
/* Replaces: task `pagerank_double_push` / `pagerank_push` =
 * ll_b_pagerank_push_ext::run (benchmark/benchmarks/pagerank.h:217-306): every
 * vertex with out-edges adds r[t]/outdeg(t) to next[w] of its out-neighbours
 * (the reference's ATOMIC_ADD), r' = (1-d)/N + d*next.  FP64 ranks and FP64
 * atomic adds; ranks_out[max_nodes] and last_diff (may be NULL) as above. */
int ll_gpu_pagerank_push(ll_gpu_graph* g, int level, int iters, double damping,
                         double* ranks_out, double* last_diff);
/* Replaces: the per-level `weight` edge property read by sssp_weighted
 * (ll_mlcsr_edge_property<float>, ll_mlcsr_properties.h:729-1135; filled by
 * the .xs1 loader, ll_load_utils.h:628-790): w[num_et_words] holds the weight
 * of every out-direction edge-table word of `level` (continuation words are
 * ignored), finite and >= 0.  Level `level` must be pushed; snapshots whose
 * levels all carry weights get them compacted with the out-CSR. */
int ll_gpu_graph_push_weights(ll_gpu_graph* g, uint32_t level, const float* w, uint64_t n);
/* Replaces: task `sssp_weighted` = ll_b_sssp_weighted::run
 * (benchmark/benchmarks/sssp.h:59-172): dist_out[max_nodes] (may be NULL) =
 * FP32 shortest-path distances from root over the snapshot's weighted
 * out-edges, INT_MAX-1 as float for unreached vertices (bit-exact with the
 * reference); reached = vertices with a finite distance; max_dist = the
 * largest one (the task's finalize under LL_SSSP_RETURNS_MAX, truncated to
 * int there); rounds = frontier rounds run.  All may be NULL. */
int ll_gpu_sssp(ll_gpu_graph* g, int level, int64_t root, float* dist_out, uint64_t* reached,
                double* max_dist, uint64_t* rounds);
/* Replaces: tasks `bc_adj` = ll_b_bc_adj::run (benchmark/benchmarks/bc_adj.h:
 * 124-181, every vertex a source) and `bc_random` = ll_b_bc_random::run
 * (bc_random.h:158-193, K sources from pick_random_node): bc_out[max_nodes] =
 * FP32 betweenness centrality (Brandes) over the snapshot's out-edges from
 * sources[0..n) (sources NULL = every vertex, n ignored), summed over sources
 * in the given order.  Path counts and dependencies are FP32 as in the
 * reference; a vertex's dependency sums its next-level out-neighbours' terms
 * in iterator order, so results match the reference bit for bit while path
 * counts stay below 2^24 (past that the reference's own atomic adds make
 * them depend on thread timing). */
int ll_gpu_bc(ll_gpu_graph* g, int level, const int64_t* sources, int64_t n, float* bc_out);
/* Replaces: task `sssp_unweighted` (benchmark/benchmarks/sssp.h:220-333) and
 * `bfs` (bfs.h:85-128): levels_out[max_nodes] (may be NULL) gets the BFS level,
 * INT_MAX-1 for unreached; reached = number of reached vertices (bfs_count);
 * edges_examined = edges the direction-optimising traversal inspected. */
int ll_gpu_bfs(ll_gpu_graph* g, int level, int64_t root, int32_t* levels_out, uint64_t* reached,
               uint64_t* edges_examined);
/* No counterpart in the reference: ll_gpu_bfs_ex for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The search works in a fixed 64 MiB
 * device array of the library (one per GPU, held until the process ends) and
 * calls no allocating routine.  levels_dev is a DEVICE pointer (int32 per
 * vertex, max_nodes entries, INT_MAX-1 for unreached) on the context's
 * device; reached and edges_examined (may be NULL) as in ll_gpu_bfs.
 * LL_GPU_ESTATE until ll_gpu_bfs has run once on `level` (it builds the
 * snapshot structures this call reads); LL_GPU_ERANGE when the working
 * memory of max_nodes vertices (about 28.4 bytes each) exceeds the array:
 * above about 2.3 million vertices.  All calls on one GPU share the array:
 * calls that overlap on one GPU (two contexts or threads) race. */
int ll_gpu_bfs_noalloc(ll_gpu_graph* g, int level, int64_t root, int32_t* levels_dev,
                       uint64_t* reached, uint64_t* edges_examined);
/* Multi-source BFS, sources in lanes of lane_width bits (64, or 0 = autotune
 * among 16/32/64/128).  levels_out (may be NULL) is n x max_nodes int32,
 * row-major by source; reached_out[n] (may be NULL) per-source reached counts. */
int ll_gpu_bfs_batch(ll_gpu_graph* g, int level, const int64_t* roots, int n, int lane_width,
                     int32_t* levels_out, uint64_t* reached_out);
/* Replaces: tasks `tc_u` / `tc_od` (benchmark/benchmarks/triangle_counting.h:343-470,
 * 682-831), extended to multi-level snapshots: triangles of the simple
 * undirected graph of live edges (out and in, self-loops and duplicates removed).
 * LL_GPU_ERANGE when one row of the degree-oriented DAG reaches 2^31 or more
 * neighbour-list entries in total (sum of d+(b) over its arcs), or the DAG
 * holds 2^48 or more arcs. */
int ll_gpu_tc(ll_gpu_graph* g, int level, uint64_t* triangles);
/* Replaces: task `tc_i` = ll_b_triangle_counting_LI::run
 * (benchmark/benchmarks/triangle_counting.h:154-336; registered with reverse
 * edges, benchmark.cc:906-908): for every pair u < v joined by a live edge,
 * once per direction it is joined in, the common neighbours w > v (out or in)
 * -- every triangle u<v<w weighs the number of directed edges between u and
 * v.  The reference needs one level with sorted lists; any snapshot level
 * works here.  Same limits as ll_gpu_tc. */
int ll_gpu_tc_i(ll_gpu_graph* g, int level, uint64_t* count);
/* Replaces: task `tarjan_scc` (benchmark/benchmarks/tarjan_scc.h:121-206) with
 * a correct SCC decomposition; labels_out[max_nodes] = smallest vertex ID of
 * each vertex's strongly connected component.  num_scc may be NULL. */
int ll_gpu_scc(ll_gpu_graph* g, int level, int64_t* labels_out, uint64_t* num_scc);
/* Batched personalized PageRank: seeds[n] single-vertex teleports,
 * r0 = e_s, r' = (1-d) e_s + d * sum_{w in in(t)} r[w]/outdeg(w), at most
 * `iters` iterations, stopping once every column's L1 change <= tol.
 * ranks_out (may be NULL) is max_nodes x n FP32 row-major (vertex-major);
 * batch_width 0 = autotune among 16/32/64/128.  iters_done may be NULL. */
int ll_gpu_ppr_batch(ll_gpu_graph* g, int level, const int64_t* seeds, int n, int iters,
                     double damping, double tol, int batch_width, float* ranks_out,
                     int* iters_done);

/* Weighted PageRank on the snapshot's out-edges, weights in the snapshot's
 * out-edge order (ll_gpu_snapshot_adjacency(..., LL_GPU_DIR_OUT, ...)).
 * Pointers below are DEVICE pointers on this context's device; the call is
 * ordered on ll_gpu_stream(ctx) and returns after it completes.
 *   forward: r0 = p, r' = (1-d) p + d * W^T r, `iters` iterations.
 *   adjoint: given dL/dr of the returned r, solves the transposed system
 *     lambda = dL/dr + d * W lambda with the same number of iterations
 *     (a pull over out-edges = the transposed push sweep of
 *     ll_b_pagerank_push_ext, pagerank.h:217-360) and returns
 *     dL/dw[e=(u,v)] = d * r[u] * lambda[v] and dL/dp = (1-d) * lambda.
 * The adjoint is that of the fixed point; it equals the derivative of the
 * iterated forward once the iteration has converged. */
int ll_gpu_wpr_forward(ll_gpu_graph* g, int level, const float* edge_w, const float* teleport,
                       int iters, double damping, float* r_out);
int ll_gpu_wpr_adjoint(ll_gpu_graph* g, int level, const float* edge_w, const float* r,
                       const float* dl_dr, int iters, double damping, float* dl_dw_out,
                       float* dl_dteleport_out);

/* --- extended forms, measurement and tuning ------------------------------------ */
/* ll_gpu_pagerank with an explicit row-block capacity (0 = option pr_cap,
 * default 2048; 512/1024/2048/4096 entries per workgroup for runs of short
 * rows). */
int ll_gpu_pagerank_ex(ll_gpu_graph* g, int level, int iters, double damping, uint32_t cap,
                       double* ranks_out, double* last_diff);
/* ll_gpu_bfs plus edges_traversed = sum of live out-degrees of the reached
 * vertices (Graph500 TEPS numerator). */
int ll_gpu_bfs_ex(ll_gpu_graph* g, int level, int64_t root, int32_t* levels_out,
                  uint64_t* reached, uint64_t* edges_examined, uint64_t* edges_traversed);
/* ll_gpu_bfs_batch plus the number of pull sweeps run (may be NULL). */
int ll_gpu_bfs_batch_ex(ll_gpu_graph* g, int level, const int64_t* roots, int n, int lane_width,
                        int32_t* levels_out, uint64_t* reached_out, uint64_t* sweeps);
/* ll_gpu_tc plus the oriented arc count and W = sum over arcs (a,b) of
 * d+(a) + d+(b) (intersection element reads). */
int ll_gpu_tc_ex(ll_gpu_graph* g, int level, uint64_t* triangles, uint64_t* arcs,
                 uint64_t* work);
/* ll_gpu_scc plus the number of full-graph sweeps (trim, propagation, BFS). */
int ll_gpu_scc_ex(ll_gpu_graph* g, int level, int64_t* labels_out, uint64_t* num_scc,
                  uint64_t* sweeps);
/* Per-kernel device time from HIP events: ll_gpu_profile(ctx, 1) clears the
 * totals and starts recording; ll_gpu_profile_get returns launches and the
 * summed milliseconds of one kernel name (e.g. "k_pr_pull"). */
int ll_gpu_profile(ll_gpu_ctx* ctx, int on);
int ll_gpu_profile_get(ll_gpu_ctx* ctx, const char* kernel, uint64_t* launches,
                       double* total_ms);
/* Autotuner decision whose key starts with `prefix` ("ppr:", "bfs_lanes:"). */
int ll_gpu_tuned(ll_gpu_ctx* ctx, const char* prefix, int* value);
/* Tuning knobs: "pr_bin_bits" (PageRank as the binned pull over the
 * out-CSR with destination bins of 2^N vertices, N in 12..14; 0 = the
 * row-block pull over the in-CSR, which the knobs below tune; -1 = automatic:
 * while the snapshot has fewer than 2^32 live edges and at most 2^28
 * vertices, 13 when the edges between two bins of 2^13 number 32 or more on
 * average, else 14; otherwise 0),
 * "bfs_reach" (1 = single-source searches whose root reaches only the
 * snapshot's reach set -- the vertices reachable from its highest-degree
 * vertex -- treat the other rows as visited; 0 = dead rows only; default 1),
 * "pr_reorder" (1 = gather PageRank contributions in
 * descending-out-degree slot order; default), "pr_cap" (default row-block
 * capacity, 2048), "pr_seg_bits" (pull over column segments of 2^N
 * contribution slots; 0 = off, -1 = automatic = off), "pr_item_sort" (sort
 * each row block's entries by slot; 1 on, 0 off, -1 automatic = off),
 * "pr_hot_bits" (hot/cold split of the pull at slot 2^N; 0 = off, -1 =
 * automatic = off), "sssp_delta_milli" (weighted SSSP near/far threshold
 * step in thousandths of the mean edge weight; 0 = no far pile, plain
 * Bellman-Ford rounds; -1 = automatic = 5), "sssp_min_queue" (once edges worth
 * half the snapshot have been relaxed, the smallest queue a near/far phase
 * starts with when the far pile holds that many -- smaller queues rejoin the
 * pile; 0 = off, -1 = automatic = max_nodes / 32), "bfs_alpha" / "bfs_beta"
 * (direction-optimising BFS switches to bottom-up when the frontier's
 * out-edges exceed the unexplored edges / alpha, back to top-down when a
 * shrinking frontier holds under max_nodes / beta vertices; <= 0 = defaults:
 * alpha 10; beta 24, or no return to top-down for a search that treats the
 * rows outside the snapshot's reach set as visited), "bfs_batch_bins"
 * (multi-source BFS sweeps over the binned out-edge layout of "pr_bin_bits":
 * -1 = automatic = where that layout exists and fits the lane width in LDS,
 * for the sweeps that do not push while the rows some search has not reached
 * hold at least "bfs_batch_bin_open" percent of the in-edges (default 40);
 * 0 = off; 1 = every sweep that does not push, with bins of 2^12 where
 * "pr_bin_bits" names none).  Results never depend on them. */
int ll_gpu_set_option(ll_gpu_ctx* ctx, const char* name, long value);
/* Drop the cached snapshot materialisations (the next call rebuilds them). */
void ll_gpu_graph_drop_snapshots(ll_gpu_graph* g);

#ifdef __cplusplus
}
#endif

#endif /* LL_GPU_H */
