/*
 * ll_host.h - C ABI of the host-side level builder and graph generator.
 *
 * The host keeps the multi-level snapshot CSR in LLAMA's read-only layout
 * (per-level vertex tables flattened from copy-on-write pages, edge tables with
 * continuation records, max-visible-level deletion bits).  Each frozen level is
 * handed to the device through ll_gpu_graph_push_level() (ll_gpu.h) as an
 * ll_gpu_level_view that points into the arrays owned here.
 *
 * Reference counterparts (LLAMA, header-only C++):
 *   ll_host_graph_add_level   ll_mlcsr_ro_graph::checkpoint
 *                             (llama/include/llama/ll_mlcsr_graph.h:1501-1831) driven by
 *                             ll_writable_graph::add_edge / delete_edge
 *                             (ll_writable_graph.h:575-650, 852-951); level 0 equals
 *                             ll_edge_list_loader::load_direct on sorted input
 *                             (loaders/ll_load_utils.h:289-880)
 *   ll_host_graph_level_view  ll_csr_base::vertex_table(l) / edge_table(l)
 *                             (ll_mlcsr_sp.h:362-476)
 *   ll_host_graph_patches     the ET words rewritten by
 *                             update_max_visible_level_lower_only (ll_mlcsr_sp.h:849-863)
 *   ll_host_gen_rmat          ll_generator_rmat (loaders/ll_gen_rmat.h:55-353), Graph500
 *                             style without per-bit noise, permuted, deduplicated
 *   ll_host_xs1_*             X-Stream type-1 edge files + .ini
 *                             (loaders/ll_load_xstream1.h:146-273)
 *
 * All functions return 0 on success or a negative LL_GPU_* status code, and
 * never abort; ll_host_last_error() describes the last failure of the calling
 * thread.
 */
#ifndef LL_HOST_H
#define LL_HOST_H

#include <stdint.h>

#include "ll_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ll_host_graph ll_host_graph;
typedef struct ll_host_edges ll_host_edges;

const char* ll_host_last_error(void);

/* --- multi-level graph ---------------------------------------------------- */
int ll_host_graph_create(uint64_t num_nodes, int deletions, ll_host_graph** out);
void ll_host_graph_destroy(ll_host_graph* g);
uint64_t ll_host_graph_num_nodes(const ll_host_graph* g);
int ll_host_graph_num_levels(const ll_host_graph* g);
int ll_host_graph_has_deletions(const ll_host_graph* g);
/* Freeze one level: inserted edges in insertion order, deleted frozen edges
 * (each must be live at the latest level; level 0 takes none). */
int ll_host_graph_add_level(ll_host_graph* g, const uint32_t* ins_src, const uint32_t* ins_dst,
                            uint64_t n_ins, const uint32_t* del_src, const uint32_t* del_dst,
                            uint64_t n_del);
/* As ll_host_graph_add_level with a weight per inserted edge (ins_w may be
 * NULL).  The first level decides whether the graph is weighted; weights
 * live per level beside the out-direction edge table, one per word (0 for
 * continuation words), like the reference's `weight` edge property
 * (ll_mlcsr_properties.h:729-1135, filled by ll_load_utils.h:684-687). */
int ll_host_graph_add_level_w(ll_host_graph* g, const uint32_t* ins_src, const uint32_t* ins_dst,
                              const float* ins_w, uint64_t n_ins, const uint32_t* del_src,
                              const uint32_t* del_dst, uint64_t n_del);
int ll_host_graph_is_weighted(const ll_host_graph* g);
/* Borrowed out-direction weights of one level (*w = NULL, *n = 0 when the
 * graph is unweighted). */
int ll_host_graph_level_weights(const ll_host_graph* g, uint32_t level, const float** w,
                                uint64_t* n);
/* Out-direction weights of the snapshot in ll_host_graph_flatten order. */
int ll_host_graph_flatten_weights(ll_host_graph* g, int level, const float** w, uint64_t* n);
/* Draw and freeze one random incremental level (uniform new edges, uniform
 * live deletions); the batches stay readable through ll_host_graph_last_batch. */
int ll_host_graph_add_random_level(ll_host_graph* g, double ins_frac, double del_frac,
                                   uint64_t seed);
/* Freeze one level confined to the vertex window [lo, hi): n_ins new edges
 * and n_del deletions of live edges (at most as many as there are), all with
 * both ends in the window (seeded as ll_host_graph_add_random_level). */
int ll_host_graph_add_window_level(ll_host_graph* g, uint64_t n_ins, uint64_t n_del, uint64_t lo,
                                   uint64_t hi, uint64_t seed);
int ll_host_graph_last_batch(const ll_host_graph* g, const ll_host_edges** ins,
                             const ll_host_edges** del);
/* Borrowed view of one frozen level of one direction (LL_GPU_DIR_OUT/IN). */
int ll_host_graph_level_view(const ll_host_graph* g, uint32_t dir, uint32_t level,
                             ll_gpu_level_view* out);
/* Copy-on-write pages of one level's vertex table, for
 * ll_gpu_graph_push_level_pages: pages[p] (num_pages = ceil((max_nodes+1)/512))
 * points into the level's table where the freeze wrote page p, NULL where it
 * did not (the page then equals the previous level's, as the reference's
 * ll_mem_array_swcow shares it, ll_mem_array.h:1090-1140); *written (may be
 * NULL) = non-NULL pages. */
int ll_host_graph_level_pages(const ll_host_graph* g, uint32_t dir, uint32_t level,
                              const void** pages, uint64_t num_pages, uint64_t* written);
/* Deletion patches written while freezing `level` into older levels of `dir`. */
int ll_host_graph_patches(const ll_host_graph* g, uint32_t dir, uint32_t level,
                          const uint64_t** edge_ids, const uint32_t** max_levels, uint64_t* n);
/* Snapshot at `level` (-1 = latest) flattened in iterator order; arrays are
 * owned by the graph until the next call. */
int ll_host_graph_flatten(ll_host_graph* g, uint32_t dir, int level, const uint64_t** offsets,
                          const uint32_t** neighbours, uint64_t* num_edges);

/* --- edge lists -------------------------------------------------------------- */
int ll_host_gen_rmat(uint32_t scale, double edge_factor, double a, double b, double c,
                     uint64_t seed, uint64_t perm_seed, ll_host_edges** out);
int ll_host_edges_create(const uint32_t* src, const uint32_t* dst, const float* weight,
                         uint64_t n, ll_host_edges** out);
void ll_host_edges_destroy(ll_host_edges* e);
uint64_t ll_host_edges_count(const ll_host_edges* e);
const uint32_t* ll_host_edges_src(const ll_host_edges* e);
const uint32_t* ll_host_edges_dst(const ll_host_edges* e);
const float* ll_host_edges_weight(const ll_host_edges* e); /* NULL when none */
int ll_host_graph_add_level_edges(ll_host_graph* g, const ll_host_edges* ins,
                                  const ll_host_edges* del);
/* Give every edge a seeded U[0, 1) weight (24 random bits each; random levels
 * of a weighted graph draw theirs the same way from the level's seed). */
int ll_host_edges_fill_weights(ll_host_edges* e, uint64_t seed);

/* --- X-Stream type-1 files: struct {u32 tail; u32 head; f32 weight} -------- */
int ll_host_xs1_write(const char* path, const ll_host_edges* e, uint64_t num_nodes);
int ll_host_xs1_read(const char* path, ll_host_edges** out, uint64_t* num_nodes);

/* --- SNAP .net text edge lists ------------------------------------------------ */
/* Replaces: ll_loader_net::net_loader (llama/include/llama/loaders/ll_load_net.h:
 * 57-375): "tail head" lines, '#'/blank lines skipped, up to 99 malformed lines
 * skipped and counted (the 100th fails), part/nparts > 0 loads the byte range
 * of part `part` (1-based) the same way; num_nodes = largest ID + 1. */
int ll_host_net_read(const char* path, int part, int nparts, ll_host_edges** out,
                     uint64_t* num_nodes, uint64_t* parse_errors);
int ll_host_net_write(const char* path, const ll_host_edges* e);

/* Replaces: utils/xs1-split.cpp: records to part 0 with probability p_large
 * (the tool's 0.8), else to one of n_small parts, with the tool's per-part
 * quotas; `undirected` keeps (u,v),(v,u) record pairs together.  parts must
 * hold n_small + 1 pointers; each part is owned by the caller. */
int ll_host_edges_split(const ll_host_edges* in, int n_small, double p_large, int undirected,
                        uint64_t seed, ll_host_edges** parts);

#ifdef __cplusplus
}
#endif

#endif /* LL_HOST_H */
