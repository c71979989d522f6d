// Graph-access overloads that let the ll_gpu tasks (ll_gpu_tasks.h) run on the
// host layout of this package (llhost::Graph, ll_layout.h).
#pragma once

#include <cstring>
#include <vector>

#include "ll_gpu.h"
#include "ll_host.h"
#include "ll_layout.h"

namespace llhost {
// The C++ graph behind a C-ABI handle (defined in ll_host.cc).
const Graph& graph_of(const ll_host_graph* h);
}  // namespace llhost

inline uint64_t ll_gpu_max_nodes(const llhost::Graph& G) { return G.num_nodes(); }
inline int ll_gpu_num_levels(const llhost::Graph& G) { return G.num_levels(); }
inline uint32_t ll_gpu_flags(const llhost::Graph& G) {
  return LL_GPU_HAS_CONT | (G.deletions() ? LL_GPU_HAS_DELETIONS : 0u);
}
inline void ll_gpu_view(const llhost::Graph& G, uint32_t dir, int level, ll_gpu_level_view* v) {
  const llhost::Level& lv = G.level(int(dir), level);
  std::memset(v, 0, sizeof(*v));
  v->dir = dir;
  v->level = uint32_t(level);
  v->num_nodes = G.num_nodes();
  v->num_vt_entries = lv.vt.size();
  v->num_et_words = lv.et.size();
  v->vt = lv.vt.data();
  v->et = lv.et.data();
  v->flags = ll_gpu_flags(G);
}
inline void ll_gpu_patches(const llhost::Graph& G, uint32_t dir, int level,
                           std::vector<uint64_t>& ids, std::vector<uint32_t>& maxlev) {
  const llhost::PatchList& p = G.patches(int(dir), level);
  ids.assign(p.edge_ids.begin(), p.edge_ids.end());
  maxlev.assign(p.maxlev.begin(), p.maxlev.end());
}
inline bool ll_gpu_weights(const llhost::Graph& G, int level, const float** w, uint64_t* n) {
  if (!G.weighted()) return false;
  const llhost::Level& lv = G.level(0, level);
  *w = lv.wt.data();
  *n = lv.wt.size();
  return true;
}
// bc_random's sources on the host layout: a seeded splitmix64 stream (the
// reference draws them with pick_random_node from libc rand(),
// ll_mlcsr_graph.h:677-682; the reference-side overload in
// ll_gpu_reference.h calls that function, so its picks are the reference's).
inline int64_t ll_gpu_pick_source(const llhost::Graph& G, uint64_t seed, int i) {
  uint64_t z = seed * 0x9e3779b97f4a7c15ull + uint64_t(i) + 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;  // splitmix64 finaliser (Steele, Lea,
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;  // Flood 2014)
  z ^= z >> 31;
  return G.num_nodes() ? int64_t(z % G.num_nodes()) : 0;
}
