// oracle.cc - CPU restatement of LLAMA's read-only analytics path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
// and never as the thing measured or shipped.
//
// It reads the multi-level CSR through borrowed plain arrays (per direction and
// level: the 16-byte vertex table and the 64-bit edge table) and restates the
// reference algorithms on them:
//   iterator          ll_mlcsr_core::iter_begin / iter_next / iter_descend
//                     (llama/include/llama/ll_mlcsr_sp.h:1994-2196), deletion test
//                     is_edge_deleted (:1290-1310), LL_VALUE_* (ll_mlcsr_helpers.h:90-107)
//   pagerank_double   ll_b_pagerank_pull_ext<Graph,double>::run
//                     (benchmark/benchmarks/pagerank.h:98-182), OpenMP
//                     schedule(dynamic,4096), per-edge divide by out_degree
//   bfs levels        sssp_unweighted (benchmark/benchmarks/sssp.h:220-333): dist =
//                     BFS level, INT_MAX-1 unreached; bfs_count (bfs.h:117-127)
//   triangles         tc_u count_for / run (triangle_counting.h:373-459) on the
//                     symmetrised, deduplicated, self-loop-free snapshot
//   scc               textbook Tarjan (the reference tarjan_scc.h:55-206 is not a
//                     correct SCC algorithm), labels = min vertex ID per component
//   sssp_weighted     ll_b_sssp_weighted::run (benchmark/benchmarks/sssp.h:59-172):
//                     synchronous rounds, dist_nxt[s] = min(dist[n] + weight[e]) in
//                     FP32 over the out-edges e of updated vertices, weights read
//                     by edge ID from per-level arrays (the `weight` property)
//   bc                ll_b_bc_adj / ll_b_bc_random (benchmark/benchmarks/bc_adj.h,
//                     bc_random.h): Brandes in FP32, sigma along down-edges of the
//                     BFS from each source, delta summed in iterator order
//   ppr, wpr, adjoint FP64 restatements of the definitions in ll_gpu.h (no
//                     reference counterpart)
//
// Parity pins (tests/): closed forms (cycles, stars, empty graphs, K_n, DAGs),
// brute force on tiny graphs, the reference's own outputs on fixtures
// (tests/golden), and the delete-every-10th-edge count invariant of
// benchmark/tests/delete_edges.h:81-233.
#include <omp.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kBitsIndex = 38;
constexpr uint64_t kIndexMask = (uint64_t(1) << kBitsIndex) - 1;
constexpr uint32_t kLevelMask = (1u << 26) - 1;

struct Vt {
  int64_t start;
  uint32_t len;
  uint32_t degree;
};

struct Lvl {
  const Vt* vt;
  const uint64_t* et;
  uint64_t n_vt, n_et;
};

}  // namespace

struct orc_graph {
  uint64_t V;
  int deletions;
  uint32_t min_level = 0;  // LL_MIN_LEVEL window: segments of older levels are not visited
  std::vector<Lvl> lv[2];
  std::vector<const float*> wt;  // out-direction `weight` property per level (by edge index)
  std::vector<uint64_t> n_wt;
};

namespace {

// ll_edge_iterator restated (ll_mlcsr_iterator.h:53-68).
struct Iter {
  int64_t edge;
  uint64_t left;
  const uint64_t* ptr;
  uint32_t max_level;
  int64_t last_node;
};

inline uint64_t payload(const orc_graph* g, uint64_t w) { return g->deletions ? (w & kIndexMask) : w; }
inline uint32_t level_of(int64_t e) { return uint32_t(uint64_t(e) >> kBitsIndex); }
inline uint64_t index_of(int64_t e) { return uint64_t(e) & kIndexMask; }

inline bool is_deleted(const orc_graph* g, const Iter& it) {
  if (!g->deletions || it.edge == -1) return false;
  uint32_t ml = uint32_t(*it.ptr >> kBitsIndex) & kLevelMask;
  return ml <= it.max_level;
}

void iter_descend(const orc_graph* g, int dir, Iter& it) {
  uint32_t level = level_of(it.edge);
  if (level == 0) {
    it.edge = -1;
    return;
  }
  // continuation record right after the run (ll_mlcsr_sp.h:2100-2102)
  const uint64_t* c = it.ptr + 1;
  it.edge = int64_t(c[0]);
  it.left = uint32_t(c[1]);
  if (level_of(it.edge) < g->min_level) {  // LL_MIN_LEVEL, ll_mlcsr_sp.h:2107-2111
    it.left = 0;
    it.edge = -1;
  } else if (it.left == 0) {
    it.edge = -1;
  } else {
    it.ptr = g->lv[dir][level_of(it.edge)].et + index_of(it.edge);
  }
}

int64_t iter_next(const orc_graph* g, int dir, Iter& it) {
  int64_t r = it.edge;
  if (r != -1) {
    it.last_node = int64_t(payload(g, *it.ptr));
    do {
      if (it.left > 1) {
        it.left--;
        it.edge = it.edge + 1;
        it.ptr = it.ptr + 1;
      } else {
        iter_descend(g, dir, it);
      }
    } while (is_deleted(g, it));
  }
  return r;
}

void iter_begin(const orc_graph* g, int dir, int level, uint64_t n, Iter& it) {
  it.max_level = uint32_t(level);
  const Vt& b = g->lv[dir][level].vt[n];
  it.edge = b.start;
  it.left = b.len;
  it.last_node = -1;
  it.ptr = nullptr;
  if (level_of(it.edge) < g->min_level) {  // LL_MIN_LEVEL, ll_mlcsr_sp.h:2021-2027
    it.left = 0;
    it.edge = -1;
  } else if (it.left == 0) {
    it.edge = -1;
  } else {
    it.ptr = g->lv[dir][level_of(it.edge)].et + index_of(it.edge);
  }
  if (is_deleted(g, it)) {
    int64_t keep = it.last_node;
    iter_next(g, dir, it);
    it.last_node = keep;
  }
}

// Visit (edge ID, neighbour) of the live out/in edges of n in iterator order.
template <typename F>
inline void foreach_edge(const orc_graph* g, int dir, int level, uint64_t n, F&& f) {
  Iter it;
  iter_begin(g, dir, level, n, it);
  for (int64_t e = iter_next(g, dir, it); e != -1; e = iter_next(g, dir, it))
    f(uint64_t(e), uint64_t(it.last_node));
}

// Visit the live neighbours of n at `level` in iterator order.
template <typename F>
inline void foreach_nbr(const orc_graph* g, int dir, int level, uint64_t n, F&& f) {
  Iter it;
  iter_begin(g, dir, level, n, it);
  while (iter_next(g, dir, it) != -1) f(uint64_t(it.last_node));
}

inline uint32_t degree(const orc_graph* g, int dir, int level, uint64_t n) {
  if (g->min_level == 0) return g->lv[dir][level].vt[n].degree;
  // LL_MIN_LEVEL degree(n), ll_mlcsr_sp.h:1913-1925: the stored degree less the
  // one below the window; with deletions (unimplemented there) and for a
  // snapshot below the window, the live edges the window shows
  if (!g->deletions && int(g->min_level) - 1 <= level)
    return g->lv[dir][level].vt[n].degree - g->lv[dir][g->min_level - 1].vt[n].degree;
  uint32_t d = 0;
  foreach_nbr(g, dir, level, n, [&](uint64_t) { d++; });
  return d;
}

int resolve(const orc_graph* g, int level) {
  int L = int(g->lv[0].size());
  if (level < 0) level = L - 1;
  if (level < 0 || level >= L || level >= int(g->lv[1].size())) return -1;
  return level;
}

// Simple undirected snapshot: sorted unique neighbours, no self-loops.
void symmetrise(const orc_graph* g, int level, std::vector<uint64_t>& ofs,
                std::vector<uint32_t>& adj) {
  const uint64_t V = g->V;
  std::vector<std::vector<uint32_t>> nb(V);
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t u = 0; u < V; u++) {
    auto& l = nb[u];
    for (int d = 0; d < 2; d++)
      foreach_nbr(g, d, level, u, [&](uint64_t v) {
        if (v != u) l.push_back(uint32_t(v));
      });
    std::sort(l.begin(), l.end());
    l.erase(std::unique(l.begin(), l.end()), l.end());
  }
  ofs.assign(V + 1, 0);
  for (uint64_t u = 0; u < V; u++) ofs[u + 1] = ofs[u] + nb[u].size();
  adj.resize(ofs[V]);
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t u = 0; u < V; u++) std::copy(nb[u].begin(), nb[u].end(), adj.begin() + ofs[u]);
}

}  // namespace

extern "C" {

orc_graph* orc_graph_create(uint64_t V, int deletions) {
  auto* g = new orc_graph();
  g->V = V;
  g->deletions = deletions;
  return g;
}

void orc_graph_destroy(orc_graph* g) { delete g; }

// Borrow one level of one direction; levels must be added in order.
int orc_graph_add_level(orc_graph* g, uint32_t dir, const void* vt, uint64_t n_vt,
                        const uint64_t* et, uint64_t n_et) {
  if (dir > 1 || n_vt != g->V + 1) return -1;
  g->lv[dir].push_back(Lvl{static_cast<const Vt*>(vt), et, n_vt, n_et});
  return 0;
}

// Snapshot adjacency in iterator order (CSR: V+1 offsets, neighbours).
int64_t orc_flatten(const orc_graph* g, uint32_t dir, int level, uint64_t* ofs, uint32_t* adj) {
  level = resolve(g, level);
  if (level < 0) return -1;
  uint64_t o = 0;
  for (uint64_t v = 0; v < g->V; v++) {
    if (ofs) ofs[v] = o;
    foreach_nbr(g, int(dir), level, v, [&](uint64_t w) {
      if (adj) adj[o] = uint32_t(w);
      o++;
    });
  }
  if (ofs) ofs[g->V] = o;
  return int64_t(o);
}

// Number of iterated live edges vs the sum of precomputed degrees
// (the count invariant of benchmark/tests/delete_edges.h:81-233).
int orc_count_check(const orc_graph* g, uint32_t dir, int level, uint64_t* iterated,
                    uint64_t* degree_sum) {
  level = resolve(g, level);
  if (level < 0) return -1;
  uint64_t it = 0, ds = 0;
#pragma omp parallel for reduction(+ : it, ds) schedule(dynamic, 4096)
  for (uint64_t v = 0; v < g->V; v++) {
    foreach_nbr(g, int(dir), level, v, [&](uint64_t) { it++; });
    ds += degree(g, int(dir), level, v);
  }
  *iterated = it;
  *degree_sum = ds;
  return 0;
}

// pagerank_double: pagerank.h:98-182.
int orc_pagerank_double(const orc_graph* g, int level, int iters, double d, double* rank,
                        double* last_diff) {
  level = resolve(g, level);
  if (level < 0) return -1;
  const int64_t V = int64_t(g->V);
  const double N = double(V);
  std::vector<double> nxt(V);
#pragma omp parallel for
  for (int64_t t = 0; t < V; t++) rank[t] = 1 / N;
  int cnt = 0;
  double diff = 0;
  do {
    diff = 0;
#pragma omp parallel
    {
      double diff_prv = 0;
#pragma omp for nowait schedule(dynamic, 4096)
      for (int64_t t = 0; t < V; t++) {
        double s = 0;
        foreach_nbr(g, 1, level, uint64_t(t), [&](uint64_t w) {
          s = s + rank[w] / double(degree(g, 0, level, w));
        });
        double val = (1 - d) / N + d * s;
        diff_prv += std::abs(val - rank[t]);
        nxt[t] = val;
      }
#pragma omp atomic
      diff += diff_prv;
    }
#pragma omp parallel for
    for (int64_t i = 0; i < V; i++) rank[i] = nxt[i];
    cnt++;
  } while (cnt < iters);
  if (last_diff) *last_diff = diff;
  return 0;
}

// BFS levels over out-edges (sssp_unweighted semantics): level-synchronous.
int64_t orc_bfs_levels(const orc_graph* g, int level, int64_t root, int32_t* dist) {
  level = resolve(g, level);
  if (level < 0 || root < 0 || uint64_t(root) >= g->V) return -1;
  const uint64_t V = g->V;
  for (uint64_t v = 0; v < V; v++) dist[v] = INT_MAX - 1;
  std::vector<uint32_t> cur{uint32_t(root)}, nxt;
  dist[root] = 0;
  int64_t reached = 1;
  int32_t lev = 0;
  while (!cur.empty()) {
    nxt.clear();
    for (uint32_t u : cur)
      foreach_nbr(g, 0, level, u, [&](uint64_t v) {
        if (dist[v] == INT_MAX - 1) {
          dist[v] = lev + 1;
          nxt.push_back(uint32_t(v));
        }
      });
    reached += int64_t(nxt.size());
    cur.swap(nxt);
    lev++;
  }
  return reached;
}

// Triangles of the symmetrised simple snapshot, tc_u style (triangle_counting.h:373-459).
int64_t orc_triangles(const orc_graph* g, int level) {
  level = resolve(g, level);
  if (level < 0) return -1;
  std::vector<uint64_t> ofs;
  std::vector<uint32_t> adj;
  symmetrise(g, level, ofs, adj);
  // parallel over list entries (u's i-th neighbour), not vertices, so a hub's
  // merges spread over the threads; u = the row holding entry i
  int64_t T = 0;
  const uint64_t M = ofs[g->V];
#pragma omp parallel for reduction(+ : T) schedule(dynamic, 1024)
  for (uint64_t i = 0; i < M; i++) {
    const uint64_t u = uint64_t(std::upper_bound(ofs.begin(), ofs.end(), i) - ofs.begin()) - 1;
    const uint64_t v = adj[i];
    if (u >= v) continue;
    uint64_t a = ofs[u], ae = ofs[u + 1], b = ofs[v], be = ofs[v + 1];
    while (a < ae && b < be) {
      if (adj[a] < adj[b]) a++;
      else if (adj[b] < adj[a]) b++;
      else {
        if (v < adj[b]) T++;
        a++;
        b++;
      }
    }
  }
  return T;
}

// tc_i (ll_b_triangle_counting_LI, triangle_counting.h:154-336): for every
// vertex u and every out-neighbour and then every in-neighbour v > u (a pair
// joined in both directions is visited twice), count_for (:176-266) merges
// the sorted out- and in-lists of u and of v and counts the common w > v.
// The reference needs one level with sorted lists; the snapshot's lists are
// sorted here, so any level works.
int64_t orc_tc_i(const orc_graph* g, int level) {
  level = resolve(g, level);
  if (level < 0) return -1;
  std::vector<uint64_t> oo, io;
  std::vector<uint32_t> oa, ia;
  for (int d = 0; d < 2; d++) {
    auto& ofs = d == 0 ? oo : io;
    auto& adj = d == 0 ? oa : ia;
    ofs.assign(g->V + 1, 0);
    adj.resize(uint64_t(orc_flatten(g, uint32_t(d), level, nullptr, nullptr)));
    orc_flatten(g, uint32_t(d), level, ofs.data(), adj.data());
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint64_t v = 0; v < g->V; v++) std::sort(adj.begin() + ofs[v], adj.begin() + ofs[v + 1]);
  }
  // the merged walk of count_for: both lists of each endpoint advance past
  // the current minimum, so a neighbour in both lists counts once
  auto count_for = [&](uint64_t u, uint64_t v) {
    int64_t c = 0;
    uint64_t a = oo[u], ae = oo[u + 1], b = io[u], be = io[u + 1];
    uint64_t x = oo[v], xe = oo[v + 1], y = io[v], ye = io[v + 1];
    while ((a < ae || b < be) && (x < xe || y < ye)) {
      const uint64_t mu = a == ae ? ia[b] : (b == be ? oa[a] : std::min(oa[a], ia[b]));
      const uint64_t mv = x == xe ? ia[y] : (y == ye ? oa[x] : std::min(oa[x], ia[y]));
      const bool au = a < ae && oa[a] == mu, bu = b < be && ia[b] == mu;
      const bool xv = x < xe && oa[x] == mv, yv = y < ye && ia[y] == mv;
      if (mu < mv) {
        a += au;
        b += bu;
      } else if (mv < mu) {
        x += xv;
        y += yv;
      } else {
        if (v < mv) c++;
        a += au;
        b += bu;
        x += xv;
        y += yv;
      }
    }
    return c;
  };
  // parallel over list entries, out-lists then in-lists (see orc_triangles)
  int64_t T = 0;
  const uint64_t Mo = oo[g->V], Mi = io[g->V];
#pragma omp parallel for reduction(+ : T) schedule(dynamic, 1024)
  for (uint64_t j = 0; j < Mo + Mi; j++) {
    const bool out = j < Mo;
    const std::vector<uint64_t>& o = out ? oo : io;
    const uint64_t i = out ? j : j - Mo;
    const uint64_t u = uint64_t(std::upper_bound(o.begin(), o.end(), i) - o.begin()) - 1;
    const uint64_t v = out ? oa[i] : ia[i];
    if (u < v) T += count_for(u, v);
  }
  return T;
}

// Symmetrised simple snapshot as CSR (for brute-force and degree checks).
int64_t orc_symmetrise(const orc_graph* g, int level, uint64_t* ofs_out, uint32_t* adj_out) {
  level = resolve(g, level);
  if (level < 0) return -1;
  std::vector<uint64_t> ofs;
  std::vector<uint32_t> adj;
  symmetrise(g, level, ofs, adj);
  if (ofs_out) std::copy(ofs.begin(), ofs.end(), ofs_out);
  if (adj_out) std::copy(adj.begin(), adj.end(), adj_out);
  return int64_t(adj.size());
}

// Textbook iterative Tarjan over out-edges; label = min vertex ID in the SCC.
int64_t orc_scc(const orc_graph* g, int level, int64_t* label) {
  level = resolve(g, level);
  if (level < 0) return -1;
  const uint64_t V = g->V;
  std::vector<uint64_t> ofs;
  std::vector<uint32_t> adj;
  {
    ofs.assign(V + 1, 0);
    for (uint64_t v = 0; v < V; v++) ofs[v + 1] = ofs[v] + degree(g, 0, level, v);
    adj.resize(ofs[V]);
    for (uint64_t v = 0; v < V; v++) {
      uint64_t o = ofs[v];
      foreach_nbr(g, 0, level, v, [&](uint64_t w) { adj[o++] = uint32_t(w); });
      if (o != ofs[v + 1]) return -2;
    }
  }
  std::vector<int64_t> index(V, -1), low(V, 0);
  std::vector<char> on(V, 0);
  std::vector<uint32_t> stack;
  std::vector<std::pair<uint32_t, uint64_t>> call;  // (vertex, next edge)
  int64_t counter = 0, nscc = 0;
  for (uint64_t s = 0; s < V; s++) {
    if (index[s] != -1) continue;
    call.push_back({uint32_t(s), ofs[s]});
    index[s] = low[s] = counter++;
    stack.push_back(uint32_t(s));
    on[s] = 1;
    while (!call.empty()) {
      auto& fr = call.back();
      uint32_t v = fr.first;
      if (fr.second < ofs[v + 1]) {
        uint32_t w = adj[fr.second++];
        if (index[w] == -1) {
          index[w] = low[w] = counter++;
          stack.push_back(w);
          on[w] = 1;
          call.push_back({w, ofs[w]});
        } else if (on[w]) {
          low[v] = std::min(low[v], index[w]);
        }
        continue;
      }
      if (low[v] == index[v]) {
        size_t pos = stack.size();
        uint32_t mn = v;
        do {
          pos--;
          mn = std::min(mn, stack[pos]);
        } while (stack[pos] != v);
        for (size_t i = pos; i < stack.size(); i++) {
          label[stack[i]] = mn;
          on[stack[i]] = 0;
        }
        stack.resize(pos);
        nscc++;
      }
      call.pop_back();
      if (!call.empty()) {
        uint32_t p = call.back().first;
        low[p] = std::min(low[p], low[v]);
      }
    }
  }
  return nscc;
}

// Personalized PageRank, FP64: r0 = e_s, r' = (1-d) e_s + d * A^T (r / outdeg);
// out is V x n row-major.  Stops when every column's L1 change <= tol.
int orc_ppr(const orc_graph* g, int level, const int64_t* seeds, int n, int iters, double d,
            double tol, double* out, int* iters_done) {
  level = resolve(g, level);
  if (level < 0) return -1;
  const uint64_t V = g->V;
  std::vector<double> r(V * n, 0.0), nx(V * n);
  for (int j = 0; j < n; j++) r[uint64_t(seeds[j]) * n + j] = 1.0;
  int it = 0;
  while (it < iters) {
    std::vector<double> diff(n, 0.0);
#pragma omp parallel
    {
      std::vector<double> dl(n, 0.0), s(n);
#pragma omp for schedule(dynamic, 4096)
      for (uint64_t t = 0; t < V; t++) {
        std::fill(s.begin(), s.end(), 0.0);
        foreach_nbr(g, 1, level, t, [&](uint64_t w) {
          double inv = 1.0 / double(degree(g, 0, level, w));
          for (int j = 0; j < n; j++) s[j] += r[w * n + j] * inv;
        });
        for (int j = 0; j < n; j++) {
          double v = d * s[j] + (uint64_t(seeds[j]) == t ? 1 - d : 0.0);
          dl[j] += std::abs(v - r[t * n + j]);
          nx[t * n + j] = v;
        }
      }
#pragma omp critical
      for (int j = 0; j < n; j++) diff[j] += dl[j];
    }
    r.swap(nx);
    it++;
    bool done = true;
    for (int j = 0; j < n; j++) done = done && diff[j] <= tol;
    if (done) break;
  }
  std::copy(r.begin(), r.end(), out);
  if (iters_done) *iters_done = it;
  return 0;
}

// Weighted PageRank forward, FP64, on a flattened out-CSR (ofs, adj) with
// weights w in out-edge order: r0 = p, r' = (1-d) p + d * W^T r.
int orc_wpr_forward(uint64_t V, const uint64_t* ofs, const uint32_t* adj, const double* w,
                    const double* p, int iters, double d, double* r_out) {
  // r' = (1-d) p + d W^T r as a pull over the transpose (built once, heads'
  // in-edges in out-edge order), so every head is summed by one thread.
  const uint64_t E = ofs[V];
  std::vector<uint64_t> tofs(V + 1, 0);
  for (uint64_t e = 0; e < E; e++) tofs[adj[e] + 1]++;
  for (uint64_t v = 0; v < V; v++) tofs[v + 1] += tofs[v];
  std::vector<uint32_t> tsrc(E);
  std::vector<double> tw(E);
  {
    std::vector<uint64_t> cur(tofs.begin(), tofs.end() - 1);
    for (uint64_t u = 0; u < V; u++)
      for (uint64_t e = ofs[u]; e < ofs[u + 1]; e++) {
        const uint64_t k = cur[adj[e]]++;
        tsrc[k] = uint32_t(u);
        tw[k] = w[e];
      }
  }
  std::vector<double> r(p, p + V), nx(V);
  for (int it = 0; it < iters; it++) {
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint64_t v = 0; v < V; v++) {
      double s = 0;
      for (uint64_t k = tofs[v]; k < tofs[v + 1]; k++) s += tw[k] * r[tsrc[k]];
      nx[v] = (1 - d) * p[v] + d * s;
    }
    r.swap(nx);
  }
  std::copy(r.begin(), r.end(), r_out);
  return 0;
}

// Fixed-point adjoint of weighted PageRank, FP64:
// lambda = g + d * W lambda (iterated `iters` times from lambda0 = g),
// dL/dw[e=(u,v)] = d * r[u] * lambda[v], dL/dp = (1-d) * lambda.
int orc_wpr_adjoint(uint64_t V, const uint64_t* ofs, const uint32_t* adj, const double* w,
                    const double* r, const double* gr, int iters, double d, double* dw,
                    double* dp) {
  std::vector<double> lam(gr, gr + V), nx(V);
  for (int it = 0; it < iters; it++) {
#pragma omp parallel for schedule(dynamic, 4096)
    for (uint64_t u = 0; u < V; u++) {
      double s = 0;
      for (uint64_t e = ofs[u]; e < ofs[u + 1]; e++) s += w[e] * lam[adj[e]];
      nx[u] = gr[u] + d * s;
    }
    lam.swap(nx);
  }
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t u = 0; u < V; u++)
    for (uint64_t e = ofs[u]; e < ofs[u + 1]; e++) dw[e] = d * r[u] * lam[adj[e]];
#pragma omp parallel for
  for (uint64_t v = 0; v < V; v++) dp[v] = (1 - d) * lam[v];
  return 0;
}

// Borrow the out-direction weights of one level (one per edge-table word).
int orc_graph_add_weights(orc_graph* g, uint32_t level, const float* w, uint64_t n) {
  if (level >= g->lv[0].size() || n != g->lv[0][level].n_et) return -1;
  if (g->wt.size() <= level) {
    g->wt.resize(level + 1, nullptr);
    g->n_wt.resize(level + 1, 0);
  }
  g->wt[level] = w;
  g->n_wt[level] = n;
  return 0;
}

// sssp_weighted (sssp.h:59-172), restated round for round: every vertex
// updated in the previous round relaxes its out-edges into dist_nxt (FP32
// sums dist[n] + weight[e]), then dist = dist_nxt; stops when a round updates
// nothing.  dist[v] = INT_MAX-1 (as float) when unreached.  Returns the number
// of reached vertices; *max_dist (may be NULL) gets the largest finite distance
// (the task's finalize with LL_SSSP_RETURNS_MAX, truncated to int there).
int64_t orc_sssp_weighted(const orc_graph* g, int level, int64_t root, float* dist,
                          double* max_dist, int64_t* rounds) {
  level = resolve(g, level);
  if (level < 0 || root < 0 || uint64_t(root) >= g->V) return -1;
  for (int k = 0; k <= level; k++)
    if (k >= int(g->wt.size()) || !g->wt[k]) return -2;  // unweighted level
  const uint64_t V = g->V;
  const float inf = float(INT_MAX - 1);
  std::vector<float> nxt(V);
  std::vector<char> upd(V, 0), upd_nxt(V, 0);
  for (uint64_t v = 0; v < V; v++) dist[v] = nxt[v] = (int64_t(v) == root) ? 0.0f : inf;
  upd[root] = 1;
  int64_t r = 0;
  for (bool any = true; any; r++) {
    for (uint64_t n = 0; n < V; n++) {
      if (!upd[n]) continue;
      foreach_edge(g, 0, level, n, [&](uint64_t e, uint64_t s) {
        const float cand = dist[n] + g->wt[e >> kBitsIndex][e & kIndexMask];
        if (nxt[s] > cand) {
          nxt[s] = cand;
          upd_nxt[s] = 1;
        }
      });
    }
    any = false;
    for (uint64_t v = 0; v < V; v++) {
      dist[v] = nxt[v];
      upd[v] = upd_nxt[v];
      upd_nxt[v] = 0;
      any = any || upd[v];
    }
  }
  int64_t reached = 0;
  double mx = 0;
  for (uint64_t v = 0; v < V; v++)
    if (dist[v] < inf) {
      reached++;
      mx = std::max(mx, double(dist[v]));
    }
  if (max_dist) *max_dist = mx;
  if (rounds) *rounds = r;
  return reached;
}

// bc_adj (sources == NULL: every vertex, ascending) / bc_random (the given
// sources): per source a BFS over out-edges, sigma[w] += sigma[v] along
// down-edges (level[w] = level[v] + 1) in BFS order (visit_fw), then in
// reverse BFS order delta[v] = sum over v's down-edges in iterator order of
// sigma[v] / sigma[w] * (1 + delta[w]) and BC[v] += delta[v] for v != source
// (visit_rv).  All arithmetic in FP32 as the generated task code does.
int orc_bc(const orc_graph* g, int level, const int64_t* sources, int64_t n, float* bc) {
  level = resolve(g, level);
  if (level < 0) return -1;
  const uint64_t V = g->V;
  const int64_t ns = sources ? n : int64_t(V);
  std::vector<float> sigma(V), delta(V, 0.0f);
  std::vector<int32_t> lv(V);
  std::vector<uint64_t> order;
  order.reserve(V);
  for (uint64_t v = 0; v < V; v++) bc[v] = 0.0f;
  for (int64_t i = 0; i < ns; i++) {
    const uint64_t s = sources ? uint64_t(sources[i]) : uint64_t(i);
    if (s >= V) return -2;
    std::fill(lv.begin(), lv.end(), -1);
    std::fill(sigma.begin(), sigma.end(), 0.0f);
    order.clear();
    lv[s] = 0;
    sigma[s] = 1.0f;
    order.push_back(s);
    for (size_t h = 0; h < order.size(); h++) {
      const uint64_t v = order[h];
      foreach_nbr(g, 0, level, v, [&](uint64_t w) {
        if (lv[w] < 0) {
          lv[w] = lv[v] + 1;
          order.push_back(w);
        }
      });
    }
    for (uint64_t v : order)
      foreach_nbr(g, 0, level, v, [&](uint64_t w) {
        if (lv[w] == lv[v] + 1) sigma[w] += sigma[v];
      });
    for (size_t h = order.size(); h-- > 0;) {
      const uint64_t v = order[h];
      if (v == s) continue;
      float S = 0.0f;
      foreach_nbr(g, 0, level, v, [&](uint64_t w) {
        if (lv[w] == lv[v] + 1) S = S + sigma[v] / sigma[w] * (1 + delta[w]);
      });
      delta[v] = S;
      bc[v] = bc[v] + delta[v];
    }
  }
  return 0;
}

// LL_MIN_LEVEL: snapshots visit only segments of levels >= m (set_min_level,
// ll_mlcsr_sp.h:1801-1862) and degrees count the window (degree above).
int orc_graph_set_min_level(orc_graph* g, uint32_t m) {
  if (m < g->min_level) return -1;  // the window only moves forward (:1806)
  g->min_level = m;
  return 0;
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
