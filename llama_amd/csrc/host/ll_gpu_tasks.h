// C++ host side of the drop-in: benchmark tasks with the reference's task
// interface (ll_benchmark<Graph>, benchmark/benchmarks/benchmark.h:127-318:
// ctor(params), initialize(Graph*), run() timed, finalize(), print_results(FILE*))
// whose run() calls the device through the ll_gpu_* C ABI.
//
// A task is templated on the graph type, as the reference tasks are on
// ll_mlcsr_ro_graph (benchmark.cc:97-103).  The graph type must provide
//   uint64_t  ll_gpu_max_nodes(const Graph&)
//   int       ll_gpu_num_levels(const Graph&)
//   uint32_t  ll_gpu_flags(const Graph&)                 (LL_GPU_HAS_* bits)
//   void      ll_gpu_view(const Graph&, dir, level, ll_gpu_level_view*)
//   void      ll_gpu_patches(const Graph&, dir, level, std::vector<uint64_t>& ids,
//                            std::vector<uint32_t>& maxlev)
//   bool      ll_gpu_weights(const Graph&, level, const float** w, uint64_t* n)
//             (the level's `weight` property by out-edge index; false if none)
//   int64_t   ll_gpu_pick_source(const Graph&, uint64_t seed, int i)
//             (bc_random's i-th source; the reference graph's pick_random_node)
// (overloads for the host layout are in ll_gpu_host_graph.h; those for
// ll_mlcsr_ro_graph are a few lines over its public accessors).  The mirror is shared by all tasks of
// one graph and refreshed with each newly frozen level before run().
//
// Errors follow the reference convention: a failing call prints a message and
// aborts (LL_E_PRINT + abort, llama/include/llama/ll_utils.h:394-421).
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ll_gpu.h"

namespace llgpu_tasks {

[[noreturn]] inline void die(const char* what, int rc, const char* msg) {
  std::fprintf(stderr, "Error: %s failed (%d): %s\n", what, rc, msg ? msg : "");
  std::abort();
}

// Device context + mirror of one graph (one host thread at a time).
class Mirror {
 public:
  static Mirror& of(const void* graph_key, uint64_t max_nodes, uint32_t flags) {
    static std::map<const void*, std::unique_ptr<Mirror>> all;
    auto& m = all[graph_key];
    if (!m) m.reset(new Mirror(max_nodes, flags));
    return *m;
  }
  ~Mirror() {
    if (g_) ll_gpu_graph_destroy(g_);
    if (ctx_) ll_gpu_destroy(ctx_);
  }
  ll_gpu_ctx* ctx() { return ctx_; }
  ll_gpu_graph* graph() { return g_; }
  void check(int rc, const char* what) {
    if (rc != LL_GPU_OK) die(what, rc, ll_gpu_last_error(ctx_));
  }

  // Push every frozen level not mirrored yet, replaying its deletion patches.
  template <class Graph>
  void refresh(const Graph& G) {
    const int have = ll_gpu_graph_num_levels(g_, LL_GPU_DIR_OUT);
    std::vector<uint64_t> ids;
    std::vector<uint32_t> ml;
    for (int k = have; k < ll_gpu_num_levels(G); k++)
      for (uint32_t d = 0; d < 2; d++) {
        ids.clear();
        ml.clear();
        ll_gpu_patches(G, d, k, ids, ml);
        if (!ids.empty())
          check(ll_gpu_graph_patch_deletions(g_, d, ids.data(), ml.data(), ids.size()),
                "ll_gpu_graph_patch_deletions");
        ll_gpu_level_view v;
        ll_gpu_view(G, d, k, &v);
        check(ll_gpu_graph_push_level(g_, &v), "ll_gpu_graph_push_level");
      }
    for (int k = have; k < ll_gpu_num_levels(G); k++) {
      const float* w = nullptr;
      uint64_t n = 0;
      if (ll_gpu_weights(G, k, &w, &n))
        check(ll_gpu_graph_push_weights(g_, uint32_t(k), w, n), "ll_gpu_graph_push_weights");
    }
    apply_window();
  }

  // LL_MIN_LEVEL window, as the reference benchmark's `-l min-max` applies it
  // (benchmark.cc:1684-1692): set_min_level(m), then delete the levels i with
  // i + 1 < m.  Applied on the next refresh once those levels are mirrored.
  void set_window(uint32_t m) {
    window_ = m;
    apply_window();
  }

 private:
  void apply_window() {
    if (window_ == 0) return;
    const int L = ll_gpu_graph_num_levels(g_, LL_GPU_DIR_OUT);
    if (L <= 0) return;
    check(ll_gpu_graph_set_min_level(g_, window_), "ll_gpu_graph_set_min_level");
    for (uint32_t i = deleted_; i + 1 < window_ && int(i) + 1 < L; i++, deleted_++)
      check(ll_gpu_graph_delete_level(g_, i), "ll_gpu_graph_delete_level");
  }
  uint32_t window_ = 0, deleted_ = 0;

 public:

 private:
  Mirror(uint64_t max_nodes, uint32_t flags) {
    int rc = ll_gpu_create(0, &ctx_);
    if (rc != LL_GPU_OK) die("ll_gpu_create", rc, "no usable GPU");
    check(ll_gpu_graph_create(ctx_, max_nodes, flags, &g_), "ll_gpu_graph_create");
  }
  ll_gpu_ctx* ctx_ = nullptr;
  ll_gpu_graph* g_ = nullptr;
};

// Result printing in the reference format (print_results_part,
// benchmark/benchmarks/benchmark.h:566-647): first 50 values, 10 per line.
inline void print_part(FILE* f, const double* a, uint64_t n) {
  const uint64_t m = n < 50 ? n : 50;
  for (uint64_t i = 0; i < m; i++) {
    if (i % 10 == 0) std::fprintf(f, "%7lu:", (unsigned long)i);
    std::fprintf(f, " %0.10lf", a[i]);
    if (i % 10 == 9 || i + 1 == m) std::fprintf(f, "\n");
  }
  std::fprintf(f, "\n");
}
template <typename T>
inline void print_part_int(FILE* f, const T* a, uint64_t n) {
  const uint64_t m = n < 50 ? n : 50;
  for (uint64_t i = 0; i < m; i++) {
    if (i % 10 == 0) std::fprintf(f, "%7lu:", (unsigned long)i);
    std::fprintf(f, " %7ld", (long)a[i]);
    if (i % 10 == 9 || i + 1 == m) std::fprintf(f, "\n");
  }
  std::fprintf(f, "\n");
}

template <class Graph>
class ll_gpu_task {
 public:
  explicit ll_gpu_task(const char* name) : name_(name) {}
  virtual ~ll_gpu_task() = default;
  const char* name() const { return name_.c_str(); }
  virtual void initialize(Graph* G) {
    G_ = G;
    m_ = &Mirror::of(G, ll_gpu_max_nodes(*G), ll_gpu_flags(*G));
    m_->refresh(*G);
  }
  virtual double run() = 0;
  virtual double finalize() { return NAN; }
  virtual void print_results(FILE*) {}

 protected:
  std::string name_;
  Graph* G_ = nullptr;
  Mirror* m_ = nullptr;
};

// pagerank / pagerank_double (pagerank.h:98-182).
template <class Graph>
class gpu_pagerank : public ll_gpu_task<Graph> {
 public:
  explicit gpu_pagerank(int iters, double d = 0.85)
      : ll_gpu_task<Graph>("GPU PageRank"), iters_(iters), d_(d) {}
  void initialize(Graph* G) override {
    ll_gpu_task<Graph>::initialize(G);
    rank_.assign(ll_gpu_max_nodes(*G), 0.0);
  }
  double run() override {
    this->m_->check(ll_gpu_pagerank(this->m_->graph(), -1, iters_, d_, rank_.data(), &diff_),
                    "ll_gpu_pagerank");
    return 0;
  }
  double finalize() override {
    double s = 0;
    for (double r : rank_) s += r;
    return s;
  }
  void print_results(FILE* f) override { print_part(f, rank_.data(), rank_.size()); }
  const std::vector<double>& ranks() const { return rank_; }

 private:
  int iters_;
  double d_, diff_ = 0;
  std::vector<double> rank_;
};

// pagerank_push / pagerank_double_push (pagerank.h:217-306): same ranks by
// pushing contributions over out-edges; run() returns 0, finalize() the sum.
template <class Graph>
class gpu_pagerank_push : public ll_gpu_task<Graph> {
 public:
  explicit gpu_pagerank_push(int iters, double d = 0.85)
      : ll_gpu_task<Graph>("GPU PageRank - Push"), iters_(iters), d_(d) {}
  void initialize(Graph* G) override {
    ll_gpu_task<Graph>::initialize(G);
    rank_.assign(ll_gpu_max_nodes(*G), 0.0);
  }
  double run() override {
    this->m_->check(ll_gpu_pagerank_push(this->m_->graph(), -1, iters_, d_, rank_.data(), nullptr),
                    "ll_gpu_pagerank_push");
    return 0;
  }
  double finalize() override {
    double s = 0;
    for (double r : rank_) s += r;
    return s;
  }
  void print_results(FILE* f) override { print_part(f, rank_.data(), rank_.size()); }

 private:
  int iters_;
  double d_;
  std::vector<double> rank_;
};

// sssp_weighted (sssp.h:59-172): FP32 distances over the `weight` property,
// INT_MAX-1 unreached; finalize = largest finite distance truncated to int
// (LL_SSSP_RETURNS_MAX, as the reference benchmark is built).
template <class Graph>
class gpu_sssp_weighted : public ll_gpu_task<Graph> {
 public:
  explicit gpu_sssp_weighted(int64_t root) : ll_gpu_task<Graph>("GPU SSSP - Weighted"), root_(root) {}
  void initialize(Graph* G) override {
    ll_gpu_task<Graph>::initialize(G);
    dist_.assign(ll_gpu_max_nodes(*G), 0.0f);
  }
  double run() override {
    this->m_->check(ll_gpu_sssp(this->m_->graph(), -1, root_, dist_.data(), &reached_, &max_,
                                nullptr),
                    "ll_gpu_sssp");
    return 0;
  }
  double finalize() override { return double(int32_t(max_)); }
  void print_results(FILE* f) override {
    std::vector<double> d(dist_.begin(), dist_.end());
    print_part(f, d.data(), d.size());
  }
  uint64_t reached() const { return reached_; }

 private:
  int64_t root_;
  std::vector<float> dist_;
  uint64_t reached_ = 0;
  double max_ = 0;
};

// bc_adj (bc_adj.h:124-181) / bc_random (bc_random.h:158-193): FP32
// betweenness centrality from every vertex, or from K sources given by the
// graph's ll_gpu_pick_source (on the reference graph its pick_random_node,
// bc_random.h:177, so the sources are the reference task's); finalize = the
// largest value (:201-209).
template <class Graph>
class gpu_bc : public ll_gpu_task<Graph> {
 public:
  gpu_bc() : ll_gpu_task<Graph>("GPU Betweenness centrality - exact"), k_(-1) {}
  gpu_bc(int k, uint64_t seed)
      : ll_gpu_task<Graph>("GPU Betweenness centrality - randomized"), k_(k), seed_(seed) {}
  void initialize(Graph* G) override {
    // sources first: the reference's pick draws from libc rand(), and the
    // device set-up below must not come between the driver and the draws
    const uint64_t V = ll_gpu_max_nodes(*G);
    sources_.clear();
    for (int i = 0; i < k_ && V > 0; i++) sources_.push_back(ll_gpu_pick_source(*G, seed_, i));
    ll_gpu_task<Graph>::initialize(G);
    bc_.assign(V, 0.0f);
  }
  double run() override {
    this->m_->check(ll_gpu_bc(this->m_->graph(), -1, k_ < 0 ? nullptr : sources_.data(),
                              int64_t(sources_.size()), bc_.data()),
                    "ll_gpu_bc");
    return 0;
  }
  double finalize() override {
    float mx = 0;
    for (float x : bc_)
      if (x > mx) mx = x;
    return mx;
  }
  void print_results(FILE* f) override {
    std::vector<double> d(bc_.begin(), bc_.end());
    print_part(f, d.data(), d.size());
  }
  const std::vector<int64_t>& sources() const { return sources_; }

 private:
  int k_;
  uint64_t seed_ = 1;
  std::vector<int64_t> sources_;
  std::vector<float> bc_;
};

// sssp_unweighted (sssp.h:220-333): dist = BFS level, INT_MAX-1 unreached;
// finalize returns the largest level (LL_SSSP_RETURNS_MAX, sssp.h:53).
template <class Graph>
class gpu_sssp_unweighted : public ll_gpu_task<Graph> {
 public:
  explicit gpu_sssp_unweighted(int64_t root) : ll_gpu_task<Graph>("GPU BFS levels"), root_(root) {}
  void initialize(Graph* G) override {
    ll_gpu_task<Graph>::initialize(G);
    dist_.assign(ll_gpu_max_nodes(*G), 0);
  }
  double run() override {
    this->m_->check(ll_gpu_bfs(this->m_->graph(), -1, root_, dist_.data(), &reached_, &examined_),
                    "ll_gpu_bfs");
    return 0;
  }
  double finalize() override {
    int32_t mx = 0;
    for (int32_t x : dist_)
      if (x < 2147483646 && x > mx) mx = x;
    return mx;
  }
  void print_results(FILE* f) override { print_part_int(f, dist_.data(), dist_.size()); }

 private:
  int64_t root_;
  uint64_t reached_ = 0, examined_ = 0;
  std::vector<int32_t> dist_;
};

// bfs (bfs.h:85-128): run() returns the number of reached vertices.
template <class Graph>
class gpu_bfs : public ll_gpu_task<Graph> {
 public:
  explicit gpu_bfs(int64_t root) : ll_gpu_task<Graph>("GPU BFS"), root_(root) {}
  double run() override {
    uint64_t reached = 0, examined = 0;
    this->m_->check(ll_gpu_bfs(this->m_->graph(), -1, root_, nullptr, &reached, &examined),
                    "ll_gpu_bfs");
    return double(reached);
  }

 private:
  int64_t root_;
};

// tc_u / tc_od (triangle_counting.h:343-470, 682-831) on the symmetrised snapshot.
template <class Graph>
class gpu_triangle_counting : public ll_gpu_task<Graph> {
 public:
  gpu_triangle_counting() : ll_gpu_task<Graph>("GPU Triangle Counting") {}
  double run() override {
    this->m_->check(ll_gpu_tc(this->m_->graph(), -1, &t_), "ll_gpu_tc");
    return double(t_);
  }
  void print_results(FILE* f) override {
    std::fprintf(f, "Number of triangles: %llu\n", (unsigned long long)t_);
  }

 private:
  uint64_t t_ = 0;
};

// tc_i (triangle_counting.h:154-336): run() returns the count, and
// print_results writes it in the reference's words (:329-333).
template <class Graph>
class gpu_tc_i : public ll_gpu_task<Graph> {
 public:
  gpu_tc_i() : ll_gpu_task<Graph>("GPU Triangle Counting (tc_i)") {}
  double run() override {
    this->m_->check(ll_gpu_tc_i(this->m_->graph(), -1, &t_), "ll_gpu_tc_i");
    return double(t_);
  }
  void print_results(FILE* f) override {
    std::fprintf(f, "Number of triangles: %lld\n", (long long)t_);
  }

 private:
  uint64_t t_ = 0;
};

// tarjan_scc replacement (tarjan_scc.h:121-206): label = min vertex ID per SCC;
// finalize returns the largest label, as the reference's finalize does.
template <class Graph>
class gpu_scc : public ll_gpu_task<Graph> {
 public:
  gpu_scc() : ll_gpu_task<Graph>("GPU SCC") {}
  void initialize(Graph* G) override {
    ll_gpu_task<Graph>::initialize(G);
    lab_.assign(ll_gpu_max_nodes(*G), -1);
  }
  double run() override {
    this->m_->check(ll_gpu_scc(this->m_->graph(), -1, lab_.data(), &n_), "ll_gpu_scc");
    return double(n_);
  }
  double finalize() override {
    int64_t mx = -1;
    for (int64_t x : lab_) mx = x > mx ? x : mx;
    return double(mx);
  }
  void print_results(FILE* f) override { print_part_int(f, lab_.data(), lab_.size()); }

 private:
  uint64_t n_ = 0;
  std::vector<int64_t> lab_;
};

// Multi-source BFS over `roots` (lane_width 0 = autotuned); run() returns the
// total number of (source, vertex) pairs reached.
template <class Graph>
class gpu_bfs_batch : public ll_gpu_task<Graph> {
 public:
  gpu_bfs_batch(std::vector<int64_t> roots, int lane_width)
      : ll_gpu_task<Graph>("GPU multi-source BFS"), roots_(std::move(roots)), lanes_(lane_width) {}
  double run() override {
    reached_.assign(roots_.size(), 0);
    this->m_->check(ll_gpu_bfs_batch(this->m_->graph(), -1, roots_.data(), int(roots_.size()),
                                     lanes_, nullptr, reached_.data()),
                    "ll_gpu_bfs_batch");
    double t = 0;
    for (uint64_t r : reached_) t += double(r);
    return t;
  }

 private:
  std::vector<int64_t> roots_;
  int lanes_;
  std::vector<uint64_t> reached_;
};

// Batched personalized PageRank; run() returns the iterations executed.
template <class Graph>
class gpu_ppr_batch : public ll_gpu_task<Graph> {
 public:
  gpu_ppr_batch(std::vector<int64_t> seeds, int iters, double d, double tol, int width)
      : ll_gpu_task<Graph>("GPU batched PPR"), seeds_(std::move(seeds)), iters_(iters), d_(d),
        tol_(tol), width_(width) {}
  double run() override {
    int done = 0;
    this->m_->check(ll_gpu_ppr_batch(this->m_->graph(), -1, seeds_.data(), int(seeds_.size()),
                                     iters_, d_, tol_, width_, nullptr, &done),
                    "ll_gpu_ppr_batch");
    return done;
  }

 private:
  std::vector<int64_t> seeds_;
  int iters_;
  double d_, tol_;
  int width_;
};

}  // namespace llgpu_tasks
