// ll_gpu_bench - command-line driver for the GPU tasks, modelled on the
// reference driver (benchmark/benchmark.cc): a task registry
// (g_runnable_things, :871-946), one level per input .xs1 file
// (load_direct per file, :1624-1659), `-c` timed repetitions of run() bracketed
// by wall-clock time (run_benchmark, :1140-1154), mean +- c95
// (ll_utils.h:218-223), "Return value:" from finalize (:2019-2021) and
// print_results into `-o`.
//
//   ll_gpu_bench -r gpu_pagerank -n 20 -c 3 base.xs1 level1.xs1 ...
//   ll_gpu_bench -r gpu_bfs -R 7 --rmat 20 --levels 11
#include <getopt.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "ll_gpu.h"
#include "ll_gpu_host_graph.h"
#include "ll_gpu_tasks.h"
#include "ll_host.h"
#include "ll_layout.h"

using Graph = llhost::Graph;
using namespace llgpu_tasks;

namespace {

struct Entry {
  const char* id;
  const char* name;
};
const Entry kTasks[] = {
    {"gpu_pagerank", "PageRank, pull, FP32 compensated (pagerank / pagerank_double)"},
    {"gpu_pagerank_push", "PageRank, push, FP64 atomics (pagerank_push / pagerank_double_push)"},
    {"gpu_sssp_unweighted", "BFS levels from -R (sssp_unweighted)"},
    {"gpu_sssp_weighted", "Weighted SSSP from -R over the .xs1 weights (sssp_weighted)"},
    {"gpu_bfs", "BFS reached count from -R (bfs)"},
    {"gpu_bc_adj", "Betweenness centrality from every vertex (bc_adj)"},
    {"gpu_bc_random", "Betweenness centrality from --sources K random sources (bc_random)"},
    {"gpu_tc", "Triangle counting on the symmetrised snapshot (tc_u / tc_od)"},
    {"gpu_tc_i", "Triangles counted once per direction of their lowest pair (tc_i)"},
    {"gpu_scc", "Strongly connected components, min-ID labels (tarjan_scc)"},
    {"gpu_bfs_batch", "Multi-source BFS, --sources N roots, --lanes W"},
    {"gpu_ppr_batch", "Batched personalized PageRank, --sources N seeds, -n iterations"},
};

void usage() {
  std::fprintf(stderr,
               "Usage: ll_gpu_bench -r TASK [-n ITERS] [-R ROOT] [-c COUNT] [-o FILE]\n"
               "                    [-l MIN-[MAX]] [--sources N] [--lanes W] [--tol T]\n"
               "                    (FILE.xs1|FILE.net ... | --rmat SCALE [--levels L] [--seed S])\n"
               "Tasks:\n");
  for (const Entry& e : kTasks) std::fprintf(stderr, "  %-22s %s\n", e.id, e.name);
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char** argv) {
  std::string task;
  int iters = -1, count = 1, sources = 64, lanes = 0, scale = 0, levels = 1;
  bool sources_set = false;
  long root = 0;
  double tol = 1e-7;
  unsigned long seed = 1;
  const char* out_file = nullptr;
  int min_level = 0, max_level = -1;
  static option opts[] = {{"run", required_argument, 0, 'r'},
                          {"levels-window", required_argument, 0, 'l'},
                          {"num-iters", required_argument, 0, 'n'},
                          {"root", required_argument, 0, 'R'},
                          {"count", required_argument, 0, 'c'},
                          {"output", required_argument, 0, 'o'},
                          {"sources", required_argument, 0, 1},
                          {"lanes", required_argument, 0, 2},
                          {"rmat", required_argument, 0, 3},
                          {"levels", required_argument, 0, 4},
                          {"seed", required_argument, 0, 5},
                          {"tol", required_argument, 0, 6},
                          {"help", no_argument, 0, 'h'},
                          {0, 0, 0, 0}};
  int c;
  while ((c = getopt_long(argc, argv, "r:n:R:c:o:l:h", opts, nullptr)) != -1) {
    switch (c) {
      case 'r': task = optarg; break;
      case 'n': iters = std::atoi(optarg); break;
      case 'R': root = std::atol(optarg); break;
      case 'c': count = std::atoi(optarg); break;
      case 'o': out_file = optarg; break;
      case 'l': {  // "MAX" or "MIN-MAX" (benchmark.cc:1239-1250); MAX may be empty
        const char* dash = std::strchr(optarg, '-');
        if (!dash) {
          max_level = std::atoi(optarg);
        } else {
          min_level = std::atoi(optarg);
          if (dash[1]) max_level = std::atoi(dash + 1);
        }
        break;
      }
      case 1: sources = std::atoi(optarg), sources_set = true; break;
      case 2: lanes = std::atoi(optarg); break;
      case 3: scale = std::atoi(optarg); break;
      case 4: levels = std::atoi(optarg); break;
      case 5: seed = std::strtoul(optarg, nullptr, 10); break;
      case 6: tol = std::atof(optarg); break;
      case 'h': usage(); return 0;
      default: usage(); return 1;
    }
  }
  bool known = false;
  for (const Entry& e : kTasks) known = known || task == e.id;
  if (!known) {
    std::fprintf(stderr, "Error: Unknown task \"%s\"\n", task.c_str());
    usage();
    return 1;
  }

  // Load: one level per .xs1 file, or a generated multi-level R-MAT graph.
  double t0 = now_ms();
  ll_host_graph* hg = nullptr;
  if (scale > 0) {
    ll_host_edges* base = nullptr;
    // seeded U[0, 1) weights, so gpu_sssp_weighted runs on generated graphs too
    if (ll_host_gen_rmat(uint32_t(scale), 16.0, 0.57, 0.19, 0.19, seed, seed + 1, &base) ||
        ll_host_edges_fill_weights(base, seed) ||
        ll_host_graph_create(uint64_t(1) << scale, 1, &hg) ||
        ll_host_graph_add_level_edges(hg, base, nullptr)) {
      std::fprintf(stderr, "Error: %s\n", ll_host_last_error());
      return 1;
    }
    ll_host_edges_destroy(base);
    for (int k = 1; k < levels; k++)
      if (ll_host_graph_add_random_level(hg, 0.01, 0.001, seed + 2 + k)) {
        std::fprintf(stderr, "Error: %s\n", ll_host_last_error());
        return 1;
      }
  } else {
    if (optind >= argc) {
      usage();
      return 1;
    }
    std::vector<ll_host_edges*> files;
    uint64_t V = 0;
    // -l ...-MAX: the levels 0..MAX, one per file
    const int last = max_level >= 0 ? std::min(argc, optind + max_level + 1) : argc;
    for (int i = optind; i < last; i++) {
      ll_host_edges* e = nullptr;
      uint64_t n = 0;
      const std::string path = argv[i];
      const auto ends = [&](const char* x) {
        const size_t l = std::strlen(x);
        return path.size() >= l && path.compare(path.size() - l, l, x) == 0;
      };
      const int rc = ends(".net") || ends(".snap")
                         ? ll_host_net_read(argv[i], 0, 0, &e, &n, nullptr)
                         : ll_host_xs1_read(argv[i], &e, &n);
      if (rc) {
        std::fprintf(stderr, "Error: %s\n", ll_host_last_error());
        return 1;
      }
      V = std::max(V, n);
      files.push_back(e);
    }
    if (ll_host_graph_create(V, 1, &hg)) {
      std::fprintf(stderr, "Error: %s\n", ll_host_last_error());
      return 1;
    }
    for (ll_host_edges* e : files) {
      if (ll_host_graph_add_level_edges(hg, e, nullptr)) {
        std::fprintf(stderr, "Error: %s\n", ll_host_last_error());
        return 1;
      }
      ll_host_edges_destroy(e);
    }
  }
  const Graph& G = llhost::graph_of(hg);
  std::printf("LOAD\nLoad Time  : %.3f seconds\nLevels     : %d--%d\nNodes      : %lu\n\n",
              (now_ms() - t0) / 1000.0, min_level, G.num_levels() - 1,
              (unsigned long)G.num_nodes());

  std::vector<int64_t> roots;
  for (int i = 0; i < sources; i++)
    roots.push_back(int64_t(llhost::rng_u64(seed, 77, uint64_t(i)) % G.num_nodes()));
  std::unique_ptr<ll_gpu_task<Graph>> t;
  if (task == "gpu_pagerank") t.reset(new gpu_pagerank<Graph>(iters < 0 ? 10 : iters));
  else if (task == "gpu_pagerank_push")
    t.reset(new gpu_pagerank_push<Graph>(iters < 0 ? 10 : iters));
  else if (task == "gpu_sssp_unweighted") t.reset(new gpu_sssp_unweighted<Graph>(root));
  else if (task == "gpu_sssp_weighted") t.reset(new gpu_sssp_weighted<Graph>(root));
  else if (task == "gpu_bfs") t.reset(new gpu_bfs<Graph>(root));
  else if (task == "gpu_bc_adj") t.reset(new gpu_bc<Graph>());
  else if (task == "gpu_bc_random") t.reset(new gpu_bc<Graph>(sources_set ? sources : 100, seed));
  else if (task == "gpu_tc") t.reset(new gpu_triangle_counting<Graph>());
  else if (task == "gpu_tc_i") t.reset(new gpu_tc_i<Graph>());
  else if (task == "gpu_scc") t.reset(new gpu_scc<Graph>());
  else if (task == "gpu_bfs_batch") t.reset(new gpu_bfs_batch<Graph>(roots, lanes));
  else t.reset(new gpu_ppr_batch<Graph>(roots, iters < 0 ? 30 : iters, 0.85, tol, lanes));

  if (min_level > 0) {
    if (min_level >= G.num_levels()) {
      std::fprintf(stderr, "Error: -l %d-: the graph has %d levels\n", min_level, G.num_levels());
      return 1;
    }
    Mirror::of(&G, G.num_nodes(), ll_gpu_flags(G)).set_window(uint32_t(min_level));
  }
  std::vector<double> times;
  double ret = NAN;
  Graph& Gm = const_cast<Graph&>(G);
  for (int i = 0; i < count; i++) {
    t->initialize(&Gm);
    double a = now_ms();
    ret = t->run();
    times.push_back(now_ms() - a);
    double f = t->finalize();
    if (!std::isnan(f)) ret = f;
  }
  double mean = 0, sd = 0;
  for (double x : times) mean += x;
  mean /= times.size();
  for (double x : times) sd += (x - mean) * (x - mean);
  sd = times.size() > 1 ? std::sqrt(sd / (times.size() - 1)) : 0;
  const double c95 = 1.96 * sd / std::sqrt(double(times.size()));
  std::printf("TIME\nTime       : %.3f +- %.3f ms (%d runs)\n\n", mean, c95, count);
  std::printf("Return value: %lf\n", ret);
  if (out_file) {
    FILE* f = std::fopen(out_file, "w");
    if (f) {
      t->print_results(f);
      std::fclose(f);
    }
  }
  t.reset();
  ll_host_graph_destroy(hg);
  return 0;
}
