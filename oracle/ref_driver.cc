// ref_driver.cc - TEST INFRASTRUCTURE ONLY.
//
// Drives the reference LLAMA implementation (header-only, compiled in place
// from the reference tree by oracle/Makefile into oracle/_ref/) to write golden
// vectors for the tests: it freezes levels through the reference's own
// writable path (ll_writable_graph::add_edge / delete_edge / checkpoint,
// llama/include/llama/ll_writable_graph.h:652, 791-954, 1957-2020), dumps
// every level's vertex table and edge table of both directions
// (ll_csr_base::vertex_table / edge_table), and runs the reference tasks on the
// read-only graph through their own initialize/run/finalize/print_results
// (benchmark/benchmarks/benchmark.h:216-318).
//
// Input (stdin, text):
//   V L
//   then per level:  I n  followed by n lines "u v",  D m  followed by m lines "u v"
//   then task lines: "pagerank_double iters" | "sssp_unweighted root" | "bfs root" |
//                    "tarjan_scc" | "tc_u" | "tc_od"
// Output (stdout, text): "VT dir level k" blocks of "start len degree" lines,
// "ET dir level n" blocks of words, then per task "TASK name arg", its
// print_results text, and "RETURN value".  With LL_DUMP_EACH=1 in the
// environment the tables are dumped right after each level's freeze instead
// ("FREEZE k", the new level's vertex tables and every level's edge tables as
// they stand then) -- the state a device push at that freeze reads, which is
// what differs between the vertex-table containers (LL_FLAT_VT, LL_ONE_VT).
#include <llama.h>

#include <cstdio>
#include <cstdlib>
#include <unistd.h>
#include <cstring>
#include <string>
#include <vector>

#include "benchmarks/bfs.h"
#include "benchmarks/pagerank.h"
#include "benchmarks/sssp.h"
#include "benchmarks/tarjan_scc.h"
#include "benchmarks/triangle_counting.h"

template <class Task>
static void run_task(Task& t, ll_mlcsr_ro_graph& G, const char* name, long arg) {
  printf("TASK %s %ld\n", name, arg);
  t.initialize(&G);
  double r = t.run();
  double f = t.finalize();
  t.print_results(stdout);
  printf("RETURN %.17g %.17g\n", r, f);
}

static void dump_vt(ll_mlcsr_ro_graph& G, int dir, int l) {
  auto& csr = dir == 0 ? G.out() : G.in();
  auto* vt = csr.vertex_table(l);
  size_t n = vt->size();
  printf("VT %d %d %zu\n", dir, l, n + 1);
  for (size_t v = 0; v <= n; v++) {
    const ll_mlcsr_core__begin_t& e = (*vt)[v];
    printf("%lld %u %u\n", (long long)e.adj_list_start, (unsigned)e.level_length,
           (unsigned)e.degree);
  }
}

static void dump_et(ll_mlcsr_ro_graph& G, int dir, int l) {
  auto& csr = dir == 0 ? G.out() : G.in();
  auto* et = csr.edge_table(l);
  size_t w = csr.max_edges(l);
  printf("ET %d %d %zu\n", dir, l, w);
  for (size_t i = 0; i < w; i++) printf("%llu\n", (unsigned long long)(*et)[i]);
}

static void dump(ll_mlcsr_ro_graph& G) {
  for (int dir = 0; dir < 2; dir++) {
    auto& csr = dir == 0 ? G.out() : G.in();
    for (int l = 0; l < csr.num_levels(); l++) {
      auto* vt = csr.vertex_table(l);
      size_t n = vt->size();
      printf("VT %d %d %zu\n", dir, l, n + 1);
      for (size_t v = 0; v <= n; v++) {
        const ll_mlcsr_core__begin_t& e = (*vt)[v];
        printf("%lld %u %u\n", (long long)e.adj_list_start, (unsigned)e.level_length,
               (unsigned)e.degree);
      }
      auto* et = csr.edge_table(l);
      size_t w = csr.max_edges(l);
      printf("ET %d %d %zu\n", dir, l, w);
      for (size_t i = 0; i < w; i++) printf("%llu\n", (unsigned long long)(*et)[i]);
    }
  }
}

int main() {
  long V = 0, L = 0;
  const char* each_env = getenv("LL_DUMP_EACH");
  const bool each = each_env && each_env[0] == '1';
  if (scanf("%ld %ld", &V, &L) != 2) return 2;
  ll_database db;
  ll_writable_graph& W = *db.graph();
  for (long v = 0; v < V; v++) W.add_node();
  for (long k = 0; k < L; k++) {
    char tag[4];
    long n = 0;
    std::vector<std::pair<long, long>> ins, del;
    if (scanf("%3s %ld", tag, &n) != 2 || strcmp(tag, "I") != 0) return 3;
    ins.resize(n);
    for (auto& p : ins)
      if (scanf("%ld %ld", &p.first, &p.second) != 2) return 4;
    if (scanf("%3s %ld", tag, &n) != 2 || strcmp(tag, "D") != 0) return 5;
    del.resize(n);
    for (auto& p : del)
      if (scanf("%ld %ld", &p.first, &p.second) != 2) return 6;
    W.tx_begin();
    for (auto& p : del) {
      edge_t e = W.find(p.first, p.second);
      if (e == LL_NIL_EDGE) {
        fprintf(stderr, "edge %ld->%ld not found\n", p.first, p.second);
        return 7;
      }
      W.delete_edge(p.first, e);
    }
    for (auto& p : ins) W.add_edge(p.first, p.second);
    W.tx_commit();
    W.checkpoint();
    if (each) {
      ll_mlcsr_ro_graph& R = W.ro_graph();
      printf("FREEZE %ld\n", k);
      for (int dir = 0; dir < 2; dir++) {
        dump_vt(R, dir, int(k));
        for (int l = 0; l <= int(k); l++) dump_et(R, dir, l);
      }
    }
  }
  ll_mlcsr_ro_graph& G = W.ro_graph();
  printf("LEVELS %d NODES %ld\n", G.num_levels(), (long)G.max_nodes());
  if (!each) dump(G);
  char task[64];
  while (scanf("%63s", task) == 1) {
    long arg = 0;
    std::string t = task;
    if (t == "pagerank_double") {
      if (scanf("%ld", &arg) != 1) return 8;
      ll_b_pagerank_pull_double<ll_mlcsr_ro_graph> b((int32_t)arg);
      run_task(b, G, task, arg);
    } else if (t == "sssp_unweighted") {
      if (scanf("%ld", &arg) != 1) return 8;
      ll_b_sssp_unweighted_bfs<ll_mlcsr_ro_graph> b(arg);
      run_task(b, G, task, arg);
    } else if (t == "bfs") {
      if (scanf("%ld", &arg) != 1) return 8;
      ll_b_bfs<ll_mlcsr_ro_graph> b(arg);
      run_task(b, G, task, arg);
    } else if (t == "tarjan_scc") {
      ll_b_tarjan_scc<ll_mlcsr_ro_graph> b;
      run_task(b, G, task, 0);
    } else if (t == "tc_u") {
      ll_b_triangle_counting_LU<ll_mlcsr_ro_graph> b;
      run_task(b, G, task, 0);
    } else {
      fprintf(stderr, "unknown task %s\n", task);
      return 9;
    }
  }
  // Leave without running the database destructors: their teardown aborts
  // (munmap_chunk) on some multi-level graphs with deletions, after every
  // result above has been printed.
  fflush(stdout);
  _exit(0);
}
