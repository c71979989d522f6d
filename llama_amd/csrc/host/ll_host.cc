// C ABI over the host level builder, generator and .xs1 files (ll_host.h).
#include "ll_host.h"

#include <sys/types.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>

#include "ll_layout.h"

struct ll_host_edges {
  llhost::EdgeList e;
  llhost::vec<float> w;
};

struct ll_host_graph {
  std::unique_ptr<llhost::Graph> g;
  std::unique_ptr<llhost::Updater> upd;
  ll_host_edges last_ins, last_del;
  llhost::vec<uint64_t> flat_ofs;
  llhost::vec<uint32_t> flat_adj;
  llhost::vec<float> flat_w;
};

namespace {

// Seeded U[0, 1) weights with 24 random bits each (exact in float).
void uniform_weights(uint64_t n, uint64_t seed, llhost::vec<float>& w) {
  w.resize(n);
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; i++)
    w[i] = float(llhost::rng_u64(seed, 0x5757, i) >> 40) * (1.0f / 16777216.0f);
}

}  // namespace

namespace llhost {
const Graph& graph_of(const ll_host_graph* h) { return *h->g; }
}  // namespace llhost

namespace {
thread_local std::string g_err;
int fail(int code, const char* msg) noexcept {
  try {
    g_err = msg;
  } catch (...) {
  }
  return code;
}
int fail(int code, const std::string& msg) noexcept { return fail(code, msg.c_str()); }
// Status of the exception in flight inside an entry point's catch (...):
// std::bad_alloc -> LL_GPU_ENOMEM, anything else -> LL_GPU_ESTATE.  No
// exception crosses the C ABI.
int host_exception() noexcept {
  try {
    throw;
  } catch (const std::bad_alloc&) {
    return fail(LL_GPU_ENOMEM, "out of host memory");
  } catch (const std::exception& e) {
    return fail(LL_GPU_ESTATE, e.what());
  } catch (...) {
    return fail(LL_GPU_ESTATE, "unknown exception");
  }
}
}  // namespace

extern "C" {

const char* ll_host_last_error(void) { return g_err.c_str(); }

int ll_host_graph_create(uint64_t num_nodes, int deletions, ll_host_graph** out) try {
  if (!out) return fail(LL_GPU_EINVAL, "out is NULL");
  if (num_nodes == 0 || num_nodes >= (uint64_t(1) << 32))
    return fail(LL_GPU_ERANGE, "num_nodes must be in [1, 2^32)");
  auto* h = new (std::nothrow) ll_host_graph();
  if (!h) return fail(LL_GPU_ENOMEM, "out of host memory");
  h->g.reset(new llhost::Graph(num_nodes, deletions != 0));
  *out = h;
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

void ll_host_graph_destroy(ll_host_graph* g) { delete g; }

uint64_t ll_host_graph_num_nodes(const ll_host_graph* g) { return g ? g->g->num_nodes() : 0; }
int ll_host_graph_num_levels(const ll_host_graph* g) { return g ? g->g->num_levels() : 0; }
int ll_host_graph_has_deletions(const ll_host_graph* g) { return g ? int(g->g->deletions()) : 0; }

int ll_host_graph_add_level(ll_host_graph* g, const uint32_t* ins_src, const uint32_t* ins_dst,
                            uint64_t n_ins, const uint32_t* del_src, const uint32_t* del_dst,
                            uint64_t n_del) {
  return ll_host_graph_add_level_w(g, ins_src, ins_dst, nullptr, n_ins, del_src, del_dst, n_del);
}

int ll_host_graph_add_level_w(ll_host_graph* g, const uint32_t* ins_src, const uint32_t* ins_dst,
                              const float* ins_w, uint64_t n_ins, const uint32_t* del_src,
                              const uint32_t* del_dst, uint64_t n_del) {
  if (!g) return fail(LL_GPU_EINVAL, "graph is NULL");
  if ((n_ins && (!ins_src || !ins_dst)) || (n_del && (!del_src || !del_dst)))
    return fail(LL_GPU_EINVAL, "NULL edge array");
  if (ins_w && g->g->num_levels() > 0 && !g->g->weighted())
    return fail(LL_GPU_EINVAL, "weights on a level of a graph whose first level had none");
  try {
    if (!g->g->add_level(ins_src, ins_dst, n_ins, del_src, del_dst, n_del, ins_w))
      return fail(LL_GPU_EINVAL, g->g->last_error());
    g->upd.reset();  // the sampling snapshot is rebuilt lazily from the graph
  } catch (const std::bad_alloc&) {
    return fail(LL_GPU_ENOMEM, "out of host memory while freezing a level");
  }
  return LL_GPU_OK;
}

int ll_host_graph_add_level_edges(ll_host_graph* g, const ll_host_edges* ins,
                                  const ll_host_edges* del) try {
  if (!g) return fail(LL_GPU_EINVAL, "graph is NULL");
  const uint64_t ni = ins ? ins->e.src.size() : 0, nd = del ? del->e.src.size() : 0;
  const bool w = ni && ins->w.size() == ni;  // .xs1 records carry weights
  return ll_host_graph_add_level_w(g, ni ? ins->e.src.data() : nullptr,
                                   ni ? ins->e.dst.data() : nullptr, w ? ins->w.data() : nullptr,
                                   ni, nd ? del->e.src.data() : nullptr,
                                   nd ? del->e.dst.data() : nullptr, nd);
} catch (...) {
  return host_exception();
}

// Vertex-table page size of the reference's copy-on-write array
// (LL_ENTRIES_PER_PAGE, ll_common.h:361-363).
constexpr uint64_t kPageEntries = 512;

int ll_host_graph_add_random_level(ll_host_graph* g, double ins_frac, double del_frac,
                                   uint64_t seed) {
  if (!g) return fail(LL_GPU_EINVAL, "graph is NULL");
  if (g->g->num_levels() == 0) return fail(LL_GPU_ESTATE, "random levels need a base level");
  if (del_frac > 0 && !g->g->deletions())
    return fail(LL_GPU_EINVAL, "deletions need a graph created with deletions on");
  if (ins_frac < 0 || del_frac < 0 || ins_frac > 1 || del_frac > 1)
    return fail(LL_GPU_EINVAL, "fractions must be in [0, 1]");
  try {
    if (!g->upd) g->upd.reset(new llhost::Updater(*g->g));
    llhost::EdgeList ins, del;
    g->upd->next(ins_frac, del_frac, seed, ins, del);
    llhost::vec<float> w;
    if (g->g->weighted()) uniform_weights(ins.src.size(), seed, w);  // U[0, 1), seeded
    if (!g->g->add_level(ins.src.data(), ins.dst.data(), ins.src.size(), del.src.data(),
                         del.dst.data(), del.src.size(), w.empty() ? nullptr : w.data())) {
      g->upd.reset();
      return fail(LL_GPU_ESTATE, g->g->last_error());
    }
    g->last_ins.e = std::move(ins);
    g->last_ins.w = std::move(w);
    g->last_del.e = std::move(del);
    g->last_del.w.clear();
  } catch (const std::bad_alloc&) {
    return fail(LL_GPU_ENOMEM, "out of host memory while drawing a level");
  }
  return LL_GPU_OK;
}

int ll_host_graph_add_window_level(ll_host_graph* g, uint64_t n_ins, uint64_t n_del, uint64_t lo,
                                   uint64_t hi, uint64_t seed) {
  if (!g) return fail(LL_GPU_EINVAL, "graph is NULL");
  if (g->g->num_levels() == 0) return fail(LL_GPU_ESTATE, "random levels need a base level");
  if (n_del > 0 && !g->g->deletions())
    return fail(LL_GPU_EINVAL, "deletions need a graph created with deletions on");
  if (lo >= hi || hi > g->g->num_nodes() || hi - lo < 2)
    return fail(LL_GPU_EINVAL, "window must satisfy lo + 2 <= hi <= max_nodes");
  if (n_ins > (hi - lo) * (hi - lo - 1) / 2)
    return fail(LL_GPU_EINVAL, "more inserts than the window can hold");
  try {
    if (!g->upd) g->upd.reset(new llhost::Updater(*g->g));
    llhost::EdgeList ins, del;
    g->upd->next_window(n_ins, n_del, lo, hi, seed, ins, del);
    llhost::vec<float> w;
    if (g->g->weighted()) uniform_weights(ins.src.size(), seed, w);  // U[0, 1), seeded
    if (!g->g->add_level(ins.src.data(), ins.dst.data(), ins.src.size(), del.src.data(),
                         del.dst.data(), del.src.size(), w.empty() ? nullptr : w.data())) {
      g->upd.reset();
      return fail(LL_GPU_ESTATE, g->g->last_error());
    }
    g->last_ins.e = std::move(ins);
    g->last_ins.w = std::move(w);
    g->last_del.e = std::move(del);
    g->last_del.w.clear();
  } catch (const std::bad_alloc&) {
    return fail(LL_GPU_ENOMEM, "out of host memory while drawing a level");
  }
  return LL_GPU_OK;
}

int ll_host_graph_last_batch(const ll_host_graph* g, const ll_host_edges** ins,
                             const ll_host_edges** del) try {
  if (!g || !ins || !del) return fail(LL_GPU_EINVAL, "NULL argument");
  *ins = &g->last_ins;
  *del = &g->last_del;
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

int ll_host_graph_level_view(const ll_host_graph* g, uint32_t dir, uint32_t level,
                             ll_gpu_level_view* out) try {
  if (!g || !out) return fail(LL_GPU_EINVAL, "NULL argument");
  if (dir > 1) return fail(LL_GPU_EINVAL, "dir must be 0 (out) or 1 (in)");
  if (int(level) >= g->g->num_levels()) return fail(LL_GPU_EINVAL, "no such level");
  const auto& lv = g->g->level(int(dir), int(level));
  std::memset(out, 0, sizeof(*out));
  out->dir = dir;
  out->level = level;
  out->num_nodes = g->g->num_nodes();
  out->num_vt_entries = lv.vt.size();
  out->num_et_words = lv.et.size();
  out->vt = lv.vt.data();
  out->et = lv.et.data();
  out->flags = LL_GPU_HAS_CONT | (g->g->deletions() ? LL_GPU_HAS_DELETIONS : 0u);
  out->page_table = nullptr;
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

int ll_host_graph_level_pages(const ll_host_graph* g, uint32_t dir, uint32_t level,
                              const void** pages, uint64_t num_pages, uint64_t* written) try {
  if (!g || !pages) return fail(LL_GPU_EINVAL, "NULL argument");
  if (dir > 1) return fail(LL_GPU_EINVAL, "dir must be 0 (out) or 1 (in)");
  if (int(level) >= g->g->num_levels()) return fail(LL_GPU_EINVAL, "no such level");
  const auto& lv = g->g->level(int(dir), int(level));
  const uint64_t np = lv.page_written.size();
  if (num_pages != np) return fail(LL_GPU_EINVAL, "num_pages must be ceil((max_nodes + 1) / 512)");
  // the builder records the pages it writes, as the reference's copy-on-write
  // array allocates a page on its first write (ll_mem_array.h:1122-1140)
  uint64_t n = 0;
  for (uint64_t p = 0; p < np; p++) {
    pages[p] = lv.page_written[p] ? static_cast<const void*>(lv.vt.data() + p * kPageEntries)
                                  : nullptr;
    n += lv.page_written[p] ? 1 : 0;
  }
  if (written) *written = n;
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

int ll_host_edges_fill_weights(ll_host_edges* e, uint64_t seed) {
  if (!e) return fail(LL_GPU_EINVAL, "edges are NULL");
  try {
    uniform_weights(e->e.src.size(), seed, e->w);
  } catch (const std::bad_alloc&) {
    return fail(LL_GPU_ENOMEM, "out of host memory for weights");
  }
  return LL_GPU_OK;
}

int ll_host_graph_is_weighted(const ll_host_graph* g) { return g ? int(g->g->weighted()) : 0; }

int ll_host_graph_level_weights(const ll_host_graph* g, uint32_t level, const float** w,
                                uint64_t* n) try {
  if (!g || !w || !n) return fail(LL_GPU_EINVAL, "NULL argument");
  if (int(level) >= g->g->num_levels()) return fail(LL_GPU_EINVAL, "no such level");
  const auto& lv = g->g->level(0, int(level));
  *w = lv.wt.empty() ? nullptr : lv.wt.data();
  *n = lv.wt.size();
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

int ll_host_graph_flatten_weights(ll_host_graph* g, int level, const float** w, uint64_t* n) try {
  if (!g || !w || !n) return fail(LL_GPU_EINVAL, "NULL argument");
  const int L = g->g->num_levels();
  if (L == 0) return fail(LL_GPU_ESTATE, "graph has no levels");
  if (level < 0) level = L - 1;
  if (level >= L) return fail(LL_GPU_EINVAL, "no such level");
  g->g->flatten_weights(level, g->flat_w);
  *w = g->flat_w.data();
  *n = g->flat_w.size();
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

int ll_host_graph_patches(const ll_host_graph* g, uint32_t dir, uint32_t level,
                          const uint64_t** edge_ids, const uint32_t** max_levels, uint64_t* n) try {
  if (!g || !edge_ids || !max_levels || !n) return fail(LL_GPU_EINVAL, "NULL argument");
  if (dir > 1 || int(level) >= g->g->num_levels()) return fail(LL_GPU_EINVAL, "no such level");
  const auto& p = g->g->patches(int(dir), int(level));
  *n = p.edge_ids.size();
  *edge_ids = p.edge_ids.data();
  *max_levels = p.maxlev.data();
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

int ll_host_graph_flatten(ll_host_graph* g, uint32_t dir, int level, const uint64_t** offsets,
                          const uint32_t** neighbours, uint64_t* num_edges) try {
  if (!g || !offsets || !neighbours || !num_edges) return fail(LL_GPU_EINVAL, "NULL argument");
  const int L = g->g->num_levels();
  if (L == 0) return fail(LL_GPU_ESTATE, "graph has no levels");
  if (level < 0) level = L - 1;
  if (dir > 1 || level >= L) return fail(LL_GPU_EINVAL, "no such level");
  g->g->flatten(int(dir), level, g->flat_ofs, g->flat_adj);
  *offsets = g->flat_ofs.data();
  *neighbours = g->flat_adj.data();
  *num_edges = g->flat_adj.size();
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

int ll_host_gen_rmat(uint32_t scale, double edge_factor, double a, double b, double c,
                     uint64_t seed, uint64_t perm_seed, ll_host_edges** out) {
  if (!out) return fail(LL_GPU_EINVAL, "out is NULL");
  if (scale < 1 || scale > 31) return fail(LL_GPU_ERANGE, "scale must be in [1, 31]");
  if (!(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0 + 1e-12) || edge_factor <= 0)
    return fail(LL_GPU_EINVAL, "bad R-MAT parameters");
  try {
    auto* e = new ll_host_edges();
    llhost::gen_rmat(scale, edge_factor, a, b, c, seed, perm_seed, e->e);
    *out = e;
  } catch (const std::bad_alloc&) {
    return fail(LL_GPU_ENOMEM, "out of host memory in the R-MAT generator");
  }
  return LL_GPU_OK;
}

int ll_host_edges_create(const uint32_t* src, const uint32_t* dst, const float* weight,
                         uint64_t n, ll_host_edges** out) try {
  if (!out || (n && (!src || !dst))) return fail(LL_GPU_EINVAL, "NULL argument");
  auto* e = new (std::nothrow) ll_host_edges();
  if (!e) return fail(LL_GPU_ENOMEM, "out of host memory");
  e->e.src.assign(src, src + n);
  e->e.dst.assign(dst, dst + n);
  if (weight) e->w.assign(weight, weight + n);
  *out = e;
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

void ll_host_edges_destroy(ll_host_edges* e) { delete e; }
uint64_t ll_host_edges_count(const ll_host_edges* e) { return e ? e->e.src.size() : 0; }
const uint32_t* ll_host_edges_src(const ll_host_edges* e) { return e ? e->e.src.data() : nullptr; }
const uint32_t* ll_host_edges_dst(const ll_host_edges* e) { return e ? e->e.dst.data() : nullptr; }
const float* ll_host_edges_weight(const ll_host_edges* e) {
  return e && !e->w.empty() ? e->w.data() : nullptr;
}

// X-Stream type 1: packed {u32 tail, u32 head, f32 weight}; the .ini side file
// carries vertices= and edges= (loaders/ll_load_xstream1.h:146-152, 243-273).
int ll_host_xs1_write(const char* path, const ll_host_edges* e, uint64_t num_nodes) try {
  if (!path || !e) return fail(LL_GPU_EINVAL, "NULL argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return fail(LL_GPU_EINVAL, std::string("cannot open ") + path);
  const uint64_t n = e->e.src.size();
  const size_t chunk = 1 << 16;
  std::vector<unsigned char> buf(chunk * 12);
  for (uint64_t i = 0; i < n; i += chunk) {
    size_t m = size_t(std::min<uint64_t>(chunk, n - i));
    for (size_t j = 0; j < m; j++) {
      float w = e->w.empty() ? 1.0f : e->w[i + j];
      std::memcpy(&buf[j * 12 + 0], &e->e.src[i + j], 4);
      std::memcpy(&buf[j * 12 + 4], &e->e.dst[i + j], 4);
      std::memcpy(&buf[j * 12 + 8], &w, 4);
    }
    if (std::fwrite(buf.data(), 12, m, f) != m) {
      std::fclose(f);
      return fail(LL_GPU_EINVAL, "short write");
    }
  }
  std::fclose(f);
  std::string ini = std::string(path) + ".ini";
  FILE* fi = std::fopen(ini.c_str(), "w");
  if (!fi) return fail(LL_GPU_EINVAL, "cannot write .ini");
  std::fprintf(fi, "[graph]\ntype=1\nname=%s\nvertices=%llu\nedges=%llu\n", path,
               (unsigned long long)num_nodes, (unsigned long long)n);
  std::fclose(fi);
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

int ll_host_xs1_read(const char* path, ll_host_edges** out, uint64_t* num_nodes) try {
  if (!path || !out) return fail(LL_GPU_EINVAL, "NULL argument");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(LL_GPU_EINVAL, std::string("cannot open ") + path);
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (sz < 0 || sz % 12 != 0) {
    std::fclose(f);
    return fail(LL_GPU_EINVAL, "file size is not a multiple of 12 bytes");
  }
  auto* e = new (std::nothrow) ll_host_edges();
  if (!e) { std::fclose(f); return fail(LL_GPU_ENOMEM, "out of host memory"); }
  uint64_t n = uint64_t(sz) / 12, vmax = 0;
  e->e.src.resize(n);
  e->e.dst.resize(n);
  e->w.resize(n);
  std::vector<unsigned char> buf(12);
  for (uint64_t i = 0; i < n; i++) {
    if (std::fread(buf.data(), 12, 1, f) != 1) {
      std::fclose(f);
      delete e;
      return fail(LL_GPU_EINVAL, "short read");
    }
    std::memcpy(&e->e.src[i], &buf[0], 4);
    std::memcpy(&e->e.dst[i], &buf[4], 4);
    std::memcpy(&e->w[i], &buf[8], 4);
    vmax = std::max<uint64_t>(vmax, std::max(e->e.src[i], e->e.dst[i]) + uint64_t(1));
  }
  std::fclose(f);
  uint64_t v_ini = 0;
  FILE* fi = std::fopen((std::string(path) + ".ini").c_str(), "r");
  if (fi) {
    char line[512];
    while (std::fgets(line, sizeof line, fi))
      if (std::strncmp(line, "vertices=", 9) == 0) v_ini = std::strtoull(line + 9, nullptr, 10);
    std::fclose(fi);
  }
  if (num_nodes) *num_nodes = std::max(vmax, v_ini);
  *out = e;
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

}  // extern "C"

// ---- SNAP .net text edge lists --------------------------------------------------
// Restates the reference loader's line rules (llama/include/llama/loaders/
// ll_load_net.h:312-375): lines starting with '#', '\r', '\n' or NUL are
// skipped; a record is "<ws>*tail<ws>+head" followed by end of line; anything
// else is a parse error, and the 100th error fails the load.  A partial load
// of part k of N (1-based) covers the byte range [size*(k-1)/N, size*k/N): it
// starts after the first newline at or past the range start (unless that is
// byte 0) and stops after the line that reaches the range end (:180-300).
// The reference tests the range end only after a record, so a record that
// follows a skipped line reaching the end is loaded by two parts; here every
// record lands in exactly one part.  IDs must fit in 32 bits (the reference
// reads them with strtol into unsigned).

namespace {

bool net_parse(const char* line, uint32_t* tail, uint32_t* head) {
  const char* s = line;
  char* e = nullptr;
  while (std::isspace(static_cast<unsigned char>(*s))) s++;
  errno = 0;
  const long t = std::strtol(s, &e, 10);
  if (s == e || *e == '\0' || !std::isspace(static_cast<unsigned char>(*e))) return false;
  s = e + 1;
  while (std::isspace(static_cast<unsigned char>(*s))) s++;
  const long h = std::strtol(s, &e, 10);
  if (s == e || (*e != '\0' && *e != '\r' && *e != '\n')) return false;
  if (errno || t < 0 || h < 0 || t > long(UINT32_MAX) || h > long(UINT32_MAX)) return false;
  *tail = uint32_t(t);
  *head = uint32_t(h);
  return true;
}

}  // namespace

int ll_host_net_read(const char* path, int part, int nparts, ll_host_edges** out,
                     uint64_t* num_nodes, uint64_t* parse_errors) try {
  if (!path || !out) return fail(LL_GPU_EINVAL, "NULL argument");
  if (nparts < 0 || (nparts > 0 && (part < 1 || part > nparts)))
    return fail(LL_GPU_EINVAL, "the partial load part ID is out of bounds");
  FILE* f = std::fopen(path, "rb");
  if (!f) return fail(LL_GPU_EINVAL, std::string("cannot open ") + path);
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  long start = 0, stop = 0;  // stop 0 = to the end of the file
  if (nparts > 0) {
    start = long(uint64_t(size) * uint64_t(part - 1) / uint64_t(nparts));
    stop = long(uint64_t(size) * uint64_t(part) / uint64_t(nparts));
  }
  long offset = 0;
  if (start > 0) {  // resume after the line that straddles the range start
    std::fseek(f, start - 1, SEEK_SET);
    offset = start - 1;
    int c;
    while ((c = std::fgetc(f)) != EOF) {
      offset++;
      if (c == '\n') break;
    }
  }
  auto* e = new (std::nothrow) ll_host_edges();
  if (!e) {
    std::fclose(f);
    return fail(LL_GPU_ENOMEM, "out of host memory");
  }
  char* line = nullptr;
  size_t cap = 0;
  ssize_t r;
  uint64_t errors = 0, vmax = 0;
  bool done = offset >= size;
  while (!done && (r = getline(&line, &cap, f)) != -1) {
    offset += r;
    // every consumed line counts towards the range end, skipped ones too
    if (stop > 0 && offset >= stop) done = true;
    const char c0 = line[0];
    if (c0 == '\0' || c0 == '#' || c0 == '\n' || c0 == '\r') continue;
    uint32_t t, h;
    if (!net_parse(line, &t, &h)) {
      if (++errors >= 100) {
        std::free(line);
        std::fclose(f);
        delete e;
        return fail(LL_GPU_EINVAL, "reached the limit of 100 parse errors");
      }
      continue;
    }
    e->e.src.push_back(t);
    e->e.dst.push_back(h);
    vmax = std::max<uint64_t>(vmax, std::max(t, h) + uint64_t(1));
  }
  std::free(line);
  std::fclose(f);
  if (num_nodes) *num_nodes = vmax;
  if (parse_errors) *parse_errors = errors;
  *out = e;
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}

int ll_host_net_write(const char* path, const ll_host_edges* e) try {
  if (!path || !e) return fail(LL_GPU_EINVAL, "NULL argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return fail(LL_GPU_EINVAL, std::string("cannot create ") + path);
  std::fprintf(f, "# Directed graph: %lu edges\n# FromNodeId\tToNodeId\n",
               (unsigned long)e->e.src.size());
  for (size_t i = 0; i < e->e.src.size(); i++)
    std::fprintf(f, "%u\t%u\n", e->e.src[i], e->e.dst[i]);
  const bool ok = std::fclose(f) == 0;
  return ok ? LL_GPU_OK : fail(LL_GPU_EINVAL, "write failed");
} catch (...) {
  return host_exception();
}

// ---- base + N splits of an edge list (utils/xs1-split.cpp) ----------------------
// Each record (each tail/head pair of records with `undirected`) goes to part 0
// with probability p_large, else to small part 1 + floor((r - p_large) /
// p_small), p_small = (1 - p_large) / n_small; a part that already holds its
// quota (floor(p * records), part 0 takes the remainder) passes the record to
// the next part with room.  r comes from a seeded generator instead of rand().
int ll_host_edges_split(const ll_host_edges* in, int n_small, double p_large, int undirected,
                        uint64_t seed, ll_host_edges** parts) try {
  if (!in || !parts) return fail(LL_GPU_EINVAL, "NULL argument");
  if (n_small <= 0) return fail(LL_GPU_EINVAL, "invalid number of small files");
  if (!(p_large >= 0 && p_large <= 1)) return fail(LL_GPU_EINVAL, "p_large must be in [0, 1]");
  const uint64_t n = in->e.src.size();
  const bool w = in->w.size() == n;
  if (undirected && n % 2) return fail(LL_GPU_EINVAL, "odd number of edges");
  const double p_small = (1.0 - p_large) / n_small;
  std::vector<uint64_t> expect(n_small + 1), have(n_small + 1, 0);
  uint64_t sum = expect[0] = uint64_t(p_large * double(n));
  for (int i = 1; i <= n_small; i++) sum += expect[i] = uint64_t(p_small * double(n));
  if (sum > n) return fail(LL_GPU_ESTATE, "split quotas exceed the records");
  expect[0] += n - sum;
  std::vector<ll_host_edges*> out(n_small + 1, nullptr);
  for (auto& p : out) {
    p = new (std::nothrow) ll_host_edges();
    if (!p) {
      for (auto* q : out) delete q;
      return fail(LL_GPU_ENOMEM, "out of host memory");
    }
  }
  const uint64_t step = undirected ? 2 : 1;
  for (uint64_t i = 0; i < n; i += step) {
    if (undirected && (in->e.src[i + 1] != in->e.dst[i] || in->e.dst[i + 1] != in->e.src[i])) {
      for (auto* q : out) delete q;
      return fail(LL_GPU_EINVAL, "not an undirected edge pair");
    }
    const double r = double(llhost::rng_u64(seed, 0x5311, i) >> 11) * (1.0 / 9007199254740992.0);
    int t = r < p_large ? 0 : 1 + int((r - p_large) / p_small);
    if (t > n_small) t = n_small;
    for (int k = 0; k <= n_small; k++) {
      if (have[t] < expect[t]) break;
      t = (t + 1) % (n_small + 1);
    }
    for (uint64_t j = i; j < i + step; j++) {
      out[t]->e.src.push_back(in->e.src[j]);
      out[t]->e.dst.push_back(in->e.dst[j]);
      if (w) out[t]->w.push_back(in->w[j]);
      have[t]++;
    }
  }
  for (int i = 0; i <= n_small; i++) parts[i] = out[i];
  return LL_GPU_OK;
} catch (...) {
  return host_exception();
}
