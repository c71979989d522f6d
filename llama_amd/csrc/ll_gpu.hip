// Context, device mirror of the frozen levels, uploads and deletion patches.
//
// The mirror keeps, per direction and level, the flattened vertex table
// VT_k[V+1] and the edge table ET_k exactly as the host level builder froze
// them (ll_mlcsr_sp.h:624-808, 1375-1607).  A device table of per-level ET base
// pointers lets kernels follow a continuation record's edge ID
// (level<<38 | index) into any older level, as iter_descend does
// (ll_mlcsr_sp.h:2090-2134).
#include <algorithm>
#include <chrono>
#include <cstring>

#include "ll_gpu_internal.h"

using llgpu::DevBuf;
using llgpu::Error;

namespace {

constexpr size_t kStageBytes = size_t(64) << 20;  // two halves of pinned staging

int fail(ll_gpu_ctx* c, const Error& e) {
  if (c) c->err = e.msg;
  return e.code;
}

template <typename F>
int guarded(ll_gpu_ctx* c, F&& f) {
  try {
    f();
    return LL_GPU_OK;
  } catch (const Error& e) {
    return fail(c, e);
  } catch (...) {
    return llgpu::api_exception(c);
  }
}

__global__ void k_zero_flag(uint32_t* p) { *p = 0; }

// int64 neighbour IDs of a single-level CSR to the mirror's uint32 IDs.
__global__ void k_narrow_ids(const int64_t* in, uint64_t n, uint64_t V, uint32_t* out,
                             uint32_t* bad) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t x = in[i];
  if (x < 0 || uint64_t(x) >= V) {
    atomicOr(bad, 1u);
    out[i] = 0;
  } else {
    out[i] = uint32_t(x);
  }
}

__global__ void k_apply_patches(const uint64_t* const* et_table, const uint64_t* ids,
                                const uint32_t* maxlev, uint64_t n) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t id = ids[i];
  uint64_t* et = const_cast<uint64_t*>(et_table[llgpu::edge_level(id)]);
  uint64_t idx = llgpu::edge_index(id);
  uint64_t w = et[idx];
  uint32_t ml = maxlev[i];
  // lower only, as update_max_visible_level_lower_only (ll_mlcsr_sp.h:849-863)
  if (ml < llgpu::word_maxlev(w))
    et[idx] = (w & llgpu::kIndexMask) | (uint64_t(ml) << llgpu::kBitsIndex);
}

// Copy host bytes to device through the double-buffered pinned staging area.
void upload(ll_gpu_ctx* c, void* dst, const void* src, size_t bytes) {
  const char* s = static_cast<const char*>(src);
  char* d = static_cast<char*>(dst);
  const size_t half = c->pinned_bytes / 2;
  int slot = 0;
  for (size_t off = 0; off < bytes; off += half, slot ^= 1) {
    size_t m = std::min(half, bytes - off);
    LLG_HIP(hipEventSynchronize(c->pinned_ev[slot]));
    char* stage = static_cast<char*>(c->pinned) + slot * half;
    std::memcpy(stage, s + off, m);
    LLG_HIP(hipMemcpyAsync(d + off, stage, m, hipMemcpyHostToDevice, c->stream));
    LLG_HIP(hipEventRecord(c->pinned_ev[slot], c->stream));
  }
}

}  // namespace

namespace llgpu {

TimedLaunch::TimedLaunch(ll_gpu_ctx* c, const char* n) : ctx(c), name(n) {
  if (!ctx->timer.on) return;
  auto take = [&]() {
    if (!ctx->timer.pool.empty()) {
      hipEvent_t e = ctx->timer.pool.back();
      ctx->timer.pool.pop_back();
      return e;
    }
    hipEvent_t e;
    LLG_HIP(hipEventCreate(&e));
    return e;
  };
  a = take();
  b = take();
  LLG_HIP(hipEventRecord(a, ctx->stream));
}

TimedLaunch::~TimedLaunch() {
  if (!a) return;
  if (hipEventRecord(b, ctx->stream) == hipSuccess)
    ctx->timer.pending.push_back({name, a, b});
}

int resolve_level(const ll_gpu_graph* g, int level) {
  if (g->slcsr) {
    LLG_CHECK(level == -1 || level == 0, LL_GPU_EINVAL, "a single-level CSR has only level 0");
    return 0;
  }
  int L = int(g->dir[0].vt.size());
  LLG_CHECK(L > 0 && int(g->dir[1].vt.size()) == L, LL_GPU_ESTATE,
            "graph mirror needs the same number of out- and in-levels (>= 1)");
  if (level < 0) level = L - 1;
  LLG_CHECK(level < L, LL_GPU_EINVAL, "snapshot level beyond the pushed levels");
  LLG_CHECK(size_t(level) >= g->level_deleted.size() || !g->level_deleted[level], LL_GPU_ESTATE,
            "snapshot level was deleted (ll_gpu_graph_delete_level)");
  return level;
}

void drop_snapshots(ll_gpu_graph* g) {
  for (int d = 0; d < 2; d++) {
    g->deg_level[d] = -1;
    g->deg[d].release();
    g->csr[d] = Csr();
    g->rb[d].clear();
  }
  g->sym_level = -1;
  g->sym_ofs.release();
  g->sym_adj.release();
  g->sym_mult.release();
  g->sym_rows.release();
  g->dead_level = -1;
  g->dead_in.release();
  g->reach_level = -1;
  g->reach_notr.release();
  g->pr_level = -1;
  g->pr_inv.release();
  g->pr_adj.release();
  g->pr_sadj.release();
  g->pr_sperm.release();
  g->pr_sort_cap = 0;
  g->pr_seg = llgpu::PrSegments();
  g->pr_hc = llgpu::PrHotCold();
  g->pr_bins = llgpu::PrBins();
  g->w_level = -1;
  g->wcsr = Csr();
  g->wperm.release();
  g->wrb.clear();
}

}  // namespace llgpu

extern "C" {

int ll_gpu_create(int device, ll_gpu_ctx** out) {
  if (!out) return LL_GPU_EINVAL;
  *out = nullptr;
  auto* c = new (std::nothrow) ll_gpu_ctx();
  if (!c) return LL_GPU_ENOMEM;
  int rc = guarded(c, [&]() {
    int n = 0;
    LLG_HIP(hipGetDeviceCount(&n));
    LLG_CHECK(device >= 0 && device < n, LL_GPU_EINVAL, "no such device");
    LLG_HIP(hipSetDevice(device));
    c->device = device;
    c->scratch = llgpu::scratch_setup(device);
    hipDeviceProp_t prop;
    LLG_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    LLG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    // Keep memory freed by the stream-ordered allocator in the device's pool:
    // with the default release threshold (0) every synchronisation returns it
    // and the next allocation of a snapshot-sized buffer maps pages again
    // (about a millisecond each at R-MAT scale 24).
    hipMemPool_t pool = nullptr;
    LLG_HIP(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t keep = UINT64_MAX;
    LLG_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    hipError_t e = hipHostMalloc(&c->pinned, kStageBytes, hipHostMallocDefault);
    LLG_CHECK(e == hipSuccess, LL_GPU_ENOMEM, "pinned staging allocation failed");
    c->pinned_bytes = kStageBytes;
    e = hipHostMalloc(&c->readback, kReadbackBytes, hipHostMallocDefault);
    LLG_CHECK(e == hipSuccess, LL_GPU_ENOMEM, "pinned read-back allocation failed");
    for (auto& ev : c->pinned_ev) {
      LLG_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
      LLG_HIP(hipEventRecord(ev, c->stream));
    }
  });
  if (rc != LL_GPU_OK) {
    ll_gpu_destroy(c);
    return rc;
  }
  *out = c;
  return LL_GPU_OK;
}

void ll_gpu_destroy(ll_gpu_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& r : c->timer.pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : c->timer.pool) (void)hipEventDestroy(e);
  for (auto& ev : c->pinned_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->readback) (void)hipHostFree(c->readback);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  // hand the memory kept in the device's stream-ordered pool back
  hipMemPool_t pool = nullptr;
  if (hipDeviceGetDefaultMemPool(&pool, c->device) == hipSuccess && pool)
    (void)hipMemPoolTrimTo(pool, 0);
  delete c;
}

const char* ll_gpu_last_error(const ll_gpu_ctx* c) { return c ? c->err.c_str() : "NULL context"; }

int ll_gpu_sync(ll_gpu_ctx* c) {
  if (!c) return LL_GPU_EINVAL;
  return guarded(c, [&]() { LLG_HIP(hipStreamSynchronize(c->stream)); });
}

void* ll_gpu_stream(ll_gpu_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

int ll_gpu_set_option(ll_gpu_ctx* c, const char* name, long value) {
  if (!c || !name) return LL_GPU_EINVAL;
  static const char* known[] = {"pr_reorder", "pr_cap", "pr_seg_bits", "pr_item_sort",
                                "pr_hot_bits", "pr_bin_bits", "sssp_delta_milli",
                                "sssp_min_queue", "bfs_alpha", "bfs_beta", "bfs_reach",
                                "bfs_batch_bins", "bfs_batch_bin_open"};
  try {
    for (const char* k : known)
      if (std::strcmp(k, name) == 0) {
        c->options[name] = value;
        return LL_GPU_OK;
      }
    c->err = std::string("unknown option ") + name;
    return LL_GPU_EINVAL;
  } catch (...) {
    return llgpu::api_exception(c);
  }
}

int ll_gpu_profile(ll_gpu_ctx* c, int on) {
  if (!c) return LL_GPU_EINVAL;
  return guarded(c, [&]() {
    LLG_HIP(hipStreamSynchronize(c->stream));
    for (auto& r : c->timer.pending) {
      c->timer.pool.push_back(r.a);
      c->timer.pool.push_back(r.b);
    }
    c->timer.pending.clear();
    c->timer.acc.clear();
    c->timer.on = on != 0;
  });
}

int ll_gpu_profile_get(ll_gpu_ctx* c, const char* kernel, uint64_t* launches, double* total_ms) {
  if (!c || !kernel) return LL_GPU_EINVAL;
  return guarded(c, [&]() {
    LLG_HIP(hipStreamSynchronize(c->stream));
    for (auto& r : c->timer.pending) {
      float ms = 0;
      LLG_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      auto& a = c->timer.acc[r.name];
      a.first += 1;
      a.second += ms;
      c->timer.pool.push_back(r.a);
      c->timer.pool.push_back(r.b);
    }
    c->timer.pending.clear();
    auto it = c->timer.acc.find(kernel);
    if (launches) *launches = it == c->timer.acc.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == c->timer.acc.end() ? 0.0 : it->second.second;
  });
}

int ll_gpu_graph_create(ll_gpu_ctx* c, uint64_t max_nodes, uint32_t flags, ll_gpu_graph** out) {
  if (!c || !out) return LL_GPU_EINVAL;
  *out = nullptr;
  return guarded(c, [&]() {
    LLG_CHECK(max_nodes > 0 && max_nodes < (uint64_t(1) << 32), LL_GPU_ERANGE,
              "max_nodes must be in [1, 2^32) (uint32 vertex IDs on the device)");
    LLG_CHECK((flags & ~(LL_GPU_HAS_DELETIONS | LL_GPU_HAS_CONT)) == 0, LL_GPU_ENOTSUP,
              "unknown graph flags");
    auto* g = new ll_gpu_graph();
    g->ctx = c;
    g->V = max_nodes;
    g->flags = flags;
    *out = g;
  });
}

void ll_gpu_graph_destroy(ll_gpu_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  delete g;
  // stream-ordered frees were queued by the buffers' destructors
}

}  // extern "C"

namespace {

// Upload the listed host pages packed: page i of `src` lands at
// dst + i * kVtPageEntries; entries[i] of it are valid (fewer than a page
// only for the last page of a table).  Staged through the two pinned halves.
void upload_pages(ll_gpu_ctx* c, llgpu::VtEntry* dst, const std::vector<const void*>& src,
                  const std::vector<uint32_t>& entries) {
  using llgpu::kVtPageEntries;
  const size_t page_bytes = kVtPageEntries * sizeof(llgpu::VtEntry);
  const size_t per_half = (c->pinned_bytes / 2) / page_bytes;
  int slot = 0;
  for (size_t i0 = 0; i0 < src.size(); i0 += per_half, slot ^= 1) {
    const size_t n = std::min(per_half, src.size() - i0);
    LLG_HIP(hipEventSynchronize(c->pinned_ev[slot]));
    char* stage = static_cast<char*>(c->pinned) + slot * (c->pinned_bytes / 2);
    for (size_t i = 0; i < n; i++)
      std::memcpy(stage + i * page_bytes, src[i0 + i], entries[i0 + i] * sizeof(llgpu::VtEntry));
    LLG_HIP(hipMemcpyAsync(dst + i0 * kVtPageEntries, stage, n * page_bytes,
                           hipMemcpyHostToDevice, c->stream));
    LLG_HIP(hipEventRecord(c->pinned_ev[slot], c->stream));
  }
}

// Shared by the full and the page-wise push: `pages` is null for a full
// vertex table in v->vt, else one host address per page (null = the previous
// level's page).
int push_level(ll_gpu_graph* g, const ll_gpu_level_view* v, const void* const* pages,
               uint64_t num_pages) {
  using llgpu::DevBuf;
  using llgpu::kVtPageEntries;
  using llgpu::VtEntry;
  ll_gpu_ctx* c = g->ctx;
  return guarded(c, [&]() {
    LLG_HIP(hipSetDevice(c->device));
    LLG_CHECK(v->dir <= 1, LL_GPU_EINVAL, "dir must be LL_GPU_DIR_OUT or LL_GPU_DIR_IN");
    LLG_CHECK(!g->slcsr, LL_GPU_ESTATE, "the graph holds a single-level CSR");
    auto& d = g->dir[v->dir];
    LLG_CHECK(v->level == d.vt.size(), LL_GPU_ESTATE,
              "levels must be pushed in order (expected level " + std::to_string(d.vt.size()) + ")");
    LLG_CHECK(v->page_table == nullptr, LL_GPU_EINVAL, "page_table is reserved and must be NULL");
    LLG_CHECK(v->num_nodes == g->V && v->num_vt_entries == g->V + 1, LL_GPU_EINVAL,
              "vertex table must have max_nodes+1 entries");
    LLG_CHECK(v->flags == g->flags, LL_GPU_EINVAL, "level flags differ from the graph flags");
    LLG_CHECK((pages != nullptr || v->vt != nullptr) && (v->num_et_words == 0 || v->et != nullptr),
              LL_GPU_EINVAL, "NULL level arrays");
    LLG_CHECK(v->level < llgpu::kLevelMask - 2, LL_GPU_ERANGE, "level IDs exhausted");
    const uint64_t np = (g->V + 1 + kVtPageEntries - 1) / kVtPageEntries;
    std::vector<const void*> src;
    std::vector<uint32_t> ents;
    std::vector<uint64_t> slot(np, ~uint64_t(0));  // page -> packed index, or shared
    auto valid = [&](uint64_t p) {
      return uint32_t(std::min<uint64_t>(kVtPageEntries, g->V + 1 - p * kVtPageEntries));
    };
    if (pages) {
      LLG_CHECK(num_pages == np, LL_GPU_EINVAL,
                "num_pages must be ceil((max_nodes + 1) / 512) = " + std::to_string(np));
      for (uint64_t p = 0; p < np; p++) {
        if (!pages[p]) {
          LLG_CHECK(v->level > 0, LL_GPU_EINVAL, "level 0 has no previous level to share pages with");
          continue;
        }
        slot[p] = src.size();
        src.push_back(pages[p]);
        ents.push_back(valid(p));
      }
    }
    llgpu::EventPair ev;
    hipEvent_t e0 = ev.a, e1 = ev.b;
    LLG_HIP(hipEventRecord(e0, c->stream));
    DevBuf vt, et, ptab;
    const size_t et_bytes = size_t(v->num_et_words) * sizeof(uint64_t);
    size_t vt_bytes = 0;
    std::vector<const VtEntry*> table(np);
    if (pages) {
      vt.alloc(std::max<size_t>(src.size(), 1) * kVtPageEntries * sizeof(VtEntry), c->stream);
      upload_pages(c, vt.as<VtEntry>(), src, ents);
      for (uint32_t x : ents) vt_bytes += size_t(x) * sizeof(VtEntry);
      for (uint64_t p = 0; p < np; p++)
        table[p] = slot[p] != ~uint64_t(0) ? vt.as<VtEntry>() + slot[p] * kVtPageEntries
                                           : d.vt_pages.back()[p];
    } else {
      vt_bytes = size_t(v->num_vt_entries) * sizeof(VtEntry);
      vt.alloc(np * kVtPageEntries * sizeof(VtEntry), c->stream);
      upload(c, vt.p, v->vt, vt_bytes);
      for (uint64_t p = 0; p < np; p++) table[p] = vt.as<VtEntry>() + p * kVtPageEntries;
    }
    ptab.alloc(np * sizeof(void*), c->stream);
    upload(c, ptab.p, table.data(), np * sizeof(void*));
    et.alloc(std::max<size_t>(et_bytes, 8), c->stream);
    if (et_bytes) upload(c, et.p, v->et, et_bytes);
    d.vt.push_back(std::move(vt));
    d.vt_pages.push_back(std::move(table));
    d.vt_ptab.push_back(std::move(ptab));
    d.et.push_back(std::move(et));
    d.et_words.push_back(v->num_et_words);
    std::vector<const uint64_t*> etab;
    for (auto& b : d.et) etab.push_back(b.as<const uint64_t>());
    DevBuf t;
    t.alloc(etab.size() * sizeof(void*), c->stream);
    LLG_HIP(hipMemcpyAsync(t.p, etab.data(), etab.size() * sizeof(void*),
                           hipMemcpyHostToDevice, c->stream));
    d.et_table = std::move(t);
    DevBuf wt;
    wt.alloc(d.et_words.size() * sizeof(uint64_t), c->stream);
    LLG_HIP(hipMemcpyAsync(wt.p, d.et_words.data(), d.et_words.size() * sizeof(uint64_t),
                           hipMemcpyHostToDevice, c->stream));
    d.et_words_dev = std::move(wt);
    LLG_HIP(hipEventRecord(e1, c->stream));
    LLG_HIP(hipEventSynchronize(e1));  // host arrays and the tables are released on return
    float ms = 0;
    LLG_HIP(hipEventElapsedTime(&ms, e0, e1));
    g->last_upload_ms = ms;
    g->upload_bytes += vt_bytes + et_bytes;
    llgpu::drop_snapshots(g);
  });
}

}  // namespace

extern "C" {

int ll_gpu_graph_push_level(ll_gpu_graph* g, const ll_gpu_level_view* v) {
  if (!g || !v) return LL_GPU_EINVAL;
  return push_level(g, v, nullptr, 0);
}

int ll_gpu_graph_push_level_pages(ll_gpu_graph* g, const ll_gpu_level_view* v,
                                  const void* const* pages, uint64_t num_pages) {
  if (!g || !v) return LL_GPU_EINVAL;
  if (!pages) {
    g->ctx->err = "NULL page list";
    return LL_GPU_EINVAL;
  }
  return push_level(g, v, pages, num_pages);
}

int ll_gpu_graph_push_slcsr(ll_gpu_graph* g, uint32_t dir, const int64_t* begin,
                            const int64_t* values, uint64_t num_edges) {
  if (!g) return LL_GPU_EINVAL;
  ll_gpu_ctx* c = g->ctx;
  return guarded(c, [&]() {
    LLG_HIP(hipSetDevice(c->device));
    LLG_CHECK(dir <= 1, LL_GPU_EINVAL, "dir must be LL_GPU_DIR_OUT or LL_GPU_DIR_IN");
    LLG_CHECK(!(g->flags & LL_GPU_HAS_DELETIONS), LL_GPU_EINVAL,
              "a single-level CSR has no deletions (LL_SLCSR excludes LL_DELETIONS)");
    LLG_CHECK(g->dir[0].vt.empty() && g->dir[1].vt.empty(), LL_GPU_ESTATE,
              "the graph already holds multi-level levels");
    LLG_CHECK(!g->sl[dir].valid, LL_GPU_ESTATE, "this direction is already pushed");
    LLG_CHECK(begin && (num_edges == 0 || values), LL_GPU_EINVAL, "NULL CSR arrays");
    const uint64_t V = g->V;
    LLG_CHECK(begin[0] == 0 && uint64_t(begin[V]) == num_edges, LL_GPU_EINVAL,
              "begin[0] must be 0 and begin[max_nodes] the number of edges");
    for (uint64_t v = 0; v < V; v++)
      LLG_CHECK(begin[v + 1] >= begin[v], LL_GPU_EINVAL, "begin must not decrease");
    llgpu::Csr cs;
    cs.level = 0;
    cs.V = V;
    cs.E = cs.stored = num_edges;
    cs.segments = 0;
    cs.ofs.alloc((V + 1) * 8, c->stream);
    upload(c, cs.ofs.p, begin, (V + 1) * 8);  // non-negative int64 = uint64 offsets
    cs.adj.alloc(std::max<uint64_t>(num_edges, 1) * 4, c->stream);
    if (num_edges) {
      DevBuf wide, bad;
      wide.alloc(num_edges * 8, c->stream);
      bad.alloc(4, c->stream);
      upload(c, wide.p, values, num_edges * 8);
      k_zero_flag<<<1, 1, 0, c->stream>>>(bad.as<uint32_t>());
      k_narrow_ids<<<llgpu::grid_for(num_edges, 256), 256, 0, c->stream>>>(
          wide.as<int64_t>(), num_edges, V, cs.adj.as<uint32_t>(), bad.as<uint32_t>());
      LLG_HIP(hipGetLastError());
      uint32_t hbad = 0;
      LLG_HIP(hipMemcpyAsync(&hbad, bad.p, 4, hipMemcpyDeviceToHost, c->stream));
      LLG_HIP(hipStreamSynchronize(c->stream));
      LLG_CHECK(hbad == 0, LL_GPU_EINVAL, "neighbour IDs must lie in [0, max_nodes)");
    }
    LLG_HIP(hipStreamSynchronize(c->stream));
    cs.valid = true;
    g->sl[dir] = std::move(cs);
    g->slcsr = true;
    g->upload_bytes += (V + 1) * 8 + num_edges * 8;
    llgpu::drop_snapshots(g);
  });
}

int ll_gpu_graph_set_min_level(ll_gpu_graph* g, uint32_t m) {
  if (!g) return LL_GPU_EINVAL;
  return guarded(g->ctx, [&]() {
    LLG_CHECK(m >= g->min_level, LL_GPU_ERANGE, "the minimum level only moves forward");
    g->min_level = m;
    llgpu::drop_snapshots(g);
  });
}

int ll_gpu_graph_delete_level(ll_gpu_graph* g, uint32_t level) {
  if (!g) return LL_GPU_EINVAL;
  ll_gpu_ctx* c = g->ctx;
  return guarded(c, [&]() {
    LLG_HIP(hipSetDevice(c->device));
    LLG_CHECK(level + 1 < g->min_level, LL_GPU_ESTATE,
              "only levels below min_level - 1 can be deleted");
    LLG_CHECK(level + 1 < g->dir[0].vt.size(), LL_GPU_EINVAL,
              "no such level, or the newest level (later pushes share its pages)");
    if (g->level_deleted.size() < g->dir[0].vt.size()) g->level_deleted.resize(g->dir[0].vt.size());
    if (g->level_deleted[level]) return;
    for (auto& d : g->dir) {
      // no snapshot inside the window walks into this level's edge table
      d.et[level].release();
      d.et_words[level] = 0;
      if (level < d.wt.size()) d.wt[level].release();
      // the level's own pages go unless a later level's page table shares them
      const auto* lo = d.vt[level].as<const llgpu::VtEntry>();
      const auto* hi = lo + d.vt[level].bytes / sizeof(llgpu::VtEntry);
      bool shared = false;
      for (size_t j = level + 1; j < d.vt_pages.size() && !shared; j++)
        for (const llgpu::VtEntry* p : d.vt_pages[j])
          if (p >= lo && p < hi) {
            shared = true;
            break;
          }
      if (!shared) d.vt[level].release();
      std::vector<const uint64_t*> etab;
      for (auto& b : d.et) etab.push_back(b.as<const uint64_t>());
      llgpu::DevBuf t;
      t.alloc(etab.size() * sizeof(void*), c->stream);
      LLG_HIP(hipMemcpyAsync(t.p, etab.data(), etab.size() * sizeof(void*), hipMemcpyHostToDevice,
                             c->stream));
      d.et_table = std::move(t);
      llgpu::DevBuf w;
      w.alloc(d.et_words.size() * sizeof(uint64_t), c->stream);
      LLG_HIP(hipMemcpyAsync(w.p, d.et_words.data(), d.et_words.size() * sizeof(uint64_t),
                             hipMemcpyHostToDevice, c->stream));
      d.et_words_dev = std::move(w);
      if (!d.wt.empty()) {
        std::vector<const float*> wtab(d.et.size(), nullptr);
        for (size_t k = 0; k < d.wt.size(); k++) wtab[k] = d.wt[k].as<const float>();
        llgpu::DevBuf wt;
        wt.alloc(wtab.size() * sizeof(void*), c->stream);
        LLG_HIP(hipMemcpyAsync(wt.p, wtab.data(), wtab.size() * sizeof(void*),
                               hipMemcpyHostToDevice, c->stream));
        d.wt_table = std::move(wt);
      }
    }
    g->level_deleted[level] = 1;
    LLG_HIP(hipStreamSynchronize(c->stream));  // the host tables are released on return
    llgpu::drop_snapshots(g);
  });
}

int ll_gpu_graph_push_weights(ll_gpu_graph* g, uint32_t level, const float* w, uint64_t n) {
  if (!g) return LL_GPU_EINVAL;
  ll_gpu_ctx* c = g->ctx;
  return guarded(c, [&]() {
    LLG_HIP(hipSetDevice(c->device));
    auto& d = g->dir[LL_GPU_DIR_OUT];
    LLG_CHECK(level < d.et.size(), LL_GPU_ESTATE, "weights for a level that is not pushed");
    LLG_CHECK(n == d.et_words[level], LL_GPU_EINVAL,
              "one weight per out-direction edge-table word of the level");
    LLG_CHECK(n == 0 || w != nullptr, LL_GPU_EINVAL, "NULL weights");
    for (uint64_t i = 0; i < n; i++)
      LLG_CHECK(w[i] >= 0.0f && w[i] <= 3.0e38f, LL_GPU_EINVAL,
                "weights must be finite and non-negative");
    if (d.wt.size() < d.et.size()) d.wt.resize(d.et.size());
    DevBuf b;
    b.alloc(std::max<size_t>(n * 4, 4), c->stream);
    if (n) upload(c, b.p, w, n * 4);
    d.wt[level] = std::move(b);
    std::vector<const float*> table(d.et.size(), nullptr);
    for (size_t k = 0; k < d.wt.size(); k++) table[k] = d.wt[k].as<const float>();
    DevBuf t;
    t.alloc(table.size() * sizeof(void*), c->stream);
    LLG_HIP(hipMemcpyAsync(t.p, table.data(), table.size() * sizeof(void*), hipMemcpyHostToDevice,
                           c->stream));
    d.wt_table = std::move(t);
    LLG_HIP(hipStreamSynchronize(c->stream));  // the host table is released on return
    g->upload_bytes += n * 4;
    llgpu::drop_snapshots(g);
  });
}

int ll_gpu_graph_patch_deletions(ll_gpu_graph* g, uint32_t dir, const uint64_t* ids,
                                 const uint32_t* maxlev, uint64_t n) {
  if (!g) return LL_GPU_EINVAL;
  ll_gpu_ctx* c = g->ctx;
  return guarded(c, [&]() {
    LLG_CHECK(dir <= 1, LL_GPU_EINVAL, "dir must be LL_GPU_DIR_OUT or LL_GPU_DIR_IN");
    if (n == 0) return;
    LLG_CHECK(ids && maxlev, LL_GPU_EINVAL, "NULL patch arrays");
    LLG_CHECK(g->flags & LL_GPU_HAS_DELETIONS, LL_GPU_ENOTSUP,
              "graph was created without deletions");
    auto& d = g->dir[dir];
    // words of deleted levels are outside every snapshot's window: their
    // patches change nothing and are dropped
    std::vector<uint64_t> keep_ids;
    std::vector<uint32_t> keep_ml;
    const bool any_deleted =
        std::find(g->level_deleted.begin(), g->level_deleted.end(), 1) != g->level_deleted.end();
    for (uint64_t i = 0; i < n; i++) {
      uint32_t l = llgpu::edge_level(ids[i]);
      LLG_CHECK(l < d.et.size(), LL_GPU_EINVAL, "patch edge ID outside the pushed levels");
      LLG_CHECK(maxlev[i] <= llgpu::kLevelMask, LL_GPU_EINVAL, "patch level out of range");
      if (l < g->level_deleted.size() && g->level_deleted[l]) continue;
      LLG_CHECK(llgpu::edge_index(ids[i]) < d.et_words[l], LL_GPU_EINVAL,
                "patch edge ID outside the pushed levels");
      if (any_deleted) {
        keep_ids.push_back(ids[i]);
        keep_ml.push_back(maxlev[i]);
      }
    }
    if (any_deleted) {
      ids = keep_ids.data();
      maxlev = keep_ml.data();
      n = keep_ids.size();
      if (n == 0) return;
    }
    LLG_HIP(hipSetDevice(c->device));
    DevBuf di, dm;
    di.alloc(n * 8, c->stream);
    dm.alloc(n * 4, c->stream);
    upload(c, di.p, ids, n * 8);
    upload(c, dm.p, maxlev, n * 4);
    k_apply_patches<<<llgpu::grid_for(n, 256), 256, 0, c->stream>>>(
        d.et_table.as<const uint64_t* const>(), di.as<uint64_t>(), dm.as<uint32_t>(), n);
    LLG_HIP(hipGetLastError());
    LLG_HIP(hipStreamSynchronize(c->stream));
    g->upload_bytes += n * 12;
    llgpu::drop_snapshots(g);
  });
}

int ll_gpu_graph_num_levels(const ll_gpu_graph* g, uint32_t dir) {
  if (!g || dir > 1) return LL_GPU_EINVAL;
  return int(g->dir[dir].vt.size());
}

uint64_t ll_gpu_graph_max_nodes(const ll_gpu_graph* g) { return g ? g->V : 0; }

int ll_gpu_graph_upload_stats(const ll_gpu_graph* g, uint64_t* bytes, double* last_ms) {
  if (!g) return LL_GPU_EINVAL;
  if (bytes) *bytes = g->upload_bytes;
  if (last_ms) *last_ms = g->last_upload_ms;
  return LL_GPU_OK;
}

void ll_gpu_graph_drop_snapshots(ll_gpu_graph* g) {
  if (!g) return;
  (void)hipSetDevice(g->ctx->device);
  llgpu::drop_snapshots(g);
}

}  // extern "C"
