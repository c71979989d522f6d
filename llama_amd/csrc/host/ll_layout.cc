// Level builder for the host multi-level CSR (see ll_layout.h).
//
// Semantics follow LLAMA's level freeze:
//   * ll_mlcsr_core::init_node (llama/include/llama/ll_mlcsr_sp.h:1375-1550):
//     NIL / copy-previous / new-segment VT entry, degree = prev + new - deleted,
//     NIL when the degree drops to 0, 2-word continuation after every entry
//     whose start is not NIL on levels > 0 (also for deletion-only vertices);
//   * init_level_from_degrees + finish_level_vertices (:1562-1607, :786-808):
//     vertices laid out in ID order, sentinel VT[V] = {edge(k, words), 0, 0};
//   * ll_writable_graph::delete_edge on a frozen edge (ll_writable_graph.h:852-951)
//     -> update_max_visible_level_lower_only (ll_mlcsr_sp.h:849-863) on the
//     out-edge word and its in-edge twin, max level := k;
//   * ll_mlcsr_ro_graph::checkpoint (ll_mlcsr_graph.h:1501-1831): new out- and
//     in-edges are written per vertex in insertion order.
#include "ll_layout.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <numeric>

namespace llhost {

namespace {

// Parallel exclusive prefix sum; returns the total.
template <typename InV, typename OutV>
uint64_t exclusive_scan(const InV& in, OutV& out) {
  using U = typename OutV::value_type;
  const size_t n = in.size();
  out.resize(n);
  int nt = omp_get_max_threads();
  std::vector<uint64_t> part(nt + 1, 0);
#pragma omp parallel num_threads(nt)
  {
    int t = omp_get_thread_num();
    size_t lo = n * t / nt, hi = n * (t + 1) / nt;
    uint64_t s = 0;
    for (size_t i = lo; i < hi; i++) s += in[i];
    part[t + 1] = s;
#pragma omp barrier
#pragma omp single
    for (int i = 0; i < nt; i++) part[i + 1] += part[i];
    uint64_t acc = part[t];
    for (size_t i = lo; i < hi; i++) {
      out[i] = U(acc);
      acc += in[i];
    }
  }
  return part[nt];
}

// Parallel stable LSD radix sort of 64-bit keys on their low `bits` bits.
void radix_sort(vec<uint64_t>& a, int bits) {
  constexpr int R = 11;
  constexpr size_t B = size_t(1) << R;
  const size_t n = a.size();
  if (n < 2 || bits <= 0) return;
  vec<uint64_t> b(n);
  const int nt = omp_get_max_threads();
  std::vector<size_t> hist(size_t(nt) * B);
  for (int shift = 0; shift < bits; shift += R) {
#pragma omp parallel num_threads(nt)
    {
      const int t = omp_get_thread_num();
      const size_t lo = n * t / nt, hi = n * (t + 1) / nt;
      size_t* h = &hist[size_t(t) * B];
      std::fill(h, h + B, size_t(0));
      for (size_t i = lo; i < hi; i++) h[(a[i] >> shift) & (B - 1)]++;
#pragma omp barrier
#pragma omp single
      {
        size_t acc = 0;
        for (size_t d = 0; d < B; d++)
          for (int u = 0; u < nt; u++) {
            size_t c = hist[size_t(u) * B + d];
            hist[size_t(u) * B + d] = acc;
            acc += c;
          }
      }
      for (size_t i = lo; i < hi; i++) b[h[(a[i] >> shift) & (B - 1)]++] = a[i];
    }
    a.swap(b);
  }
}

int bit_width(uint64_t x) {
  int w = 0;
  while (x) { w++; x >>= 1; }
  return w;
}

}  // namespace

namespace {
// splitmix64 output function (Steele, Lea, Flood 2014)
inline uint64_t splitmix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
}  // namespace

uint64_t rng_u64(uint64_t seed, uint64_t stream, uint64_t counter) {
  return splitmix(splitmix(splitmix(seed) ^ (stream * 0xd1342543de82ef95ull)) ^ counter);
}

bool Graph::find_live_word(int dir, uint32_t u, uint32_t v, uint64_t* edge_id) const {
  const auto& lv = dir_[dir].levels;
  const int L = int(lv.size()) - 1;
  if (L < 0) return false;
  VtEntry e = lv[L].vt[u];
  while (e.start != kNilEdge && e.len > 0) {
    const uint32_t k = edge_level(e.start);
    const uint64_t idx = edge_index(e.start);
    const uint64_t* seg = lv[k].et.data() + idx;
    auto live = [&](uint64_t w) { return !deletions_ || word_maxlev(w) > uint32_t(L); };
    if (lv[k].sorted_segments) {
      const uint64_t* p = std::lower_bound(seg, seg + e.len, uint64_t(v),
          [](uint64_t w, uint64_t x) { return word_payload(w) < x; });
      for (; p < seg + e.len && word_payload(*p) == v; ++p)
        if (live(*p)) { *edge_id = edge_create(k, idx + uint64_t(p - seg)); return true; }
    } else {
      for (uint32_t j = 0; j < e.len; j++)
        if (word_payload(seg[j]) == v && live(seg[j])) {
          *edge_id = edge_create(k, idx + j);
          return true;
        }
    }
    if (k == 0) break;
    const uint64_t* cont = seg + e.len;
    e.start = int64_t(cont[0]);
    e.len = uint32_t(cont[1]);
    e.degree = uint32_t(cont[1] >> 32);
  }
  return false;
}

bool Graph::add_level(const uint32_t* ins_src, const uint32_t* ins_dst, uint64_t n_ins,
                      const uint32_t* del_src, const uint32_t* del_dst, uint64_t n_del,
                      const float* ins_w) {
  const uint32_t k = uint32_t(num_levels());
  if (k == 0) weighted_ = ins_w != nullptr;
  if (k >= kMaxLevel) { err_ = "level IDs exhausted"; return false; }
  if (k == 0 && n_del > 0) { err_ = "level 0 cannot delete edges"; return false; }
  if (n_del > 0 && !deletions_) { err_ = "deletions need a graph created with deletions on"; return false; }
  for (uint64_t i = 0; i < n_ins; i++)
    if (ins_src[i] >= V_ || ins_dst[i] >= V_) { err_ = "inserted edge endpoint out of range"; return false; }
  for (uint64_t i = 0; i < n_del; i++)
    if (del_src[i] >= V_ || del_dst[i] >= V_) { err_ = "deleted edge endpoint out of range"; return false; }
  if (uint64_t(n_ins) >= (uint64_t(1) << 32) && k > 0) { err_ = "too many inserts for one level"; return false; }

  // Locate the frozen words to delete (out-edge word and its in-edge twin).
  std::vector<uint64_t> eo(n_del), ei(n_del);
  std::vector<char> ok(n_del, 1);
#pragma omp parallel for schedule(dynamic, 1024)
  for (uint64_t i = 0; i < n_del; i++)
    ok[i] = find_live_word(0, del_src[i], del_dst[i], &eo[i]) &&
            find_live_word(1, del_dst[i], del_src[i], &ei[i]);
  for (uint64_t i = 0; i < n_del; i++)
    if (!ok[i]) {
      err_ = "deleted edge (" + std::to_string(del_src[i]) + "," + std::to_string(del_dst[i]) +
             ") is not live at the latest level";
      return false;
    }
  {
    std::vector<uint64_t> s(eo);
    std::sort(s.begin(), s.end());
    if (std::adjacent_find(s.begin(), s.end()) != s.end()) {
      err_ = "the same edge is deleted twice in one level";
      return false;
    }
  }

  vec<uint32_t> new_out(V_), new_in(V_), del_out(V_), del_in(V_);
#pragma omp parallel for schedule(static)
  for (uint64_t v = 0; v < V_; v++) new_out[v] = new_in[v] = del_out[v] = del_in[v] = 0;
  for (uint64_t i = 0; i < n_ins; i++) { new_out[ins_src[i]]++; new_in[ins_dst[i]]++; }
  for (uint64_t i = 0; i < n_del; i++) { del_out[del_src[i]]++; del_in[del_dst[i]]++; }

  // Rewrite the max-visible level of the deleted words in their old levels.
  patches_[0].emplace_back();
  patches_[1].emplace_back();
  for (int d = 0; d < 2; d++) {
    const auto& ids = d == 0 ? eo : ei;
    auto& out = patches_[d].back();
    out.edge_ids.reserve(n_del);
    out.maxlev.reserve(n_del);
    for (uint64_t i = 0; i < n_del; i++) {
      uint64_t& w = dir_[d].levels[edge_level(ids[i])].et[edge_index(ids[i])];
      if (k < word_maxlev(w)) {
        w = word_create(word_payload(w), k);
        out.edge_ids.push_back(ids[i]);
        out.maxlev.push_back(k);
      }
    }
  }

  bool src_sorted = true;
  for (uint64_t i = 1; i < n_ins && src_sorted; i++)
    src_sorted = ins_src[i - 1] <= ins_src[i];
  std::vector<float> ones;
  if (weighted_ && !ins_w) {  // an unweighted level of a weighted graph: weight 1
    ones.assign(n_ins, 1.0f);
    ins_w = ones.data();
  }
  if (!build_direction(0, k, new_out, del_out, ins_src, ins_dst, n_ins, src_sorted,
                       weighted_ ? ins_w : nullptr))
    return false;
  if (!build_direction(1, k, new_in, del_in, ins_dst, ins_src, n_ins, false)) return false;
  return true;
}

// Group `val` by `key` keeping insertion order within a group, then lay out
// level k of direction `dir` exactly as init_level_from_degrees/init_node do.
bool Graph::build_direction(int dir, uint32_t k, const vec<uint32_t>& new_cnt,
                            const vec<uint32_t>& del_cnt, const uint32_t* key,
                            const uint32_t* val, uint64_t n, bool key_sorted,
                            const float* wts) {
  vec<uint64_t> gofs;
  exclusive_scan(new_cnt, gofs);
  vec<uint32_t> grouped;
  vec<float> gw;
  const uint32_t* gval = val;
  const float* gwt = wts;
  if (!key_sorted) {
    grouped.resize(n);
    if (wts) gw.resize(n);
    if (n > (1u << 16)) {
      // stable group-by: sort (key << 32 | position), positions break ties
      vec<uint64_t> kp(n);
#pragma omp parallel for schedule(static)
      for (uint64_t i = 0; i < n; i++) kp[i] = (uint64_t(key[i]) << 32) | i;
      radix_sort(kp, 32 + bit_width(V_));
#pragma omp parallel for schedule(static)
      for (uint64_t i = 0; i < n; i++) {
        grouped[i] = val[kp[i] & 0xffffffffull];
        if (wts) gw[i] = wts[kp[i] & 0xffffffffull];
      }
    } else {
      std::vector<uint64_t> cur(gofs.begin(), gofs.end());
      for (uint64_t i = 0; i < n; i++) {
        const uint64_t p = cur[key[i]]++;
        grouped[p] = val[i];
        if (wts) gw[p] = wts[i];
      }
    }
    gval = grouped.data();
    if (wts) gwt = gw.data();
  }

  auto& levels = dir_[dir].levels;
  const VtEntry* prev = k > 0 ? levels[k - 1].vt.data() : nullptr;
  Level lv;
  lv.vt.resize(V_ + 1);

  // Pass 1: per-vertex entry and number of ET words it writes.
  vec<uint64_t> delta(V_);
  vec<char> touched(V_);
  bool bad_degree = false;
#pragma omp parallel for schedule(static) reduction(|| : bad_degree)
  for (uint64_t v = 0; v < V_; v++) {
    const uint32_t nw = new_cnt[v], nd = del_cnt[v];
    VtEntry e{kNilEdge, 0, 0};
    delta[v] = 0;
    touched[v] = 0;
    if (k > 0 && nw == 0 && nd == 0) {
      lv.vt[v] = prev[v];  // unchanged: LLAMA shares the COW page
      continue;
    }
    touched[v] = 1;
    if (k == 0 || nw > 0) {
      e.start = nw == 0 ? kNilEdge : 0;  // index filled in pass 2
      e.len = nw;
    } else {
      e = prev[v];
    }
    int64_t deg = int64_t(nw) - int64_t(nd);
    if (k > 0 && prev[v].start != kNilEdge) deg += prev[v].degree;
    if (deg < 0) { bad_degree = true; deg = 0; }
    e.degree = uint32_t(deg);
    if (deg == 0) { e.start = kNilEdge; e.len = 0; }
    uint64_t words = nw;
    if (k > 0 && e.start != kNilEdge) words += 2;  // continuation record
    lv.vt[v] = e;
    delta[v] = words;
  }
  if (bad_degree) { err_ = "negative degree while freezing a level"; return false; }

  vec<uint64_t> pos;
  const uint64_t total = exclusive_scan(delta, pos);
  lv.et.resize(total);
  if (gwt) lv.wt.resize(total);
  const bool dels = deletions_;

  // Pass 2: write segments, continuation records and final starts.
  bool sorted = true;
#pragma omp parallel for schedule(dynamic, 4096) reduction(&& : sorted)
  for (uint64_t v = 0; v < V_; v++) {
    if (!touched[v]) continue;
    const uint32_t nw = new_cnt[v];
    VtEntry& e = lv.vt[v];
    uint64_t w = pos[v];
    const uint32_t* src = gval + gofs[v];
    for (uint32_t j = 0; j < nw; j++) {
      lv.et[w + j] = dels ? word_create(src[j], kNeverDeleted) : uint64_t(src[j]);
      if (gwt) lv.wt[w + j] = gwt[gofs[v] + j];
      if (j > 0 && src[j - 1] > src[j]) sorted = false;
    }
    if (nw > 0 && e.start != kNilEdge) e.start = int64_t(edge_create(k, w));
    if (k > 0 && e.start != kNilEdge) {
      const VtEntry b = prev[v];
      lv.et[w + nw] = uint64_t(b.start);
      lv.et[w + nw + 1] = uint64_t(b.len) | (uint64_t(b.degree) << 32);
      if (gwt) lv.wt[w + nw] = lv.wt[w + nw + 1] = 0.0f;
    }
  }
  lv.vt[V_] = VtEntry{int64_t(edge_create(k, total)), 0, 0};
  lv.sorted_segments = sorted;
  // pages of VT written by this level (the copy-on-write pages): any touched
  // vertex, and the sentinel's page, which every level rewrites
  const uint64_t np = (V_ + 1 + kVtPageEntries - 1) / kVtPageEntries;
  lv.page_written.assign(np, k == 0 ? 1 : 0);
  if (k > 0) {
#pragma omp parallel for schedule(static)
    for (uint64_t p = 0; p < np; p++) {
      const uint64_t lo = p * kVtPageEntries, hi = std::min<uint64_t>(V_, lo + kVtPageEntries);
      char any = 0;
      for (uint64_t v = lo; v < hi && !any; v++) any = touched[v];
      lv.page_written[p] = any;
    }
    lv.page_written[V_ / kVtPageEntries] = 1;
  }
  levels.push_back(std::move(lv));
  return true;
}

void Graph::flatten_weights(int level, vec<float>& w) const {
  const auto& lv = dir_[0].levels;
  vec<uint64_t> ofs;
  vec<uint32_t> adj;
  flatten(0, level, ofs, adj);
  w.resize(adj.size());
  if (!weighted_) {
#pragma omp parallel for schedule(static)
    for (uint64_t i = 0; i < adj.size(); i++) w[i] = 1.0f;
    return;
  }
  const bool dels = deletions_;
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t v = 0; v < V_; v++) {
    uint64_t o = ofs[v], end = ofs[v + 1];
    VtEntry e = lv[level].vt[v];
    while (e.start != kNilEdge && e.len > 0 && o < end) {
      const uint32_t k = edge_level(e.start);
      const uint64_t* seg = lv[k].et.data() + edge_index(e.start);
      const float* sw = lv[k].wt.data() + edge_index(e.start);
      for (uint32_t j = 0; j < e.len; j++)
        if (!dels || word_maxlev(seg[j]) > uint32_t(level)) w[o++] = sw[j];
      if (k == 0) break;
      e.start = int64_t(seg[e.len]);
      e.len = uint32_t(seg[e.len + 1]);
    }
  }
}

void Graph::flatten(int dir, int level, vec<uint64_t>& ofs, vec<uint32_t>& adj) const {
  const auto& lv = dir_[dir].levels;
  vec<uint32_t> deg(V_);
#pragma omp parallel for
  for (uint64_t v = 0; v < V_; v++) deg[v] = lv[level].vt[v].degree;
  ofs.resize(V_ + 1);
  uint64_t total = exclusive_scan(deg, ofs);
  ofs.resize(V_ + 1);
  ofs[V_] = total;
  adj.resize(total);
  const bool dels = deletions_;
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t v = 0; v < V_; v++) {
    uint64_t o = ofs[v], end = ofs[v + 1];
    VtEntry e = lv[level].vt[v];
    while (e.start != kNilEdge && e.len > 0 && o < end) {
      const uint32_t k = edge_level(e.start);
      const uint64_t* seg = lv[k].et.data() + edge_index(e.start);
      for (uint32_t j = 0; j < e.len; j++)
        if (!dels || word_maxlev(seg[j]) > uint32_t(level)) adj[o++] = uint32_t(word_payload(seg[j]));
      if (k == 0) break;
      e.start = int64_t(seg[e.len]);
      e.len = uint32_t(seg[e.len + 1]);
    }
  }
}

void gen_rmat(uint32_t scale, double edge_factor, double a, double b, double c,
              uint64_t seed, uint64_t perm_seed, EdgeList& out) {
  const uint64_t V = uint64_t(1) << scale;
  const uint64_t n = uint64_t(std::llround(edge_factor * double(V)));
  // Seeded Fisher-Yates permutation of vertex labels.
  vec<uint32_t> perm(V);
  std::iota(perm.begin(), perm.end(), 0u);
  for (uint64_t i = V - 1; i > 0; i--) {
    uint64_t j = rng_u64(perm_seed, 0, i) % (i + 1);
    std::swap(perm[i], perm[j]);
  }
  const uint64_t ta = uint64_t(a * 4294967296.0);
  const uint64_t tab = uint64_t((a + b) * 4294967296.0);
  const uint64_t tabc = uint64_t((a + b + c) * 4294967296.0);
  vec<uint32_t> es(n), ed(n);
  const uint64_t sbase = rng_u64(seed, 0x52u, 0);
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; i++) {
    uint64_t u = 0, v = 0;
    // splitmix64 stream per edge: state = hash(seed, i) + j * golden
    const uint64_t st = splitmix(sbase ^ (i * 0xd1342543de82ef95ull));
    for (uint32_t l = 0; l < scale; l += 2) {
      uint64_t r = splitmix(st + uint64_t(l >> 1) * 0x9e3779b97f4a7c15ull);
      for (uint32_t h = 0; h < 2 && l + h < scale; h++) {
        uint64_t x = (r >> (32 * h)) & 0xffffffffull;
        uint64_t bu = x >= tab ? 1 : 0;
        uint64_t bv = (x >= ta && x < tab) || x >= tabc ? 1 : 0;
        u = (u << 1) | bu;
        v = (v << 1) | bv;
      }
    }
    es[i] = perm[u];
    ed[i] = perm[v];
  }
  // Sort (src, dst) keys, drop self-loops and duplicates.
  vec<uint64_t> key(n);
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; i++) key[i] = (uint64_t(es[i]) << scale) | ed[i];
  vec<uint32_t>().swap(es);
  vec<uint32_t>().swap(ed);
  radix_sort(key, 2 * int(scale));
  const uint64_t vmask = V - 1;
  vec<uint64_t> keep(n);
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; i++)
    keep[i] = ((key[i] >> scale) != (key[i] & vmask)) && (i == 0 || key[i] != key[i - 1]);
  vec<uint64_t> pos;
  const uint64_t m = exclusive_scan(keep, pos);
  out.src.resize(m);
  out.dst.resize(m);
#pragma omp parallel for schedule(static)
  for (uint64_t i = 0; i < n; i++)
    if (keep[i]) {
      out.src[pos[i]] = uint32_t(key[i] >> scale);
      out.dst[pos[i]] = uint32_t(key[i] & vmask);
    }
}

Updater::Updater(const Graph& g) : V_(g.num_nodes()) {
  if (g.num_levels() == 0) {
    ofs_.assign(V_ + 1, 0);
    return;
  }
  g.flatten(0, g.num_levels() - 1, ofs_, adj_);
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t v = 0; v < V_; v++) std::sort(adj_.begin() + ofs_[v], adj_.begin() + ofs_[v + 1]);
}

void Updater::next(double ins_frac, double del_frac, uint64_t seed, EdgeList& ins, EdgeList& del) {
  const uint64_t E = adj_.size();
  next_window(uint64_t(ins_frac * double(E)), std::min<uint64_t>(E, uint64_t(del_frac * double(E))),
              0, V_, seed, ins, del);
}

void Updater::next_window(uint64_t n_ins, uint64_t n_del, uint64_t lo, uint64_t hi,
                          uint64_t seed, EdgeList& ins, EdgeList& del) {
  // deletions and inserts among edges with both ends in [lo, hi)
  const uint64_t e_lo = ofs_[lo], E = ofs_[hi] - e_lo;
  const uint64_t W = hi - lo;
  const bool whole = lo == 0 && hi == V_;
  auto eligible = [&](uint64_t idx) { return whole || (adj_[idx] >= lo && adj_[idx] < hi); };
  uint64_t n_eligible = E;
  if (!whole) {
    n_eligible = 0;
    for (uint64_t i = e_lo; i < e_lo + E; i++) n_eligible += eligible(i) ? 1 : 0;
  }
  n_del = std::min(n_del, n_eligible);
  // Deletions: the first n_del distinct draws of a uniform live-edge index
  // (whole graph); in a window, a seeded partial shuffle of its edges.
  std::vector<uint64_t> pick;
  if (!whole) {
    std::vector<uint64_t> elig;
    elig.reserve(n_eligible);
    for (uint64_t i = e_lo; i < e_lo + E; i++)
      if (eligible(i)) elig.push_back(i);
    for (uint64_t i = 0; i < n_del; i++) {
      const uint64_t j = i + rng_u64(seed, 1, i) % (elig.size() - i);
      std::swap(elig[i], elig[j]);
    }
    pick.assign(elig.begin(), elig.begin() + n_del);
    std::sort(pick.begin(), pick.end());
  } else {
    std::vector<std::pair<uint64_t, uint64_t>> cand;  // (index, draw number)
    uint64_t drawn = 0;
    while (true) {
      uint64_t want = n_del + n_del / 8 + 16;
      for (uint64_t j = 0; j < want; j++, drawn++)
        cand.push_back({e_lo + rng_u64(seed, 1, drawn) % std::max<uint64_t>(E, 1), drawn});
      std::vector<std::pair<uint64_t, uint64_t>> s(cand);
      std::sort(s.begin(), s.end());
      std::vector<std::pair<uint64_t, uint64_t>> first;  // (draw number, index)
      for (size_t i = 0; i < s.size(); i++)
        if (i == 0 || s[i].first != s[i - 1].first) first.push_back({s[i].second, s[i].first});
      if (first.size() >= n_del || E == 0) {
        std::sort(first.begin(), first.end());
        pick.clear();
        for (uint64_t i = 0; i < n_del && i < first.size(); i++) pick.push_back(first[i].second);
        break;
      }
    }
    std::sort(pick.begin(), pick.end());
  }
  del.src.resize(pick.size());
  del.dst.resize(pick.size());
#pragma omp parallel for
  for (uint64_t i = 0; i < pick.size(); i++) {
    uint64_t idx = pick[i];
    uint64_t u = uint64_t(std::upper_bound(ofs_.begin(), ofs_.end(), idx) - ofs_.begin()) - 1;
    del.src[i] = uint32_t(u);
    del.dst[i] = adj_[idx];
  }
  // Inserts: uniform (u, v), u != v, absent at the latest level, distinct.
  std::vector<uint64_t> keys;
  uint64_t drawn = 0;
  while (keys.size() < n_ins) {
    uint64_t want = (n_ins - keys.size()) + (n_ins - keys.size()) / 8 + 16;
    std::vector<uint64_t> cand(want, ~0ull);
    const uint64_t base = drawn;
#pragma omp parallel for
    for (uint64_t j = 0; j < want; j++) {
      uint64_t u = lo + rng_u64(seed, 2, base + j) % W;
      uint64_t v = lo + rng_u64(seed, 3, base + j) % W;
      if (u == v) continue;
      auto b = adj_.begin() + ofs_[u], e = adj_.begin() + ofs_[u + 1];
      if (std::binary_search(b, e, uint32_t(v))) continue;
      cand[j] = (u << 32) | v;
    }
    drawn += want;
    for (uint64_t j = 0; j < want && keys.size() < n_ins; j++) {
      if (cand[j] == ~0ull) continue;
      keys.push_back(cand[j]);
    }
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  }
  ins.src.resize(keys.size());
  ins.dst.resize(keys.size());
  for (uint64_t i = 0; i < keys.size(); i++) {
    ins.src[i] = uint32_t(keys[i] >> 32);
    ins.dst[i] = uint32_t(keys[i] & 0xffffffffull);
  }
  apply(ins, del);
}

void Updater::apply(const EdgeList& ins, const EdgeList& del) {
  // Both batches are sorted by (src, dst).
  vec<uint64_t> iofs(V_ + 1), dofs(V_ + 1);
#pragma omp parallel for schedule(static)
  for (uint64_t v = 0; v <= V_; v++) iofs[v] = dofs[v] = 0;
  for (uint32_t s : ins.src) iofs[s + 1]++;
  for (uint32_t s : del.src) dofs[s + 1]++;
  for (uint64_t v = 0; v < V_; v++) { iofs[v + 1] += iofs[v]; dofs[v + 1] += dofs[v]; }
  vec<uint64_t> deg(V_), nofs;
#pragma omp parallel for
  for (uint64_t v = 0; v < V_; v++)
    deg[v] = (ofs_[v + 1] - ofs_[v]) - (dofs[v + 1] - dofs[v]) + (iofs[v + 1] - iofs[v]);
  uint64_t total = exclusive_scan(deg, nofs);
  nofs.resize(V_ + 1);
  nofs[V_] = total;
  vec<uint32_t> nadj(total);
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t v = 0; v < V_; v++) {
    uint64_t o = nofs[v];
    uint64_t a = ofs_[v], ae = ofs_[v + 1];
    uint64_t d = dofs[v], de = dofs[v + 1];
    uint64_t i = iofs[v], ie = iofs[v + 1];
    while (a < ae || i < ie) {
      if (a < ae && d < de && adj_[a] == del.dst[d]) { a++; d++; continue; }
      if (i >= ie || (a < ae && adj_[a] < ins.dst[i])) nadj[o++] = adj_[a++];
      else nadj[o++] = ins.dst[i++];
    }
  }
  ofs_.swap(nofs);
  adj_.swap(nadj);
}

}  // namespace llhost
