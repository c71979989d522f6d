"""Host-side graph objects: the LLAMA-layout multi-level graph and its HBM mirror.

``HostGraph`` wraps the native level builder (include/ll_host.h); it plays the
role of ``ll_database`` / ``ll_mlcsr_ro_graph`` ownership of the frozen levels
(llama/include/llama/ll_database.h:61-80, ll_mlcsr_graph.h:180-309).
``DeviceGraph`` is the HBM mirror (include/ll_gpu.h) kept current by pushing
each newly frozen level and replaying its deletion patches, the device side of
the level-freeze event ll_mlcsr_ro_graph::checkpoint (ll_mlcsr_graph.h:1501-1831).
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from . import _native as N

DIR_OUT, DIR_IN = 0, 1
HAS_DELETIONS, HAS_CONT = 0x1, 0x2
UNREACHED = 2**31 - 2  # INT_MAX-1, sssp.h:291


def _u32p(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


@dataclass
class LevelArrays:
    """Borrowed numpy views of one frozen level of one direction."""

    vt: np.ndarray      # structured (start int64, len uint32, degree uint32), V+1 rows
    et: np.ndarray      # uint64 words
    flags: int


VT_DTYPE = np.dtype([("start", "<i8"), ("len", "<u4"), ("degree", "<u4")])


class Edges:
    """Owned native edge list (ll_host_edges)."""

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)

    @classmethod
    def from_arrays(cls, src, dst, weight=None) -> "Edges":
        s = np.ascontiguousarray(src, dtype=np.uint32)
        d = np.ascontiguousarray(dst, dtype=np.uint32)
        if s.shape != d.shape:
            raise ValueError("src and dst differ in length")
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.float32)
        h = C.c_void_p()
        N.check_host(N.host().ll_host_edges_create(
            _u32p(s), _u32p(d),
            None if w is None else w.ctypes.data_as(C.POINTER(C.c_float)), s.size, C.byref(h)))
        return cls(h.value)

    @classmethod
    def rmat(cls, scale: int, edge_factor: float = 16.0, a: float = 0.57, b: float = 0.19,
             c: float = 0.19, seed: int = 1, perm_seed: int = 2) -> "Edges":
        """Seeded Graph500-style R-MAT, permuted, no self-loops, deduplicated."""
        h = C.c_void_p()
        N.check_host(N.host().ll_host_gen_rmat(scale, edge_factor, a, b, c, seed, perm_seed,
                                               C.byref(h)))
        return cls(h.value)

    def __len__(self) -> int:
        return int(N.host().ll_host_edges_count(self._h))

    def arrays(self) -> tuple[np.ndarray, np.ndarray]:
        n = len(self)
        if n == 0:
            return np.zeros(0, np.uint32), np.zeros(0, np.uint32)
        s = np.ctypeslib.as_array(N.host().ll_host_edges_src(self._h), shape=(n,))
        d = np.ctypeslib.as_array(N.host().ll_host_edges_dst(self._h), shape=(n,))
        return s.copy(), d.copy()

    def fill_weights(self, seed: int) -> "Edges":
        """Seeded U[0, 1) weight per edge (in place)."""
        N.check_host(N.host().ll_host_edges_fill_weights(self._h, seed))
        return self

    @classmethod
    def read_net(cls, path: str, part: int = 0, nparts: int = 0) -> tuple["Edges", int, int]:
        """SNAP .net edge list: (edges, num_nodes, parse errors skipped)."""
        h, nn, errs = C.c_void_p(), C.c_uint64(), C.c_uint64()
        N.check_host(N.host().ll_host_net_read(path.encode(), part, nparts, C.byref(h),
                                               C.byref(nn), C.byref(errs)))
        return cls(h.value), int(nn.value), int(errs.value)

    def write_net(self, path: str) -> None:
        N.check_host(N.host().ll_host_net_write(path.encode(), self._h))

    def split(self, n_small: int = 10, p_large: float = 0.8, undirected: bool = False,
              seed: int = 1) -> list["Edges"]:
        """Base + n_small parts the way utils/xs1-split.cpp cuts a file."""
        arr = (C.c_void_p * (n_small + 1))()
        N.check_host(N.host().ll_host_edges_split(self._h, n_small, p_large, int(undirected),
                                                  seed, arr))
        return [Edges(p) for p in arr]

    def write_xs1(self, path: str, num_nodes: int) -> None:
        N.check_host(N.host().ll_host_xs1_write(path.encode(), self._h, num_nodes))

    @classmethod
    def read_xs1(cls, path: str) -> tuple["Edges", int]:
        h, nn = C.c_void_p(), C.c_uint64()
        N.check_host(N.host().ll_host_xs1_read(path.encode(), C.byref(h), C.byref(nn)))
        return cls(h.value), int(nn.value)

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            N.host().ll_host_edges_destroy(self._h)
            self._h = C.c_void_p()


class HostGraph:
    """Multi-level snapshot CSR in LLAMA's read-only layout (host memory)."""

    def __init__(self, num_nodes: int, deletions: bool = True):
        h = C.c_void_p()
        N.check_host(N.host().ll_host_graph_create(num_nodes, int(deletions), C.byref(h)))
        self._h = h

    # ---- building --------------------------------------------------------------
    def add_level(self, ins_src=(), ins_dst=(), del_src=(), del_dst=(), weights=None) -> int:
        """Freeze one level (inserts in insertion order, deletions of live edges);
        `weights` (one per inserted edge) makes the graph weighted when given
        on its first level."""
        s = np.ascontiguousarray(ins_src, dtype=np.uint32)
        d = np.ascontiguousarray(ins_dst, dtype=np.uint32)
        ds = np.ascontiguousarray(del_src, dtype=np.uint32)
        dd = np.ascontiguousarray(del_dst, dtype=np.uint32)
        if s.shape != d.shape or ds.shape != dd.shape:
            raise ValueError("edge arrays differ in length")
        w = None
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.float32)
            if w.shape != s.shape:
                raise ValueError("one weight per inserted edge")
        N.check_host(N.host().ll_host_graph_add_level_w(
            self._h, _u32p(s), _u32p(d),
            w.ctypes.data_as(C.POINTER(C.c_float)) if w is not None else None, s.size,
            _u32p(ds), _u32p(dd), ds.size))
        return self.num_levels - 1

    @property
    def weighted(self) -> bool:
        return bool(N.host().ll_host_graph_is_weighted(self._h))

    def level_weights(self, level: int) -> np.ndarray | None:
        """Out-direction weights of one level, one per edge-table word."""
        p, n = C.POINTER(C.c_float)(), C.c_uint64()
        N.check_host(N.host().ll_host_graph_level_weights(self._h, level, C.byref(p), C.byref(n)))
        return np.ctypeslib.as_array(p, shape=(n.value,)) if n.value else None

    def flatten_weights(self, level: int = -1) -> np.ndarray:
        """Out-direction weights in flatten(DIR_OUT, level) order."""
        p, n = C.POINTER(C.c_float)(), C.c_uint64()
        N.check_host(N.host().ll_host_graph_flatten_weights(self._h, level, C.byref(p),
                                                            C.byref(n)))
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy() if n.value else \
            np.zeros(0, np.float32)

    def add_level_edges(self, ins: Edges | None, dels: Edges | None = None) -> int:
        N.check_host(N.host().ll_host_graph_add_level_edges(
            self._h, ins._h if ins is not None else None, dels._h if dels is not None else None))
        return self.num_levels - 1

    def add_random_level(self, ins_frac: float, del_frac: float, seed: int) -> int:
        N.check_host(N.host().ll_host_graph_add_random_level(self._h, ins_frac, del_frac, seed))
        return self.num_levels - 1

    def last_batch(self) -> tuple[tuple[np.ndarray, np.ndarray], tuple[np.ndarray, np.ndarray]]:
        pi, pd = C.c_void_p(), C.c_void_p()
        N.check_host(N.host().ll_host_graph_last_batch(self._h, C.byref(pi), C.byref(pd)))
        out = []
        for p in (pi, pd):
            n = int(N.host().ll_host_edges_count(p))
            if n == 0:
                out.append((np.zeros(0, np.uint32), np.zeros(0, np.uint32)))
                continue
            s = np.ctypeslib.as_array(N.host().ll_host_edges_src(p), shape=(n,)).copy()
            d = np.ctypeslib.as_array(N.host().ll_host_edges_dst(p), shape=(n,)).copy()
            out.append((s, d))
        return out[0], out[1]

    # ---- reading ---------------------------------------------------------------
    @property
    def num_nodes(self) -> int:
        return int(N.host().ll_host_graph_num_nodes(self._h))

    @property
    def num_levels(self) -> int:
        return int(N.host().ll_host_graph_num_levels(self._h))

    @property
    def deletions(self) -> bool:
        return bool(N.host().ll_host_graph_has_deletions(self._h))

    def level_view(self, direction: int, level: int) -> N.LevelView:
        v = N.LevelView()
        N.check_host(N.host().ll_host_graph_level_view(self._h, direction, level, C.byref(v)))
        return v

    def add_window_level(self, n_ins: int, n_del: int, lo: int, hi: int, seed: int) -> None:
        """Freeze a seeded level confined to vertex IDs [lo, hi)."""
        N.check_host(N.host().ll_host_graph_add_window_level(self._h, n_ins, n_del, lo, hi, seed))

    def level_pages(self, direction: int, level: int):
        """Copy-on-write pages of a level's vertex table: (ctypes array of host
        page addresses, None where the page equals the previous level's;
        number of written pages)."""
        np_ = (self.num_nodes + 1 + 511) // 512
        arr = (C.c_void_p * np_)()
        n = C.c_uint64()
        N.check_host(N.host().ll_host_graph_level_pages(self._h, direction, level, arr, np_,
                                                        C.byref(n)))
        return arr, int(n.value)

    def level_arrays(self, direction: int, level: int) -> LevelArrays:
        v = self.level_view(direction, level)
        vt_raw = (C.c_uint8 * (v.num_vt_entries * 16)).from_address(v.vt)
        vt = np.frombuffer(vt_raw, dtype=VT_DTYPE)
        if v.num_et_words:
            et = np.ctypeslib.as_array(v.et, shape=(v.num_et_words,))
        else:
            et = np.zeros(0, np.uint64)
        return LevelArrays(vt=vt, et=et, flags=v.flags)

    def patches(self, direction: int, level: int) -> tuple[np.ndarray, np.ndarray]:
        ids, ml, n = C.POINTER(C.c_uint64)(), C.POINTER(C.c_uint32)(), C.c_uint64()
        N.check_host(N.host().ll_host_graph_patches(self._h, direction, level, C.byref(ids),
                                                    C.byref(ml), C.byref(n)))
        if n.value == 0:
            return np.zeros(0, np.uint64), np.zeros(0, np.uint32)
        return (np.ctypeslib.as_array(ids, shape=(n.value,)).copy(),
                np.ctypeslib.as_array(ml, shape=(n.value,)).copy())

    def flatten(self, direction: int, level: int = -1) -> tuple[np.ndarray, np.ndarray]:
        """Snapshot CSR in iterator order (host restatement, for update sampling)."""
        o, a, e = C.POINTER(C.c_uint64)(), C.POINTER(C.c_uint32)(), C.c_uint64()
        N.check_host(N.host().ll_host_graph_flatten(self._h, direction, level, C.byref(o),
                                                    C.byref(a), C.byref(e)))
        ofs = np.ctypeslib.as_array(o, shape=(self.num_nodes + 1,)).copy()
        adj = (np.ctypeslib.as_array(a, shape=(e.value,)).copy() if e.value
               else np.zeros(0, np.uint32))
        return ofs, adj

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            N.host().ll_host_graph_destroy(self._h)
            self._h = C.c_void_p()


def build_multilevel(scale: int, levels: int = 1, edge_factor: float = 16.0,
                     ins_frac: float = 0.01, del_frac: float = 0.001, seed: int = 1,
                     deletions: bool = True, rmat=(0.57, 0.19, 0.19),
                     weighted: bool = False) -> HostGraph:
    """R-MAT base level (seed, seed+1 permutation) plus `levels-1` random levels
    (seed+2+k), the multi-level snapshot configuration; `weighted` gives every
    edge a seeded U[0, 1) weight (the levels' `weight` property)."""
    base = Edges.rmat(scale, edge_factor, *rmat, seed=seed, perm_seed=seed + 1)
    if weighted:
        base.fill_weights(seed)
    g = HostGraph(1 << scale, deletions=deletions)
    g.add_level_edges(base)
    del base
    for k in range(1, levels):
        g.add_random_level(ins_frac, del_frac if deletions else 0.0, seed + 2 + k)
    return g


class GpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class GpuContext:
    """One HIP device + stream (ll_gpu_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = N.gpu()
        h = C.c_void_p()
        rc = self.lib.ll_gpu_create(device, C.byref(h))
        if rc != 0:
            raise GpuError(rc, f"ll_gpu_create({device}) failed")
        self._h = h
        self._graphs = weakref.WeakSet()

    def check(self, rc: int) -> None:
        if rc != 0:
            raise GpuError(rc, self.lib.ll_gpu_last_error(self._h).decode(errors="replace"))

    def sync(self) -> None:
        self.check(self.lib.ll_gpu_sync(self._h))

    @property
    def stream(self) -> int:
        return int(self.lib.ll_gpu_stream(self._h) or 0)

    def set_option(self, name: str, value: int) -> None:
        self.check(self.lib.ll_gpu_set_option(self._h, name.encode(), int(value)))

    def profile(self, on: bool) -> None:
        self.check(self.lib.ll_gpu_profile(self._h, int(on)))

    def kernel_time(self, name: str) -> tuple[int, float]:
        n, ms = C.c_uint64(), C.c_double()
        self.check(self.lib.ll_gpu_profile_get(self._h, name.encode(), C.byref(n), C.byref(ms)))
        return int(n.value), float(ms.value)

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            for gr in list(getattr(self, "_graphs", ())):
                gr.close()  # graphs free their buffers on this context's stream
            self.lib.ll_gpu_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()


class DeviceGraph:
    """HBM mirror of a HostGraph (ll_gpu_graph)."""

    def __init__(self, ctx: GpuContext, host: HostGraph, cow: bool = False):
        """cow: push levels after the first as copy-on-write vertex-table pages
        (only the pages a level wrote travel), else whole tables."""
        self.cow = cow
        self.ctx = ctx
        self.lib = ctx.lib
        self.host = host
        flags = HAS_CONT | (HAS_DELETIONS if host.deletions else 0)
        h = C.c_void_p()
        ctx.check(self.lib.ll_gpu_graph_create(ctx._h, host.num_nodes, flags, C.byref(h)))
        self._h = h
        self.V = host.num_nodes
        ctx._graphs.add(self)
        self.sync_levels()

    def sync_levels(self) -> None:
        """Push every frozen host level not yet mirrored, with its patches."""
        have = self.lib.ll_gpu_graph_num_levels(self._h, DIR_OUT)
        for k in range(have, self.host.num_levels):
            self.push_prepared(self.prepare_level(k))

    def prepare_level(self, k: int) -> list:
        """Everything level k's push reads from the host graph (patches, level
        views, copy-on-write page lists, weights), collected up front: the
        push can then run on another thread while the host freezes level k+1
        (frozen levels are not moved or rewritten by later freezes)."""
        items = []
        for d in (DIR_OUT, DIR_IN):
            ids, ml = self.host.patches(d, k)
            if ids.size:
                items.append(("patch", d, ids, ml))
            v = self.host.level_view(d, k)
            pages = self.host.level_pages(d, k)[0] if self.cow else None
            items.append(("level", v, pages))
        w = self.host.level_weights(k)  # the level's `weight` property, if any
        if w is not None:
            items.append(("weights", k, w))
        return items

    def push_prepared(self, items: list) -> None:
        for it in items:
            if it[0] == "patch":
                _, d, ids, ml = it
                self.ctx.check(self.lib.ll_gpu_graph_patch_deletions(
                    self._h, d, ids.ctypes.data_as(C.POINTER(C.c_uint64)),
                    ml.ctypes.data_as(C.POINTER(C.c_uint32)), ids.size))
            elif it[0] == "level":
                _, v, pages = it
                if pages is not None:
                    self.ctx.check(self.lib.ll_gpu_graph_push_level_pages(
                        self._h, C.byref(v), pages, len(pages)))
                else:
                    self.ctx.check(self.lib.ll_gpu_graph_push_level(self._h, C.byref(v)))
            else:
                _, k, w = it
                self.ctx.check(self.lib.ll_gpu_graph_push_weights(
                    self._h, k, w.ctypes.data_as(C.POINTER(C.c_float)), w.size))

    @property
    def num_levels(self) -> int:
        return int(self.lib.ll_gpu_graph_num_levels(self._h, DIR_OUT))

    def set_min_level(self, m: int) -> None:
        """LL_MIN_LEVEL window: snapshots visit only segments of levels >= m."""
        self.ctx.check(self.lib.ll_gpu_graph_set_min_level(self._h, m))

    def delete_level(self, level: int) -> None:
        """Free a level below the window (level + 1 < min level)."""
        self.ctx.check(self.lib.ll_gpu_graph_delete_level(self._h, level))

    def drop_snapshots(self) -> None:
        self.lib.ll_gpu_graph_drop_snapshots(self._h)

    def upload_stats(self) -> tuple[int, float]:
        b, ms = C.c_uint64(), C.c_double()
        self.ctx.check(self.lib.ll_gpu_graph_upload_stats(self._h, C.byref(b), C.byref(ms)))
        return int(b.value), float(ms.value)

    def snapshot_info(self, level: int = -1) -> N.SnapshotInfo:
        info = N.SnapshotInfo()
        self.ctx.check(self.lib.ll_gpu_snapshot_stats(self._h, level, C.byref(info)))
        return info

    def adjacency(self, direction: int, level: int = -1) -> tuple[np.ndarray, np.ndarray]:
        info = self.snapshot_info(level)
        ofs = np.zeros(self.V + 1, np.uint64)
        # in- and out-directions hold the same number of live edges
        adj = np.zeros(max(int(info.live_edges), 1), np.uint32)
        self.ctx.check(self.lib.ll_gpu_snapshot_adjacency(
            self._h, level, direction, ofs.ctypes.data_as(C.POINTER(C.c_uint64)),
            adj.ctypes.data_as(C.POINTER(C.c_uint32))))
        return ofs, adj[: int(ofs[-1])]

    def degrees(self, direction: int, level: int = -1) -> np.ndarray:
        d = np.zeros(self.V, np.uint32)
        self.ctx.check(self.lib.ll_gpu_snapshot_degrees(self._h, level, direction,
                                                        d.ctypes.data_as(C.POINTER(C.c_uint32))))
        return d

    # ---- algorithms --------------------------------------------------------------
    def pagerank(self, iters: int = 10, damping: float = 0.85, level: int = -1,
                 cap: int = 0, want_ranks: bool = True) -> tuple[np.ndarray | None, float]:
        out = np.zeros(self.V, np.float64) if want_ranks else None
        diff = C.c_double()
        self.ctx.check(self.lib.ll_gpu_pagerank_ex(
            self._h, level, iters, damping, cap,
            out.ctypes.data_as(C.POINTER(C.c_double)) if out is not None else None,
            C.byref(diff)))
        return out, float(diff.value)

    def pagerank_push(self, iters: int = 10, damping: float = 0.85,
                      level: int = -1) -> tuple[np.ndarray, float]:
        """Push PageRank (FP64 atomics over out-edges): (ranks, last L1 change)."""
        out = np.zeros(self.V, np.float64)
        diff = C.c_double()
        self.ctx.check(self.lib.ll_gpu_pagerank_push(
            self._h, level, iters, damping, out.ctypes.data_as(C.POINTER(C.c_double)),
            C.byref(diff)))
        return out, float(diff.value)

    def sssp(self, root: int, level: int = -1, want_dist: bool = True):
        """Weighted SSSP: (dist float32[V] or None, reached, max distance, rounds)."""
        dist = np.zeros(self.V, np.float32) if want_dist else None
        reached, mx, rounds = C.c_uint64(), C.c_double(), C.c_uint64()
        self.ctx.check(self.lib.ll_gpu_sssp(
            self._h, level, root,
            dist.ctypes.data_as(C.POINTER(C.c_float)) if dist is not None else None,
            C.byref(reached), C.byref(mx), C.byref(rounds)))
        return dist, int(reached.value), float(mx.value), int(rounds.value)

    def bc(self, sources=None, level: int = -1) -> np.ndarray:
        """Betweenness centrality, FP32[V]: every vertex a source (bc_adj) when
        sources is None, else the given sources in order (bc_random)."""
        out = np.zeros(self.V, np.float32)
        if sources is None:
            src, n = None, 0
        else:
            src = np.ascontiguousarray(sources, dtype=np.int64)
            n = src.size
            if n == 0:
                return out
        self.ctx.check(self.lib.ll_gpu_bc(
            self._h, level, src.ctypes.data_as(C.POINTER(C.c_int64)) if src is not None else None,
            n, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def bfs(self, root: int, level: int = -1, want_levels: bool = True):
        """Direction-optimising BFS: (levels or None, reached, edges examined)."""
        lv = np.zeros(self.V, np.int32) if want_levels else None
        reached, exam, trav = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.ctx.check(self.lib.ll_gpu_bfs_ex(
            self._h, level, root, lv.ctypes.data_as(C.POINTER(C.c_int32)) if lv is not None
            else None, C.byref(reached), C.byref(exam), C.byref(trav)))
        self.last_bfs_traversed = int(trav.value)
        return lv, int(reached.value), int(exam.value)

    def bfs_batch(self, roots, lane_width: int = 64, level: int = -1, want_levels: bool = True):
        """Multi-source BFS: (levels [n x V] or None, reached [n], pull sweeps)."""
        r = np.ascontiguousarray(roots, dtype=np.int64)
        n = r.size
        lv = np.zeros((n, self.V), np.int32) if want_levels else None
        reached = np.zeros(n, np.uint64)
        sweeps = C.c_uint64()
        self.ctx.check(self.lib.ll_gpu_bfs_batch_ex(
            self._h, level, r.ctypes.data_as(C.POINTER(C.c_int64)), n, lane_width,
            lv.ctypes.data_as(C.POINTER(C.c_int32)) if lv is not None else None,
            reached.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(sweeps)))
        return lv, reached, int(sweeps.value)

    def tc_i(self, level: int = -1) -> int:
        """tc_i: triangles u<v<w counted once per directed edge between u and v."""
        n = C.c_uint64()
        self.ctx.check(self.lib.ll_gpu_tc_i(self._h, level, C.byref(n)))
        return int(n.value)

    def triangles(self, level: int = -1, stats: bool = False):
        t, arcs, work = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self.ctx.check(self.lib.ll_gpu_tc_ex(self._h, level, C.byref(t), C.byref(arcs),
                                             C.byref(work)))
        if stats:
            return int(t.value), int(arcs.value), int(work.value)
        return int(t.value)

    def scc(self, level: int = -1, want_labels: bool = True):
        """Canonical SCC labels (min vertex ID per component) and the count."""
        lab = np.zeros(self.V, np.int64) if want_labels else None
        n, sweeps = C.c_uint64(), C.c_uint64()
        self.ctx.check(self.lib.ll_gpu_scc_ex(
            self._h, level, lab.ctypes.data_as(C.POINTER(C.c_int64)) if lab is not None else None,
            C.byref(n), C.byref(sweeps)))
        self.last_scc_sweeps = int(sweeps.value)
        return lab, int(n.value)

    def ppr_batch(self, seeds, iters: int = 30, damping: float = 0.85, tol: float = 1e-7,
                  batch_width: int = 0, level: int = -1, want_ranks: bool = True):
        """Batched personalized PageRank: (ranks [V x n] FP32 or None, iterations run)."""
        sd = np.ascontiguousarray(seeds, dtype=np.int64)
        out = np.zeros((self.V, sd.size), np.float32) if want_ranks else None
        done = C.c_int()
        self.ctx.check(self.lib.ll_gpu_ppr_batch(
            self._h, level, sd.ctypes.data_as(C.POINTER(C.c_int64)), sd.size, iters, damping,
            tol, batch_width, out.ctypes.data_as(C.POINTER(C.c_float)) if out is not None
            else None, C.byref(done)))
        return out, int(done.value)

    def tuned(self, prefix: str) -> int | None:
        v = C.c_int()
        rc = self.lib.ll_gpu_tuned(self.ctx._h, prefix.encode(), C.byref(v))
        return int(v.value) if rc == 0 else None

    def wpr_forward(self, w_ptr: int, p_ptr: int, r_ptr: int, iters: int, damping: float,
                    level: int = -1) -> None:
        """Weighted PageRank forward on device pointers (see ll_gpu.h)."""
        self.ctx.check(self.lib.ll_gpu_wpr_forward(self._h, level, w_ptr, p_ptr, iters, damping,
                                                   r_ptr))

    def wpr_adjoint(self, w_ptr: int, r_ptr: int, g_ptr: int, iters: int, damping: float,
                    dw_ptr: int, dp_ptr: int, level: int = -1) -> None:
        self.ctx.check(self.lib.ll_gpu_wpr_adjoint(self._h, level, w_ptr, r_ptr, g_ptr, iters,
                                                   damping, dw_ptr, dp_ptr))

    def close(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            self.lib.ll_gpu_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()


class SlcsrGraph(DeviceGraph):
    """HBM mirror of a single-level CSR (the LL_SLCSR configuration): one
    (begin[V+1], values[E]) pair per direction, snapshot level 0 only."""

    def __init__(self, ctx: GpuContext, num_nodes: int, out_csr, in_csr=None):
        self.cow = False
        self.ctx = ctx
        self.lib = ctx.lib
        self.host = None
        h = C.c_void_p()
        ctx.check(self.lib.ll_gpu_graph_create(ctx._h, num_nodes, 0, C.byref(h)))
        self._h = h
        self.V = num_nodes
        ctx._graphs.add(self)
        for d, csr in ((DIR_OUT, out_csr), (DIR_IN, in_csr)):
            if csr is None:
                continue
            b = np.ascontiguousarray(csr[0], dtype=np.int64)
            v = np.ascontiguousarray(csr[1], dtype=np.int64)
            ctx.check(self.lib.ll_gpu_graph_push_slcsr(
                self._h, d, b.ctypes.data_as(C.POINTER(C.c_int64)),
                v.ctypes.data_as(C.POINTER(C.c_int64)), v.size))

    def sync_levels(self) -> None:
        pass


def csr_of(num_nodes: int, src, dst):
    """(begin, values) of the edges grouped by src in the given order."""
    src = np.asarray(src, dtype=np.int64)
    dst = np.asarray(dst, dtype=np.int64)
    order = np.argsort(src, kind="stable")
    begin = np.zeros(num_nodes + 1, np.int64)
    np.add.at(begin, src + 1, 1)
    return np.cumsum(begin), dst[order]
