"""CPU oracle for the LLAMA read-only analytics path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; never by the product package.

``OracleGraph`` borrows the per-level vertex/edge tables of a host graph and
runs the C++ restatement in oracle.cc (built into _build/liboracle.so by
``build()``).  Which reference file each function follows, and how each is
pinned, is stated in oracle.cc's header.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

_lib = None


def build() -> Path:
    (HERE / "_build").mkdir(exist_ok=True)
    tmp = HERE / "_build" / "liboracle.so.tmp"
    subprocess.run(["g++", "-O3", "-std=c++17", "-fopenmp", "-fPIC", "-shared", "-DNDEBUG",
                    "-Wall", "-Wextra", str(HERE / "oracle.cc"), "-o", str(tmp)], check=True)
    tmp.replace(LIB)
    return LIB


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P, vp = C.POINTER, C.c_void_p
        u32, u64, i32, i64, f64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_int64, C.c_double
        sig = {
            "orc_graph_create": (vp, [u64, C.c_int]),
            "orc_graph_destroy": (None, [vp]),
            "orc_graph_add_level": (C.c_int, [vp, u32, vp, u64, P(u64), u64]),
            "orc_flatten": (i64, [vp, u32, C.c_int, P(u64), P(u32)]),
            "orc_count_check": (C.c_int, [vp, u32, C.c_int, P(u64), P(u64)]),
            "orc_pagerank_double": (C.c_int, [vp, C.c_int, C.c_int, f64, P(f64), P(f64)]),
            "orc_bfs_levels": (i64, [vp, C.c_int, i64, P(i32)]),
            "orc_triangles": (i64, [vp, C.c_int]),
            "orc_tc_i": (i64, [vp, C.c_int]),
            "orc_symmetrise": (i64, [vp, C.c_int, P(u64), P(u32)]),
            "orc_scc": (i64, [vp, C.c_int, P(i64)]),
            "orc_ppr": (C.c_int, [vp, C.c_int, P(i64), C.c_int, C.c_int, f64, f64, P(f64),
                                  P(C.c_int)]),
            "orc_wpr_forward": (C.c_int, [u64, P(u64), P(u32), P(f64), P(f64), C.c_int, f64,
                                          P(f64)]),
            "orc_wpr_adjoint": (C.c_int, [u64, P(u64), P(u32), P(f64), P(f64), P(f64), C.c_int,
                                          f64, P(f64), P(f64)]),
            "orc_graph_add_weights": (C.c_int, [vp, u32, P(C.c_float), u64]),
            "orc_sssp_weighted": (i64, [vp, C.c_int, i64, P(C.c_float), P(f64), P(i64)]),
            "orc_bc": (C.c_int, [vp, C.c_int, P(i64), i64, P(C.c_float)]),
            "orc_graph_set_min_level": (C.c_int, [vp, u32]),
            "orc_num_threads": (C.c_int, []),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _lib = L
    return _lib


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class OracleGraph:
    """Borrowed view of a multi-level host graph for the CPU restatement.

    ``levels`` is a list over levels of (vt_out, et_out, vt_in, et_in) numpy
    arrays (vt as raw 16-byte records), or use ``from_host``.
    """

    def __init__(self, num_nodes: int, deletions: bool, levels):
        self.V = num_nodes
        self._keep = []
        self._h = C.c_void_p(lib().orc_graph_create(num_nodes, int(deletions)))
        for lv in levels:
            for d in (0, 1):
                vt, et = lv[2 * d], lv[2 * d + 1]
                vt = np.ascontiguousarray(vt)
                et = np.ascontiguousarray(et, dtype=np.uint64)
                self._keep += [vt, et]
                rc = lib().orc_graph_add_level(self._h, d, vt.ctypes.data, vt.size,
                                               _p(et, C.c_uint64), et.size)
                if rc != 0:
                    raise ValueError("oracle rejected a level")

    @classmethod
    def from_host(cls, hg) -> "OracleGraph":
        levels = []
        for k in range(hg.num_levels):
            lo, li = hg.level_arrays(0, k), hg.level_arrays(1, k)
            levels.append((lo.vt, lo.et, li.vt, li.et))
        o = cls(hg.num_nodes, hg.deletions, levels)
        o._host = hg  # keep the borrowed arrays alive
        return o

    def set_min_level(self, m: int) -> None:
        """LL_MIN_LEVEL window: later snapshots skip segments of levels < m."""
        if lib().orc_graph_set_min_level(self._h, m) != 0:
            raise ValueError("the minimum level only moves forward")

    def flatten(self, direction: int, level: int = -1):
        n = lib().orc_flatten(self._h, direction, level, None, None)
        ofs = np.zeros(self.V + 1, np.uint64)
        adj = np.zeros(max(n, 1), np.uint32)
        lib().orc_flatten(self._h, direction, level, _p(ofs, C.c_uint64), _p(adj, C.c_uint32))
        return ofs, adj[:n]

    def count_check(self, direction: int, level: int = -1) -> tuple[int, int]:
        a, b = C.c_uint64(), C.c_uint64()
        lib().orc_count_check(self._h, direction, level, C.byref(a), C.byref(b))
        return int(a.value), int(b.value)

    def pagerank(self, iters: int, damping: float = 0.85, level: int = -1):
        r = np.zeros(self.V, np.float64)
        diff = C.c_double()
        rc = lib().orc_pagerank_double(self._h, level, iters, damping, _p(r, C.c_double),
                                       C.byref(diff))
        assert rc == 0
        return r, diff.value

    def add_weights(self, hg) -> None:
        """Borrow the out-direction weights of every level of host graph hg."""
        for k in range(hg.num_levels):
            w = hg.level_weights(k)
            if w is None:
                raise ValueError("host graph is unweighted")
            self._keep.append(w)
            if lib().orc_graph_add_weights(self._h, k, _p(w, C.c_float), w.size) != 0:
                raise ValueError("weights do not match the level's edge table")

    def sssp_weighted(self, root: int, level: int = -1):
        """sssp_weighted restated: (dist float32[V], reached, max distance, rounds)."""
        dist = np.zeros(self.V, np.float32)
        mx, rounds = C.c_double(), C.c_int64()
        n = lib().orc_sssp_weighted(self._h, level, root, _p(dist, C.c_float), C.byref(mx),
                                    C.byref(rounds))
        if n < 0:
            raise ValueError("sssp_weighted: bad level/root or unweighted graph")
        return dist, int(n), float(mx.value), int(rounds.value)

    def bc(self, sources=None, level: int = -1) -> np.ndarray:
        """Betweenness centrality in FP32 (bc_adj when sources is None, else the
        given sources as bc_random uses its random picks)."""
        out = np.zeros(self.V, np.float32)
        if sources is None:
            rc = lib().orc_bc(self._h, level, None, -1, _p(out, C.c_float))
        else:
            src = np.ascontiguousarray(sources, dtype=np.int64)
            rc = lib().orc_bc(self._h, level, _p(src, C.c_int64), src.size, _p(out, C.c_float))
        if rc != 0:
            raise ValueError("bc: bad level or source")
        return out

    def bfs(self, root: int, level: int = -1):
        dist = np.zeros(self.V, np.int32)
        n = lib().orc_bfs_levels(self._h, level, root, _p(dist, C.c_int32))
        return dist, int(n)

    def triangles(self, level: int = -1) -> int:
        return int(lib().orc_triangles(self._h, level))

    def tc_i(self, level: int = -1) -> int:
        """tc_i restated: triangles u<v<w counted once per direction of (u, v)."""
        return int(lib().orc_tc_i(self._h, level))

    def symmetrise(self, level: int = -1):
        n = lib().orc_symmetrise(self._h, level, None, None)
        ofs = np.zeros(self.V + 1, np.uint64)
        adj = np.zeros(max(n, 1), np.uint32)
        lib().orc_symmetrise(self._h, level, _p(ofs, C.c_uint64), _p(adj, C.c_uint32))
        return ofs, adj[:n]

    def scc(self, level: int = -1):
        lab = np.zeros(self.V, np.int64)
        n = lib().orc_scc(self._h, level, _p(lab, C.c_int64))
        return lab, int(n)

    def ppr(self, seeds, iters: int, damping: float = 0.85, tol: float = 0.0, level: int = -1):
        s = np.ascontiguousarray(seeds, dtype=np.int64)
        out = np.zeros((self.V, s.size), np.float64)
        done = C.c_int()
        lib().orc_ppr(self._h, level, _p(s, C.c_int64), s.size, iters, damping, tol,
                      _p(out, C.c_double), C.byref(done))
        return out, done.value

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().orc_graph_destroy(self._h)
            self._h = C.c_void_p()


def wpr_forward(ofs, adj, w, p, iters: int, damping: float):
    V = ofs.size - 1
    r = np.zeros(V, np.float64)
    ofs = np.ascontiguousarray(ofs, np.uint64)
    adj = np.ascontiguousarray(adj, np.uint32)
    w = np.ascontiguousarray(w, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    lib().orc_wpr_forward(V, _p(ofs, C.c_uint64), _p(adj, C.c_uint32), _p(w, C.c_double),
                          _p(p, C.c_double), iters, damping, _p(r, C.c_double))
    return r


def wpr_adjoint(ofs, adj, w, r, g, iters: int, damping: float):
    V = ofs.size - 1
    ofs = np.ascontiguousarray(ofs, np.uint64)
    adj = np.ascontiguousarray(adj, np.uint32)
    w = np.ascontiguousarray(w, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    g = np.ascontiguousarray(g, np.float64)
    dw = np.zeros(adj.size, np.float64)
    dp = np.zeros(V, np.float64)
    lib().orc_wpr_adjoint(V, _p(ofs, C.c_uint64), _p(adj, C.c_uint32), _p(w, C.c_double),
                          _p(r, C.c_double), _p(g, C.c_double), iters, damping,
                          _p(dw, C.c_double), _p(dp, C.c_double))
    return dw, dp


def num_threads() -> int:
    return int(lib().orc_num_threads())
