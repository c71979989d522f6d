"""ctypes bindings of libll_gpu.so and libll_host.so (signatures from include/*.h).

The libraries are loaded from this package directory.  There is no fallback:
when a library is missing, loading raises with the build command to run.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent

u32, u64, i32, i64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_int64
f32, f64, vp, cp = C.c_float, C.c_double, C.c_void_p, C.c_char_p
P = C.POINTER


class LevelView(C.Structure):
    """ll_gpu_level_view (include/ll_gpu.h)."""

    _fields_ = [
        ("dir", u32), ("level", u32), ("num_nodes", u64), ("num_vt_entries", u64),
        ("num_et_words", u64), ("vt", vp), ("et", P(u64)), ("flags", u32),
        ("page_table", P(u32)),
    ]


class SnapshotInfo(C.Structure):
    """ll_gpu_snapshot_info (include/ll_gpu.h)."""

    _fields_ = [
        ("num_nodes", u64), ("live_edges", u64), ("stored_words", u64), ("segments", u64),
        ("continuations", u64), ("materialize_ms", f64),
    ]


GPU_SIGNATURES = {
    "ll_gpu_create": (C.c_int, [C.c_int, P(vp)]),
    "ll_gpu_destroy": (None, [vp]),
    "ll_gpu_last_error": (cp, [vp]),
    "ll_gpu_sync": (C.c_int, [vp]),
    "ll_gpu_stream": (vp, [vp]),
    "ll_gpu_profile": (C.c_int, [vp, C.c_int]),
    "ll_gpu_profile_get": (C.c_int, [vp, cp, P(u64), P(f64)]),
    "ll_gpu_graph_create": (C.c_int, [vp, u64, u32, P(vp)]),
    "ll_gpu_graph_destroy": (None, [vp]),
    "ll_gpu_graph_push_level": (C.c_int, [vp, P(LevelView)]),
    "ll_gpu_graph_push_level_pages": (C.c_int, [vp, P(LevelView), P(vp), u64]),
    "ll_gpu_graph_set_min_level": (C.c_int, [vp, u32]),
    "ll_gpu_graph_push_slcsr": (C.c_int, [vp, u32, P(i64), P(i64), u64]),
    "ll_gpu_graph_delete_level": (C.c_int, [vp, u32]),
    "ll_gpu_graph_patch_deletions": (C.c_int, [vp, u32, P(u64), P(u32), u64]),
    "ll_gpu_graph_num_levels": (C.c_int, [vp, u32]),
    "ll_gpu_graph_max_nodes": (u64, [vp]),
    "ll_gpu_graph_upload_stats": (C.c_int, [vp, P(u64), P(f64)]),
    "ll_gpu_graph_drop_snapshots": (None, [vp]),
    "ll_gpu_snapshot_stats": (C.c_int, [vp, C.c_int, P(SnapshotInfo)]),
    "ll_gpu_snapshot_adjacency": (C.c_int, [vp, C.c_int, u32, P(u64), P(u32)]),
    "ll_gpu_snapshot_degrees": (C.c_int, [vp, C.c_int, u32, P(u32)]),
    "ll_gpu_pagerank": (C.c_int, [vp, C.c_int, C.c_int, f64, P(f64), P(f64)]),
    "ll_gpu_pagerank_ex": (C.c_int, [vp, C.c_int, C.c_int, f64, u32, P(f64), P(f64)]),
    "ll_gpu_pagerank_push": (C.c_int, [vp, C.c_int, C.c_int, f64, P(f64), P(f64)]),
    "ll_gpu_graph_push_weights": (C.c_int, [vp, u32, P(f32), u64]),
    "ll_gpu_sssp": (C.c_int, [vp, C.c_int, i64, P(f32), P(u64), P(f64), P(u64)]),
    "ll_gpu_bc": (C.c_int, [vp, C.c_int, P(i64), i64, P(f32)]),
    "ll_gpu_bfs": (C.c_int, [vp, C.c_int, i64, P(i32), P(u64), P(u64)]),
    "ll_gpu_bfs_ex": (C.c_int, [vp, C.c_int, i64, P(i32), P(u64), P(u64), P(u64)]),
    "ll_gpu_bfs_noalloc": (C.c_int, [vp, C.c_int, i64, vp, P(u64), P(u64)]),
    "ll_gpu_bfs_batch_ex": (C.c_int, [vp, C.c_int, P(i64), C.c_int, C.c_int, P(i32), P(u64),
                                      P(u64)]),
    "ll_gpu_tc_ex": (C.c_int, [vp, C.c_int, P(u64), P(u64), P(u64)]),
    "ll_gpu_scc_ex": (C.c_int, [vp, C.c_int, P(i64), P(u64), P(u64)]),
    "ll_gpu_tuned": (C.c_int, [vp, cp, P(C.c_int)]),
    "ll_gpu_set_option": (C.c_int, [vp, cp, C.c_long]),
    "ll_gpu_bfs_batch": (C.c_int, [vp, C.c_int, P(i64), C.c_int, C.c_int, P(i32), P(u64)]),
    "ll_gpu_tc": (C.c_int, [vp, C.c_int, P(u64)]),
    "ll_gpu_tc_i": (C.c_int, [vp, C.c_int, P(u64)]),
    "ll_gpu_scc": (C.c_int, [vp, C.c_int, P(i64), P(u64)]),
    "ll_gpu_ppr_batch": (C.c_int, [vp, C.c_int, P(i64), C.c_int, C.c_int, f64, f64, C.c_int,
                                   P(f32), P(C.c_int)]),
    "ll_gpu_wpr_forward": (C.c_int, [vp, C.c_int, vp, vp, C.c_int, f64, vp]),
    "ll_gpu_wpr_adjoint": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_int, f64, vp, vp]),
}

HOST_SIGNATURES = {
    "ll_host_last_error": (cp, []),
    "ll_host_graph_create": (C.c_int, [u64, C.c_int, P(vp)]),
    "ll_host_graph_destroy": (None, [vp]),
    "ll_host_graph_num_nodes": (u64, [vp]),
    "ll_host_graph_num_levels": (C.c_int, [vp]),
    "ll_host_graph_has_deletions": (C.c_int, [vp]),
    "ll_host_graph_add_level": (C.c_int, [vp, P(u32), P(u32), u64, P(u32), P(u32), u64]),
    "ll_host_graph_add_level_w": (C.c_int, [vp, P(u32), P(u32), P(f32), u64, P(u32), P(u32),
                                            u64]),
    "ll_host_graph_is_weighted": (C.c_int, [vp]),
    "ll_host_edges_fill_weights": (C.c_int, [vp, u64]),
    "ll_host_net_read": (C.c_int, [cp, C.c_int, C.c_int, P(vp), P(u64), P(u64)]),
    "ll_host_net_write": (C.c_int, [cp, vp]),
    "ll_host_edges_split": (C.c_int, [vp, C.c_int, f64, C.c_int, u64, P(vp)]),
    "ll_host_graph_level_weights": (C.c_int, [vp, u32, P(P(f32)), P(u64)]),
    "ll_host_graph_flatten_weights": (C.c_int, [vp, C.c_int, P(P(f32)), P(u64)]),
    "ll_host_graph_add_random_level": (C.c_int, [vp, f64, f64, u64]),
    "ll_host_graph_last_batch": (C.c_int, [vp, P(vp), P(vp)]),
    "ll_host_graph_level_view": (C.c_int, [vp, u32, u32, P(LevelView)]),
    "ll_host_graph_level_pages": (C.c_int, [vp, u32, u32, P(vp), u64, P(u64)]),
    "ll_host_graph_add_window_level": (C.c_int, [vp, u64, u64, u64, u64, u64]),
    "ll_host_graph_patches": (C.c_int, [vp, u32, u32, P(P(u64)), P(P(u32)), P(u64)]),
    "ll_host_graph_flatten": (C.c_int, [vp, u32, C.c_int, P(P(u64)), P(P(u32)), P(u64)]),
    "ll_host_gen_rmat": (C.c_int, [u32, f64, f64, f64, f64, u64, u64, P(vp)]),
    "ll_host_edges_create": (C.c_int, [P(u32), P(u32), P(f32), u64, P(vp)]),
    "ll_host_edges_destroy": (None, [vp]),
    "ll_host_edges_count": (u64, [vp]),
    "ll_host_edges_src": (P(u32), [vp]),
    "ll_host_edges_dst": (P(u32), [vp]),
    "ll_host_edges_weight": (P(f32), [vp]),
    "ll_host_graph_add_level_edges": (C.c_int, [vp, vp, vp]),
    "ll_host_xs1_write": (C.c_int, [cp, vp, u64]),
    "ll_host_xs1_read": (C.c_int, [cp, P(vp), P(u64)]),
}


def _load(name: str, sigs: dict, optional_ok: bool = False) -> C.CDLL:
    path = PKG / name
    if not path.exists():
        raise ImportError(
            f"{path} is not built; run `python -m llama_amd.build` "
            "(there is no non-native fallback)")
    lib = C.CDLL(str(path))
    for fn, (res, args) in sigs.items():
        f = getattr(lib, fn, None)
        if f is None:
            if optional_ok:
                continue
            raise ImportError(f"{name} does not export {fn}")
        f.restype = res
        f.argtypes = args
    return lib


_gpu = None
_host = None


def gpu() -> C.CDLL:
    """libll_gpu.so.

    When PyTorch is installed it is imported first: torch bundles its own HIP
    and HSA runtimes with the same SONAMEs (libamdhip64.so.7,
    libhsa-runtime64.so.1) as /opt/rocm, so loading torch first makes this
    library bind to the runtime torch uses.  Loading them the other way round
    puts two HSA runtimes in one process and torch then finds no GPU.
    """
    global _gpu
    if _gpu is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _gpu = _load("libll_gpu.so", GPU_SIGNATURES)
    return _gpu


def host() -> C.CDLL:
    global _host
    if _host is None:
        _host = _load("libll_host.so", HOST_SIGNATURES)
    return _host


class NativeError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def check_host(rc: int) -> None:
    if rc != 0:
        raise NativeError(rc, host().ll_host_last_error().decode(errors="replace"))
