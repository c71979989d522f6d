"""ll_gpu_bfs_noalloc (the BFS entry point that allocates nothing) against
ll_gpu_bfs_ex on the same snapshots, through the C ABI: levels bit-exact,
reached and examined-edge counts equal, on two calls in a row (the second on a
smaller graph, so stale contents of the shared device array would show), plus
the host-side refusals (unprepared snapshot, graph too large, missing output)."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_case
from llama_amd import _native as N
from llama_amd.graph import DIR_OUT, DeviceGraph, HostGraph, build_multilevel

pytestmark = pytest.mark.gpu


def _roots(g, k, seed):
    o, _ = g.flatten(DIR_OUT)
    cand = np.flatnonzero(np.diff(o) > 0)
    return [int(x) for x in np.random.default_rng(seed).choice(cand, k, replace=False)]


def _noalloc(dg, root, level=-1):
    import torch
    lv = torch.full((dg.V,), -7, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    reached, exam = C.c_uint64(), C.c_uint64()
    rc = N.gpu().ll_gpu_bfs_noalloc(dg._h, level, root, C.c_void_p(lv.data_ptr()),
                                    C.byref(reached), C.byref(exam))
    return rc, lv.cpu().numpy(), int(reached.value), int(exam.value)


def test_noalloc_bfs_matches_bfs_ex_twice(ctx):
    big = build_multilevel(15, levels=4, ins_frac=0.02, del_frac=0.005, seed=31)
    small = build_multilevel(10, levels=2, ins_frac=0.05, del_frac=0.01, seed=32)
    dbig, dsmall = DeviceGraph(ctx, big), DeviceGraph(ctx, small)
    cases = [(dbig, r) for r in _roots(big, 2, 1)] + [(dsmall, r) for r in _roots(small, 2, 2)]
    for dg, root in cases:
        gpu_case(f"bfs_noalloc V={dg.V} root={root}")
        ref, n_ref, ex_ref = dg.bfs(root)  # existing entry point (also prepares)
        rc, lv, n, ex = _noalloc(dg, root)
        assert rc == 0, ctx.lib.ll_gpu_last_error(ctx._h)
        assert np.array_equal(lv, ref)
        assert n == n_ref and ex == ex_ref
    dbig.close()
    dsmall.close()


def test_noalloc_bfs_refusals(ctx):
    g = build_multilevel(10, levels=2, ins_frac=0.05, del_frac=0.01, seed=33)
    dg = DeviceGraph(ctx, g)
    root = _roots(g, 1, 3)[0]
    gpu_case("bfs_noalloc refusals")
    rc, *_ = _noalloc(dg, root)  # snapshot structures not built yet
    assert rc == -4
    dg.bfs(root)
    rc, *_ = _noalloc(dg, root, level=0)  # another level: not built either
    assert rc == -4
    reached = C.c_uint64()
    assert N.gpu().ll_gpu_bfs_noalloc(dg._h, -1, root, None, C.byref(reached), None) == -1
    assert N.gpu().ll_gpu_bfs_noalloc(dg._h, -1, dg.V, C.c_void_p(1), None, None) == -1
    dg.close()
    # working memory beyond the fixed array: refused before any device work
    h = C.c_void_p()
    assert N.gpu().ll_gpu_graph_create(ctx._h, 1 << 23, 1, C.byref(h)) == 0
    assert N.gpu().ll_gpu_bfs_noalloc(h, -1, 0, C.c_void_p(1), None, None) == -6
    N.gpu().ll_gpu_graph_destroy(h)
