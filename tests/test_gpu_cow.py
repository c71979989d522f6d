"""Copy-on-write level upload (ll_gpu_graph_push_level_pages): a mirror fed
only the pages each level wrote gives the same snapshots, adjacency and
analytics as one fed whole vertex tables, and ships fewer bytes when levels
touch few pages."""
import ctypes as C

import numpy as np
import pytest

import oracle
from conftest import gpu_case
from llama_amd.graph import DIR_IN, DIR_OUT, DeviceGraph, GpuError
from pages_util import local_multilevel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("deletions", [True, False])
def test_cow_mirror_matches_full_mirror(ctx, deletions):
    gpu_case(f"cow mirror deletions={deletions}")
    g = local_multilevel(13, uniform_levels=2, window_levels=4, seed=7, deletions=deletions)
    full = DeviceGraph(ctx, g)
    cow = DeviceGraph(ctx, g, cow=True)
    o = oracle.OracleGraph.from_host(g)
    for L in range(g.num_levels):
        for d in (DIR_OUT, DIR_IN):
            a_ofs, a_adj = full.adjacency(d, L)
            b_ofs, b_adj = cow.adjacency(d, L)
            r_ofs, r_adj = o.flatten(d, L)
            assert np.array_equal(b_ofs, r_ofs) and np.array_equal(b_adj, r_adj), (L, d)
            assert np.array_equal(a_ofs, b_ofs) and np.array_equal(a_adj, b_adj), (L, d)
    ra, _ = full.pagerank(10)
    rb, _ = cow.pagerank(10)
    assert np.array_equal(ra, rb)
    la, na, _ = full.bfs(3)
    lb, nb, _ = cow.bfs(3)
    assert np.array_equal(la, lb) and na == nb
    fb, _ = full.upload_stats()
    cb, _ = cow.upload_stats()
    assert cb < fb


def test_cow_rejects_bad_page_lists(ctx):
    gpu_case("cow bad page lists")
    g = local_multilevel(10, uniform_levels=0, window_levels=1, seed=8)
    dg = DeviceGraph(ctx, g, cow=True)  # valid: both levels pushed
    h = C.c_void_p()
    ctx.check(ctx.lib.ll_gpu_graph_create(ctx._h, g.num_nodes, 0x3 if g.deletions else 0x2,
                                          C.byref(h)))
    try:
        v = g.level_view(DIR_OUT, 0)
        pages, _ = g.level_pages(DIR_OUT, 0)
        pages[0] = None  # level 0 cannot share
        with pytest.raises(GpuError):
            ctx.check(ctx.lib.ll_gpu_graph_push_level_pages(h, C.byref(v), pages, len(pages)))
        pages, _ = g.level_pages(DIR_OUT, 0)
        with pytest.raises(GpuError):
            ctx.check(ctx.lib.ll_gpu_graph_push_level_pages(h, C.byref(v), pages, len(pages) - 1))
        # a valid level 0 after the refusals
        ctx.check(ctx.lib.ll_gpu_graph_push_level_pages(h, C.byref(v), pages, len(pages)))
    finally:
        ctx.lib.ll_gpu_graph_destroy(h)
    assert dg.num_levels == g.num_levels
