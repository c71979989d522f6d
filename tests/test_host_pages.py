"""Copy-on-write vertex-table pages of the host layout (the pages the
reference's ll_mem_array_swcow shares between levels, ll_mem_array.h:918-1140):
a page is listed exactly when it differs from the previous level's."""
import ctypes as C

import numpy as np
import pytest

from llama_amd._native import NativeError
from llama_amd.graph import DIR_IN, DIR_OUT
from pages_util import local_multilevel


def page_bytes(g, d, k):
    vt = g.level_arrays(d, k).vt.view(np.uint8)
    pad = (-vt.size) % (512 * 16)
    return np.concatenate([vt, np.zeros(pad, np.uint8)]).reshape(-1, 512 * 16)


@pytest.mark.parametrize("d", [DIR_OUT, DIR_IN])
def test_level_pages_are_the_changed_pages(d):
    g = local_multilevel(12, uniform_levels=2, window_levels=3, seed=4)
    np_ = (g.num_nodes + 1 + 511) // 512
    for k in range(g.num_levels):
        pages, n = g.level_pages(d, k)
        assert len(pages) == np_
        addr = np.array([p or 0 for p in pages], np.uint64)
        cur = page_bytes(g, d, k)
        base = g.level_arrays(d, k).vt.ctypes.data
        if k == 0:
            assert n == np_ and np.all(addr == base + np.arange(np_, dtype=np.uint64) * 8192)
            continue
        changed = np.any(cur != page_bytes(g, d, k - 1), axis=1)
        assert np.array_equal(addr != 0, changed) and n == int(changed.sum())
        assert np.all(addr[changed] == base + np.nonzero(changed)[0].astype(np.uint64) * 8192)
        if k >= 3:  # window levels: the window's (at most 2) pages and the sentinel's
            assert n <= 3
        assert addr[-1] != 0  # the sentinel entry (VT[V]) moves every level


def test_level_pages_rejects_wrong_count():
    g = local_multilevel(10, uniform_levels=0, window_levels=1, seed=5)
    np_ = (g.num_nodes + 1 + 511) // 512
    arr = (C.c_void_p * (np_ + 1))()
    from llama_amd import _native as N
    with pytest.raises(NativeError):
        N.check_host(N.host().ll_host_graph_level_pages(g._h, 0, 1, arr, np_ + 1, None))
