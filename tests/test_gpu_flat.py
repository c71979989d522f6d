"""The device mirror fed from the reference's flat vertex-table containers
(LL_FLAT_VT, LL_ONE_VT): at every freeze the level's tables, as the flat
container holds them then, go through ll_gpu_graph_push_level, and the older
words that freeze rewrote through ll_gpu_graph_patch_deletions
(tests/golden/{flat,onevt}_*.json).  PageRank is within 1e-6 of the
reference's pagerank_double on that container, BFS levels and reached counts
equal its sssp_unweighted and bfs results, and the snapshot at every level
equals the one of the copy-on-write fixture's layout."""
import ctypes as C

import numpy as np
import pytest

from conftest import gpu_case
from flat_util import FLAT, freezes, load
from llama_amd import _native as N
from llama_amd.graph import DIR_OUT, DeviceGraph
from test_golden import build, load as load_cow

pytestmark = pytest.mark.gpu


class PushedGraph(DeviceGraph):
    """A device graph fed level by level from raw tables (no host graph)."""

    def __init__(self, ctx, V: int, levels):
        self.cow = False
        self.ctx = ctx
        self.lib = ctx.lib
        self.host = None
        h = C.c_void_p()
        ctx.check(self.lib.ll_gpu_graph_create(ctx._h, V, 0x3, C.byref(h)))  # cont + deletions
        self._h = h
        self.V = V
        ctx._graphs.add(self)
        self._keep = []
        for k, per_dir in enumerate(levels):
            for d, (vt, et, ids, ml) in enumerate(per_dir):
                if ids.size:
                    ctx.check(self.lib.ll_gpu_graph_patch_deletions(
                        h, d, ids.ctypes.data_as(C.POINTER(C.c_uint64)),
                        ml.ctypes.data_as(C.POINTER(C.c_uint32)), ids.size))
                v = N.LevelView()
                v.dir, v.level, v.num_nodes, v.num_vt_entries = d, k, V, vt.size
                v.num_et_words = et.size
                v.vt = vt.ctypes.data
                v.et = et.ctypes.data_as(C.POINTER(C.c_uint64))
                v.flags = 0x3
                v.page_table = None
                self._keep += [vt, et]
                ctx.check(self.lib.ll_gpu_graph_push_level(h, C.byref(v)))

    def sync_levels(self) -> None:
        pass


@pytest.mark.parametrize("name", FLAT)
def test_flat_container_push(ctx, name):
    gpu_case(f"flat vertex tables {name}")
    doc = load(name)
    V = doc["num_nodes"]
    dg = PushedGraph(ctx, V, freezes(doc))
    tasks = doc["reference"]["tasks"]
    r, _ = dg.pagerank(10)
    ref = np.array(tasks["pagerank_double:10"]["values"])
    assert np.abs(r - ref).sum() / ref.sum() <= 1e-6
    for key, t in tasks.items():
        task, arg = key.split(":")
        if task == "sssp_unweighted":
            lv, _, _ = dg.bfs(int(arg))
            assert lv.tolist() == [int(x) for x in t["values"]], key
        elif task == "bfs":
            _, n, _ = dg.bfs(int(arg))
            assert n == int(t["return"][0]), key
    cow = DeviceGraph(ctx, build(load_cow(doc["source_fixture"][:-5])))
    for L in range(cow.num_levels):
        for d in (0, 1):
            a = dg.adjacency(d, level=L)
            b = cow.adjacency(d, level=L)
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (L, d)
    dg.close()
    cow.close()
