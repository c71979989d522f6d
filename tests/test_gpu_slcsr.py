"""The single-level CSR configuration on the device (ll_gpu_graph_push_slcsr,
the reference's LL_SLCSR): against the reference's own outputs, and against a
multi-level mirror holding the same edges as one level (same CSR, so the
same results bit for bit)."""
import json
from pathlib import Path

import numpy as np
import pytest

from conftest import gpu_case
from llama_amd.graph import (DIR_IN, DIR_OUT, DeviceGraph, GpuError, HostGraph, SlcsrGraph,
                             build_multilevel, csr_of)

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
PR_TOL = 1e-6


def l1rel(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def slcsr_of(ctx, n, src, dst):
    return SlcsrGraph(ctx, n, csr_of(n, src, dst), csr_of(n, dst, src))


@pytest.mark.parametrize("name", sorted(p.stem for p in GOLDEN.glob("slcsr_*.json")))
def test_slcsr_matches_reference(ctx, name):
    gpu_case(f"slcsr golden {name}")
    doc = json.loads((GOLDEN / f"{name}.json").read_text())
    n, e, ref = doc["num_nodes"], doc["edges"], doc["reference"]
    dg = slcsr_of(ctx, n, e["src"], e["dst"])
    r, _ = dg.pagerank(10)
    assert l1rel(r, np.array(ref["pagerank_double"]["values"])) <= PR_TOL
    for key, val in ref.items():
        root = key.rsplit("_", 1)[1]
        if key.startswith("sssp_unweighted_"):
            lv, _, _ = dg.bfs(int(root))
            assert np.array_equal(lv, np.array(val["values"], np.int64)), key
        elif key.startswith("bfs_"):
            _, cnt, _ = dg.bfs(int(root))
            assert cnt == int(val["return"]), key


def test_slcsr_matches_single_level_mirror(ctx):
    gpu_case("slcsr vs single-level mirror, rmat s12")
    base = build_multilevel(12, levels=1, seed=17, deletions=False)
    ofs, adj = base.flatten(DIR_OUT)
    V = base.num_nodes
    src = np.repeat(np.arange(V, dtype=np.int64), np.diff(ofs.astype(np.int64)))
    g = HostGraph(V, deletions=False)
    g.add_level(src, adj.astype(np.int64))
    ml = DeviceGraph(ctx, g)
    iofs, iadj = ml.adjacency(DIR_IN)
    isrc = np.repeat(np.arange(V, dtype=np.int64), np.diff(iofs.astype(np.int64)))
    sl = SlcsrGraph(ctx, V, (ofs.astype(np.int64), adj.astype(np.int64)),
                    (iofs.astype(np.int64), iadj.astype(np.int64)))
    for d in (DIR_OUT, DIR_IN):
        a = ml.adjacency(d)
        b = sl.adjacency(d)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(ml.pagerank(10)[0], sl.pagerank(10)[0])
    assert np.array_equal(ml.bfs(5)[0], sl.bfs(5)[0])
    assert ml.triangles() == sl.triangles()
    assert np.array_equal(ml.scc()[0], sl.scc()[0])
    assert np.array_equal(ml.bc([0, 3]), sl.bc([0, 3]))
    assert isrc.size == adj.size


def test_slcsr_rejects_bad_input(ctx):
    gpu_case("slcsr bad input")
    n = 8
    good = csr_of(n, [0, 1, 2], [1, 2, 3])
    with pytest.raises(GpuError):  # neighbour out of range
        SlcsrGraph(ctx, n, csr_of(n, [0, 1], [1, 8]))
    with pytest.raises(GpuError):  # decreasing begin
        SlcsrGraph(ctx, n, (np.array([0, 2, 1, 3, 3, 3, 3, 3, 3]), np.array([1, 2, 3])))
    with pytest.raises(GpuError):  # begin[V] != number of edges
        SlcsrGraph(ctx, n, (good[0] + 0, good[1][:2]))
    sl = SlcsrGraph(ctx, n, good, csr_of(n, [1, 2, 3], [0, 1, 2]))
    with pytest.raises(GpuError):
        sl.adjacency(DIR_OUT, 1)  # only level 0
    g = HostGraph(n, deletions=False)
    g.add_level([0], [1])
    with pytest.raises(GpuError):  # no multi-level pushes onto a single-level CSR
        v = g.level_view(DIR_OUT, 0)
        import ctypes as C
        ctx.check(ctx.lib.ll_gpu_graph_push_level(sl._h, C.byref(v)))
