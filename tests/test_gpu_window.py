"""LL_MIN_LEVEL windows on the device (ll_gpu_graph_set_min_level /
ll_gpu_graph_delete_level) against the oracle's window: adjacency
bit-exact, PageRank to the FP32 bar, BFS levels exact, also after the levels
below the window are deleted and on a copy-on-write mirror whose pages are
shared with deleted levels."""
from pathlib import Path

import numpy as np
import pytest

import oracle
from conftest import gpu_case
from llama_amd.graph import DIR_IN, DIR_OUT, DeviceGraph, GpuError, build_multilevel
from pages_util import local_multilevel
from window_util import batched_graph

pytestmark = pytest.mark.gpu
PR_TOL = 1e-6


def l1rel(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def check(dg, o, levels):
    for L in levels:
        for d in (DIR_OUT, DIR_IN):
            a_ofs, a_adj = dg.adjacency(d, L)
            r_ofs, r_adj = o.flatten(d, L)
            assert np.array_equal(a_ofs, r_ofs) and np.array_equal(a_adj, r_adj), (L, d)
    r, _ = dg.pagerank(10)
    ref, _ = o.pagerank(10)
    assert l1rel(r, ref) <= PR_TOL
    lv, n, _ = dg.bfs(1)
    rl, rn = o.bfs(1)
    assert np.array_equal(lv, rl) and n == rn


@pytest.mark.parametrize("maker", ["batched", "rmat"])
def test_window_matches_oracle_then_delete(ctx, maker):
    gpu_case(f"window {maker}")
    g = batched_graph(300, 7, seed=21)[0] if maker == "batched" else \
        build_multilevel(12, levels=7, ins_frac=0.05, del_frac=0.01, seed=9)
    dg = DeviceGraph(ctx, g)
    o = oracle.OracleGraph.from_host(g)
    check(dg, o, [g.num_levels - 1])
    for m in (2, 4, 6):
        dg.set_min_level(m)
        o.set_min_level(m)
        check(dg, o, [m - 1, m, g.num_levels - 1])
    for k in range(5):  # levels below min - 1 = 5
        dg.delete_level(k)
    check(dg, o, [5, 6])
    with pytest.raises(GpuError):
        dg.adjacency(DIR_OUT, 2)  # a deleted level
    with pytest.raises(GpuError):
        dg.set_min_level(3)  # the window only moves forward
    with pytest.raises(GpuError):
        dg.delete_level(5)  # not below min - 1


def test_window_on_cow_mirror_with_shared_pages(ctx):
    gpu_case("window cow mirror")
    g = local_multilevel(12, uniform_levels=1, window_levels=5, seed=31)
    dg = DeviceGraph(ctx, g, cow=True)
    o = oracle.OracleGraph.from_host(g)
    dg.set_min_level(4)
    o.set_min_level(4)
    for k in range(3):  # pages of levels 0-2 are shared by the later tables
        dg.delete_level(k)
    check(dg, o, [3, 4, g.num_levels - 1])
    # the next push still shares pages with the newest level
    g.add_window_level(500, 20, 64, 64 + 64, 77)
    o2 = oracle.OracleGraph.from_host(g)
    o2.set_min_level(4)
    dg.sync_levels()
    check(dg, o2, [g.num_levels - 1])


def test_cli_window_option(tmp_path):
    """ll_gpu_bench -l MIN-: the window the reference benchmark applies with
    -l (set_min_level, then delete_level below it), on .xs1 levels."""
    import subprocess
    from llama_amd.graph import Edges, HostGraph
    gpu_case("ll_gpu_bench -l window")
    root = Path(__file__).resolve().parent.parent
    exe = root / "llama_amd" / "ll_gpu_bench"
    g, batches = batched_graph(200, 4, seed=8, deletions=False)
    files = []
    for k, (ins, _) in enumerate(batches):
        f = str(tmp_path / f"l{k}.xs1")
        Edges.from_arrays([e[0] for e in ins], [e[1] for e in ins]).write_xs1(f, 200)
        files.append(f)
    out = str(tmp_path / "pr.txt")
    r = subprocess.run([str(exe), "-r", "gpu_pagerank", "-n", "10", "-l", "2-", "-o", out] + files,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    h = HostGraph(200, deletions=False)
    for ins, _ in batches:
        h.add_level([e[0] for e in ins], [e[1] for e in ins])
    o = oracle.OracleGraph.from_host(h)
    o.set_min_level(2)
    ref, _ = o.pagerank(10)
    vals = []
    for line in open(out):
        if ":" in line:
            vals += [float(x) for x in line.split(":")[1].split()]
    assert np.allclose(vals, ref[:50], atol=1e-8)
    assert "Levels     : 2--3" in r.stdout
