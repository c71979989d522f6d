"""Parity at the full benchmark size: R-MAT scale 24, edge factor 16, plus 10
levels of +1% / -0.1% (deletion vector on, U[0, 1) weights) -- the graph
bench.py measures.  The oracle's per-edge restatement is too slow for the
algorithms at this size, so each result is pinned by a property that fixes it
exactly: the snapshot equals the oracle's iterator bit for bit; BFS levels
satisfy the two edge conditions that characterise BFS distances; SSSP
distances are the FP32 fixed point on every edge; PageRank agrees between the
pull (FP32 compensated) and push (FP64 atomics) paths; Brandes dependencies
sum to the number of interior vertices of the shortest paths; every lane of
the batched BFS equals the single-source run; SCC labels equal scipy's strong
components (canonical smallest-ID labels)."""
import numpy as np
import pytest

import oracle
from conftest import gpu_case
from llama_amd.graph import DIR_IN, DIR_OUT, UNREACHED, DeviceGraph, build_multilevel

pytestmark = pytest.mark.gpu
SCALE, LEVELS = 24, 11


@pytest.fixture(scope="module")
def big(ctx):
    gpu_case(f"full size: building R-MAT s{SCALE} x {LEVELS} levels")
    g = build_multilevel(SCALE, levels=LEVELS, seed=1, weighted=True)
    dg = DeviceGraph(ctx, g)
    ofs, adj = dg.adjacency(DIR_OUT)
    src = np.repeat(np.arange(g.num_nodes, dtype=np.uint32), np.diff(ofs.astype(np.int64)))
    rng = np.random.default_rng(5)
    deg = np.diff(ofs.astype(np.int64))
    roots = [int(x) for x in rng.choice(np.nonzero(deg)[0], 4, replace=False)]
    yield g, dg, ofs, adj, src, roots
    dg.close()


def test_full_snapshot_equals_oracle(big):
    g, dg, ofs, adj, _, _ = big
    gpu_case("full size: snapshot vs oracle iterator")
    o = oracle.OracleGraph.from_host(g)
    r_ofs, r_adj = o.flatten(DIR_OUT)
    assert np.array_equal(ofs, r_ofs) and np.array_equal(adj, r_adj)
    i_ofs, i_adj = dg.adjacency(DIR_IN)
    r_ofs, r_adj = o.flatten(DIR_IN)
    assert np.array_equal(i_ofs, r_ofs) and np.array_equal(i_adj, r_adj)


def test_full_bfs_levels_are_distances(big):
    g, dg, ofs, adj, src, roots = big
    for root in roots:
        gpu_case(f"full size: BFS root {root}")
        lv, reached, _ = dg.bfs(root)
        lv = lv.astype(np.int64)
        assert lv[root] == 0
        live = lv[src] != UNREACHED
        # no edge skips a level, and every reached vertex has a parent one level up
        assert (lv[adj[live]] <= lv[src[live]] + 1).all()
        parent = np.zeros(g.num_nodes, bool)
        tight = live & (lv[src] + 1 == lv[adj])
        parent[adj[tight]] = True
        hit = lv != UNREACHED
        hit[root] = False
        assert parent[hit].all()
        assert reached == int((lv != UNREACHED).sum())


def test_full_sssp_is_the_fixed_point(big):
    g, dg, ofs, adj, src, roots = big
    gpu_case("full size: weighted SSSP")
    w = g.flatten_weights()
    assert w.size == adj.size
    root = roots[0]
    dist, reached, mx, _ = dg.sssp(root)
    unreached = np.float32(2**31 - 2)
    live = dist[src] != unreached
    cand = dist[src[live]] + w[live]  # float32 sums, rounded as the kernel rounds
    assert (dist[adj[live]] <= cand).all()
    best = np.zeros(g.num_nodes, bool)
    best[adj[live][dist[adj[live]] == cand]] = True
    hit = dist != unreached
    hit[root] = False
    assert best[hit].all() and dist[root] == 0
    assert reached == int((dist != unreached).sum())
    assert mx == float(dist[dist != unreached].max())


def test_full_pagerank_pull_equals_push(big):
    _, dg, _, _, _, _ = big
    gpu_case("full size: PageRank pull vs push")
    r, _ = dg.pagerank(20)
    p, _ = dg.pagerank_push(20)
    assert float(np.abs(r - p).sum() / np.abs(p).sum()) <= 1e-6


def test_full_bc_dependency_identity(big):
    _, dg, _, _, _, roots = big
    for s in roots[:2]:
        gpu_case(f"full size: BC source {s}")
        b = dg.bc([s]).astype(np.float64)
        lv, _, _ = dg.bfs(s)
        lv = lv.astype(np.int64)
        t = (lv != UNREACHED) & (lv > 0)
        # each shortest s-t path has dist(s, t) - 1 interior vertices
        assert np.isclose(b.sum(), float((lv[t] - 1).sum()), rtol=1e-4)


def test_full_batched_bfs_lanes(big):
    _, dg, _, _, _, roots = big
    gpu_case("full size: batched BFS lanes")
    lvs, reached, _ = dg.bfs_batch(roots, lane_width=64)
    for i, r in enumerate(roots):
        lv, n, _ = dg.bfs(r)
        assert np.array_equal(lvs[i], lv) and int(reached[i]) == n


def test_full_scc_matches_scipy(big):
    g, dg, ofs, adj, _, _ = big
    gpu_case("full size: SCC vs scipy strong components")
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import connected_components
    V = g.num_nodes
    A = csr_matrix((np.ones(adj.size, np.int8), adj.astype(np.int32), ofs.astype(np.int64)),
                   shape=(V, V))
    n, lab = connected_components(A, directed=True, connection="strong")
    # canonical labels: the smallest vertex ID of each component
    low = np.full(n, V, np.int64)
    np.minimum.at(low, lab, np.arange(V, dtype=np.int64))
    mine, count = dg.scc()
    assert count == n
    assert np.array_equal(mine, low[lab])
