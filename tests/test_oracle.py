"""Pins of the CPU oracle (oracle/oracle.cc) against closed forms and
independent formulations (dense numpy algebra, scipy.sparse.csgraph, brute
force), chosen so that a dropped term, a wrong index or a transposed operand
fails at least one of them."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse import csgraph

import oracle
from llama_amd.graph import DIR_IN, DIR_OUT, HostGraph, build_multilevel

D = 0.85


def og(n, src, dst, deletions=True):
    g = HostGraph(n, deletions=deletions)
    g.add_level(src, dst)
    return g, oracle.OracleGraph.from_host(g)


def edge_set(g, level=-1):
    o, a = g.flatten(DIR_OUT, level)
    src = np.repeat(np.arange(g.num_nodes), np.diff(o).astype(np.int64))
    return src, a.astype(np.int64)


def dense_pagerank(n, src, dst, iters, d=D):
    A = np.zeros((n, n))
    A[src, dst] = 1.0
    outdeg = A.sum(1)
    r = np.full(n, 1.0 / n)
    for _ in range(iters):
        contrib = np.divide(r, outdeg, out=np.zeros(n), where=outdeg > 0)
        r = (1 - d) / n + d * (A.T @ contrib)
    return r


# ---- PageRank (pagerank.h:98-182) ---------------------------------------------
def test_pagerank_cycle_is_uniform():
    n = 9
    _, o = og(n, np.arange(n), (np.arange(n) + 1) % n)
    r, diff = o.pagerank(7)
    assert np.allclose(r, 1 / n, rtol=0, atol=1e-15) and diff < 1e-15


def test_pagerank_empty_graph_one_iteration():
    g = HostGraph(6)
    g.add_level([], [])
    o = oracle.OracleGraph.from_host(g)
    r, _ = o.pagerank(1)
    assert np.allclose(r, (1 - D) / 6, rtol=1e-15)


def test_pagerank_star_closed_form():
    k = 5
    n = k + 1
    _, o = og(n, np.arange(1, n), np.zeros(k, int))
    r, _ = o.pagerank(3)
    leaf = (1 - D) / n
    assert np.allclose(r[1:], leaf, rtol=1e-15)
    assert np.isclose(r[0], (1 - D) / n + D * k * leaf, rtol=1e-14)
    assert r.sum() < 1  # dangling mass is not redistributed


def test_pagerank_matches_dense_power_iteration_multilevel():
    g = build_multilevel(8, levels=4, ins_frac=0.05, del_frac=0.02, seed=11)
    o = oracle.OracleGraph.from_host(g)
    for L in (0, 2, 3):
        s, d = edge_set(g, L)
        ref = dense_pagerank(g.num_nodes, s, d, 10)
        r, _ = o.pagerank(10, level=L)
        assert np.allclose(r, ref, rtol=1e-12, atol=1e-16), L


# ---- BFS levels (sssp.h:220-333) ----------------------------------------------
def test_bfs_matches_scipy_shortest_paths():
    g = build_multilevel(9, levels=3, ins_frac=0.02, del_frac=0.01, seed=4, edge_factor=4)
    o = oracle.OracleGraph.from_host(g)
    s, d = edge_set(g)
    A = sp.csr_matrix((np.ones(s.size), (s, d)), shape=(g.num_nodes,) * 2)
    for root in (0, 7, 100):
        dist, reached = o.bfs(root)
        ref = csgraph.shortest_path(A, unweighted=True, indices=root)
        exp = np.where(np.isinf(ref), 2**31 - 2, ref).astype(np.int32)
        assert np.array_equal(dist, exp)
        assert reached == int(np.isfinite(ref).sum())


# ---- triangles (triangle_counting.h:343-470) -------------------------------------
def test_triangles_brute_force_and_closed_forms():
    g = build_multilevel(7, levels=3, ins_frac=0.1, del_frac=0.05, seed=8, edge_factor=6)
    o = oracle.OracleGraph.from_host(g)
    s, d = edge_set(g)
    n = g.num_nodes
    A = np.zeros((n, n), dtype=np.int64)
    A[s, d] = 1
    A = ((A + A.T) > 0).astype(np.int64)
    np.fill_diagonal(A, 0)
    assert o.triangles() == int(np.trace(A @ A @ A) // 6)
    # K_n and a bipartite graph
    m = 7
    pairs = [(i, j) for i in range(m) for j in range(m) if i != j]
    _, ok = og(m, [p[0] for p in pairs], [p[1] for p in pairs])
    assert ok.triangles() == m * (m - 1) * (m - 2) // 6
    bip = [(i, j) for i in range(4) for j in range(4, 9)]
    _, ob = og(9, [p[0] for p in bip], [p[1] for p in bip])
    assert ob.triangles() == 0


# ---- SCC ------------------------------------------------------------------------
def canon(labels):
    out = np.empty_like(labels)
    for c in np.unique(labels):
        m = labels == c
        out[m] = np.flatnonzero(m).min()
    return out


def test_scc_matches_scipy_strong_components():
    g = build_multilevel(9, levels=3, ins_frac=0.05, del_frac=0.02, seed=21, edge_factor=3)
    o = oracle.OracleGraph.from_host(g)
    s, d = edge_set(g)
    A = sp.csr_matrix((np.ones(s.size), (s, d)), shape=(g.num_nodes,) * 2)
    n, lab = csgraph.connected_components(A, directed=True, connection="strong")
    mine, nscc = o.scc()
    assert nscc == n
    assert np.array_equal(mine, canon(lab))


def test_scc_reference_counterexamples():
    # graphs on which the reference tarjan_scc mislabels (it is not an SCC algorithm)
    _, o = og(3, [0, 2, 1], [2, 1, 2])
    assert o.scc()[0].tolist() == [0, 1, 1]
    _, o = og(2, [1], [0])
    assert o.scc()[0].tolist() == [0, 1]
    _, o = og(5, [0, 1, 2, 3], [1, 2, 3, 4])  # DAG: singletons
    assert o.scc()[0].tolist() == [0, 1, 2, 3, 4]


# ---- personalized PageRank --------------------------------------------------------
def test_ppr_dense_and_uniform_reduction():
    g = build_multilevel(7, levels=2, ins_frac=0.05, del_frac=0.02, seed=3, edge_factor=6)
    o = oracle.OracleGraph.from_host(g)
    s, d = edge_set(g)
    n = g.num_nodes
    A = np.zeros((n, n))
    A[s, d] = 1
    od = A.sum(1)
    seeds = [0, 5, 17]
    R, it = o.ppr(seeds, 12)
    assert it == 12
    for j, sd in enumerate(seeds):
        e = np.zeros(n)
        e[sd] = 1
        r = e.copy()
        for _ in range(12):
            r = (1 - D) * e + D * (A.T @ np.divide(r, od, out=np.zeros(n), where=od > 0))
        assert np.allclose(R[:, j], r, rtol=1e-12, atol=1e-16)
    # linearity: the average over all single-vertex seeds is the uniform-teleport PR
    Rall, _ = o.ppr(np.arange(n), 6)
    pr, _ = o.pagerank(6)
    assert np.allclose(Rall.mean(axis=1), pr, rtol=1e-11)


def test_ppr_tolerance_stops_early():
    _, o = og(4, [0, 1, 2, 3], [1, 2, 3, 0])
    _, it = o.ppr([0], 1000, tol=1e-3)
    assert 1 < it < 1000


# ---- weighted PageRank and its adjoint -------------------------------------------
def _wgraph(seed=0, n=60, m=300):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m)
    dst = rng.integers(0, n, m)
    keep = src != dst
    key = np.unique(src[keep] * n + dst[keep])
    src, dst = key // n, key % n
    ofs = np.zeros(n + 1, np.uint64)
    np.add.at(ofs, src + 1, 1)
    ofs = np.cumsum(ofs).astype(np.uint64)
    w = rng.uniform(0.1, 1.0, src.size)
    rs = np.bincount(src, weights=w, minlength=n)
    w = w / rs[src]
    p = np.full(n, 1.0 / n)
    return n, ofs, dst.astype(np.uint32), src, w, p


def test_wpr_forward_matches_dense_and_adjoint_matches_fd():
    n, ofs, adj, src, w, p = _wgraph()
    K = 200
    r = oracle.wpr_forward(ofs, adj, w, p, K, D)
    Wm = np.zeros((n, n))
    Wm[src, adj] = w
    exact = np.linalg.solve(np.eye(n) - D * Wm.T, (1 - D) * p)
    assert np.allclose(r, exact, rtol=1e-10)
    tgt = np.arange(0, n, 3)
    g = np.zeros(n)
    g[tgt] = 2 * r[tgt]           # L = sum_{t in T} r_t^2
    dw, dp = oracle.wpr_adjoint(ofs, adj, w, r, g, K, D)
    lam = np.linalg.solve(np.eye(n) - D * Wm, g)
    assert np.allclose(dp, (1 - D) * lam, rtol=1e-9)
    h = 1e-6
    rng = np.random.default_rng(1)
    for e in rng.choice(adj.size, 12, replace=False):
        wp, wm = w.copy(), w.copy()
        wp[e] += h
        wm[e] -= h
        Lp = (oracle.wpr_forward(ofs, adj, wp, p, K, D)[tgt] ** 2).sum()
        Lm = (oracle.wpr_forward(ofs, adj, wm, p, K, D)[tgt] ** 2).sum()
        assert np.isclose(dw[e], (Lp - Lm) / (2 * h), rtol=1e-5, atol=1e-12)
    # directional derivative identity along the teleport
    dvec = rng.normal(size=n)
    Lp = (oracle.wpr_forward(ofs, adj, w, p + h * dvec, K, D)[tgt] ** 2).sum()
    Lm = (oracle.wpr_forward(ofs, adj, w, p - h * dvec, K, D)[tgt] ** 2).sum()
    assert np.isclose(dp @ dvec, (Lp - Lm) / (2 * h), rtol=1e-5)


# ---- iterator / snapshot ------------------------------------------------------------
def test_oracle_iterator_matches_host_flatten_all_levels():
    g = build_multilevel(9, levels=5, ins_frac=0.05, del_frac=0.03, seed=31)
    o = oracle.OracleGraph.from_host(g)
    for L in range(g.num_levels):
        for d in (DIR_OUT, DIR_IN):
            ho, ha = g.flatten(d, L)
            oo, oa = o.flatten(d, L)
            assert np.array_equal(ho, oo) and np.array_equal(ha, oa)
            it, ds = o.count_check(d, L)
            assert it == ds == ha.size


# ---- weighted SSSP (sssp.h:59-172) ---------------------------------------------------
INF32 = np.float32(2**31 - 2)


def test_sssp_weighted_path_closed_form():
    """A chain 0->1->...->k: distances are the float32 running sums in path
    order; the vertex after a gap is never reached."""
    k = 12
    w = np.random.default_rng(3).uniform(0.1, 3.0, k).astype(np.float32)
    g = HostGraph(k + 2)
    g.add_level(np.arange(k), np.arange(1, k + 1), weights=w)
    o = oracle.OracleGraph.from_host(g)
    o.add_weights(g)
    dist, reached, mx, rounds = o.sssp_weighted(0)
    exp = np.zeros(k + 1, np.float32)
    for i in range(k):
        exp[i + 1] = np.float32(exp[i] + w[i])
    assert np.array_equal(dist[:k + 1], exp) and dist[k + 1] == INF32
    assert reached == k + 1 and mx == float(exp[-1]) and rounds == k + 1


def test_sssp_weighted_matches_scipy_dijkstra_multilevel():
    """Random weighted levels with deletions: against scipy's FP64 Dijkstra on
    the snapshot's weighted edges (FP32 sums: relative 1e-5)."""
    rng = np.random.default_rng(8)
    n = 300
    g = HostGraph(n, deletions=True)
    live = set()
    for k in range(4):
        cnt = 1500 if k == 0 else 120
        ins = []
        while len(ins) < cnt:
            e = (int(rng.integers(0, n)), int(rng.integers(0, n)))
            if e[0] != e[1] and e not in live and e not in ins:
                ins.append(e)
        dels = [] if k == 0 else [e for e in sorted(live) if rng.random() < 0.03]
        w = rng.uniform(0.0, 2.0, len(ins)).astype(np.float32)
        g.add_level([e[0] for e in ins], [e[1] for e in ins],
                    [e[0] for e in dels], [e[1] for e in dels], weights=w)
        live |= set(ins)
        live -= set(dels)
    o = oracle.OracleGraph.from_host(g)
    o.add_weights(g)
    for level in (0, g.num_levels - 1):
        ofs, adj = g.flatten(DIR_OUT, level)
        wts = g.flatten_weights(level).astype(np.float64)
        src = np.repeat(np.arange(n), np.diff(ofs).astype(np.int64))
        A = sp.csr_matrix((wts + 1e-300, (src, adj)), shape=(n, n))  # keep zero weights
        for root in (0, 17, 200):
            dist, reached, _, _ = o.sssp_weighted(root, level=level)
            ref = csgraph.dijkstra(A, indices=root)
            fin = np.isfinite(ref)
            assert np.array_equal(dist[~fin], np.full((~fin).sum(), INF32))
            assert np.allclose(dist[fin], ref[fin], rtol=1e-5, atol=1e-6)
            assert reached == int(fin.sum())
