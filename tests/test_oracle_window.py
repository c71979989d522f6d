"""LL_MIN_LEVEL windows in the oracle (iter_begin / descend stop at segments
of levels below the minimum, ll_mlcsr_sp.h:2021-2027, 2107-2117).  The
reference's LL_MIN_LEVEL build does not compile, so the restatement is pinned
by a brute force over the insert/delete batches and by the iterator's prefix
property."""
import numpy as np
import pytest

import oracle
from llama_amd.graph import DIR_IN, DIR_OUT
from window_util import batched_graph, windowed_edges


@pytest.mark.parametrize("deletions", [True, False])
def test_window_matches_batches_and_is_a_prefix(deletions):
    g, batches = batched_graph(60, 6, seed=13, deletions=deletions)
    full = oracle.OracleGraph.from_host(g)
    for m in (1, 2, 4, 5, 6):
        o = oracle.OracleGraph.from_host(g)
        o.set_min_level(m)
        for L in range(g.num_levels):
            want = windowed_edges(batches, m, L)
            for d in (DIR_OUT, DIR_IN):
                ofs, adj = o.flatten(d, L)
                fofs, fadj = full.flatten(d, L)
                got = set()
                for v in range(g.num_nodes):
                    row = adj[ofs[v]:ofs[v + 1]].tolist()
                    frow = fadj[fofs[v]:fofs[v + 1]].tolist()
                    assert row == frow[:len(row)], (m, L, d, v)  # newest segments first
                    got |= {(v, w) if d == DIR_OUT else (w, v) for w in row}
                assert got == want, (m, L, d)


@pytest.mark.parametrize("deletions", [False, True])
def test_window_degrees_and_pagerank(deletions):
    """Inside the window a vertex's degree is its windowed row length -- without
    deletions the reference's VT_L.degree - VT_(m-1).degree -- and PageRank
    divides by it (checked against a dense power iteration)."""
    g, _ = batched_graph(50, 5, seed=3, deletions=deletions)
    o = oracle.OracleGraph.from_host(g)
    o.set_min_level(2)
    with pytest.raises(ValueError):
        o.set_min_level(1)
    ofs, adj = o.flatten(DIR_OUT)
    rows = np.diff(ofs.astype(np.int64))
    if not deletions:
        top = g.level_arrays(DIR_OUT, 4).vt["degree"][:50].astype(np.int64)
        below = g.level_arrays(DIR_OUT, 1).vt["degree"][:50].astype(np.int64)
        assert np.array_equal(rows, top - below)
    n, d, it = 50, 0.85, 8
    r = np.full(n, 1.0 / n)
    for _ in range(it):
        nxt = np.full(n, (1 - d) / n)
        for u in range(n):
            for v in adj[ofs[u]:ofs[u + 1]]:
                nxt[v] += d * r[u] / rows[u]
        r = nxt
    ref, _ = o.pagerank(it)
    assert np.allclose(ref, r, rtol=0, atol=1e-12)
