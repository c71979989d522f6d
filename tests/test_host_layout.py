"""Host level builder: LLAMA layout and deletion semantics (CPU only).

The hand-derived expectations below follow ll_mlcsr_core::init_node
(llama/include/llama/ll_mlcsr_sp.h:1375-1550): new segment / copied entry,
degree = prev + new - deleted, continuation record = copy of VT_{k-1}[v]
(also for deletion-only vertices), sentinel VT[V] = {edge(k, words), 0, 0},
and update_max_visible_level_lower_only (:849-863) on both twins.
"""
import numpy as np
import pytest

from llama_amd.graph import DIR_IN, DIR_OUT, Edges, HostGraph, build_multilevel

NEVER = (1 << 26) - 1  # LL_MAX_LEVEL + 2


def E(level, index):
    return (level << 38) | index


def W(payload, maxlev=NEVER):
    return payload | (maxlev << 38)


def vt_rows(h, d, k):
    a = h.level_arrays(d, k).vt
    return [(int(r["start"]), int(r["len"]), int(r["degree"])) for r in a]


def test_two_level_layout_by_hand():
    g = HostGraph(4, deletions=True)
    g.add_level([0, 0, 1, 2], [1, 2, 2, 0])
    assert vt_rows(g, DIR_OUT, 0) == [(E(0, 0), 2, 2), (E(0, 2), 1, 1), (E(0, 3), 1, 1),
                                      (-1, 0, 0), (E(0, 4), 0, 0)]
    assert list(g.level_arrays(DIR_OUT, 0).et) == [W(1), W(2), W(2), W(0)]
    assert vt_rows(g, DIR_IN, 0) == [(E(0, 0), 1, 1), (E(0, 1), 1, 1), (E(0, 2), 2, 2),
                                     (-1, 0, 0), (E(0, 4), 0, 0)]
    assert list(g.level_arrays(DIR_IN, 0).et) == [W(2), W(0), W(0), W(1)]

    g.add_level([0, 3], [3, 0], [0], [2])
    # deletion rewrote the old words in place (both twins)
    assert list(g.level_arrays(DIR_OUT, 0).et) == [W(1), W(2, 1), W(2), W(0)]
    assert list(g.level_arrays(DIR_IN, 0).et) == [W(2), W(0), W(0, 1), W(1)]
    u64 = lambda x: x & 0xFFFFFFFFFFFFFFFF
    assert vt_rows(g, DIR_OUT, 1) == [(E(1, 0), 1, 2), (E(0, 2), 1, 1), (E(0, 3), 1, 1),
                                      (E(1, 3), 1, 1), (E(1, 6), 0, 0)]
    assert [int(x) for x in g.level_arrays(DIR_OUT, 1).et] == [
        W(3), E(0, 0), 2 | (2 << 32), W(0), u64(-1), 0]
    # vertex 2 only lost an in-edge: entry copied, degree lowered, orphan continuation
    assert vt_rows(g, DIR_IN, 1) == [(E(1, 0), 1, 2), (E(0, 1), 1, 1), (E(0, 2), 2, 1),
                                     (E(1, 5), 1, 1), (E(1, 8), 0, 0)]
    assert [int(x) for x in g.level_arrays(DIR_IN, 1).et] == [
        W(3), E(0, 0), 1 | (1 << 32), E(0, 2), 2 | (2 << 32), W(0), u64(-1), 0]
    ids, ml = g.patches(DIR_OUT, 1)
    assert list(ids) == [E(0, 1)] and list(ml) == [1]
    ids, ml = g.patches(DIR_IN, 1)
    assert list(ids) == [E(0, 2)] and list(ml) == [1]
    # snapshots: level 0 still sees (0,2); level 1 does not
    o0, a0 = g.flatten(DIR_OUT, 0)
    o1, a1 = g.flatten(DIR_OUT, 1)
    assert list(a0[o0[0]:o0[1]]) == [1, 2]
    assert list(a1[o1[0]:o1[1]]) == [3, 1]  # newest segment first


def test_degree_drops_to_zero_gives_nil():
    g = HostGraph(3, deletions=True)
    g.add_level([0], [1])
    g.add_level([], [], [0], [1])
    assert vt_rows(g, DIR_OUT, 1)[0] == (-1, 0, 0)
    assert vt_rows(g, DIR_IN, 1)[1] == (-1, 0, 0)
    assert g.level_arrays(DIR_OUT, 1).et.size == 0  # NIL entries write no continuation
    g.add_level([0], [2])
    assert [int(x) for x in g.level_arrays(DIR_OUT, 2).et] == [W(2), 0xFFFFFFFFFFFFFFFF, 0]


def test_errors_are_reported_not_aborted():
    g = HostGraph(3, deletions=True)
    with pytest.raises(RuntimeError):
        g.add_level([0], [1], [0], [1])  # level 0 cannot delete
    g.add_level([0], [1])
    with pytest.raises(RuntimeError, match="not live"):
        g.add_level([], [], [1], [0])
    with pytest.raises(RuntimeError, match="out of range"):
        g.add_level([5], [0])
    nd = HostGraph(3, deletions=False)
    nd.add_level([0], [1])
    with pytest.raises(RuntimeError, match="deletions"):
        nd.add_level([], [], [0], [1])


def _snapshot_sets(g, level):
    out = []
    for d in (DIR_OUT, DIR_IN):
        o, a = g.flatten(d, level)
        s = set()
        for v in range(g.num_nodes):
            for w in a[o[v]:o[v + 1]]:
                s.add((v, int(w)) if d == DIR_OUT else (int(w), v))
        out.append(s)
    return out


def test_random_levels_match_set_model_at_every_snapshot():
    base = Edges.rmat(9, 8, seed=5, perm_seed=6)
    bs, bd = base.arrays()
    g = HostGraph(1 << 9, deletions=True)
    g.add_level_edges(base)
    model = [set(zip(bs.tolist(), bd.tolist()))]
    for k in range(1, 6):
        g.add_random_level(0.05, 0.03, 100 + k)
        (is_, id_), (ds, dd) = g.last_batch()
        cur = set(model[-1])
        dels = set(zip(ds.tolist(), dd.tolist()))
        ins = set(zip(is_.tolist(), id_.tolist()))
        assert dels <= cur and not (ins & cur) and all(u != v for u, v in ins)
        assert len(ins) == int(0.05 * len(cur)) and len(dels) == int(0.03 * len(cur))
        model.append((cur - dels) | ins)
    for L in range(g.num_levels):
        out_s, in_s = _snapshot_sets(g, L)
        assert out_s == model[L] and in_s == model[L], L
        # stored degrees equal iterated counts
        for d in (DIR_OUT, DIR_IN):
            o, _ = g.flatten(d, L)
            deg = g.level_arrays(d, L).vt["degree"][:-1]
            assert np.array_equal(np.diff(o), deg.astype(np.uint64))


def test_delete_every_tenth_edge_counts():
    """benchmark/tests/delete_edges.h:81-233: delete edges with ID = 0 mod 10,
    then iterated count == sum of degrees == original - deleted."""
    g = build_multilevel(10, levels=2, ins_frac=0.02, del_frac=0.0, seed=9)
    o, a = g.flatten(DIR_OUT)
    src = np.repeat(np.arange(g.num_nodes, dtype=np.uint32), np.diff(o).astype(np.int64))
    sel = np.arange(a.size) % 10 == 0
    g.add_level([], [], src[sel], a[sel])
    for d in (DIR_OUT, DIR_IN):
        o2, a2 = g.flatten(d)
        deg = g.level_arrays(d, g.num_levels - 1).vt["degree"][:-1].astype(np.int64)
        assert a2.size == deg.sum() == a.size - sel.sum()


def test_rmat_generator_properties():
    e = Edges.rmat(10, 16, seed=1, perm_seed=2)
    s, d = e.arrays()
    assert np.all(s != d)
    key = s.astype(np.uint64) << np.uint64(32) | d.astype(np.uint64)
    assert np.all(np.diff(key) > 0)  # sorted, unique
    assert 0.5 * 16 * 1024 < s.size <= 16 * 1024
    # deterministic
    s2, d2 = Edges.rmat(10, 16, seed=1, perm_seed=2).arrays()
    assert np.array_equal(s, s2) and np.array_equal(d, d2)
    # skew: the top 1% of vertices by out-degree hold a large share of edges
    deg = np.bincount(s, minlength=1024)
    assert np.sort(deg)[-10:].sum() > 0.05 * s.size


def test_xs1_roundtrip(tmp_path):
    e = Edges.from_arrays([0, 3, 2], [1, 0, 3], [0.5, 1.0, 2.0])
    p = str(tmp_path / "g.xs1")
    e.write_xs1(p, 5)
    raw = np.fromfile(p, dtype=np.dtype([("t", "<u4"), ("h", "<u4"), ("w", "<f4")]))
    assert raw["t"].tolist() == [0, 3, 2] and raw["h"].tolist() == [1, 0, 3]
    assert raw["w"].tolist() == [0.5, 1.0, 2.0]
    e2, n = Edges.read_xs1(p)
    assert n == 5
    assert [a.tolist() for a in e2.arrays()] == [[0, 3, 2], [1, 0, 3]]
    assert "vertices=5" in open(p + ".ini").read()
