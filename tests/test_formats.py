"""Edge-list formats: the SNAP .net reader (ll_load_net.h) against the
reference loader's own results, its partial loads, and the xs1-split-style
base + N splits (utils/xs1-split.cpp)."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from llama_amd.graph import Edges, HostGraph, GpuError  # noqa: F401

GOLDEN = Path(__file__).resolve().parent / "golden"


def net_fixture(tmp_path):
    doc = json.loads((GOLDEN / "net_tricky.json").read_text())
    f = tmp_path / "g.net"
    f.write_text(doc["text"])
    return doc, str(f)


def test_net_reader_matches_reference_loader(tmp_path):
    doc, f = net_fixture(tmp_path)
    ref = doc["reference"]
    e, n, errors = Edges.read_net(f)
    s, d = e.arrays()
    assert n == ref["nodes"] and s.size == ref["edges"] and errors > 0
    g = HostGraph(n)
    g.add_level_edges(e)
    o = oracle.OracleGraph.from_host(g)
    for root in (0, 9):
        lv, _ = o.bfs(root)
        assert lv.tolist() == [int(x) for x in ref[f"sssp_unweighted:{root}"]]
    pr, _ = o.pagerank(5)
    assert np.abs(pr - np.array(ref["pagerank_double:5"])).max() <= 6e-11


@pytest.mark.parametrize("nparts", [1, 2, 3, 5, 8])
def test_net_partial_loads_cover_the_file_once(tmp_path, nparts):
    doc, f = net_fixture(tmp_path)
    full, _, full_err = Edges.read_net(f)
    fs, fd = full.arrays()
    ss, dd, errs = [], [], 0
    for k in range(1, nparts + 1):
        p, _, pe = Edges.read_net(f, k, nparts)
        s, d = p.arrays()
        ss.append(s)
        dd.append(d)
        errs += pe
    assert np.array_equal(np.concatenate(ss), fs) and np.array_equal(np.concatenate(dd), fd)
    assert errs == full_err


def test_net_rules_and_round_trip(tmp_path):
    f = tmp_path / "t.net"
    f.write_text("# c\n0 1\n0\t2\r\n\n   2   3\n7 x\n4 5 6\n3 4\r\n\r\n")
    e, n, errors = Edges.read_net(str(f))
    s, d = e.arrays()
    assert list(zip(s.tolist(), d.tolist())) == [(0, 1), (0, 2), (2, 3), (3, 4)]
    assert n == 5 and errors == 2
    g = tmp_path / "w.net"
    e.write_net(str(g))
    e2, n2, err2 = Edges.read_net(str(g))
    assert np.array_equal(e2.arrays()[0], s) and np.array_equal(e2.arrays()[1], d) and err2 == 0
    bad = tmp_path / "bad.net"
    bad.write_text("x y\n" * 100)
    with pytest.raises(Exception):
        Edges.read_net(str(bad))
    with pytest.raises(Exception):
        Edges.read_net(str(f), 3, 2)


def test_split_quotas_and_pairs():
    rng = np.random.default_rng(3)
    n = 10007
    e = Edges.from_arrays(rng.integers(0, 999, n), rng.integers(0, 999, n),
                          rng.random(n).astype(np.float32))
    parts = e.split(n_small=10, p_large=0.8, seed=4)
    sizes = [p.arrays()[0].size for p in parts]
    small = int((1 - 0.8) / 10 * n)
    assert sizes[1:] == [small] * 10 and sizes[0] == n - 10 * small
    s, d = e.arrays()
    got = sorted(zip(np.concatenate([p.arrays()[0] for p in parts]).tolist(),
                     np.concatenate([p.arrays()[1] for p in parts]).tolist()))
    assert got == sorted(zip(s.tolist(), d.tolist()))
    again = e.split(n_small=10, p_large=0.8, seed=4)
    assert all(np.array_equal(a.arrays()[0], b.arrays()[0]) for a, b in zip(parts, again))
    # undirected: (u,v),(v,u) record pairs stay together
    u, v = rng.integers(0, 500, 300), rng.integers(0, 500, 300)
    und = Edges.from_arrays(np.stack([u, v], 1).ravel(), np.stack([v, u], 1).ravel())
    for p in und.split(n_small=4, undirected=True, seed=1):
        ps, pd = p.arrays()
        assert ps.size % 2 == 0 and np.array_equal(ps[0::2], pd[1::2])
    with pytest.raises(Exception):
        Edges.from_arrays([0, 1], [1, 2]).split(n_small=2, undirected=True)
