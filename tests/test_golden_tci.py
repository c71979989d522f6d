"""tc_i (ll_b_triangle_counting_LI, benchmark/benchmarks/triangle_counting.h:
154-336) in the oracle: against the reference's own tc_i output
(tests/golden/make_golden_tci.py) and against its closed form -- every
triangle u<v<w counted once per directed edge between u and v -- on
multi-level snapshots with deletions."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from llama_amd.graph import DIR_OUT, HostGraph, build_multilevel

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = sorted(p.stem for p in GOLDEN.glob("tci_*.json"))


def closed_form(n, arcs):
    N = [set() for _ in range(n)]
    for u, v in arcs:
        if u != v:
            N[u].add(v)
            N[v].add(u)
    T = 0
    for u in range(n):
        for v in N[u]:
            if v > u:
                k = sum(1 for w in N[u] & N[v] if w > v)
                T += k * (((u, v) in arcs) + ((v, u) in arcs))
    return T


def test_tci_fixtures_present():
    assert len(FIXTURES) >= 4


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_tci_matches_reference(name):
    doc = json.loads((GOLDEN / f"{name}.json").read_text())
    e = doc["edges"]
    g = HostGraph(doc["num_nodes"], deletions=False)
    g.add_level(e["src"], e["dst"])
    t = oracle.OracleGraph.from_host(g).tc_i()
    assert t == int(doc["reference"]["return"])
    assert t == closed_form(doc["num_nodes"], set(zip(e["src"], e["dst"])))


def test_oracle_tci_closed_form_multilevel():
    g = build_multilevel(9, levels=4, edge_factor=6, ins_frac=0.05, del_frac=0.02, seed=12)
    o = oracle.OracleGraph.from_host(g)
    ofs, adj = o.flatten(DIR_OUT)
    arcs = set()
    for u in range(g.num_nodes):
        arcs.update((u, int(v)) for v in adj[ofs[u]:ofs[u + 1]])
    assert o.tc_i() == closed_form(g.num_nodes, arcs)
    # with no reverse pairs tc_i is the simple triangle count
    one_way = {(u, v) for (u, v) in arcs if (v, u) not in arcs or u < v}
    h = HostGraph(g.num_nodes, deletions=False)
    srt = sorted(one_way)
    h.add_level([a for a, _ in srt], [b for _, b in srt])
    oh = oracle.OracleGraph.from_host(h)
    assert oh.tc_i() == oh.triangles()
