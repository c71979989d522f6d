"""Betweenness centrality oracle (bc_adj / bc_random,
benchmark/benchmarks/bc_adj.h, bc_random.h) against the reference's own
bc_adj output (tests/golden/make_golden_bc.py), bit-exact, and against
networkx's exact directed betweenness in FP64."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from llama_amd.graph import DIR_OUT, HostGraph, build_multilevel

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = sorted(p.stem for p in GOLDEN.glob("bc_*.json"))


def build(doc) -> HostGraph:
    g = HostGraph(doc["num_nodes"])
    for lv in doc["levels"]:
        g.add_level(lv["src"], lv["dst"])
    return g


def test_bc_fixtures_present():
    assert len(FIXTURES) >= 3


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_bc_matches_reference(name):
    doc = json.loads((GOLDEN / f"{name}.json").read_text())
    b = oracle.OracleGraph.from_host(build(doc)).bc()
    ref = np.array(doc["reference"]["values"], np.float64).astype(np.float32)
    assert np.array_equal(b, ref)
    assert np.isclose(float(b.max()), doc["reference"]["return"], rtol=1e-6)


def test_oracle_bc_matches_networkx_multilevel():
    nx = pytest.importorskip("networkx")
    g = build_multilevel(8, levels=3, ins_frac=0.05, del_frac=0.02, seed=3, edge_factor=4)
    o = oracle.OracleGraph.from_host(g)
    ofs, adj = g.flatten(DIR_OUT)
    G = nx.DiGraph()
    G.add_nodes_from(range(g.num_nodes))
    for u in range(g.num_nodes):
        G.add_edges_from((u, int(v)) for v in adj[ofs[u]:ofs[u + 1]])
    exact = nx.betweenness_centrality(G, normalized=False)
    ref = np.array([exact[i] for i in range(g.num_nodes)])
    assert np.allclose(o.bc(), ref, rtol=2e-6, atol=1e-3)
    # sources subset = sum of the single-source contributions
    src = [3, 17, 40]
    parts = sum(o.bc([s]).astype(np.float64) for s in src)
    assert np.allclose(o.bc(src), parts, rtol=1e-5, atol=1e-4)
