"""The single-level CSR configuration (LL_SLCSR): the reference's own outputs
from oracle/_ref/benchmark-slcsr (tests/golden/make_golden_slcsr.py) against
the oracle on the same edges frozen as one level -- a single-level snapshot
iterates exactly begin[n]..begin[n+1]-1, so the two configurations agree."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from llama_amd.graph import HostGraph

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = sorted(p.stem for p in GOLDEN.glob("slcsr_*.json"))


def test_slcsr_fixtures_present():
    assert len(FIXTURES) >= 2


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_matches_reference_slcsr(name):
    doc = json.loads((GOLDEN / f"{name}.json").read_text())
    n, e, ref = doc["num_nodes"], doc["edges"], doc["reference"]
    g = HostGraph(n, deletions=False)
    g.add_level(e["src"], e["dst"])
    o = oracle.OracleGraph.from_host(g)
    r, _ = o.pagerank(10)
    # printed with 10 decimals
    assert np.abs(r - np.array(ref["pagerank_double"]["values"])).max() <= 6e-11
    for key, val in ref.items():
        if key.startswith("sssp_unweighted_"):
            lv, _ = o.bfs(int(key.rsplit("_", 1)[1]))
            assert np.array_equal(lv, np.array(val["values"], np.int64)), key
        elif key.startswith("bfs_"):
            _, cnt = o.bfs(int(key.rsplit("_", 1)[1]))
            assert cnt == int(val["return"]), key
