"""Weighted SSSP oracle (sssp_weighted, benchmark/benchmarks/sssp.h:59-172)
against the reference's own outputs (tests/golden/make_golden_sssp.py).

Distances are bit-exact: the reference prints 10 decimals of each FP32
distance, which rounds back to the same float32.  One fixture keeps unsorted
level files, for which the reference loader gives every edge of the level
the weight of the file's last record (ll_load_utils.h:750-755); the oracle
reproduces that output only with the same substitution."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from llama_amd.graph import HostGraph

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURES = sorted(p.stem for p in GOLDEN.glob("sssp_*.json"))


def load(name):
    return json.loads((GOLDEN / f"{name}.json").read_text())


def build(doc, loader_weights: bool) -> HostGraph:
    """The fixture's levels; loader_weights = weights as the reference loader
    assigns them (last record's weight for every edge of an unsorted level)."""
    g = HostGraph(doc["num_nodes"])
    for k, lv in enumerate(doc["levels"]):
        w = lv["w"]
        rows = list(zip(lv["src"], lv["dst"]))
        if loader_weights and k > 0 and rows != sorted(rows):
            w = [lv["w"][-1]] * len(w)
        g.add_level(lv["src"], lv["dst"], weights=w)
    return g


def reference_dist(ref):
    return np.array(ref["values"], np.float64).astype(np.float32)


def test_sssp_fixtures_present():
    assert len(FIXTURES) >= 4


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_sssp_matches_reference(name):
    doc = load(name)
    g = build(doc, loader_weights=True)
    o = oracle.OracleGraph.from_host(g)
    o.add_weights(g)
    for root, ref in doc["reference"].items():
        dist, reached, mx, _ = o.sssp_weighted(int(root))
        assert np.array_equal(dist, reference_dist(ref)), (name, root)
        assert int(mx) == int(ref["return"]), (name, root)  # LL_SSSP_RETURNS_MAX


def test_reference_loader_weight_quirk_changes_distances():
    """With the weights the files hold, the unsorted fixture's distances differ
    from the reference's: the discrepancy is the loader's, not the task's."""
    doc = load("sssp_unsorted_40_3level")
    assert not doc["levels_sorted"]
    g = build(doc, loader_weights=False)
    o = oracle.OracleGraph.from_host(g)
    o.add_weights(g)
    dist, _, _, _ = o.sssp_weighted(0)
    assert not np.array_equal(dist, reference_dist(doc["reference"]["0"]))
