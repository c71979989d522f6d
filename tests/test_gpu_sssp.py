"""Weighted SSSP on the GPU (sssp_weighted, benchmark/benchmarks/sssp.h:59-172)
against the oracle and the reference's own outputs: distances bit-exact."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from conftest import gpu_case
from llama_amd.graph import DeviceGraph, GpuError, HostGraph

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def weighted_multilevel(n, m0, per_level, levels, seed, deletions=True):
    rng = np.random.default_rng(seed)
    g = HostGraph(n, deletions=deletions)
    live = set()
    for k in range(levels):
        cnt = m0 if k == 0 else per_level
        ins = []
        seen = set()
        while len(ins) < cnt:
            e = (int(rng.integers(0, n)), int(rng.integers(0, n)))
            if e[0] != e[1] and e not in live and e not in seen:
                seen.add(e)
                ins.append(e)
        dels = [] if k == 0 or not deletions else [e for e in sorted(live) if rng.random() < 0.02]
        w = rng.uniform(0.0, 4.0, len(ins)).astype(np.float32)
        g.add_level([e[0] for e in ins], [e[1] for e in ins],
                    [e[0] for e in dels], [e[1] for e in dels], weights=w)
        live |= set(ins)
        live -= set(dels)
    return g


def test_sssp_matches_oracle_multilevel(ctx):
    g = weighted_multilevel(4096, 40000, 2000, 5, seed=21)
    o = oracle.OracleGraph.from_host(g)
    o.add_weights(g)
    dg = DeviceGraph(ctx, g)
    for level in (0, g.num_levels - 1):
        for root in (0, 5, 1000):
            gpu_case(f"sssp level={level} root={root}")
            ref, n_ref, mx_ref, _ = o.sssp_weighted(root, level=level)
            d, n, mx, rounds = dg.sssp(root, level=level)
            assert np.array_equal(d, ref), (level, root)
            assert n == n_ref and mx == mx_ref and rounds > 0


@pytest.mark.parametrize("name", sorted(p.stem for p in GOLDEN.glob("sssp_*.json")))
def test_sssp_matches_reference(ctx, name):
    gpu_case(f"sssp golden {name}")
    doc = json.loads((GOLDEN / f"{name}.json").read_text())
    g = HostGraph(doc["num_nodes"])
    for k, lv in enumerate(doc["levels"]):
        rows = list(zip(lv["src"], lv["dst"]))
        w = lv["w"] if k == 0 or rows == sorted(rows) else [lv["w"][-1]] * len(rows)
        g.add_level(lv["src"], lv["dst"], weights=w)  # weights as the reference loaded them
    dg = DeviceGraph(ctx, g)
    for root, ref in doc["reference"].items():
        d, n, mx, _ = dg.sssp(int(root))
        assert np.array_equal(d, np.array(ref["values"], np.float64).astype(np.float32))
        assert int(mx) == int(ref["return"])


def chain_with_shortcuts(n, seed):
    """A long weighted chain plus a few forward shortcuts: distances spread
    over thousands of threshold steps with a handful of vertices each."""
    rng = np.random.default_rng(seed)
    src = list(range(n - 1))
    dst = list(range(1, n))
    for _ in range(n // 20):
        a = int(rng.integers(0, n - 10))
        src.append(a)
        dst.append(a + int(rng.integers(2, 10)))
    pairs = sorted(set(zip(src, dst)))
    w = rng.uniform(0.0, 4.0, len(pairs)).astype(np.float32)
    g = HostGraph(n)
    g.add_level([p[0] for p in pairs], [p[1] for p in pairs], weights=w)
    return g


@pytest.mark.parametrize("milli,minq", [(0, 0), (60, 0), (60, -1), (30, 1), (250, 64), (1000, 0),
                                        (10, 1 << 30), (1, -1)])
def test_sssp_schedules_match_oracle(ctx, milli, minq):
    """Every near/far step and queue floor gives the same distances (the
    schedule only changes how often edges are relaxed)."""
    gpu_case(f"sssp schedule delta={milli}/1000 min_queue={minq}")
    graphs = [weighted_multilevel(4096, 40000, 2000, 3, seed=23), chain_with_shortcuts(3000, 5)]
    try:
        ctx.set_option("sssp_delta_milli", milli)
        ctx.set_option("sssp_min_queue", minq)
        for g in graphs:
            o = oracle.OracleGraph.from_host(g)
            o.add_weights(g)
            dg = DeviceGraph(ctx, g)
            for root in (0, 7, 1500):
                ref, n_ref, mx_ref, _ = o.sssp_weighted(root)
                d, n, mx, _ = dg.sssp(root)
                assert np.array_equal(d, ref), (milli, minq, root)
                assert n == n_ref and mx == mx_ref
    finally:
        ctx.set_option("sssp_delta_milli", -1)
        ctx.set_option("sssp_min_queue", -1)


def test_sssp_edge_cases(ctx):
    gpu_case("sssp edge cases")
    g = HostGraph(6)
    g.add_level([0, 1, 1], [1, 2, 3], weights=[0.0, 2.5, 1.0])
    dg = DeviceGraph(ctx, g)
    d, n, mx, _ = dg.sssp(5)  # root without out-edges: only itself
    assert n == 1 and d[5] == 0 and mx == 0
    d, n, mx, _ = dg.sssp(0)  # zero weight
    assert d[1] == 0 and d[2] == np.float32(2.5) and d[3] == 1 and n == 4
    with pytest.raises(GpuError):
        dg.sssp(6)
    u = HostGraph(3)
    u.add_level([0], [1])
    with pytest.raises(GpuError):  # unweighted snapshot
        DeviceGraph(ctx, u).sssp(0)


def test_push_weights_validation(ctx):
    gpu_case("push weights validation")
    g = HostGraph(3)
    g.add_level([0], [1], weights=[1.0])
    dg = DeviceGraph(ctx, g)
    import ctypes as C
    bad = np.array([-1.0], np.float32)
    with pytest.raises(GpuError):
        ctx.check(ctx.lib.ll_gpu_graph_push_weights(dg._h, 0, bad.ctypes.data_as(C.POINTER(C.c_float)), 1))
    two = np.array([1.0, 2.0], np.float32)
    with pytest.raises(GpuError):
        ctx.check(ctx.lib.ll_gpu_graph_push_weights(dg._h, 0, two.ctypes.data_as(C.POINTER(C.c_float)), 2))
