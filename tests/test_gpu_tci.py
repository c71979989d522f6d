"""tc_i on the device (ll_gpu_tc_i, benchmark/benchmarks/triangle_counting.h:
154-336) against the reference's own outputs and the oracle's restatement of
its merge loop, on single-level fixtures with reciprocal edges and on
multi-level snapshots with deletions."""
import json
from pathlib import Path

import pytest

import oracle
from conftest import gpu_case
from llama_amd.graph import DIR_OUT, DeviceGraph, HostGraph, build_multilevel

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.mark.parametrize("name", sorted(p.stem for p in GOLDEN.glob("tci_*.json")))
def test_tci_matches_reference(ctx, name):
    gpu_case(f"tc_i golden {name}")
    doc = json.loads((GOLDEN / f"{name}.json").read_text())
    e = doc["edges"]
    g = HostGraph(doc["num_nodes"], deletions=False)
    g.add_level(e["src"], e["dst"])
    assert DeviceGraph(ctx, g).tc_i() == int(doc["reference"]["return"])


@pytest.mark.parametrize("scale,levels,seed", [(10, 4, 3), (12, 5, 4), (13, 1, 5)])
def test_tci_matches_oracle_multilevel(ctx, scale, levels, seed):
    gpu_case(f"tc_i rmat s{scale} x {levels}")
    g = build_multilevel(scale, levels=levels, edge_factor=8, ins_frac=0.05, del_frac=0.01,
                         seed=seed)
    dg = DeviceGraph(ctx, g)
    o = oracle.OracleGraph.from_host(g)
    for level in range(g.num_levels):
        assert dg.tc_i(level) == o.tc_i(level), level
    assert dg.tc_i() >= dg.triangles()


def test_tc_and_tci_rmat_s18_match_oracle(ctx):
    """R-MAT s18 EF16 with two incremental levels: the oriented DAG holds rows
    of a few hundred arcs (415 at most for the single-level s18 graph), so
    both the wave rows (<= 256 arcs) and the workgroup rows (257..4096) count
    at a realistic degree skew; the oracle's merge-based tc_u and tc_i agree
    exactly."""
    gpu_case("tc / tc_i rmat s18 x 3")
    g = build_multilevel(18, levels=3, edge_factor=16, ins_frac=0.02, del_frac=0.005, seed=11)
    dg = DeviceGraph(ctx, g)
    o = oracle.OracleGraph.from_host(g)
    t, arcs, work = dg.triangles(stats=True)
    assert t == o.triangles()
    assert dg.tc_i() == o.tc_i()
    assert arcs > 0 and work > 2 * arcs


def test_tci_without_reverse_pairs_is_the_triangle_count(ctx):
    gpu_case("tc_i one-way graph")
    base = build_multilevel(11, levels=1, edge_factor=8, seed=6)
    o = oracle.OracleGraph.from_host(base)
    ofs, adj = o.flatten(DIR_OUT)
    arcs = set()
    for u in range(base.num_nodes):
        arcs.update((u, int(v)) for v in adj[ofs[u]:ofs[u + 1]])
    one_way = sorted((u, v) for (u, v) in arcs if (v, u) not in arcs or u < v)
    g = HostGraph(base.num_nodes, deletions=False)
    g.add_level([a for a, _ in one_way], [b for _, b in one_way])
    dg = DeviceGraph(ctx, g)
    assert dg.tc_i() == dg.triangles() == oracle.OracleGraph.from_host(g).triangles()


def test_cli_tc_i(tmp_path):
    """ll_gpu_bench -r gpu_tc_i on an .xs1 level prints the reference's count."""
    import subprocess
    from llama_amd.graph import Edges
    gpu_case("ll_gpu_bench -r gpu_tc_i")
    doc = json.loads((GOLDEN / "tci_dense_30.json").read_text())
    e = doc["edges"]
    f = str(tmp_path / "g.xs1")
    Edges.from_arrays(e["src"], e["dst"]).write_xs1(f, doc["num_nodes"])
    exe = Path(__file__).resolve().parent.parent / "llama_amd" / "ll_gpu_bench"
    r = subprocess.run([str(exe), "-r", "gpu_tc_i", f], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    ret = [l for l in r.stdout.splitlines() if l.startswith("Return value:")][0]
    assert int(float(ret.split(":")[1])) == int(doc["reference"]["return"])
