"""HIP path against the reference's own outputs (tests/golden/*.json): the
device snapshot of every fixture level, PageRank, BFS levels and counts, and
triangle counts of the symmetric single-level fixtures."""
import numpy as np
import pytest

from conftest import gpu_case
from llama_amd.graph import DIR_IN, DIR_OUT, DeviceGraph
from test_golden import FIXTURES, build, load

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", FIXTURES)
def test_gpu_matches_reference(ctx, name):
    gpu_case(f"golden {name}")
    doc = load(name)
    tasks = doc["reference"]["tasks"]
    g = build(doc)
    dg = DeviceGraph(ctx, g)
    for it in (1, 10):
        r, _ = dg.pagerank(it)
        ref = np.array(tasks[f"pagerank_double:{it}"]["values"])
        assert np.abs(r - ref).sum() / ref.sum() <= 1e-6, (name, it)
        # push variant: same terms in FP64; the bar is the fixture's printed
        # precision (10 decimal places), per element
        rp, _ = dg.pagerank_push(it)
        assert np.abs(rp - ref).max() <= 6e-11, (name, it, "push")
    for key, t in tasks.items():
        task, arg = key.split(":")
        if task == "sssp_unweighted":
            lv, n, _ = dg.bfs(int(arg))
            assert lv.tolist() == [int(x) for x in t["values"]], (name, key)
        elif task == "bfs":
            _, n, _ = dg.bfs(int(arg))
            assert n == int(t["return"][0]), (name, key)
        elif task == "tc_u" and name.endswith("_sym"):
            assert dg.triangles() == int(t["return"][0]), name
    roots = [int(k.split(":")[1]) for k in tasks if k.startswith("sssp_unweighted")]
    lv, _, _ = dg.bfs_batch(roots, lane_width=16)
    for j, r in enumerate(roots):
        assert lv[j].tolist() == [int(x) for x in tasks[f"sssp_unweighted:{r}"]["values"]]
