"""The reference's own driver with the GPU tasks registered
(oracle/_ref/benchmark-memory-wd-gpu: benchmark/benchmark.cc plus the
insertions of oracle/gpu_binding.py, built against libll_gpu.so) runs each GPU
task on levels the reference's loader froze from .xs1 files, and its results
are checked against the reference's CPU task in the same binary on the same
files: BFS levels, reached counts, weighted SSSP distances, betweenness
centrality (bc_adj and bc_random, whose sources both draw from the reference's
pick_random_node) and triangle counts bit-exact; push PageRank (FP64) to the
printed digits; pull PageRank (FP32) within 1e-6 L1 of pagerank_double over the
printed vertices, and the rank sums to the printed digits."""
import pytest

from binding_util import GPU_DRIVER, run, values, write_levels
from conftest import gpu_case

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not GPU_DRIVER.exists(),
                                 reason="oracle/_ref/benchmark-memory-wd-gpu is not built "
                                        "(it needs the reference tree at build time)")]


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    return write_levels(tmp_path_factory.mktemp("xs1"), 10, 3, seed=7)


def both(files, tmp_path, cpu_task, gpu_task, *extra):
    gpu_case(f"reference driver: {gpu_task} vs {cpu_task} {' '.join(extra)}")
    c = run(GPU_DRIVER, cpu_task, files, tmp_path / "cpu.txt", *extra)
    g = run(GPU_DRIVER, gpu_task, files, tmp_path / "gpu.txt", *extra)
    return c, g


def test_bfs_levels_and_counts(files, tmp_path):
    for root in ("3", "17"):
        c, g = both(files, tmp_path, "sssp_unweighted", "gpu_sssp_unweighted", "-R", root)
        assert c == g
        c, g = both(files, tmp_path, "bfs_count", "gpu_bfs_count", "-R", root)
        assert c[0] == g[0]


def test_sssp_weighted(files, tmp_path):
    c, g = both(files, tmp_path, "sssp_weighted", "gpu_sssp_weighted", "-R", "5")
    assert c == g


def test_betweenness_centrality(files, tmp_path):
    c, g = both(files, tmp_path, "bc_random", "gpu_bc_random")
    assert values(c[1]) == values(g[1]) and c[0] == g[0]
    c, g = both(files, tmp_path, "bc_adj", "gpu_bc_adj")
    assert values(c[1]) == values(g[1]) and c[0] == g[0]


def test_pagerank(files, tmp_path):
    c, g = both(files, tmp_path, "pagerank_double_push", "gpu_pagerank_push", "-n", "10")
    assert c == g
    c, g = both(files, tmp_path, "pagerank_double", "gpu_pagerank", "-n", "10")
    a, b = values(c[1]), values(g[1])
    assert len(a) == len(b) == 50
    assert sum(abs(x - y) for x, y in zip(a, b)) <= 1e-6 * sum(a)
    assert c[0] == g[0]


def test_triangles_single_level(tmp_path):
    # tc_u needs one level loaded with -U (the loader adds each edge's
    # reverse); the input keeps one direction of every pair, so the loaded
    # lists are simple.  gpu_tc symmetrises the same loaded graph itself.
    one = write_levels(tmp_path, 10, 1, seed=9, oneway=True)
    gpu_case("reference driver: gpu_tc vs tc_u (-U)")
    c = run(GPU_DRIVER, "tc_u", one, tmp_path / "cpu.txt", "-U")
    g = run(GPU_DRIVER, "gpu_tc", one, tmp_path / "gpu.txt", "-U")
    assert c[0] == g[0] and float(g[0]) > 0
