"""HIP path vs the oracle for BFS, multi-source BFS, triangle counting, SCC,
batched personalized PageRank and weighted PageRank + adjoint.

Bars: BFS levels, reached counts, triangle counts and canonical SCC labels are
bit-exact; FP32 PPR within 1e-5 L1 relative per column of the FP64 oracle;
weighted PageRank within 1e-5 and its adjoint within 1e-4 relative."""
import numpy as np
import pytest

import oracle
from conftest import gpu_case
from llama_amd.graph import DIR_IN, DIR_OUT, DeviceGraph, HostGraph, build_multilevel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ml():
    g = build_multilevel(12, levels=5, ins_frac=0.03, del_frac=0.01, seed=13, edge_factor=8)
    return g, oracle.OracleGraph.from_host(g)


def roots_with_edges(g, k, seed=0):
    o, _ = g.flatten(DIR_OUT)
    deg = np.diff(o)
    cand = np.flatnonzero(deg > 0)
    return np.random.default_rng(seed).choice(cand, k, replace=False)


# ---- BFS --------------------------------------------------------------------------
def test_bfs_levels_bit_exact(ctx, ml):
    g, o = ml
    dg = DeviceGraph(ctx, g)
    for root in list(roots_with_edges(g, 6)) + [0]:
        gpu_case(f"bfs root={root}")
        ref, n = o.bfs(int(root))
        lv, reached, exam = dg.bfs(int(root))
        assert np.array_equal(lv, ref)
        assert reached == n and exam > 0


def test_bfs_older_snapshot_and_isolated_root(ctx, ml):
    g, o = ml
    dg = DeviceGraph(ctx, g)
    gpu_case("bfs older level")
    for L in (0, 2):
        ref, n = o.bfs(5, level=L)
        lv, reached, _ = dg.bfs(5, level=L)
        assert np.array_equal(lv, ref) and reached == n
    h = HostGraph(10)
    h.add_level([0, 1], [1, 2])
    d2 = DeviceGraph(ctx, h)
    lv, reached, _ = d2.bfs(7)  # no out-edges
    assert reached == 1 and lv[7] == 0 and (lv != 0).sum() == 9


def test_bfs_batch_matches_single_source(ctx, ml):
    g, o = ml
    dg = DeviceGraph(ctx, g)
    roots = np.concatenate([roots_with_edges(g, 70, seed=3), [3, 3]])  # repeats allowed
    for lanes in (16, 32, 64, 128):
        gpu_case(f"bfs_batch lanes={lanes}")
        lv, reached, sweeps = dg.bfs_batch(roots, lane_width=lanes)
        for j in (0, 1, 15, 16, 63, 64, 69, 70, 71):
            ref, n = o.bfs(int(roots[j]))
            assert np.array_equal(lv[j], ref), (lanes, j)
            assert reached[j] == n
        assert sweeps > 0
    gpu_case("bfs_batch autotune")
    _, reached0, _ = dg.bfs_batch(roots[:40], lane_width=0, want_levels=False)
    assert dg.tuned("bfs_lanes:") in (16, 32, 64, 128)
    assert np.array_equal(reached0, reached[:40])


def test_bfs_batch_binned_sweeps(ctx):
    """Sweeps over the binned out-edge layout (bins of 2^12 over 2^14
    vertices: several bins, chunks of every size) give every lane the
    single-source levels: forced for every sweep that does not push (1), in
    the automatic mix of push, binned and pull sweeps (-1), and off (0)."""
    g = build_multilevel(14, levels=3, ins_frac=0.02, del_frac=0.01, seed=21, edge_factor=8)
    o = oracle.OracleGraph.from_host(g)
    dg = DeviceGraph(ctx, g)
    roots = np.concatenate([roots_with_edges(g, 130, seed=5), [7]])
    refs = {j: o.bfs(int(roots[j])) for j in (0, 1, 17, 63, 64, 100, 127, 128, 130)}
    ctx.set_option("pr_bin_bits", 12)
    try:
        for mode in (1, -1, 0):
            ctx.set_option("bfs_batch_bins", mode)
            for lanes in (16, 32, 64, 128):
                gpu_case(f"bfs_batch bins mode={mode} lanes={lanes}")
                ctx.profile(True)
                lv, reached, _ = dg.bfs_batch(roots, lane_width=lanes)
                binned, _ = ctx.kernel_time("k_msb_sweep")
                ctx.profile(False)
                assert (binned > 0) == (mode != 0), (mode, lanes, binned)
                for j, (ref, n) in refs.items():
                    assert np.array_equal(lv[j], ref), (mode, lanes, j)
                    assert reached[j] == n
    finally:
        ctx.set_option("bfs_batch_bins", -1)
        ctx.set_option("pr_bin_bits", -1)


# ---- triangles ---------------------------------------------------------------------
def test_triangles_match_oracle(ctx, ml):
    g, o = ml
    dg = DeviceGraph(ctx, g)
    for L in (0, g.num_levels - 1):
        gpu_case(f"tc level {L}")
        assert dg.triangles(level=L) == o.triangles(level=L)


def test_triangles_closed_forms(ctx):
    gpu_case("tc K_n and hub rows")
    m = 40  # K_40, arcs per row up to 39; plus reciprocal duplicates
    pairs = [(i, j) for i in range(m) for j in range(m) if i != j]
    h = HostGraph(m)
    h.add_level([p[0] for p in pairs], [p[1] for p in pairs])
    assert DeviceGraph(ctx, h).triangles() == m * (m - 1) * (m - 2) // 6
    # a dense core that forces the workgroup and global-memory row paths
    core = 700
    rng = np.random.default_rng(5)
    src, dst = np.nonzero(rng.random((core, core)) < 0.9)
    keep = src != dst
    h2 = HostGraph(core)
    h2.add_level(src[keep], dst[keep])
    o2 = oracle.OracleGraph.from_host(h2)
    t, arcs, work = DeviceGraph(ctx, h2).triangles(stats=True)
    assert t == o2.triangles() and work > 0


def test_triangles_global_memory_rows(ctx):
    """K_n with both directions of every pair: all degrees tie, so the oriented
    DAG ranks by ID and vertex i keeps n-1-i arcs; at n = 4100 the three
    lowest rows exceed the 4096-entry LDS row and take the global-memory
    prefix path.  Closed forms: C(n,3) triangles, tc_i twice that (two directed
    edges join every pair)."""
    n = 4100
    gpu_case(f"tc K_{n} global-memory rows")
    iu, ju = np.nonzero(~np.eye(n, dtype=bool))
    h = HostGraph(n)
    h.add_level(iu.astype(np.uint32), ju.astype(np.uint32))
    dg = DeviceGraph(ctx, h)
    t, arcs, _ = dg.triangles(stats=True)
    c3 = n * (n - 1) * (n - 2) // 6
    assert arcs == n * (n - 1) // 2
    assert t == c3
    assert dg.tc_i() == 2 * c3


# ---- SCC -----------------------------------------------------------------------------
def test_scc_labels_bit_exact(ctx, ml):
    g, o = ml
    dg = DeviceGraph(ctx, g)
    for L in (0, g.num_levels - 1):
        gpu_case(f"scc level {L}")
        ref, n = o.scc(level=L)
        lab, m = dg.scc(level=L)
        assert np.array_equal(lab, ref) and m == n


def test_scc_small_cases(ctx):
    gpu_case("scc small graphs")
    cases = [
        (3, [0, 2, 1], [2, 1, 2]),            # the reference tarjan_scc mislabels this
        (2, [1], [0]),
        (6, [0, 1, 2, 3, 4, 5], [1, 2, 0, 4, 5, 3]),  # two 3-cycles
        (5, [0, 1, 2, 3], [1, 2, 3, 4]),     # DAG
        (4, [0, 1], [0, 1]),                 # self-loops
    ]
    for n, s, d in cases:
        h = HostGraph(n)
        h.add_level(s, d)
        ref, k = oracle.OracleGraph.from_host(h).scc()
        lab, m = DeviceGraph(ctx, h).scc()
        assert np.array_equal(lab, ref) and m == k, (n, s, d)


# ---- batched personalized PageRank ---------------------------------------------------
def test_ppr_batch_matches_fp64(ctx, ml):
    g, o = ml
    dg = DeviceGraph(ctx, g)
    seeds = roots_with_edges(g, 40, seed=9)
    ref, it_ref = o.ppr(seeds, 15)
    for bw in (16, 32, 64, 128):
        gpu_case(f"ppr batch width {bw}")
        R, it = dg.ppr_batch(seeds, iters=15, tol=0.0, batch_width=bw)
        assert it == 15
        err = np.abs(R - ref).sum(axis=0) / np.abs(ref).sum(axis=0)
        print(f"bw {bw} max L1rel {err.max():.2e}")
        assert err.max() <= 1e-5


def test_ppr_tolerance_and_autotune(ctx, ml):
    g, o = ml
    dg = DeviceGraph(ctx, g)
    seeds = roots_with_edges(g, 8, seed=2)
    gpu_case("ppr tolerance stop + autotune")
    R, it = dg.ppr_batch(seeds, iters=200, tol=1e-4, batch_width=0)
    ref, it_ref = o.ppr(seeds, 200, tol=1e-4)
    assert abs(it - it_ref) <= 1 and it < 200
    assert dg.tuned("ppr:") in (16, 32, 64, 128)
    ref2, _ = o.ppr(seeds, it)
    err = np.abs(R - ref2).sum(axis=0) / np.abs(ref2).sum(axis=0)
    assert err.max() <= 1e-5


def test_ppr_uniform_teleport_reduces_to_pagerank(ctx):
    """Averaging PPR over every single-vertex seed gives the reference PR."""
    gpu_case("ppr linearity vs pagerank")
    g = build_multilevel(8, levels=3, ins_frac=0.05, del_frac=0.02, seed=4, edge_factor=6)
    dg = DeviceGraph(ctx, g)
    R, _ = dg.ppr_batch(np.arange(g.num_nodes), iters=10, tol=0.0, batch_width=64)
    pr, _ = dg.pagerank(10)
    assert np.abs(R.astype(np.float64).mean(axis=1) - pr).sum() / pr.sum() <= 1e-5


# ---- weighted PageRank + adjoint --------------------------------------------------------
def _weights(dg, seed=0):
    ofs, adj = dg.adjacency(DIR_OUT)
    rng = np.random.default_rng(seed)
    w = rng.uniform(0.1, 1.0, adj.size)
    src = np.repeat(np.arange(dg.V), np.diff(ofs).astype(np.int64))
    rs = np.bincount(src, weights=w, minlength=dg.V)
    w = w / rs[src]
    return ofs, adj, src, w


def test_wpr_forward_and_adjoint_vs_fp64(ctx):
    torch = pytest.importorskip("torch")
    from llama_amd.autograd import weighted_pagerank
    gpu_case("wpr forward/adjoint s11")
    g = build_multilevel(11, levels=3, ins_frac=0.03, del_frac=0.01, seed=17, edge_factor=8)
    dg = DeviceGraph(ctx, g)
    ofs, adj, src, w = _weights(dg)
    V = dg.V
    p = np.full(V, 1.0 / V)
    K = 60
    dev = torch.device("cuda:0")
    wt = torch.tensor(w, dtype=torch.float32, device=dev, requires_grad=True)
    pt = torch.tensor(p, dtype=torch.float32, device=dev, requires_grad=True)
    r = weighted_pagerank(wt, pt, dg, K, 0.85)
    r_ref = oracle.wpr_forward(ofs, adj, w, p, K, 0.85)
    rr = r.detach().cpu().numpy().astype(np.float64)
    assert np.abs(rr - r_ref).sum() / np.abs(r_ref).sum() <= 1e-5
    tgt = np.random.default_rng(1).choice(V, 50, replace=False)
    loss = (r[torch.tensor(tgt, device=dev)] ** 2).sum()
    loss.backward()
    gr = np.zeros(V)
    gr[tgt] = 2 * r_ref[tgt]
    dw_ref, dp_ref = oracle.wpr_adjoint(ofs, adj, w, r_ref, gr, K, 0.85)
    dw = wt.grad.cpu().numpy().astype(np.float64)
    dp = pt.grad.cpu().numpy().astype(np.float64)
    assert np.abs(dw - dw_ref).sum() / np.abs(dw_ref).sum() <= 1e-4
    assert np.abs(dp - dp_ref).sum() / np.abs(dp_ref).sum() <= 1e-4


# ---- C++ task driver (ll_gpu_bench) ----------------------------------------------------
def test_cpp_task_driver_matches_oracle(tmp_path):
    """The C++ task classes (reference task interface) through the CLI driver:
    one level per .xs1 file, results printed in the reference format."""
    import subprocess
    from pathlib import Path
    from llama_amd.graph import Edges
    gpu_case("ll_gpu_bench cli")
    root = Path(__file__).resolve().parent.parent
    exe = root / "llama_amd" / "ll_gpu_bench"
    e0 = Edges.rmat(8, 8, seed=3, perm_seed=4)
    s, d = e0.arrays()
    f0 = str(tmp_path / "l0.xs1")
    e0.write_xs1(f0, 256)
    e1 = Edges.from_arrays([1, 2, 3], [200, 100, 50])
    f1 = str(tmp_path / "l1.xs1")
    e1.write_xs1(f1, 256)
    out = str(tmp_path / "pr.txt")
    r = subprocess.run([str(exe), "-r", "gpu_pagerank", "-n", "10", "-c", "2", "-o", out, f0, f1],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = HostGraph(256)
    g.add_level(s, d)
    g.add_level([1, 2, 3], [200, 100, 50])
    ref, _ = oracle.OracleGraph.from_host(g).pagerank(10)
    vals = []
    for line in open(out):
        if ":" in line:
            vals += [float(x) for x in line.split(":")[1].split()]
    assert np.allclose(vals, ref[:50], atol=1e-8)
    ret = [l for l in r.stdout.splitlines() if l.startswith("Return value:")][0]
    assert abs(float(ret.split(":")[1]) - ref.sum()) < 1e-5
    r2 = subprocess.run([str(exe), "-r", "gpu_sssp_unweighted", "-R", "3", f0, f1],
                        capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0, r2.stderr
    lv, _ = oracle.OracleGraph.from_host(g).bfs(3)
    ret2 = [l for l in r2.stdout.splitlines() if l.startswith("Return value:")][0]
    assert int(float(ret2.split(":")[1])) == int(lv[lv < 2**31 - 2].max())
    # weighted SSSP over the .xs1 records' weights
    rng = np.random.default_rng(6)
    w0 = rng.uniform(0.1, 2.0, s.size).astype(np.float32)
    fw0, fw1 = str(tmp_path / "w0.xs1"), str(tmp_path / "w1.xs1")
    Edges.from_arrays(s, d, w0).write_xs1(fw0, 256)
    Edges.from_arrays([1, 2, 3], [200, 100, 50], np.float32([0.5, 0.25, 4.0])).write_xs1(fw1, 256)
    out3 = str(tmp_path / "sssp.txt")
    r3 = subprocess.run([str(exe), "-r", "gpu_sssp_weighted", "-R", "3", "-o", out3, fw0, fw1],
                        capture_output=True, text=True, timeout=120)
    assert r3.returncode == 0, r3.stderr
    gw = HostGraph(256)
    gw.add_level(s, d, weights=w0)
    gw.add_level([1, 2, 3], [200, 100, 50], weights=[0.5, 0.25, 4.0])
    ow = oracle.OracleGraph.from_host(gw)
    ow.add_weights(gw)
    dist, _, mx, _ = ow.sssp_weighted(3)
    vals = []
    for line in open(out3):
        if ":" in line:
            vals += [float(x) for x in line.split(":")[1].split()]
    assert np.array_equal(np.array(vals, np.float64).astype(np.float32), dist[:50])
    ret3 = [l for l in r3.stdout.splitlines() if l.startswith("Return value:")][0]
    assert int(float(ret3.split(":")[1])) == int(mx)
    # betweenness centrality from every vertex
    out4 = str(tmp_path / "bc.txt")
    r4 = subprocess.run([str(exe), "-r", "gpu_bc_adj", "-o", out4, f0, f1],
                        capture_output=True, text=True, timeout=120)
    assert r4.returncode == 0, r4.stderr
    bc = oracle.OracleGraph.from_host(g).bc()
    vals = []
    for line in open(out4):
        if ":" in line:
            vals += [float(x) for x in line.split(":")[1].split()]
    assert np.allclose(vals, bc[:50], rtol=1e-6)
    ret4 = [l for l in r4.stdout.splitlines() if l.startswith("Return value:")][0]
    assert np.isclose(float(ret4.split(":")[1]), float(bc.max()), rtol=1e-6)
