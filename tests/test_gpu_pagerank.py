"""HIP path vs the oracle: snapshot materialisation (bit-exact) and PageRank
(L1 relative error <= 1e-6 against FP64 pagerank_double), for the binned pull
(the default; exact fixed-point row sums, so results are bit-identical for
every bin size and run) and for the row-block pull and its variants (option
pr_bin_bits = 0)."""
import numpy as np
import pytest

import oracle
from conftest import gpu_case
from llama_amd.graph import DIR_IN, DIR_OUT, DeviceGraph, HostGraph, build_multilevel

pytestmark = pytest.mark.gpu

PR_TOL = 1e-6  # L1 relative, FP32 + compensated sums vs FP64


def l1rel(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


@pytest.fixture
def rowpull(ctx):
    """The row-block pull (binned pull off) for one test."""
    ctx.set_option("pr_bin_bits", 0)
    yield
    ctx.set_option("pr_bin_bits", -1)


@pytest.fixture
def seg_bits(ctx, rowpull):
    """Set the PageRank column-segment option of the row-block pull for one
    test, then restore the automatic default (-1)."""
    def set_bits(n):
        ctx.set_option("pr_seg_bits", n)
    yield set_bits
    ctx.set_option("pr_seg_bits", -1)


@pytest.fixture
def bin_bits(ctx):
    def set_bits(n):
        ctx.set_option("pr_bin_bits", n)
    yield set_bits
    ctx.set_option("pr_bin_bits", -1)


@pytest.fixture(scope="module")
def ml():
    g = build_multilevel(12, levels=6, ins_frac=0.03, del_frac=0.01, seed=7)
    return g, oracle.OracleGraph.from_host(g)


def test_snapshot_adjacency_bit_exact_every_level(ctx, ml):
    g, o = ml
    gpu_case("snapshot adjacency s12 x6 levels")
    dg = DeviceGraph(ctx, g)
    for L in range(g.num_levels):
        for d in (DIR_OUT, DIR_IN):
            oo, oa = o.flatten(d, L)
            go, ga = dg.adjacency(d, L)
            assert np.array_equal(go, oo), (L, d)
            assert np.array_equal(ga, oa), (L, d)
            deg = g.level_arrays(d, L).vt["degree"][:-1]
            assert np.array_equal(dg.degrees(d, L), deg)
    info = dg.snapshot_info()
    assert info.live_edges == oa.size and info.stored_words >= info.live_edges


def test_incremental_push_matches_full_push(ctx):
    gpu_case("incremental level push")
    g = build_multilevel(10, levels=2, ins_frac=0.05, del_frac=0.02, seed=3)
    dg = DeviceGraph(ctx, g)
    for k in range(3):
        g.add_random_level(0.04, 0.02, 50 + k)
        dg.sync_levels()  # push the new level + replay its deletion patches
        o = oracle.OracleGraph.from_host(g)
        for d in (DIR_OUT, DIR_IN):
            oo, oa = o.flatten(d)
            go, ga = dg.adjacency(d)
            assert np.array_equal(ga, oa) and np.array_equal(go, oo)


def test_missing_deletion_patches_are_reported(ctx):
    """A level pushed without the deletion patches it made leaves older words
    live that its degrees no longer count: materialising the snapshot reports
    the mismatch (and writes nothing past the row offsets) instead of
    returning a wrong adjacency.  Pushed with the patches, the same level is
    the oracle's."""
    from llama_amd.graph import GpuError
    g = build_multilevel(10, levels=2, ins_frac=0.05, del_frac=0.02, seed=7)
    dg = DeviceGraph(ctx, g)
    g.add_random_level(0.04, 0.03, 61)
    items = dg.prepare_level(g.num_levels - 1)
    assert any(it[0] == "patch" for it in items)
    dg.push_prepared([it for it in items if it[0] != "patch"])
    gpu_case("snapshot without the level's deletion patches")
    with pytest.raises(GpuError, match="missing deletion patches"):
        dg.adjacency(DIR_OUT)
    dg.close()
    dg = DeviceGraph(ctx, g)
    o = oracle.OracleGraph.from_host(g)
    for d in (DIR_OUT, DIR_IN):
        oo, oa = o.flatten(d)
        go, ga = dg.adjacency(d)
        assert np.array_equal(ga, oa) and np.array_equal(go, oo)
    dg.close()


@pytest.mark.parametrize("seg", [0, 6, 9])
@pytest.mark.parametrize("cap", [512, 1024, 2048, 4096])
def test_pagerank_multilevel(ctx, ml, cap, seg, seg_bits):
    """seg = column segments of 2^seg slots (6: 64 segments, the maximum;
    9: 8 segments; 0: plain pull)."""
    g, o = ml
    gpu_case(f"pagerank s12 x6 cap={cap} seg_bits={seg}")
    seg_bits(seg)
    dg = DeviceGraph(ctx, g)
    for L in (0, g.num_levels - 1):
        ref, ref_diff = o.pagerank(20, level=L)
        r, diff = dg.pagerank(20, level=L, cap=cap)
        err = l1rel(r, ref)
        print(f"level {L} cap {cap} L1rel {err:.3e}")
        assert err <= PR_TOL
        assert abs(diff - ref_diff) <= 1e-3 * ref_diff + 1e-9


def test_pagerank_single_level_s16(ctx):
    gpu_case("pagerank single level s16")
    g = build_multilevel(16, levels=1, seed=1, deletions=False)
    o = oracle.OracleGraph.from_host(g)
    ref, _ = o.pagerank(10)
    dg = DeviceGraph(ctx, g)
    r, _ = dg.pagerank(10)
    err = l1rel(r, ref)
    print(f"s16 L1rel {err:.3e}")
    assert err <= PR_TOL


def test_pagerank_hub_split_over_pieces(ctx, seg_bits):
    """In-degree 20000 > one piece: the partials are combined in order (plain
    pull; segmented with one segment, where the hub's single pair is split;
    and with 40 segments)."""
    gpu_case("pagerank hub star 20000")
    k = 20000
    n = k + 3
    src = np.concatenate([np.arange(1, k + 1), [0, n - 1]])
    dst = np.concatenate([np.zeros(k, int), [n - 2, 0]])
    g = HostGraph(n)
    g.add_level(src, dst)
    o = oracle.OracleGraph.from_host(g)
    ref, _ = o.pagerank(5)
    dg = DeviceGraph(ctx, g)
    for seg in (0, 15, 9):
        seg_bits(seg)
        for cap in (512, 4096):
            r, _ = dg.pagerank(5, cap=cap)
            assert l1rel(r, ref) <= PR_TOL, (seg, cap)


@pytest.mark.parametrize("seg", [0, 2])
def test_pagerank_empty_and_tiny(ctx, seg, seg_bits):
    gpu_case(f"pagerank empty graph seg_bits={seg}")
    seg_bits(seg)
    g = HostGraph(5)
    g.add_level([], [])
    dg = DeviceGraph(ctx, g)
    r, _ = dg.pagerank(1)
    assert np.allclose(r, 0.15 / 5, rtol=1e-6)
    g2 = HostGraph(1)
    g2.add_level([], [])
    r2, _ = DeviceGraph(ctx, g2).pagerank(3)
    assert np.allclose(r2, [0.15], rtol=1e-6)
    g3 = HostGraph(300)  # one edge; rows beyond the first merge block
    g3.add_level([7], [290])
    r3, _ = DeviceGraph(ctx, g3).pagerank(2)
    ref3, _ = oracle.OracleGraph.from_host(g3).pagerank(2)
    assert l1rel(r3, ref3) <= PR_TOL


@pytest.mark.parametrize("seg", [0, 7])
def test_pagerank_deterministic(ctx, ml, seg, seg_bits):
    g, _ = ml
    gpu_case(f"pagerank determinism seg_bits={seg}")
    seg_bits(seg)
    dg = DeviceGraph(ctx, g)
    a, _ = dg.pagerank(10)
    dg.drop_snapshots()
    b, _ = dg.pagerank(10)
    assert np.array_equal(a, b)


def test_mirror_errors(ctx):
    gpu_case("mirror error paths")
    from llama_amd.graph import GpuError
    g = HostGraph(4)
    g.add_level([0], [1])
    dg = DeviceGraph(ctx, g)
    v = g.level_view(DIR_OUT, 0)
    with pytest.raises(GpuError):  # level pushed twice
        ctx.check(ctx.lib.ll_gpu_graph_push_level(dg._h, v))
    with pytest.raises(GpuError):
        dg.pagerank(0)


def test_pagerank_degree_ordered_contributions(ctx, ml, seg_bits):
    """Option pr_reorder gathers contributions in descending-out-degree slot
    order; results stay within the bar and match the default layout."""
    g, o = ml
    gpu_case("pagerank pr_reorder")
    seg_bits(0)
    dg = DeviceGraph(ctx, g)
    ref, _ = o.pagerank(20)
    try:
        ctx.set_option("pr_reorder", 0)
        base, _ = dg.pagerank(20)
        ctx.set_option("pr_reorder", 1)
        dg.drop_snapshots()
        r, _ = dg.pagerank(20)
    finally:
        ctx.set_option("pr_reorder", 1)  # the default
    assert l1rel(r, ref) <= PR_TOL and l1rel(base, ref) <= PR_TOL
    # same terms in the same order: only the slots of the contributions differ
    assert np.array_equal(r, base)


def test_pagerank_segments_by_vertex_id(ctx, ml, seg_bits):
    """Column segments over vertex IDs (no degree order) and over slots agree
    with the oracle; switching the option rebuilds the cached layout."""
    g, o = ml
    gpu_case("pagerank segments without reorder")
    ref, _ = o.pagerank(20)
    dg = DeviceGraph(ctx, g)
    seg_bits(8)
    try:
        ctx.set_option("pr_reorder", 0)
        a, _ = dg.pagerank(20)
        ctx.set_option("pr_reorder", 1)
        b, _ = dg.pagerank(20)
    finally:
        ctx.set_option("pr_reorder", 1)
    assert l1rel(a, ref) <= PR_TOL and l1rel(b, ref) <= PR_TOL
    from llama_amd.graph import GpuError
    seg_bits(40)
    with pytest.raises(GpuError):
        dg.pagerank(1)


@pytest.mark.parametrize("cap", [512, 4096])
def test_pagerank_item_sorted_gathers(ctx, ml, cap, rowpull):
    """Option pr_item_sort: every row-block item's entries sorted by slot
    (LDS restores row order; pieces of long rows summed in slot order)."""
    g, o = ml
    gpu_case(f"pagerank pr_item_sort cap={cap}")
    dg = DeviceGraph(ctx, g)
    ref, _ = o.pagerank(20)
    try:
        ctx.set_option("pr_item_sort", 1)
        a, _ = dg.pagerank(20, cap=cap)
        dg.drop_snapshots()
        b, _ = dg.pagerank(20, cap=cap)
    finally:
        ctx.set_option("pr_item_sort", -1)
    assert l1rel(a, ref) <= PR_TOL
    assert np.array_equal(a, b)


def test_pagerank_item_sorted_hub(ctx, rowpull):
    gpu_case("pagerank pr_item_sort hub 20000")
    k = 20000
    n = k + 3
    rng = np.random.default_rng(2)
    src = np.concatenate([rng.permutation(np.arange(1, k + 1)), [0, n - 1]])
    dst = np.concatenate([np.zeros(k, int), [n - 2, 0]])
    g = HostGraph(n)
    g.add_level(src, dst)
    ref, _ = oracle.OracleGraph.from_host(g).pagerank(5)
    dg = DeviceGraph(ctx, g)
    try:
        ctx.set_option("pr_item_sort", 1)
        r, _ = dg.pagerank(5, cap=1024)
    finally:
        ctx.set_option("pr_item_sort", -1)
    assert l1rel(r, ref) <= PR_TOL


@pytest.mark.parametrize("cap", [512, 4096])
@pytest.mark.parametrize("hot", [6, 9])
def test_pagerank_hot_cold_split(ctx, ml, cap, hot, rowpull):
    """Option pr_hot_bits: a hot pass over in-edges from slots < 2^hot writes
    per-vertex partials, the cold pass continues from them and finalises."""
    g, o = ml
    gpu_case(f"pagerank hot/cold hot_bits={hot} cap={cap}")
    dg = DeviceGraph(ctx, g)
    try:
        ctx.set_option("pr_hot_bits", hot)
        for L in (0, g.num_levels - 1):
            ref, ref_diff = o.pagerank(20, level=L)
            r, diff = dg.pagerank(20, level=L, cap=cap)
            assert l1rel(r, ref) <= PR_TOL
            assert abs(diff - ref_diff) <= 1e-3 * ref_diff + 1e-9
        a, _ = dg.pagerank(10, cap=cap)
        dg.drop_snapshots()
        b, _ = dg.pagerank(10, cap=cap)
        assert np.array_equal(a, b)
    finally:
        ctx.set_option("pr_hot_bits", -1)


def test_pagerank_hot_cold_hub_and_empty(ctx, rowpull):
    gpu_case("pagerank hot/cold hub 20000 + empty")
    k = 20000
    n = k + 3
    rng = np.random.default_rng(3)
    src = np.concatenate([rng.permutation(np.arange(1, k + 1)), [0, n - 1]])
    dst = np.concatenate([np.zeros(k, int), [n - 2, 0]])
    g = HostGraph(n)
    g.add_level(src, dst)
    ref, _ = oracle.OracleGraph.from_host(g).pagerank(5)
    e = HostGraph(5)
    e.add_level([], [])
    try:
        for hot in (4, 14, 20):  # hub split across both passes; all hot
            ctx.set_option("pr_hot_bits", hot)
            r, _ = DeviceGraph(ctx, g).pagerank(5, cap=1024)
            assert l1rel(r, ref) <= PR_TOL, hot
        r0, _ = DeviceGraph(ctx, e).pagerank(1)
        assert np.allclose(r0, 0.15 / 5, rtol=1e-6)
    finally:
        ctx.set_option("pr_hot_bits", -1)


@pytest.fixture(scope="module")
def ml15():
    g = build_multilevel(15, levels=5, ins_frac=0.03, del_frac=0.01, seed=11)
    return g, oracle.OracleGraph.from_host(g)


def test_pagerank_bins_every_bin_size(ctx, ml15, bin_bits):
    """Binned pull with bins of 2^12, 2^13 and 2^14 destinations (8, 4 and 2
    bins at s15): within the bar at the first and last level, the final
    iteration's L1 change as the oracle's, and bit-identical ranks for every
    bin size and after a rebuild of the layout (the row sums are exact)."""
    g, o = ml15
    dg = DeviceGraph(ctx, g)
    for L in (0, g.num_levels - 1):
        ref, ref_diff = o.pagerank(20, level=L)
        first = None
        for bb in (12, 13, 14):
            gpu_case(f"pagerank bins bb={bb} level {L}")
            bin_bits(bb)
            r, diff = dg.pagerank(20, level=L)
            err = l1rel(r, ref)
            print(f"bins bb {bb} level {L} L1rel {err:.3e}")
            assert err <= PR_TOL
            assert abs(diff - ref_diff) <= 1e-3 * ref_diff + 1e-9
            if first is None:
                first = r
            assert np.array_equal(r, first)
    dg.drop_snapshots()
    r2, _ = dg.pagerank(20)
    assert np.array_equal(r2, first)


def test_pagerank_bins_ragged_and_tiny(ctx, bin_bits):
    """Vertex counts that are not a multiple of the bin (last bin partial),
    smaller than one bin, isolated vertices, an empty graph, one vertex, a hub
    with in-degree 20000 inside one bin, and a source bin whose out-edges are
    cut into several build pieces (one cut inside a 300000-edge row)."""
    bin_bits(12)
    rng = np.random.default_rng(4)
    V = 20000 + 37
    src = rng.integers(0, V, 150000)
    dst = rng.integers(0, V, 150000)
    keep = src != dst
    g = HostGraph(V)
    g.add_level(src[keep], dst[keep])
    g.add_random_level(0.05, 0.02, 3)
    ref, _ = oracle.OracleGraph.from_host(g).pagerank(15)
    gpu_case("pagerank bins ragged V=20037")
    r, _ = DeviceGraph(ctx, g).pagerank(15)
    assert l1rel(r, ref) <= PR_TOL
    k = 20000
    n = k + 3
    hs = np.concatenate([rng.permutation(np.arange(1, k + 1)), [0, n - 1]])
    hd = np.concatenate([np.zeros(k, int), [n - 2, 0]])
    h = HostGraph(n)
    h.add_level(hs, hd)
    ref, _ = oracle.OracleGraph.from_host(h).pagerank(5)
    gpu_case("pagerank bins hub 20000")
    r, _ = DeviceGraph(ctx, h).pagerank(5)
    assert l1rel(r, ref) <= PR_TOL
    # a source bin of more than 2^18 out-edges is built in several pieces, one
    # of them starting inside the out-row of vertex 0 (300000 out-edges)
    n = 6000
    ps = np.concatenate([np.zeros(300000, int), rng.integers(1, 4096, 120000)])
    pd = rng.integers(1, n, ps.size)
    p = HostGraph(n)
    p.add_level(ps, pd)
    ref, _ = oracle.OracleGraph.from_host(p).pagerank(10)
    dp = DeviceGraph(ctx, p)
    gpu_case("pagerank bins source bin in two build pieces")
    r12, _ = dp.pagerank(10)
    assert l1rel(r12, ref) <= PR_TOL
    bin_bits(13)
    r13, _ = dp.pagerank(10)
    assert np.array_equal(r12, r13)
    bin_bits(12)
    gpu_case("pagerank bins empty / one vertex / one edge")
    e = HostGraph(5)
    e.add_level([], [])
    r, d = DeviceGraph(ctx, e).pagerank(1)
    assert np.allclose(r, 0.15 / 5, rtol=1e-6)
    one = HostGraph(1)
    one.add_level([], [])
    r, _ = DeviceGraph(ctx, one).pagerank(3)
    assert np.allclose(r, [0.15], rtol=1e-6)
    t = HostGraph(300)
    t.add_level([7], [290])
    ref, _ = oracle.OracleGraph.from_host(t).pagerank(2)
    r, _ = DeviceGraph(ctx, t).pagerank(2)
    assert l1rel(r, ref) <= PR_TOL


def test_pagerank_bins_equal_row_pull(ctx, ml, bin_bits):
    """Binned and row-block pulls on the same snapshot: both within the bar,
    and within 2e-7 of each other (different FP32 roundings of the same sums)."""
    g, o = ml
    dg = DeviceGraph(ctx, g)
    gpu_case("pagerank bins vs row-block pull")
    bin_bits(13)
    a, _ = dg.pagerank(20)
    bin_bits(0)
    b, _ = dg.pagerank(20)
    ref, _ = o.pagerank(20)
    assert l1rel(a, ref) <= PR_TOL and l1rel(b, ref) <= PR_TOL
    assert l1rel(a, b) <= 2e-7


def test_pagerank_push_matches_fp64(ctx, ml):
    """Push variant (FP64 atomics over out-edges) against pagerank_double."""
    g, o = ml
    dg = DeviceGraph(ctx, g)
    for L in (0, g.num_levels - 1):
        gpu_case(f"pagerank push level {L}")
        ref, ref_diff = o.pagerank(20, level=L)
        r, diff = dg.pagerank_push(20, level=L)
        assert l1rel(r, ref) <= 1e-12
        assert abs(diff - ref_diff) <= 1e-9 * ref_diff + 1e-15
    gpu_case("pagerank push hub + empty")
    h = HostGraph(20003)
    h.add_level(np.arange(1, 20001), np.zeros(20000, int))
    ref, _ = oracle.OracleGraph.from_host(h).pagerank(3)
    r, _ = DeviceGraph(ctx, h).pagerank_push(3)
    assert l1rel(r, ref) <= 1e-12
    e = HostGraph(4)
    e.add_level([], [])
    r, _ = DeviceGraph(ctx, e).pagerank_push(2)
    assert np.allclose(r, 0.15 / 4, rtol=1e-12)
