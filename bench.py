#!/usr/bin/env python3
"""Benchmark: PageRank over the multi-level snapshot on one MI355X.

Workload (default): R-MAT scale 24, edge factor 16 (a,b,c,d = .57,.19,.19,.05),
permuted IDs, no self-loops/duplicates, base level + 10 incremental levels of
+1% uniform inserts / -0.1% uniform deletes (deletion vector on), out- and
in-edges mirrored in HBM.  One step = one full PageRank run from the HBM
mirror: snapshot materialisation of the out-edges (the multi-level iterator,
run on the device), the binned layout built from them, and 20 pull
iterations at d = 0.85 (task pagerank, benchmark/benchmarks/pagerank.h:98-182).
GTEPS = E_live * 20 / step time.

Prints one JSON line.  The CPU baseline is the oracle's OpenMP restatement of
pagerank_double walking the same multi-level layout, on a bounded number of
iterations.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "GTEPS (PageRank iter, BFS) R-MAT s24 11 levels; % HBM peak; batch & adjoint"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pick_sources(g, k: int, seed: int):
    """k distinct uniform vertices with out+in degree > 0 (node_exists,
    ll_mlcsr_graph.h:691-695)."""
    import numpy as np
    L = g.num_levels - 1
    deg = (g.level_arrays(0, L).vt["degree"][:-1].astype(np.int64) +
           g.level_arrays(1, L).vt["degree"][:-1])
    cand = np.flatnonzero(deg > 0)
    return np.random.default_rng(seed).choice(cand, k, replace=False)


def bench_bfs(dg, g, E: int, V: int, args) -> dict:
    """64 single-source direction-optimising BFS runs (levels kept on device)."""
    roots = pick_sources(g, args.bfs_sources, args.seed + 3)
    dg.bfs(int(roots[0]), want_levels=False)  # warm the snapshot
    ctx = dg.ctx
    ctx.sync()
    trav = exam = reached = 0
    levels_total = 0
    t0 = time.perf_counter()
    for r in roots:
        _, n, x = dg.bfs(int(r), want_levels=False)
        trav += dg.last_bfs_traversed
        exam += x
        reached += n
    ctx.sync()
    t = time.perf_counter() - t0
    # the kernel breakdown from a second pass with launch timers on (they
    # add host work per launch, so the timed pass runs without them)
    ctx.profile(True)
    for r in roots:
        dg.bfs(int(r), want_levels=False)
    ctx.sync()
    n_td, ms_td = ctx.kernel_time("k_td_expand")
    # bottom-up level launches, including up to three launched past a search's
    # last level (bottom-up levels go out four per read-back once a search
    # stays bottom-up; those find nothing)
    n_bu, ms_bu = ctx.kernel_time("k_bu_step")
    ctx.profile(False)
    levels_total = n_td + n_bu
    b_bfs = 4 * exam + 12 * reached + 2 * ((V + 7) // 8) * levels_total
    log(f"BFS x{len(roots)}: {t * 1e3:.1f} ms, {trav / t / 1e9:.2f} GTEPS")
    return {
        "sources": len(roots), "seconds": round(t, 4),
        "gteps": round(trav / t / 1e9, 4), "edges_traversed": trav,
        "edges_examined": exam, "reached": reached, "level_launches": levels_total,
        "alg_bytes": b_bfs, "alg_gbs": round(b_bfs / t / 1e9, 2),
        "frac_hbm": round(b_bfs / t / 1e9 / HBM_PEAK_GBS, 4),
        "ms_top_down": round(ms_td, 3), "ms_bottom_up": round(ms_bu, 3),
    }


def measured_copy_gbs(mib: int = 2048, reps: int = 5) -> dict:
    """STREAM-like device copy of `mib` MiB (read + write bytes / time), timed
    with events on torch's current stream, the stream the copy runs on."""
    import torch
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)  # warm
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        dst.copy_(src)
    t1.record()
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / reps
    del src, dst
    torch.cuda.empty_cache()
    return {"bytes": 2 * n, "ms": round(ms, 4), "gbs": round(2 * n / (ms * 1e-3) / 1e9, 1)}


def cpu_model() -> str:
    """The host CPU's model name (first /proc/cpuinfo entry)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cpu_layout() -> dict:
    """Sockets, cores and threads of the host (lscpu) and the CPUs this
    process may run on (its affinity mask)."""
    import subprocess
    out = {"affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
                     "Model name", "NUMA node(s)"):
                out[k] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return out


def tuned_key(kind: str, dg, E: int | None, n: int) -> str:
    """Key of the autotuner decision for a request of n sources/seeds on the
    latest snapshot: kind:V:E:level:bucket, bucket = smallest power of two
    >= min(n, 128) (at least 16)."""
    if E is None:
        E = dg.snapshot_info().live_edges
    b = 16
    while b < min(n, 128):
        b <<= 1
    return f"{kind}:{dg.V}:{E}:{dg.num_levels - 1}:{b}"


def bench_sssp(dg, g, E: int, V: int, args) -> dict:
    """Weighted SSSP (sssp_weighted) from a few sources over the levels'
    U[0, 1) weights; distances stay on the device.  --sssp-sweep times the
    listed near/far threshold steps (thousandths of the mean weight) as well."""
    roots = pick_sources(g, args.sssp_sources, args.seed + 7)
    ctx = dg.ctx

    def timed(pct: int, minq: int = -1) -> dict:
        ctx.set_option("sssp_delta_milli", pct)
        ctx.set_option("sssp_min_queue", minq)
        dg.sssp(int(roots[0]), want_dist=False)  # warm the weighted snapshot
        ctx.sync()
        reached = rounds = 0
        t0 = time.perf_counter()
        for r in roots:
            _, n, _, k = dg.sssp(int(r), want_dist=False)
            reached += n
            rounds += k
        ctx.sync()
        t = time.perf_counter() - t0
        # relax kernel time from a second pass with launch timers on
        ctx.profile(True)
        for r in roots:
            dg.sssp(int(r), want_dist=False)
        ctx.sync()
        n_rx, ms_rx = ctx.kernel_time("k_relax")
        ctx.profile(False)
        # algorithmic floor per source: every live edge relaxed once (4 B ID,
        # 4 B weight, 4 B distance) and one 4 B distance per vertex
        alg = (12 * E + 4 * V) * len(roots)
        return {"sources": len(roots), "seconds": round(t, 4),
                "ms_per_source": round(t * 1e3 / len(roots), 2),
                "gteps": round(E * len(roots) / t / 1e9, 3), "reached": reached,
                "rounds": rounds, "relax_kernel_ms": round(ms_rx, 3),
                "alg_bytes": alg, "alg_gbs": round(alg / t / 1e9, 2),
                "frac_hbm": round(alg / t / 1e9 / HBM_PEAK_GBS, 4),
                "relax_alg_gbs": round(alg / max(ms_rx, 1e-9) / 1e6, 2)}

    sweep = {}
    for item in [x.strip() for x in args.sssp_sweep.split(",") if x.strip()]:
        pct, _, minq = item.partition(":")
        sweep[item] = timed(int(pct), int(minq) if minq else -1)
        log(f"SSSP delta {pct}/1000 min queue {minq or 'auto'}: {sweep[item]}")
    out = timed(args.sssp_delta_milli)
    out["delta_milli"] = args.sssp_delta_milli
    if sweep:
        out["delta_sweep_ms_per_source"] = {k: v["ms_per_source"] for k, v in sweep.items()}
    ctx.set_option("sssp_delta_milli", -1)
    ctx.set_option("sssp_min_queue", -1)
    log(f"SSSP: {out}")
    return out


def bench_cow_upload(ctx, args) -> dict:
    """The level-freeze -> device pipeline on an R-MAT base plus levels
    confined to windows of V/64 vertex IDs (1% of the base's edges inserted,
    0.01% deleted, both ends in the window): per level the host freeze, the
    push of whole vertex tables and the push of the copy-on-write pages the
    level wrote; then freezes overlapped with page pushes on a second thread."""
    from concurrent.futures import ThreadPoolExecutor

    from llama_amd.graph import DeviceGraph, build_multilevel
    g = build_multilevel(args.cow_scale, levels=1, seed=args.seed + 11)
    V = g.num_nodes
    e0 = int(g.flatten(0)[1].size)
    w = V // 64
    whole = DeviceGraph(ctx, g)
    cow = DeviceGraph(ctx, g, cow=True)
    ctx.sync()

    def window(k):
        lo = (k * 7 * w) % (V - w)
        g.add_window_level(e0 // 100, e0 // 10000, lo, lo + w, args.seed + 200 + k)

    t = {"freeze": 0.0, "whole": 0.0, "cow": 0.0}
    b0 = {"whole": whole.upload_stats()[0], "cow": cow.upload_stats()[0]}
    for k in range(args.cow_levels):
        t0 = time.perf_counter()
        window(k)
        t["freeze"] += time.perf_counter() - t0
        for name, dg in (("whole", whole), ("cow", cow)):
            t0 = time.perf_counter()
            dg.sync_levels()  # the cow mirror lists the pages the freeze wrote
            ctx.sync()
            t[name] += time.perf_counter() - t0
    n = args.cow_levels
    b1 = {"whole": whole.upload_stats()[0] - b0["whole"], "cow": cow.upload_stats()[0] - b0["cow"]}
    pages = [g.level_pages(d, g.num_levels - 1)[1] for d in (0, 1)]
    ra, _ = whole.pagerank(3)
    rb, _ = cow.pagerank(3)
    out = {"graph": f"R-MAT s{args.cow_scale} EF16 base + {n} window levels "
                    f"(window V/64, +1%/-0.01% of the base edges each)",
           "vt_pages_per_level": (V + 1 + 511) // 512,
           "written_pages_per_level_out_in": pages,
           "freeze_ms_per_level": round(t["freeze"] * 1e3 / n, 2),
           "whole_tables": {"bytes_per_level": b1["whole"] // n,
                            "ms_per_level": round(t["whole"] * 1e3 / n, 2),
                            "gb_per_s": round(b1["whole"] / t["whole"] / 1e9, 2)},
           "cow_pages": {"bytes_per_level": b1["cow"] // n,
                         "ms_per_level": round(t["cow"] * 1e3 / n, 2),
                         "gb_per_s": round(b1["cow"] / t["cow"] / 1e9, 2)},
           "identical_pagerank": bool((ra == rb).all())}
    whole.close()
    # the pipeline: level k's pages travel while the host freezes level k+1
    pipe = DeviceGraph(ctx, g, cow=True)
    ctx.sync()
    t0 = time.perf_counter()
    with ThreadPoolExecutor(1) as ex:
        fut = None
        for k in range(n, 2 * n):
            window(k)
            items = pipe.prepare_level(g.num_levels - 1)
            if fut is not None:
                fut.result()
            fut = ex.submit(pipe.push_prepared, items)
        fut.result()
    ctx.sync()
    t_pipe = time.perf_counter() - t0
    out["overlapped_ms_per_level"] = round(t_pipe * 1e3 / n, 2)
    out["sequential_ms_per_level"] = round((t["freeze"] + t["cow"]) * 1e3 / n, 2)
    rc, _ = pipe.pagerank(3)
    cow.sync_levels()
    rd, _ = cow.pagerank(3)
    out["pipelined_identical_pagerank"] = bool((rc == rd).all())
    pipe.close()
    cow.close()
    log(f"COW upload: {out}")
    return out


def bench_bc(dg, g, E: int, V: int, args) -> dict:
    """Betweenness centrality (bc_random's Brandes passes) from a few sources;
    per source a BFS, a path-count pass and a dependency pass over the edges."""
    roots = pick_sources(g, args.bc_sources, args.seed + 8)
    dg.bc([int(roots[0])])  # warm the snapshot and row blocks
    ctx = dg.ctx
    ctx.sync()
    t0 = time.perf_counter()
    b = dg.bc([int(r) for r in roots])
    t = time.perf_counter() - t0
    # pass times from a second run with launch timers on
    ctx.profile(True)
    dg.bc([int(r) for r in roots])
    ctx.sync()
    n_p, ms_p = ctx.kernel_time("k_bc_pass")
    n_b, ms_b = ctx.kernel_time("k_bu_step")
    ctx.profile(False)
    # algorithmic floor per source of the two Brandes passes: every live edge
    # read once as an in-edge and once as an out-edge (4 B ID + 4 B level each)
    alg = 16 * E * len(roots)
    out = {"sources": len(roots), "seconds": round(t, 4),
           "ms_per_source": round(t * 1e3 / len(roots), 2),
           "gteps": round(E * len(roots) / t / 1e9, 3),
           "pass_launches": n_p, "pass_kernel_ms": round(ms_p, 3),
           "bfs_bottom_up_ms": round(ms_b, 3), "max_bc": float(b.max()),
           "pass_alg_bytes": alg, "pass_alg_gbs": round(alg / max(ms_p, 1e-9) / 1e6, 2),
           "pass_frac_hbm": round(alg / max(ms_p, 1e-9) / 1e6 / HBM_PEAK_GBS, 4)}
    log(f"BC: {out}")
    return out


def bench_tc(dg, E: int, V: int, args) -> dict:
    """Triangle counting (tc_u over the symmetrised snapshot) and tc_i.  The
    first call builds the degree-oriented DAG; later calls reuse it."""
    ctx = dg.ctx
    ctx.sync()
    t0 = time.perf_counter()
    tri, arcs, work = dg.triangles(stats=True)
    t_first = time.perf_counter() - t0
    ctx.profile(True)
    t0 = time.perf_counter()
    for _ in range(args.tc_reps):
        tri2 = dg.triangles()
    ctx.sync()
    t = (time.perf_counter() - t0) / args.tc_reps
    _, ms_w = ctx.kernel_time("k_tc_wave")
    _, ms_b = ctx.kernel_time("k_tc_block")
    _, ms_g = ctx.kernel_time("k_tc_global")
    ctx.profile(False)
    t0 = time.perf_counter()
    tci = dg.tc_i()
    t_tci_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    tci2 = dg.tc_i()
    t_tci = time.perf_counter() - t0
    # B_TC: every intersection element read once (4 B) and one offset pair per vertex
    alg = 4 * work + 8 * V
    k_ms = (ms_w + ms_b + ms_g) / args.tc_reps
    out = {"triangles": tri, "repeat_equal": tri2 == tri, "dag_arcs": arcs, "work": work,
           "first_call_s": round(t_first, 4), "count_s": round(t, 4),
           "kernel_ms": round(k_ms, 3),
           "kernel_ms_by_row_class": {"wave": round(ms_w / args.tc_reps, 3),
                                      "workgroup": round(ms_b / args.tc_reps, 3),
                                      "global": round(ms_g / args.tc_reps, 3)},
           "alg_bytes": alg,
           "alg_gbs": round(alg / max(k_ms, 1e-9) / 1e6, 2),
           "frac_hbm": round(alg / max(k_ms, 1e-9) / 1e6 / HBM_PEAK_GBS, 4),
           "triangles_per_s": round(tri / t), "work_per_s": round(work / t),
           "tc_i": tci, "tc_i_repeat_equal": tci2 == tci,
           "tc_i_first_call_s": round(t_tci_first, 4), "tc_i_count_s": round(t_tci, 4)}
    log(f"TC: {out}")
    return out


def bench_scc(dg, E: int, V: int, args) -> dict:
    """Strongly connected components of the snapshot (labels stay on the device)."""
    ctx = dg.ctx
    _, n0 = dg.scc(want_labels=False)  # warm both CSR directions
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.scc_reps):
        _, n = dg.scc(want_labels=False)
    ctx.sync()
    t = (time.perf_counter() - t0) / args.scc_reps
    # B_SCC: every full-graph sweep (trim, propagation, BFS) reads the live
    # edges and writes per-vertex state: sweeps x (8 B per edge + 16 B per vertex)
    sweeps = dg.last_scc_sweeps
    per_sweep = 8 * E + 16 * V
    alg = sweeps * per_sweep
    out = {"num_scc": n, "repeat_equal": n == n0, "seconds": round(t, 4),
           "sweeps": sweeps, "gteps": round(E / t / 1e9, 3),
           "alg_bytes": alg, "alg_gbs": round(alg / t / 1e9, 2),
           "frac_hbm": round(alg / t / 1e9 / HBM_PEAK_GBS, 4),
           "floor_one_sweep_frac_hbm": round(per_sweep / t / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"SCC: {out}")
    return out


def bench_bfs_batch(dg, g, V: int, args, single: dict | None) -> dict:
    """Multi-source BFS over the batch sources in 64-bit lanes (and the
    autotuned width), against the per-source cost of single-source runs.
    Sweeps: pushes from small frontiers, binned sweeps while the rows some
    search has not reached hold many edges, pulls otherwise."""
    import numpy as np
    roots = pick_sources(g, args.batch_sources, args.seed + 4)
    ctx = dg.ctx
    out = {"sources": len(roots)}
    for lanes in (64, 0):
        dg.bfs_batch(roots[:128], lane_width=lanes, want_levels=False)  # warm + tune
        ctx.sync()
        t0 = time.perf_counter()
        _, reached, sweeps = dg.bfs_batch(roots, lane_width=lanes, want_levels=False)
        ctx.sync()
        t = time.perf_counter() - t0
        # sweep times from a second run with launch timers on
        ctx.profile(True)
        dg.bfs_batch(roots, lane_width=lanes, want_levels=False)
        ctx.sync()
        kt = {k: ctx.kernel_time(k) for k in ("k_ms_pull", "k_msb_sweep", "k_ms_push")}
        ctx.profile(False)
        key = "lanes64" if lanes == 64 else \
            f"autotuned_lanes{dg.tuned(tuned_key('bfs_lanes', dg, None, len(roots)))}"
        out[key] = {"seconds": round(t, 4), "sweeps": sweeps,
                    "pull_sweeps": kt["k_ms_pull"][0], "pull_ms": round(kt["k_ms_pull"][1], 3),
                    "binned_sweeps": kt["k_msb_sweep"][0],
                    "binned_ms": round(kt["k_msb_sweep"][1], 3),
                    "push_kernel_ms": round(kt["k_ms_push"][1], 3),
                    "reached_total": int(reached.sum())}
    # The same sources as sequential single-source runs, timed: they also give
    # the Graph500 numerator per source, E_comp(s) = sum of live out-degrees
    # of the vertices s reaches, and their reached counts cross-check the
    # batched lanes
    e_comp, same_reach = 0, True
    dg.bfs(int(roots[0]), want_levels=False)  # prepared snapshot, as for the batches
    ctx.sync()
    t0 = time.perf_counter()
    for i, r in enumerate(roots):
        _, n_r, _ = dg.bfs(int(r), want_levels=False)
        e_comp += dg.last_bfs_traversed
        same_reach &= n_r == int(reached[i])
    ctx.sync()
    seq = time.perf_counter() - t0
    out["edges_traversed_total"] = e_comp
    out["reached_equal_single_source"] = bool(same_reach)
    out["sequential_seconds"] = round(seq, 4)
    out["sequential_gteps"] = round(e_comp / seq / 1e9, 2)
    for k, v in list(out.items()):
        if isinstance(v, dict) and "seconds" in v:
            v["aggregate_gteps"] = round(e_comp / v["seconds"] / 1e9, 2)
    out["speedup_vs_sequential_lanes64"] = round(seq / out["lanes64"]["seconds"], 2)
    best = min(v["seconds"] for k, v in out.items() if isinstance(v, dict))
    out["speedup_vs_sequential_best"] = round(seq / best, 2)
    log(f"batched BFS x{len(roots)}: {out}")
    return out


def bench_ppr(dg, g, E: int, V: int, args) -> dict:
    """Batched personalized PageRank, 30 iterations, tol 1e-7, autotuned width."""
    seeds = pick_sources(g, args.ppr_seeds, args.seed + 5)
    ctx = dg.ctx
    # autotune on a request of the same size bucket (min(n, 128) seeds), then
    # time the product path (batch_width=0 reuses the decision)
    dg.ppr_batch(seeds[:128], iters=2, tol=0.0, batch_width=0, want_ranks=False)
    width = dg.tuned(tuned_key("ppr", dg, E, len(seeds)))
    ctx.sync()
    t0 = time.perf_counter()
    _, it = dg.ppr_batch(seeds, iters=30, tol=1e-7, batch_width=0, want_ranks=False)
    ctx.sync()
    t = time.perf_counter() - t0
    b = len(seeds)
    alg = (4 * E + 4 * b * E + 8 * b * V + 8 * V) * it
    out = {"seeds": b, "iterations": it, "batch_width": width, "seconds": round(t, 4),
           "gteps": round(E * b * it / t / 1e9, 3), "alg_gbs": round(alg / t / 1e9, 1),
           "frac_hbm": round(alg / t / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"PPR: {out}")
    return out


def bench_wpr(ctx, args) -> dict:
    """Weighted PageRank forward vs adjoint on R-MAT (single level), weights
    U(0.1, 1) row-normalised, uniform teleport, 50 iterations each."""
    import numpy as np
    import torch
    from llama_amd.graph import DIR_OUT, DeviceGraph, build_multilevel
    g = build_multilevel(args.wpr_scale, levels=1, seed=args.seed + 5, deletions=False)
    dg = DeviceGraph(ctx, g)
    ofs, adj = dg.adjacency(DIR_OUT)
    V, E = dg.V, adj.size
    rng = np.random.default_rng(args.seed + 6)
    w = rng.uniform(0.1, 1.0, E).astype(np.float32)
    src = np.repeat(np.arange(V), np.diff(ofs).astype(np.int64))
    rs = np.bincount(src, weights=w, minlength=V)
    w = (w / rs[src]).astype(np.float32)
    dev = torch.device("cuda:0")
    wt = torch.from_numpy(w).to(dev)
    pt = torch.full((V,), 1.0 / V, dtype=torch.float32, device=dev)
    r = torch.empty_like(pt)
    gr = torch.zeros_like(pt)
    tgt = torch.from_numpy(rng.choice(V, 1000, replace=False)).to(dev)
    dw = torch.empty_like(wt)
    dp = torch.empty_like(pt)
    K = 50
    torch.cuda.synchronize()
    dg.wpr_forward(wt.data_ptr(), pt.data_ptr(), r.data_ptr(), K, 0.85)  # warm
    t0 = time.perf_counter()
    dg.wpr_forward(wt.data_ptr(), pt.data_ptr(), r.data_ptr(), K, 0.85)
    t_fwd = time.perf_counter() - t0
    gr[tgt] = 2 * r[tgt]
    torch.cuda.synchronize()
    dg.wpr_adjoint(wt.data_ptr(), r.data_ptr(), gr.data_ptr(), K, 0.85, dw.data_ptr(),
                   dp.data_ptr())
    t0 = time.perf_counter()
    dg.wpr_adjoint(wt.data_ptr(), r.data_ptr(), gr.data_ptr(), K, 0.85, dw.data_ptr(),
                   dp.data_ptr())
    t_adj = time.perf_counter() - t0
    out = {"graph": f"R-MAT s{args.wpr_scale} EF16, 1 level", "edges": int(E),
           "iterations": K, "forward_s": round(t_fwd, 4), "adjoint_s": round(t_adj, 4),
           "adjoint_over_forward": round(t_adj / t_fwd, 3),
           "forward_alg_gbs": round((12 * E + 16 * V) * K / t_fwd / 1e9, 1)}
    log(f"WPR: {out}")
    dg.close()
    return out


# one PageRank pull iteration: pull (per vertex, per (row, segment) pair, or
# the hot pass of a hot/cold split), the cold pass of a split, split-row
# fix-ups, and the segment merge when the pull is segmented
PR_ITERATION_KERNELS = ("k_pb_iter", "k_pr_pull", "k_pr_pull_cold", "k_pr_fixup", "k_pr_merge")


# committed rocprofv3 FETCH_SIZE / WRITE_SIZE summaries, one per iteration kernel
TRAFFIC_PROFILES = {
    "k_pb_iter": ROOT / "profiles" / "k_pb_iter_traffic_rmat_s24_l11.json",
    "k_pr_pull": ROOT / "profiles" / "k_pr_pull_traffic_rmat_s24_l11.json",
}


def measured_traffic(args, V: int, E: int, used: list) -> dict | None:
    """HBM bytes per dispatch of the PageRank iteration kernel from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes (each scaled by what the
    counter reads for a 1 GiB device copy in the same run), when this run is
    the workload and PageRank variant they were measured on; else None."""
    if used == ["k_pb_iter"]:
        kernel, variant = "k_pb_iter", args.bin_bits in (-1, 13)
    elif used == ["k_pr_pull", "k_pr_fixup"]:
        kernel = "k_pr_pull"
        variant = (args.reorder == 1 and args.seg_bits in (-1, 0) and args.hot_bits in (-1, 0)
                   and args.item_sort in (-1, 0) and args.cap in (0, 2048) and args.bin_bits == 0)
    else:
        return None
    path = TRAFFIC_PROFILES[kernel]
    try:
        prof = json.loads(path.read_text())
    except (OSError, ValueError):
        return None
    w = prof["workload"]
    if not (variant and w["num_nodes"] == V and w["live_edges"] == E and w["seed"] == args.seed):
        return None
    return {"bytes": int(prof["traffic_bytes_per_dispatch"]),
            "source": f"profiles/{path.name}: FETCH_SIZE x{prof['fetch_scale']:.2f} + "
                      f"WRITE_SIZE x{prof['write_scale']:.2f} KiB per {kernel} dispatch"}


def iteration_ms(ctx, iterations):
    """Device time of one PageRank iteration: the iteration kernels' recorded
    time (HIP events on the library stream) over `iterations`."""
    return sum(ctx.kernel_time(k)[1] for k in PR_ITERATION_KERNELS) / max(iterations, 1)


def bound_env() -> dict:
    """Environment of the CPU baseline processes: OpenMP threads bound close,
    one per core, as the reference's CPU runs are measured (unless the caller
    set these variables)."""
    env = dict(os.environ)
    env.setdefault("OMP_PROC_BIND", "close")
    env.setdefault("OMP_PLACES", "cores")
    return env


def cpu_child(argv: list) -> dict:
    """Run oracle/cpu_baseline.py in a process of its own (OpenMP bound, no GPU
    runtime or torch loaded there); its progress lines pass through to stderr,
    its JSON line is returned."""
    import subprocess
    r = subprocess.run([sys.executable, str(ROOT / "oracle" / "cpu_baseline.py"), *argv],
                       stdout=subprocess.PIPE, text=True, timeout=900, env=bound_env())
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline process failed (exit {r.returncode})")
    return json.loads(r.stdout.strip().splitlines()[-1])


def reference_calibration(ctx, args) -> dict:
    """The reference's own CPU-runnable configuration: plain PageRank
    (pagerank_double, 10 iterations) on a single-level R-MAT s20 written as
    .xs1, timed with the reference benchmark built in place under
    oracle/_ref (its `Time` line, mean +- c95 of -c 3), beside the oracle
    restatement and this build's GPU PageRank on the same graph."""
    import re
    import subprocess
    import tempfile

    import numpy as np

    import oracle
    from llama_amd.graph import DeviceGraph, Edges, HostGraph
    ref = ROOT / "oracle" / "_ref" / "benchmark-memory"
    if not ref.exists():
        return {"skipped": "oracle/_ref/benchmark-memory is not built"}
    cores = int(os.environ.get("OMP_NUM_THREADS", oracle.num_threads()))
    e = Edges.rmat(20, 16.0, seed=args.seed, perm_seed=args.seed + 1)
    out = {"graph": "R-MAT s20 EF16, 1 level, no deletions", "edges": len(e), "iterations": 10,
           "cores": cores}
    with tempfile.TemporaryDirectory() as tmp:
        f = str(Path(tmp) / "s20.xs1")
        e.write_xs1(f, 1 << 20)
        env = dict(bound_env(), OMP_NUM_THREADS=str(cores))
        r = subprocess.run([str(ref), "-r", "pagerank_double", "-n", "10", "-c", "3", f],
                           capture_output=True, text=True, timeout=600, cwd=tmp, env=env)
        m = re.search(r"^Time\s*:\s*([0-9.]+) \+- ([0-9.]+) seconds", r.stdout, re.M)
        if r.returncode != 0 or not m:
            return {"skipped": "reference run failed: " + (r.stderr or r.stdout)[-300:]}
        ref_s = float(m.group(1))
        out["reference"] = {"seconds": ref_s, "c95": float(m.group(2)),
                            "gteps": round(len(e) * 10 / ref_s / 1e9, 3)}
        # the oracle restatement on the same graph, in a bound process of its own
        rfile = str(Path(tmp) / "port_ranks.npy")
        port = cpu_child(["calibration", "--seed", str(args.seed), "--ranks-out", rfile])
        rp = np.load(rfile)
    ts = port["seconds"]
    out["port"] = {"seconds": round(float(np.mean(ts)), 4),
                   "over_reference": round(float(np.mean(ts)) / ref_s, 3), "omp": port["omp"]}
    out["omp"] = {k: env.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}
    g = HostGraph(1 << 20, deletions=False)
    g.add_level_edges(e)
    dg = DeviceGraph(ctx, g)
    dg.pagerank(10)
    ctx.sync()
    ts = []
    for _ in range(3):
        t = time.perf_counter()
        rg, _ = dg.pagerank(10)
        ts.append(time.perf_counter() - t)
    ret = re.search(r"^Return value:\s*([0-9.eE+-]+)", r.stdout, re.M)
    out["gpu"] = {"seconds": round(float(np.mean(ts)), 5),
                  "speedup_over_reference": round(ref_s / float(np.mean(ts)), 1),
                  "l1rel_vs_port": float(np.abs(rg - rp).sum() / np.abs(rp).sum()),
                  "rank_sum": float(rg.sum()),
                  "reference_return_value": float(ret.group(1)) if ret else None}
    dg.close()
    log(f"reference calibration: {out}")
    return out


def bench_large(args) -> int:
    """The large single-GPU configuration: R-MAT s27 EF16 plus 20 levels of
    +0.5% / -0.05%, in- and out-CSR resident in HBM; each level's host freeze
    and upload timed separately from traversal; PageRank 20 iterations, DO-BFS
    from 16 sources and a 64-lane batched BFS.  Prints its own JSON line."""
    import numpy as np

    from llama_amd.graph import DeviceGraph, Edges, GpuContext, HostGraph
    scale, levels = args.large_scale, args.large_levels
    t0 = time.time()
    log(f"large: generating R-MAT s{scale} EF16")
    base = Edges.rmat(scale, 16.0, seed=args.seed, perm_seed=args.seed + 1)
    g = HostGraph(1 << scale, deletions=True)
    g.add_level_edges(base)
    del base
    log(f"large: base level frozen in {time.time() - t0:.1f}s")
    ctx = GpuContext(0)
    t1 = time.time()
    dg = DeviceGraph(ctx, g)
    ctx.sync()
    b0 = dg.upload_stats()[0]
    log(f"large: base level uploaded, {b0 / 1e9:.2f} GB in {time.time() - t1:.2f}s")
    freeze, upload, up_bytes = [], [], []
    e_expect = int(dg.snapshot_info().live_edges)
    for k in range(1, levels):
        t = time.time()
        g.add_random_level(0.005, 0.0005, args.seed + 2 + k)
        freeze.append(time.time() - t)
        (ins, _), (dels, _) = g.last_batch()
        e_expect += int(ins.size) - int(dels.size)
        b = dg.upload_stats()[0]
        t = time.time()
        dg.sync_levels()
        ctx.sync()
        upload.append(time.time() - t)
        up_bytes.append(dg.upload_stats()[0] - b)
        log(f"large: level {k}: freeze {freeze[-1]:.2f}s, upload {up_bytes[-1] / 1e9:.2f} GB in "
            f"{upload[-1]:.3f}s")
    info = dg.snapshot_info()
    E, V = int(info.live_edges), g.num_nodes
    ctx.sync()
    # PageRank: bins of 2^13 (the automatic choice) and 2^14 destinations; a
    # first call materialises the snapshot and builds the layout, the second
    # is timed with the layout in place
    by_bins = {}
    for bb in (-1, 14):
        ctx.set_option("pr_bin_bits", bb)
        dg.drop_snapshots()
        ctx.sync()
        t = time.perf_counter()
        dg.pagerank(20)
        ctx.sync()
        t_first = time.perf_counter() - t
        ctx.profile(True)
        t = time.perf_counter()
        dg.pagerank(20)
        ctx.sync()
        t_run = time.perf_counter() - t
        it_ms = iteration_ms(ctx, 20)
        ctx.profile(False)
        by_bins["auto" if bb < 0 else str(bb)] = {
            "seconds": round(t_run, 4), "gteps": round(E * 20 / t_run / 1e9, 2),
            "first_call_seconds_incl_materialisation_and_layout": round(t_first, 4),
            "iteration_ms": round(it_ms, 3),
            "iteration_alg_gbs": round((8 * E + 16 * V) / (it_ms * 1e-3) / 1e9, 1)}
        log(f"large: PageRank bins {bb}: {by_bins['auto' if bb < 0 else str(bb)]}")
    ctx.set_option("pr_bin_bits", args.bin_bits)
    best = min(by_bins.values(), key=lambda x: x["seconds"])
    t_pr = best["seconds"]
    roots = pick_sources(g, 16, args.seed + 3)
    dg.bfs(int(roots[0]), want_levels=False)
    ctx.sync()
    t = time.perf_counter()
    trav = 0
    for r in roots:
        dg.bfs(int(r), want_levels=False)
        trav += dg.last_bfs_traversed
    ctx.sync()
    t_bfs = time.perf_counter() - t
    log(f"large: BFS x16 {t_bfs * 1e3:.1f} ms ({trav / t_bfs / 1e9:.1f} GTEPS)")
    roots64 = pick_sources(g, 64, args.seed + 4)
    dg.bfs_batch(roots64, lane_width=64, want_levels=False)
    ctx.sync()
    t = time.perf_counter()
    _, reached, sweeps = dg.bfs_batch(roots64, lane_width=64, want_levels=False)
    ctx.sync()
    t_batch = time.perf_counter() - t
    log(f"large: batched BFS x64 {t_batch * 1e3:.1f} ms")
    # checks that would catch 32-bit index overflow past 2^31 edges
    lv_b, _, _ = dg.bfs_batch(roots64[:2], lane_width=64)
    lv0, n0, _ = dg.bfs(int(roots64[0]))
    p, _ = dg.pagerank_push(20)
    r, _ = dg.pagerank(20)
    checks = {"live_edges_equal_base_plus_inserts_minus_deletes": E == e_expect,
              "batched_lane_equals_single_source": bool(np.array_equal(lv_b[0], lv0)),
              "bfs_reached_matches_levels": n0 == int((lv0 != 2**31 - 2).sum()),
              "pagerank_pull_vs_push_l1rel": float(np.abs(r - p).sum() / np.abs(p).sum())}
    log(f"large: checks {checks}")
    tc = None
    if args.large_tc:
        # the PageRank orderings and both CSRs go first: the DAG build's sort
        # needs ~ 4 x 8 B per live edge beside the mirror
        dg.drop_snapshots()
        ctx.sync()
        t = time.perf_counter()
        tri, arcs, work = dg.triangles(stats=True)
        t_tc = time.perf_counter() - t
        t = time.perf_counter()
        tci = dg.tc_i()
        t_tci = time.perf_counter() - t
        tc = {"triangles": tri, "dag_arcs": arcs, "work": work,
              "seconds_incl_dag_build": round(t_tc, 3), "tc_i": tci,
              "tc_i_seconds": round(t_tci, 3)}
        # each DAG arc joins its ends by one or two directed live edges
        checks["tc_le_tc_i_le_2tc"] = tri <= tci <= 2 * tri
        log(f"large: TC {tc}")
    print(json.dumps({
        "metric": "large single-GPU scale (PageRank, DO-BFS, batched BFS; per-level upload)",
        "config": {"graph": f"R-MAT s{scale} EF16, {levels} levels (+0.5%/-0.05% per level), "
                            "deletion vector on", "num_nodes": V, "live_edges": E,
                   "stored_words": int(info.stored_words)},
        "data": "synthetic (seeded R-MAT + uniform incremental levels)",
        "upload": {"base_bytes": b0, "per_level_bytes_mean": int(np.mean(up_bytes)),
                   "per_level_upload_s_mean": round(float(np.mean(upload)), 4),
                   "per_level_freeze_s_mean": round(float(np.mean(freeze)), 3),
                   "upload_gb_per_s": round(sum(up_bytes) / sum(upload) / 1e9, 2)},
        "pagerank": {"iterations": 20, "seconds": round(t_pr, 4),
                     "gteps": round(E * 20 / t_pr / 1e9, 2), "by_bin_bits": by_bins},
        "bfs": {"sources": 16, "seconds": round(t_bfs, 4), "gteps": round(trav / t_bfs / 1e9, 2)},
        "bfs_batch": {"sources": 64, "lanes": 64, "seconds": round(t_batch, 4), "sweeps": sweeps,
                      "reached_total": int(np.asarray(reached).sum())},
        "triangle_counting": tc,
        "checks": checks,
    }), flush=True)
    dg.close()
    ctx.close()
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--levels", type=int, default=11)
    ap.add_argument("--edge-factor", type=float, default=16.0)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--cap", type=int, default=0, help="row-block capacity (0 = default)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="run a plain device copy of this size first (calibrates rocprof "
                         "FETCH_SIZE/WRITE_SIZE against known bytes)")
    ap.add_argument("--sweep-pr", action="store_true",
                    help="time one PageRank iteration for every row-block capacity x "
                         "contribution order x column segmentation (report on stderr) "
                         "before the timed steps")
    ap.add_argument("--sweep-configs", default="",
                    help="with --sweep-pr: comma-separated reorder:seg_bits:cap[:item_sort[:hot_bits]] "
                         "(default: the full grid)")
    ap.add_argument("--bin-bits", type=int, default=-1,
                    help="binned PageRank bins of 2^N destinations (0 = row-block pull, "
                         "-1 = automatic)")
    ap.add_argument("--reorder", type=int, default=1,
                    help="1 = degree-ordered contribution slots in PageRank (default)")
    ap.add_argument("--seg-bits", type=int, default=-1,
                    help="PageRank column segments of 2^N slots (0 = off, -1 = automatic)")
    ap.add_argument("--item-sort", type=int, default=-1,
                    help="PageRank items sorted by slot (1), not (0), automatic (-1)")
    ap.add_argument("--hot-bits", type=int, default=-1,
                    help="PageRank hot/cold split at slot 2^N (0 = off, -1 = automatic)")
    ap.add_argument("--cpu-iters", type=int, default=20,
                    help="PageRank iterations timed for the CPU baseline (0 = skip)")
    ap.add_argument("--cpu-runs", type=int, default=3,
                    help="timed repetitions of the CPU baseline sample (mean +- c95)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--bfs-sources", type=int, default=64,
                    help="single-source DO-BFS runs timed (0 = skip)")
    ap.add_argument("--batch-sources", type=int, default=512,
                    help="multi-source BFS sources (0 = skip)")
    ap.add_argument("--ppr-seeds", type=int, default=128, help="batched PPR seeds (0 = skip)")
    ap.add_argument("--sssp-sources", type=int, default=8,
                    help="weighted SSSP sources on the same graph (0 = skip)")
    ap.add_argument("--sssp-delta-milli", type=int, default=-1,
                    help="weighted SSSP near/far step, thousandths of the mean weight (-1 = auto)")
    ap.add_argument("--sssp-sweep", default="",
                    help="comma list of near/far steps (thousandths[:min queue]) also timed, "
                         "e.g. 0,10,10:0,20:65536")
    ap.add_argument("--bc-sources", type=int, default=8,
                    help="betweenness centrality sources on the same graph (0 = skip)")
    ap.add_argument("--tc-reps", type=int, default=2,
                    help="timed triangle counts after the DAG build (0 = skip)")
    ap.add_argument("--scc-reps", type=int, default=2, help="timed SCC runs (0 = skip)")
    ap.add_argument("--ref-calibration", type=int, default=1,
                    help="time the reference build on its own s20 PageRank configuration "
                         "beside the oracle and the GPU (0 = skip)")
    ap.add_argument("--large-scale", type=int, default=0,
                    help="run only the large single-GPU configuration at this R-MAT scale "
                         "(27 for the configuration; prints its own JSON line)")
    ap.add_argument("--large-tc", type=int, default=1,
                    help="with --large-scale: also count triangles (tc_u, tc_i)")
    ap.add_argument("--large-levels", type=int, default=20,
                    help="levels of the large configuration (base included)")
    ap.add_argument("--cow-scale", type=int, default=24,
                    help="copy-on-write upload leg: base graph scale (0 = skip)")
    ap.add_argument("--cow-levels", type=int, default=10,
                    help="copy-on-write upload leg: window levels")
    ap.add_argument("--wpr-scale", type=int, default=23,
                    help="weighted PageRank + adjoint graph scale, single level (0 = skip)")
    args = ap.parse_args()


    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(multi-GPU sharding is design-only)"}))
        return 2

    if args.large_scale > 0:
        return bench_large(args)

    import numpy as np

    from llama_amd.graph import DeviceGraph, GpuContext, build_multilevel

    t0 = time.time()
    log(f"building R-MAT s{args.scale} x {args.levels} levels on the host")
    g = build_multilevel(args.scale, levels=args.levels, edge_factor=args.edge_factor,
                         ins_frac=0.01, del_frac=0.001, seed=args.seed,
                         weighted=args.sssp_sources > 0)
    log(f"host graph built in {time.time() - t0:.1f}s")

    ctx = GpuContext(0)
    ctx.set_option("pr_bin_bits", args.bin_bits)
    ctx.set_option("pr_reorder", args.reorder)
    ctx.set_option("pr_seg_bits", args.seg_bits)
    ctx.set_option("pr_item_sort", args.item_sort)
    ctx.set_option("pr_hot_bits", args.hot_bits)
    t1 = time.time()
    dg = DeviceGraph(ctx, g)
    ctx.sync()
    up_bytes, _ = dg.upload_stats()
    up_s = time.time() - t1
    log(f"mirror uploaded: {up_bytes / 1e9:.2f} GB in {up_s:.2f}s")
    info = dg.snapshot_info()
    E, V = int(info.live_edges), dg.V
    log(f"snapshot: V={V} E_live={E} stored={info.stored_words} segments={info.segments} "
        f"continuations={info.continuations} materialize={info.materialize_ms:.2f}ms")

    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: 2 x {args.copy_probe_mib} MiB device copy")
        del src, dst

    if args.sweep_pr:
        if args.sweep_configs:
            configs = [tuple(int(x) for x in (c + ":0:0").split(":")[:5])
                       for c in args.sweep_configs.split(",")]
        else:
            configs = [(r, sb, c, so, hb)
                       for r, sb, so, hb in ((0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0),
                                             (1, 19, 0, 0), (1, 0, 0, 18), (1, 0, 0, 19),
                                             (1, 0, 0, 20), (1, 0, 0, 21))
                       for c in (1024, 2048, 4096)]
        for reorder, seg, cap, srt, hb in configs:
            ctx.set_option("pr_reorder", reorder)
            ctx.set_option("pr_seg_bits", seg)
            ctx.set_option("pr_item_sort", srt)
            ctx.set_option("pr_hot_bits", hb)
            if True:
                dg.drop_snapshots()
                ctx.profile(True)
                dg.pagerank(2, 0.85, cap=cap, want_ranks=False)
                ctx.sync()
                build_ms = sum(ctx.kernel_time(k)[1]
                               for k in ("pr_reorder", "pr_segments", "pr_item_sort",
                                         "pr_hotcold"))
                ctx.profile(False)
                ctx.profile(True)
                t0 = time.perf_counter()
                dg.pagerank(args.iters, 0.85, cap=cap, want_ranks=False)
                ctx.sync()
                t = time.perf_counter() - t0
                it_ms = iteration_ms(ctx, args.iters)
                ctx.profile(False)
                log(f"sweep reorder={reorder} seg_bits={seg} cap={cap} item_sort={srt} "
                    f"hot_bits={hb}: "
                    f"iteration {it_ms:.4f} ms "
                    f"({(8 * E + 16 * V) / it_ms / 1e6:.1f} GB/s), run {t * 1e3:.2f} ms, "
                    f"layout build {build_ms:.2f} ms")
        ctx.set_option("pr_reorder", args.reorder)
        ctx.set_option("pr_seg_bits", args.seg_bits)
        ctx.set_option("pr_item_sort", args.item_sort)
        ctx.set_option("pr_hot_bits", args.hot_bits)

    def step():
        dg.drop_snapshots()
        dg.pagerank(args.iters, 0.85, cap=args.cap, want_ranks=False)

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")
    ctx.sync()
    ctx.profile(True)
    ts = time.perf_counter()
    for i in range(args.steps):
        step()
    ctx.sync()
    te = time.perf_counter()
    n_pr, ms_pr = ctx.kernel_time("k_pb_iter")
    if n_pr == 0:
        n_pr, ms_pr = ctx.kernel_time("k_pr_pull")
    it_kernels = {k: ctx.kernel_time(k) for k in PR_ITERATION_KERNELS}
    mat = {k: ctx.kernel_time(k) for k in ("k_seg_count", "k_seg_fill", "k_compact",
                                           "k_degrees", "pr_bins_build", "k_pb_init",
                                           "pr_reorder", "pr_segments", "pr_item_sort",
                                           "pr_hotcold")}
    avg_it_ms = iteration_ms(ctx, args.iters * args.steps)
    ctx.profile(False)
    log(f"timed {args.steps} steps done")
    step_ms = (te - ts) * 1e3 / args.steps
    gteps = E * args.iters / (step_ms * 1e-3) / 1e9
    b_pr = 8 * E + 16 * V  # algorithmic bytes of one pull iteration
    achieved = b_pr / (avg_it_ms * 1e-3) / 1e9
    used = [k for k, v in it_kernels.items() if v[0] > 0]
    traffic = measured_traffic(args, V, E, used)

    # the headline configuration's ranks, for the parity check against FP64
    r_head, _ = dg.pagerank(args.iters, 0.85, cap=args.cap)

    copy_bw = measured_copy_gbs()
    log(f"device copy: {copy_bw['gbs']} GB/s (read + write)")

    secondary = {}
    if args.bfs_sources > 0:
        secondary["bfs"] = bench_bfs(dg, g, E, V, args)
    if args.batch_sources > 0:
        secondary["bfs_batch"] = bench_bfs_batch(dg, g, V, args, secondary.get("bfs"))
    if args.sssp_sources > 0:
        secondary["sssp_weighted"] = bench_sssp(dg, g, E, V, args)
    if args.bc_sources > 0:
        secondary["bc_random"] = bench_bc(dg, g, E, V, args)
    if args.tc_reps > 0:
        secondary["tc"] = bench_tc(dg, E, V, args)
    if args.scc_reps > 0:
        secondary["scc"] = bench_scc(dg, E, V, args)
    if args.cow_scale > 0:
        secondary["cow_upload"] = bench_cow_upload(ctx, args)
    if args.ppr_seeds > 0:
        secondary["ppr_batch"] = bench_ppr(dg, g, E, V, args)

    if args.wpr_scale > 0:
        dg.close()
        secondary["wpr_adjoint"] = bench_wpr(ctx, args)

    cpu = None
    parity = None
    if args.cpu_iters > 0:
        import tempfile

        import oracle
        log(f"CPU baseline: oracle pagerank_double, {args.cpu_iters} iterations, "
            f"{oracle.num_threads()} threads, in a process of its own")
        with tempfile.TemporaryDirectory() as tmp:
            rfile = str(Path(tmp) / "gpu_ranks.npy")
            np.save(rfile, r_head)
            argv = ["pagerank", "--scale", str(args.scale), "--levels", str(args.levels),
                    "--edge-factor", str(args.edge_factor), "--seed", str(args.seed),
                    "--iters", str(args.cpu_iters), "--runs", str(args.cpu_runs),
                    "--ranks", rfile]
            if args.sssp_sources > 0:
                argv.append("--weighted")  # the same generator calls as the graph above
            child = cpu_child(argv)
        if child["live_edges"] != E or child["num_nodes"] != V:
            raise RuntimeError(f"CPU baseline graph differs: {child['num_nodes']} vertices, "
                               f"{child['live_edges']} live edges")
        runs = child["runs"]
        if args.cpu_iters == args.iters:
            # the headline's ranks against the FP64 restatement on the same graph
            parity = {"l1_rel_vs_fp64": child["l1_rel_vs_fp64"],
                      "bar": 1e-6,
                      "against": f"oracle pagerank_double (FP64), {args.iters} iterations, "
                                 f"same graph"}
            log(f"headline PageRank L1 relative error vs FP64: {parity['l1_rel_vs_fp64']:.3e}")
        cpu_s = sum(runs) / len(runs)
        # the reference's ll_c95: 1.96 * population stdev / sqrt(n)
        # (llama/include/llama/ll_utils.h:200-221)
        sd = (sum((x - cpu_s) ** 2 for x in runs) / len(runs)) ** 0.5
        cpu = {"value": E * args.cpu_iters / cpu_s / 1e9, "unit": "GTEPS",
               "cores": child["threads"], "kind": "port",
               "sample": f"{args.cpu_iters} pull iterations of pagerank_double (FP64, "
                         f"OpenMP dynamic,4096) over the same {args.levels}-level layout, "
                         f"mean {cpu_s:.2f}s +- {1.96 * sd / len(runs) ** 0.5:.2f}s (c95) "
                         f"over {len(runs)} runs, in a process of its own",
               "seconds_runs": [round(x, 3) for x in runs],
               "cpu_model": cpu_model(),
               "host": host_cpu_layout(),
               "omp": child["omp"]}
        log(f"CPU baseline {cpu['value']:.3f} GTEPS, mean of {len(runs)} runs of {cpu_s:.1f}s")
        if args.ref_calibration:
            cal = reference_calibration(ctx, args)
            cpu["reference_calibration"] = cal
            if "port" in cal:
                # the reference build ran the same algorithm port/ref times faster
                # on the calibration graph: scale the port's rate by that ratio
                q = cal["port"]["over_reference"]
                cpu["reference_equivalent"] = {
                    "value": round(cpu["value"] * q, 4), "unit": "GTEPS",
                    "how": f"port rate x {q} (port time / reference time on the "
                           f"reference's own s20 case, same box, same threads)"}

    out = {
        "metric": METRIC,
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded R-MAT + uniform incremental levels)",
        "config": {
            "workload": f"pagerank {args.iters} iterations, d=0.85, from the HBM mirror "
                        f"(snapshot materialisation included)",
            "graph": f"R-MAT s{args.scale} EF{args.edge_factor:g}, {args.levels} levels "
                     f"(+1%/-0.1% per level), deletion vector on",
            "num_nodes": V, "live_edges": E, "stored_words": int(info.stored_words),
            "continuations": int(info.continuations), "seed": args.seed,
        },
        "roofline": {
            "bound": "hbm", "kernel": "+".join(used),
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic["bytes"] if traffic else None,
            "traffic_source": traffic["source"] if traffic else None,
            "bytes_per_launch": b_pr, "avg_launch_ms": round(avg_it_ms, 5),
            "launches": n_pr,
            "per_kernel_avg_ms": {k: round(v[1] / max(n_pr, 1), 5)
                                  for k, v in it_kernels.items()},
            "measured_copy": copy_bw,
        },
        "cpu_baseline": cpu,
        "parity": parity,
        "breakdown_ms_per_step": {
            **{k: round(v[1] / args.steps, 4) for k, v in it_kernels.items()},
            **{k: round(v[1] / args.steps, 4) for k, v in mat.items()},
        },
        "upload": {"bytes": up_bytes, "seconds": round(up_s, 3)},
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)
    dg.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
