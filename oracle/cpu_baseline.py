"""CPU baseline legs of bench.py, run in a process of their own.

TEST INFRASTRUCTURE: times the oracle's OpenMP restatement of pagerank_double
(benchmark/benchmarks/pagerank.h:98-182) on the host cores; it is the baseline
the GPU figures are reported beside, never the thing measured or shipped.

bench.py starts this script with the OpenMP threads bound (OMP_PROC_BIND=close,
OMP_PLACES=cores).  No GPU runtime and no torch are loaded here: a runtime that
binds the main thread before this process's own OpenMP pool starts would leave
that pool a single core, so the baseline gets a process to itself.  Progress
goes to stderr, one line per timed run; the result is one JSON line on stdout.

  pagerank     the benchmark graph (same generator, sizes and seed as bench.py),
               `--runs` timings of `--iters` pull iterations (stopped once the
               sample holds `--max-seconds`), and with `--ranks` the L1
               relative error of those FP32 ranks against the FP64 ranks
  calibration  the reference's own CPU case (single-level R-MAT s20, 10
               iterations), 3 timings, FP64 ranks saved to `--ranks-out`
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))


def log(msg: str) -> None:
    print(f"[cpu-baseline] {msg}", file=sys.stderr, flush=True)


def omp_env() -> dict:
    return {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}


def run_pagerank(args) -> dict:
    import numpy as np

    import oracle
    from llama_amd.graph import build_multilevel
    t0 = time.time()
    g = build_multilevel(args.scale, levels=args.levels, edge_factor=args.edge_factor,
                         ins_frac=0.01, del_frac=0.001, seed=args.seed, weighted=args.weighted)
    log(f"graph R-MAT s{args.scale} x {args.levels} levels built in {time.time() - t0:.1f}s")
    o = oracle.OracleGraph.from_host(g)
    # live edges of the last level: the sum of its vertex-table degrees
    vt = g.level_arrays(0, g.num_levels - 1).vt
    E = int(vt["degree"][:g.num_nodes].sum(dtype=np.uint64))
    runs = []
    ranks = None
    for i in range(args.runs):
        t = time.perf_counter()
        ranks, _ = o.pagerank(args.iters)
        runs.append(time.perf_counter() - t)
        log(f"run {i + 1}/{args.runs}: {args.iters} iterations in {runs[-1]:.2f}s")
        if sum(runs) > args.max_seconds:
            log(f"sample bounded at {args.max_seconds:.0f} s, no further runs")
            break
    out = {"runs": [round(x, 4) for x in runs], "iterations": args.iters,
           "threads": oracle.num_threads(), "num_nodes": g.num_nodes, "live_edges": E,
           "omp": omp_env()}
    if args.ranks:
        r = np.load(args.ranks)
        out["l1_rel_vs_fp64"] = float(np.abs(r - ranks).sum() / np.abs(ranks).sum())
    return out


def run_calibration(args) -> dict:
    import numpy as np

    import oracle
    from llama_amd.graph import Edges, HostGraph
    e = Edges.rmat(20, 16.0, seed=args.seed, perm_seed=args.seed + 1)
    g = HostGraph(1 << 20, deletions=False)
    g.add_level_edges(e)
    o = oracle.OracleGraph.from_host(g)
    ts = []
    rp = None
    for i in range(3):
        t = time.perf_counter()
        rp, _ = o.pagerank(10)
        ts.append(time.perf_counter() - t)
        log(f"calibration run {i + 1}/3: {ts[-1]:.3f}s")
    if args.ranks_out:
        np.save(args.ranks_out, rp)
    return {"seconds": [round(x, 4) for x in ts], "edges": len(e), "threads": oracle.num_threads(),
            "omp": omp_env()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=("pagerank", "calibration"))
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--levels", type=int, default=11)
    ap.add_argument("--edge-factor", type=float, default=16.0)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--weighted", action="store_true")
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--runs", type=int, default=3)
    ap.add_argument("--max-seconds", type=float, default=45.0)
    ap.add_argument("--ranks", default="")
    ap.add_argument("--ranks-out", default="")
    args = ap.parse_args()
    out = run_pagerank(args) if args.mode == "pagerank" else run_calibration(args)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
