"""Write tests/golden/tc_rmat_s24_l11.json: the triangle counts of the
full-size multi-level graph (R-MAT scale 24, edge factor 16, seed 1, plus 10
levels of +1% / -0.1%; the graph tests/test_gpu_xfullsize.py and bench.py
build), computed by the oracle's restatement of tc_u
(benchmark/benchmarks/triangle_counting.h:343-470: per vertex u, the merge of
N+(u) with N+(v) for every neighbour v > u, on the symmetrised simple graph of
live edges) and of tc_i (:154-336) on the CPU.  The oracle's tc_u / tc_i are
pinned against the reference's own outputs on the golden fixtures
(tests/test_golden.py, tests/test_golden_tci.py).

Run from the repository root: python tests/golden/make_golden_tc_s24.py
(about an hour on 8 cores)."""
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import oracle  # noqa: E402
from llama_amd.graph import DIR_OUT, build_multilevel  # noqa: E402

SCALE, LEVELS, SEED = 24, 11, 1


def main() -> None:
    t0 = time.time()
    g = build_multilevel(SCALE, levels=LEVELS, seed=SEED, weighted=True)
    o = oracle.OracleGraph.from_host(g)
    ofs, _ = o.flatten(DIR_OUT)
    live = int(ofs[-1])
    t1 = time.time()
    tc_u = o.triangles()
    t2 = time.time()
    tc_i = o.tc_i()
    t3 = time.time()
    out = {
        "graph": f"R-MAT s{SCALE} EF16 (a,b,c,d = .57,.19,.19,.05), seed {SEED}, plus "
                 f"{LEVELS - 1} levels of +1% / -0.1%, deletion vector on",
        "scale": SCALE, "levels": LEVELS, "seed": SEED,
        "num_nodes": g.num_nodes, "live_edges": live,
        "tc_u": tc_u, "tc_i": tc_i,
        "oracle_threads": oracle.num_threads(),
        "seconds": {"build": round(t1 - t0, 1), "tc_u": round(t2 - t1, 1),
                    "tc_i": round(t3 - t2, 1)},
        "script": "tests/golden/make_golden_tc_s24.py",
    }
    path = ROOT / "tests" / "golden" / "tc_rmat_s24_l11.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    os.environ.setdefault("OMP_PROC_BIND", "close")
    main()
