"""Write golden vectors for task tc_i (ll_b_triangle_counting_LI,
benchmark/benchmarks/triangle_counting.h:154-336) from the reference build.

Runs oracle/_ref/benchmark-memory (compiled in place by oracle/Makefile) as
`-r tc_i` on one X-Stream type-1 file -- the task needs exactly one level,
sorted adjacency and reverse edges (it is registered with
needs_reverse_edges) -- whose records are sorted by (tail, head) without
repeats; a share of the edges have their reverse edge too, which tc_i counts
once per direction.  Keeps the task's return value (the triangle count).

Run from the repository root in a container that has the reference tree:
    python tests/golden/make_golden_tci.py
"""
from __future__ import annotations

import json
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

OUT = Path(__file__).resolve().parent
BENCH = ROOT / "oracle" / "_ref" / "benchmark-memory"


def directed(n, m, recip, seed, hubs=0):
    rng = np.random.default_rng(seed)
    seen = set()
    while len(seen) < m:
        u = int(rng.integers(0, n))
        v = int(rng.integers(0, hubs)) if hubs and rng.random() < 0.3 else int(rng.integers(0, n))
        if u == v:
            continue
        seen.add((u, v))
        if rng.random() < recip:
            seen.add((v, u))
    e = sorted(seen)
    return {"src": [p[0] for p in e], "dst": [p[1] for p in e]}


def run(n, e, tmp):
    from llama_amd.graph import Edges
    f = str(Path(tmp) / "g.xs1")
    Edges.from_arrays(e["src"], e["dst"], np.zeros(len(e["src"]), np.float32)).write_xs1(f, n)
    r = subprocess.run([str(BENCH), "-r", "tc_i", f], capture_output=True, text=True,
                       timeout=120, cwd=tmp)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    ret = [l for l in r.stdout.splitlines() if l.startswith("Return value:")]
    return {"return": float(ret[0].split(":")[1])}


def main() -> None:
    if not BENCH.exists():
        sys.exit("oracle/_ref/benchmark-memory is not built (make -C oracle)")
    fixtures = {
        "tci_random_60": (60, directed(60, 420, 0.3, seed=51)),
        "tci_dense_30": (30, directed(30, 300, 0.5, seed=52)),
        "tci_hubs_80": (80, directed(80, 500, 0.2, seed=53, hubs=5)),
        "tci_oneway_50": (50, directed(50, 350, 0.0, seed=54)),
    }
    for name, (n, e) in fixtures.items():
        with tempfile.TemporaryDirectory() as tmp:
            doc = {"num_nodes": n, "edges": e, "reference": run(n, e, tmp)}
        (OUT / f"{name}.json").write_text(json.dumps(doc))
        print("wrote", name, doc["reference"])


if __name__ == "__main__":
    main()
