"""Write golden vectors for the single-level CSR configuration (LL_SLCSR).

Runs oracle/_ref/benchmark-slcsr (the reference benchmark compiled in place
with -DLL_SLCSR by oracle/Makefile) on one X-Stream type-1 file whose records
are sorted by (tail, head) without repeats: `-r pagerank_double -n 10` (the
printed FP64 ranks, 10 decimals), `-r sssp_unweighted -R <root>` (BFS levels)
and `-r bfs_count -R <root>` (the reached count, its return value; it prints
no results); every vertex is printed (<= 50).

Run from the repository root in a container that has the reference tree:
    python tests/golden/make_golden_slcsr.py
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
BENCH = ROOT / "oracle" / "_ref" / "benchmark-slcsr"


def edges_of(n, m, seed, hubs=0):
    rng = np.random.default_rng(seed)
    seen = set()
    while len(seen) < m:
        u = int(rng.integers(0, n))
        v = int(rng.integers(0, hubs)) if hubs and rng.random() < 0.3 else int(rng.integers(0, n))
        if u != v:
            seen.add((u, v))
    e = sorted(seen)
    return {"src": [p[0] for p in e], "dst": [p[1] for p in e]}


def run(args, n, f, tmp, printed=True):
    out = str(Path(tmp) / "out.txt")
    r = subprocess.run([str(BENCH)] + args + ["-o", out, f], capture_output=True, text=True,
                       timeout=120, cwd=tmp)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    ret = [l for l in r.stdout.splitlines() if l.startswith("Return value:")]
    values = []
    for line in open(out):
        if ":" in line:
            values += [float(x) for x in line.split(":", 1)[1].split()]
    if printed and len(values) != n:
        raise RuntimeError(f"expected {n} printed values, got {len(values)}")
    return {"values": values, "return": float(ret[0].split(":")[1])}


def main() -> None:
    if not BENCH.exists():
        sys.exit("oracle/_ref/benchmark-slcsr is not built (make -C oracle)")
    from llama_amd.graph import Edges
    fixtures = {
        "slcsr_random_50": (50, edges_of(50, 260, seed=41), [0, 9]),
        "slcsr_hubs_48": (48, edges_of(48, 200, seed=42, hubs=4), [1, 30]),
    }
    for name, (n, e, roots) in fixtures.items():
        doc = {"num_nodes": n, "edges": e, "reference": {}}
        with tempfile.TemporaryDirectory() as tmp:
            f = str(Path(tmp) / "g.xs1")
            Edges.from_arrays(e["src"], e["dst"], np.zeros(len(e["src"]), np.float32)).write_xs1(f, n)
            doc["reference"]["pagerank_double"] = run(["-r", "pagerank_double", "-n", "10"], n, f,
                                                      tmp)
            for r in roots:
                doc["reference"][f"sssp_unweighted_{r}"] = run(
                    ["-r", "sssp_unweighted", "-R", str(r)], n, f, tmp)
                doc["reference"][f"bfs_{r}"] = run(["-r", "bfs_count", "-R", str(r)], n, f, tmp,
                                                   printed=False)
        (OUT / f"{name}.json").write_text(json.dumps(doc))
        print("wrote", name)


if __name__ == "__main__":
    main()
