"""Write golden vectors for betweenness centrality from the reference build.

Runs oracle/_ref/benchmark-memory as `-r bc_adj -o <file>` (exact BC,
benchmark/benchmarks/bc_adj.h) on multi-level X-Stream type-1 files, one file
per level, every level sorted by (tail, head) and without repeated edges (the
build freezes the same levels from the same batches), and keeps the printed
scores (10 decimals; each float32 score is recovered exactly by rounding the
parsed value to float32) and the task's return value (the largest score).
Path counts stay far below 2^24, so the reference's unordered atomic sigma
additions are exact and its output is deterministic.  <= 50 vertices, so
print_results prints every vertex.

Run from the repository root in a container that has the reference tree:
    python tests/golden/make_golden_bc.py
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


def levels_of(n, m0, per_level, levels, seed, hubs=0):
    rng = np.random.default_rng(seed)
    seen, out = set(), []
    for k in range(levels):
        cnt = m0 if k == 0 else per_level
        new = []
        while len(new) < cnt:
            u = int(rng.integers(0, n))
            v = int(rng.integers(0, hubs)) if hubs and rng.random() < 0.3 else int(rng.integers(0, n))
            if u != v and (u, v) not in seen:
                seen.add((u, v))
                new.append((u, v))
        new.sort()
        out.append({"src": [p[0] for p in new], "dst": [p[1] for p in new]})
    return out


def run_reference(n, levels, tmp):
    from llama_amd.graph import Edges
    files = []
    for k, lv in enumerate(levels):
        f = str(Path(tmp) / f"level{k}.xs1")
        Edges.from_arrays(lv["src"], lv["dst"], np.zeros(len(lv["src"]), np.float32)).write_xs1(f, n)
        files.append(f)
    out = str(Path(tmp) / "out.txt")
    r = subprocess.run([str(BENCH), "-r", "bc_adj", "-o", out] + files,
                       capture_output=True, text=True, timeout=120, cwd=tmp)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    ret = [l for l in r.stdout.splitlines() if l.startswith("Return value:")]
    values = []
    for line in open(out):
        if ":" in line:
            values += [float(x) for x in line.split(":", 1)[1].split()]
    if len(values) != n:
        raise RuntimeError(f"expected {n} printed values, got {len(values)}")
    return {"values": values, "return": float(ret[0].split(":")[1])}


def main() -> None:
    if not BENCH.exists():
        sys.exit("oracle/_ref/benchmark-memory is not built (make -C oracle ref)")
    fixtures = {
        "bc_random_40_3level": (40, levels_of(40, 110, 12, 3, seed=31)),
        "bc_hubs_48_4level": (48, levels_of(48, 140, 10, 4, seed=32, hubs=3)),
        "bc_sparse_50_2level": (50, levels_of(50, 60, 20, 2, seed=33)),
    }
    for name, (n, levels) in fixtures.items():
        with tempfile.TemporaryDirectory() as tmp:
            doc = {"num_nodes": n, "levels": levels, "reference": run_reference(n, levels, tmp)}
        (OUT / f"{name}.json").write_text(json.dumps(doc))
        print("wrote", name)


if __name__ == "__main__":
    main()
