"""Write golden vectors for weighted SSSP from the reference LLAMA build.

Runs oracle/_ref/benchmark-memory (the reference benchmark compiled in place
by oracle/Makefile) as `-r sssp_weighted -R <root> -o <file>` on weighted
X-Stream type-1 files, one file per level: the reference loader freezes the
first file with load_direct and each further file as a new level, and stores
every record's weight in the `weight` edge property (ll_load_utils.h:628-790).
The levels have no repeated edges, so the snapshot is the same graph the
build freezes from the same batches.  Every level's records are sorted by
(tail, head) except in the `unsorted` fixture: for a file that is not sorted,
the reference loader's sorted pass writes `e.weight` -- the last record read
in the first pass -- as the weight of every edge of that level
(ll_load_utils.h:750-755 vs :684-687), so that fixture pins the discrepancy.
Each fixture keeps the batches with their weights and, per root, the printed
distances (10 decimals; a float32 distance is
recovered exactly by rounding the parsed value to float32) and the task's
return value (the largest finite distance as an int, LL_SSSP_RETURNS_MAX).
Graphs stay at <= 50 vertices so print_results prints every vertex.

Run from the repository root in a container that has the reference tree:
    python tests/golden/make_golden_sssp.py
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


def random_levels(n, m0, per_level, levels, wlo, whi, seed, hubs=0, sort_all=True):
    rng = np.random.default_rng(seed)
    seen = set()
    out = []
    for k in range(levels):
        cnt = m0 if k == 0 else per_level
        new = []
        while len(new) < cnt:
            u = int(rng.integers(0, n))
            v = int(rng.integers(0, hubs)) if hubs and rng.random() < 0.3 else int(rng.integers(0, n))
            if u != v and (u, v) not in seen:
                seen.add((u, v))
                new.append((u, v))
        if k == 0 or sort_all:
            new.sort()
        w = rng.uniform(wlo, whi, len(new)).astype(np.float32)
        out.append({"src": [p[0] for p in new], "dst": [p[1] for p in new],
                    "w": [float(x) for x in w]})
    return out


def hand_levels():
    # 0 -> 1 -> 2 -> 3 chain with a shortcut added later, a vertex reachable
    # only through level 1, two vertices never reached (5 has no in-edges from
    # the reachable part, 6 is isolated)
    l0 = {"src": [0, 0, 1, 2, 5], "dst": [1, 3, 2, 3, 4],
          "w": [1.0, 10.0, 2.5, 0.75, 0.5]}
    l1 = {"src": [0, 3], "dst": [2, 7], "w": [0.125, 3.0]}
    return 8, [l0, l1]


def run_reference(n, levels, root, tmp):
    from llama_amd.graph import Edges
    files = []
    for k, lv in enumerate(levels):
        f = str(Path(tmp) / f"level{k}.xs1")
        Edges.from_arrays(lv["src"], lv["dst"], np.array(lv["w"], np.float32)).write_xs1(f, n)
        files.append(f)
    out = str(Path(tmp) / "out.txt")
    r = subprocess.run([str(BENCH), "-r", "sssp_weighted", "-R", str(root), "-o", out] + files,
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
        "sssp_hand_2level": (*hand_levels(), [0, 5, 6]),
        "sssp_random_40_3level": (40, random_levels(40, 120, 10, 3, 0.1, 5.0, seed=11), [0, 7, 19]),
        "sssp_hubs_48_4level": (48, random_levels(48, 150, 12, 4, 0.0, 1.0, seed=12, hubs=4),
                                [1, 2, 30]),
        "sssp_unsorted_40_3level": (40, random_levels(40, 120, 10, 3, 0.1, 5.0, seed=11,
                                                      sort_all=False), [0, 7]),
    }
    for name, (n, levels, roots) in fixtures.items():
        doc = {"num_nodes": n, "levels": levels, "reference": {},
               "levels_sorted": all(sorted(zip(l["src"], l["dst"])) == list(zip(l["src"], l["dst"]))
                                    for l in levels)}
        with tempfile.TemporaryDirectory() as tmp:
            for r in roots:
                doc["reference"][str(r)] = run_reference(n, levels, r, tmp)
        (OUT / f"{name}.json").write_text(json.dumps(doc))
        print("wrote", name)


if __name__ == "__main__":
    main()
