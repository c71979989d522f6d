"""Write golden vectors for the SNAP .net reader from the reference LLAMA build.

Writes a .net text file with comments, blank lines, CRLF endings, tabs,
leading blanks and malformed records, loads it with the reference's own
loader (oracle/_ref/benchmark-memory; ll_loader_net,
llama/include/llama/loaders/ll_load_net.h) and keeps the printed node/edge
counts, sssp_unweighted levels from two roots and pagerank_double after 5
iterations.  The fixture holds the file's text, so the test reads the same
bytes with the build's reader.  No edge repeats and there are <= 50 vertices.

Run from the repository root in a container that has the reference tree:
    python tests/golden/make_golden_net.py
"""
from __future__ import annotations

import json
import re
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
OUT = Path(__file__).resolve().parent
BENCH = ROOT / "oracle" / "_ref" / "benchmark-memory"


def text(seed=5, n=45, m=160):
    rng = np.random.default_rng(seed)
    seen, lines = set(), ["# Directed graph (each unordered pair of nodes is saved once)",
                          "# FromNodeId\tToNodeId"]
    bad = ["7 x", "4 5 6", "12", "-", "3\t", " # not a comment: leading blank"]
    while len(seen) < m:
        u, v = int(rng.integers(0, n)), int(rng.integers(0, n))
        if (u, v) in seen:
            continue
        seen.add((u, v))
        sep = ["\t", " ", "   ", " \t "][int(rng.integers(0, 4))]
        lead = "  " if rng.random() < 0.1 else ""
        end = "\r" if rng.random() < 0.2 else ""
        lines.append(f"{lead}{u}{sep}{v}{end}")
        r = rng.random()
        if r < 0.05:
            lines.append("")
        elif r < 0.08:
            lines.append("# interleaved comment")
        elif r < 0.11:
            lines.append(bad[int(rng.integers(0, len(bad)))])
    return "\n".join(lines) + "\n"


def run(path, task, extra, tmp):
    out = str(Path(tmp) / "out.txt")
    r = subprocess.run([str(BENCH), "-r", task] + extra + ["-o", out, path],
                       capture_output=True, text=True, timeout=120, cwd=tmp)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    vals = []
    for line in open(out):
        if ":" in line:
            vals += [float(x) for x in line.split(":", 1)[1].split()]
    nodes = int(re.search(r"# Nodes\s*:\s*(\d+)", r.stdout).group(1))
    edges = int(re.search(r"# Edges\s*:\s*(\d+)", r.stdout).group(1))
    return vals, nodes, edges


def main() -> None:
    if not BENCH.exists():
        sys.exit("oracle/_ref/benchmark-memory is not built (make -C oracle ref)")
    t = text()
    doc = {"text": t, "reference": {}}
    with tempfile.TemporaryDirectory() as tmp:
        f = str(Path(tmp) / "g.net")
        Path(f).write_text(t)
        for root in (0, 9):
            v, n, e = run(f, "sssp_unweighted", ["-R", str(root)], tmp)
            doc["reference"][f"sssp_unweighted:{root}"] = v
        v, n, e = run(f, "pagerank_double", ["-n", "5"], tmp)
        doc["reference"]["pagerank_double:5"] = v
        doc["reference"]["nodes"] = n
        doc["reference"]["edges"] = e
    (OUT / "net_tricky.json").write_text(json.dumps(doc))
    print("wrote net_tricky", n, e)


if __name__ == "__main__":
    main()
