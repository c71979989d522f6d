"""Write golden vectors from the reference LLAMA implementation.

Runs oracle/_ref/ref_driver (the reference headers compiled in place by
oracle/Makefile; see oracle/ref_driver.cc) on small multi-level graphs whose
levels are frozen through the reference's own writable path, and stores, per
fixture, the input batches, the reference's per-level vertex and edge tables
of both directions, and the reference tasks' printed results
(pagerank_double, sssp_unweighted, bfs, tarjan_scc, tc_u).  The graphs stay
at <= 50 vertices so print_results (benchmark/benchmarks/benchmark.h:566-647)
prints every vertex.

Run from the repository root in a container that has the reference tree:
    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

OUT = Path(__file__).resolve().parent
DRIVER = ROOT / "oracle" / "_ref" / "ref_driver"


def rmat_batches(scale, ef, levels, ins_frac, del_frac, seed):
    """Base R-MAT level plus random levels, from the build's own generator."""
    from llama_amd.graph import HostGraph, Edges
    base = Edges.rmat(scale, ef, seed=seed, perm_seed=seed + 1)
    s, d = base.arrays()
    g = HostGraph(1 << scale, deletions=True)
    g.add_level_edges(base)
    out = [{"ins": np.stack([s, d], 1).tolist(), "del": []}]
    for k in range(1, levels):
        g.add_random_level(ins_frac, del_frac, seed + 2 + k)
        (is_, id_), (ds, dd) = g.last_batch()
        out.append({"ins": np.stack([is_, id_], 1).tolist(),
                    "del": np.stack([ds, dd], 1).tolist()})
    return 1 << scale, out


def random_batches(n, m0, levels, ins, dels, seed):
    """Arbitrary insertion order (not sorted), reciprocal pairs, hubs."""
    rng = np.random.default_rng(seed)
    live = set()
    out = []
    for k in range(levels):
        cnt = m0 if k == 0 else ins
        new = []
        while len(new) < cnt:
            u = int(rng.integers(0, n))
            v = int(rng.integers(0, n)) if rng.random() < 0.7 else int(rng.integers(0, 4))
            if u != v and (u, v) not in live and (u, v) not in new:
                new.append((u, v))
        dl = []
        if k > 0 and dels:
            pool = sorted(live)
            idx = rng.choice(len(pool), min(dels, len(pool)), replace=False)
            dl = [pool[i] for i in sorted(idx)]
        live -= set(dl)
        live |= set(new)
        out.append({"ins": [list(p) for p in new], "del": [list(p) for p in dl]})
    return n, out


def cover_deletions(V, batches, seed):
    """The reference's level freeze sizes the new edge table for vertices with
    new edges only (values_length, ll_mlcsr_sp.h:2450-2469) but init_node also
    writes a continuation record for vertices that only lose edges
    (:1472-1497), overrunning the allocation.  Give every vertex that loses an
    out- or in-edge in a level a new edge in that level too, so the
    reference's writes stay inside its buffer."""
    rng = np.random.default_rng(seed)
    live = set()
    out = []
    for b in batches:
        ins = [tuple(p) for p in b["ins"]]
        dels = [tuple(p) for p in b["del"]]
        after = (live - set(dels)) | set(ins)
        srcs = {u for u, _ in ins}
        dsts = {v for _, v in ins}
        for u, v in dels:
            for need_src in (True, False):
                if (u in srcs) if need_src else (v in dsts):
                    continue
                for _ in range(1000):
                    x = int(rng.integers(0, V))
                    e = (u, x) if need_src else (x, v)
                    if e[0] != e[1] and e not in after and e not in live:
                        ins.append(e)
                        after.add(e)
                        srcs.add(e[0])
                        dsts.add(e[1])
                        break
        live = after
        out.append({"ins": [list(p) for p in ins], "del": [list(p) for p in dels]})
    return out


def symmetrised_single_level(V, batches):
    live = set()
    for b in batches:
        live -= {tuple(p) for p in b["del"]}
        live |= {tuple(p) for p in b["ins"]}
    und = sorted({(u, v) for (a, b) in live for (u, v) in ((a, b), (b, a)) if a != b})
    return V, [{"ins": [list(p) for p in und], "del": []}]


def run_driver(V, batches, tasks):
    lines = [f"{V} {len(batches)}"]
    for b in batches:
        lines.append(f"I {len(b['ins'])}")
        lines += [f"{u} {v}" for u, v in b["ins"]]
        lines.append(f"D {len(b['del'])}")
        lines += [f"{u} {v}" for u, v in b["del"]]
    lines += tasks
    r = subprocess.run([str(DRIVER)], input="\n".join(lines) + "\n", capture_output=True,
                       text=True, check=True, env={"OMP_NUM_THREADS": "1"})
    return parse(r.stdout)


def parse(text):
    it = iter(text.splitlines())
    res = {"vt": {}, "et": {}, "tasks": {}}
    cur_task = None
    for line in it:
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "LEVELS":
            res["levels"] = int(tok[1])
        elif tok[0] == "VT":
            d, lv, n = int(tok[1]), int(tok[2]), int(tok[3])
            res["vt"][f"{d}:{lv}"] = [[int(x) for x in next(it).split()] for _ in range(n)]
        elif tok[0] == "ET":
            d, lv, n = int(tok[1]), int(tok[2]), int(tok[3])
            res["et"][f"{d}:{lv}"] = [int(next(it)) for _ in range(n)]
        elif tok[0] == "TASK":
            cur_task = f"{tok[1]}:{tok[2]}"
            res["tasks"][cur_task] = {"values": []}
        elif tok[0] == "RETURN":
            res["tasks"][cur_task]["return"] = [float(tok[1]), float(tok[2])]
            cur_task = None
        elif cur_task is not None and tok[0].endswith(":"):
            res["tasks"][cur_task]["values"] += [float(x) for x in tok[1:]]
        elif cur_task is not None:
            res["tasks"][cur_task]["text"] = line
    return res


def main():
    if not DRIVER.exists():
        sys.exit("build oracle/_ref first: make -C oracle")
    fixtures = {
        "hand_2level": (5, [
            {"ins": [[0, 1], [0, 2], [1, 2], [2, 1], [3, 0]], "del": []},
            {"ins": [[1, 3], [4, 0]], "del": [[0, 2]]}]),
        "rmat_s5_4level": rmat_batches(5, 4, 4, 0.15, 0.08, 11),
        "rmat_s5_single": rmat_batches(5, 8, 1, 0, 0, 3),
        "random_48_5level": random_batches(48, 150, 5, 25, 12, 7),
    }
    for name, (V, batches) in list(fixtures.items()):
        if name != "hand_2level":
            fixtures[name] = (V, cover_deletions(V, batches, 99))
    for name, (V, batches) in list(fixtures.items()):
        fixtures[name + "_sym"] = symmetrised_single_level(V, batches)
    for name, (V, batches) in fixtures.items():
        roots = [0, 1, V // 2, V - 1]
        tasks = ["pagerank_double 10", "pagerank_double 1"]
        tasks += [f"sssp_unweighted {r}" for r in roots] + [f"bfs {r}" for r in roots]
        tasks += ["tarjan_scc"]
        if name.endswith("_sym") or len(batches) == 1:
            tasks += ["tc_u"]
        res = run_driver(V, batches, tasks)
        doc = {
            "generator": "tests/golden/make_golden.py via oracle/_ref/ref_driver "
                         "(reference LLAMA, LL_MEMORY_ONLY + LL_DELETIONS)",
            "num_nodes": V, "batches": batches, "reference": res,
        }
        (OUT / f"{name}.json").write_text(json.dumps(doc, separators=(",", ":")))
        print(name, V, "levels", res["levels"], "tasks", len(res["tasks"]))


if __name__ == "__main__":
    main()
