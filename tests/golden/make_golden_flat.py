"""Write golden vectors for the flat vertex-table containers of the reference:
LL_FLAT_VT (one flat table per level) and LL_ONE_VT (one flat table shared by
all levels), ll_common.h:311-323 and ll_mem_array.h:554-904.

Runs oracle/_ref/ref_driver_flat and oracle/_ref/ref_driver_onevt (the
reference compiled in place with those flags by oracle/Makefile) with
LL_DUMP_EACH=1 on the input batches of three existing fixtures
(tests/golden/{hand_2level,rmat_s5_4level,random_48_5level}.json): after each
level's freeze it records the new level's vertex tables and every level's edge
tables as they stand then -- what a device push at that freeze reads -- and
the reference tasks' results on the final graph.

Run from the repository root in a container that has the reference tree:
    python tests/golden/make_golden_flat.py
"""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
OUT = Path(__file__).resolve().parent
FIXTURES = ["hand_2level", "rmat_s5_4level", "random_48_5level"]
FLAVOURS = {"flat": "ref_driver_flat", "onevt": "ref_driver_onevt"}


def parse(text: str) -> dict:
    it = iter(text.splitlines())
    res = {"freeze": [], "tasks": {}}
    cur = None
    for line in it:
        tok = line.split()
        if not tok:
            continue
        if tok[0] == "FREEZE":
            res["freeze"].append({"vt": {}, "et": {}})
        elif tok[0] == "LEVELS":
            res["levels"] = int(tok[1])
        elif tok[0] in ("VT", "ET"):
            d, lv, n = int(tok[1]), int(tok[2]), int(tok[3])
            rows = [next(it) for _ in range(n)]
            tab = res["freeze"][-1]["vt" if tok[0] == "VT" else "et"]
            tab[f"{d}:{lv}"] = ([[int(x) for x in r.split()] for r in rows] if tok[0] == "VT"
                                else [int(r) for r in rows])
        elif tok[0] == "TASK":
            cur = f"{tok[1]}:{tok[2]}"
            res["tasks"][cur] = {"values": []}
        elif tok[0] == "RETURN":
            res["tasks"][cur]["return"] = [float(tok[1]), float(tok[2])]
            cur = None
        elif cur is not None and tok[0].endswith(":"):
            res["tasks"][cur]["values"] += [float(x) for x in tok[1:]]
    return res


def main() -> None:
    for flavour, exe in FLAVOURS.items():
        driver = ROOT / "oracle" / "_ref" / exe
        if not driver.exists():
            sys.exit(f"build {driver} first: make -C oracle")
        for name in FIXTURES:
            src = json.loads((OUT / f"{name}.json").read_text())
            V, batches = src["num_nodes"], src["batches"]
            roots = [0, 1, V // 2, V - 1]
            tasks = ["pagerank_double 10"] + [f"sssp_unweighted {r}" for r in roots]
            tasks += [f"bfs {r}" for r in roots]
            lines = [f"{V} {len(batches)}"]
            for b in batches:
                lines.append(f"I {len(b['ins'])}")
                lines += [f"{u} {v}" for u, v in b["ins"]]
                lines.append(f"D {len(b['del'])}")
                lines += [f"{u} {v}" for u, v in b["del"]]
            lines += tasks
            r = subprocess.run([str(driver)], input="\n".join(lines) + "\n", capture_output=True,
                               text=True, check=True,
                               env={"OMP_NUM_THREADS": "1", "LL_DUMP_EACH": "1"})
            doc = {"generator": f"tests/golden/make_golden_flat.py via oracle/_ref/{exe} "
                                f"(reference LLAMA, LL_MEMORY_ONLY + LL_DELETIONS + "
                                f"{'LL_FLAT_VT' if flavour == 'flat' else 'LL_ONE_VT'})",
                   "source_fixture": f"{name}.json", "num_nodes": V,
                   "reference": parse(r.stdout)}
            (OUT / f"{flavour}_{name}.json").write_text(json.dumps(doc, separators=(",", ":")))
            print(flavour, name, "freezes", len(doc["reference"]["freeze"]))


if __name__ == "__main__":
    main()
