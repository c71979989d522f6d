"""The per-freeze tables of the flat vertex-table golden fixtures
(tests/golden/{flat,onevt}_*.json, made by tests/golden/make_golden_flat.py):
level views and the deletion patches a push at each freeze would send."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
FLAT = sorted(p.stem for p in GOLDEN.glob("flat_*.json")) + \
    sorted(p.stem for p in GOLDEN.glob("onevt_*.json"))
VT_DTYPE = np.dtype([("start", "<i8"), ("len", "<u4"), ("degree", "<u4")])
BITS_INDEX = 38
LEVEL_MASK = (1 << 26) - 1


def load(name: str) -> dict:
    return json.loads((GOLDEN / f"{name}.json").read_text())


def maxlev(w: int) -> int:
    return (w >> BITS_INDEX) & LEVEL_MASK


def freezes(doc: dict):
    """Per freeze k and direction d: (vt [V+1] records, et words of level k,
    patch edge IDs, patch max levels) -- the patches are the older words the
    freeze of k rewrote (max level now k, before not)."""
    fr = doc["reference"]["freeze"]
    out = []
    for k, f in enumerate(fr):
        per_dir = []
        for d in (0, 1):
            vt = np.array([tuple(r) for r in f["vt"][f"{d}:{k}"]], dtype=VT_DTYPE)
            et = np.array(f["et"][f"{d}:{k}"], dtype=np.uint64)
            ids, ml = [], []
            for j in range(k):
                now = f["et"][f"{d}:{j}"]
                before = fr[k - 1]["et"][f"{d}:{j}"]
                for i, (a, b) in enumerate(zip(now, before)):
                    if a != b:
                        assert maxlev(a) == k, "a freeze lowers older words to its own level"
                        ids.append((j << BITS_INDEX) | i)
                        ml.append(k)
            per_dir.append((vt, et, np.array(ids, dtype=np.uint64), np.array(ml, dtype=np.uint32)))
        out.append(per_dir)
    return out
