"""Helpers for the tests of the reference driver built with the GPU tasks
registered (oracle/_ref/benchmark-memory-wd-gpu, see oracle/Makefile)."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
GPU_DRIVER = ROOT / "oracle" / "_ref" / "benchmark-memory-wd-gpu"
CPU_DRIVER = ROOT / "oracle" / "_ref" / "benchmark-memory-wd"


def write_levels(tmp: Path, scale: int, levels: int, seed: int, oneway: bool = False) -> list[str]:
    """A seeded R-MAT base level plus insert-only levels as weighted .xs1
    files (the reference loader makes one level per file).  oneway: keep one
    direction of each reciprocal pair of the base level (the input tc_u's -U
    loading expects: it adds the reverse of every edge itself)."""
    import numpy as np
    from llama_amd.graph import Edges
    base = Edges.rmat(scale, 8.0, seed=seed, perm_seed=seed + 1).fill_weights(seed)
    s, d = base.arrays()
    V = 1 << scale
    if oneway:
        pairs = set(zip(s.tolist(), d.tolist()))
        keep = np.array([not (u > v and (v, u) in pairs) for u, v in zip(s.tolist(), d.tolist())],
                        dtype=bool)
        base = Edges.from_arrays(s[keep], d[keep]).fill_weights(seed)
        s, d = base.arrays()
    rng = np.random.default_rng(seed)
    have = set(zip(s.tolist(), d.tolist()))
    files = [str(tmp / "l0.xs1")]
    base.write_xs1(files[0], V)
    for k in range(1, levels):
        new = []
        while len(new) < max(1, len(s) // 20):
            u, v = int(rng.integers(0, V)), int(rng.integers(0, V))
            if u != v and (u, v) not in have:
                have.add((u, v))
                new.append((u, v))
        new.sort()
        a = np.array(new, dtype=np.uint32)
        w = rng.uniform(0, 1, len(new)).astype(np.float32)
        e = Edges.from_arrays(a[:, 0], a[:, 1], w)
        files.append(str(tmp / f"l{k}.xs1"))
        e.write_xs1(files[-1], V)
    return files


def run(driver: Path, task: str, files: list[str], out: Path, *extra: str) -> tuple[str, str]:
    """Run one task on one thread (the reference's betweenness and PageRank
    tasks add FP32 values with atomics, so their sums depend on the thread
    schedule otherwise); returns (the `Return value` text, the -o results
    text)."""
    import os
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([str(driver), "-r", task, "-o", str(out), *extra, *files],
                       capture_output=True, text=True, timeout=300, env=env)
    if r.returncode != 0:
        raise RuntimeError(f"{driver.name} -r {task} failed ({r.returncode}): "
                           f"{(r.stderr or r.stdout)[-600:]}")
    m = re.search(r"^Return value:\s*(\S+)", r.stdout, re.M)
    return (m.group(1) if m else ""), out.read_text()


def values(text: str) -> list[float]:
    """Numbers of a print_results block (the `%7ld:` row labels dropped)."""
    vals = []
    for ln in text.splitlines():
        if ":" in ln:
            vals += [float(x) for x in ln.split(":", 1)[1].split()]
    return vals
