"""Multi-level graphs with known insert/delete batches (shared by the
LL_MIN_LEVEL window tests)."""
import numpy as np

from llama_amd.graph import HostGraph


def batched_graph(n, levels, seed, deletions=True):
    """Random levels with reinserts of deleted edges; returns the host graph
    and, per level, (inserted edges, deleted edges)."""
    rng = np.random.default_rng(seed)
    g = HostGraph(n, deletions=deletions)
    live, batches = set(), []
    for k in range(levels):
        ins = set()
        while len(ins) < (6 * n if k == 0 else n):
            e = (int(rng.integers(0, n)), int(rng.integers(0, n)))
            if e[0] != e[1] and e not in live:
                ins.add(e)
        dels = sorted(e for e in live if rng.random() < 0.15) if k and deletions else []
        ins = sorted(ins)
        g.add_level([e[0] for e in ins], [e[1] for e in ins],
                    [e[0] for e in dels], [e[1] for e in dels])
        live = (live - set(dels)) | set(ins)
        batches.append((ins, dels))
    return g, batches


def windowed_edges(batches, m, L):
    """Edges inserted at a level in [m, L] and still live at L."""
    born = {}
    for k in range(L + 1):
        ins, dels = batches[k]
        for e in dels:
            born.pop(e, None)
        for e in ins:
            born[e] = k
    return {e for e, k in born.items() if k >= m}
