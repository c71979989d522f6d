"""Graphs whose later levels touch few vertex-table pages (shared by the
copy-on-write page tests)."""
from llama_amd.graph import HostGraph, build_multilevel


def local_multilevel(scale: int, uniform_levels: int, window_levels: int, seed: int,
                     deletions: bool = True) -> HostGraph:
    """R-MAT base, some uniform random levels, then levels confined to
    windows of V/64 vertex IDs (inserts and deletions with both ends inside)."""
    g = build_multilevel(scale, levels=1 + uniform_levels, edge_factor=8, ins_frac=0.02,
                         del_frac=0.005 if deletions else 0.0, seed=seed, deletions=deletions)
    V = g.num_nodes
    w = max(V // 64, 64)
    for k in range(window_levels):
        lo = (k * 7 * w) % (V - w)
        g.add_window_level(4 * w, w // 4 if deletions else 0, lo, lo + w, seed + 100 + k)
    return g
