"""Write the reference benchmark driver with the GPU tasks registered, the
edit a maintainer of the reference makes to run them (TEST INFRASTRUCTURE:
the output goes to oracle/_ref/, and oracle/Makefile compiles it against the
reference tree and libll_gpu.so to prove the drop-in boundary).

Three insertions into benchmark/benchmark.cc, each after or before a line
found by its text: the binding header after the last task header include,
the GPU rows (LL_GPU_REFERENCE_ROWS, llama_amd/csrc/host/ll_gpu_reference.h)
before the registry's NULL row, and their create lines
(LL_GPU_REFERENCE_CREATE) before the end of the task switch.

Usage: python oracle/gpu_binding.py REFERENCE_BENCHMARK_CC OUTPUT_CC
"""
from __future__ import annotations

import re
import sys


def patch(src: str) -> str:
    lines = src.split("\n")
    out: list[str] = []
    in_table = False
    rows = 0
    done = {"include": False, "rows": False, "create": False}
    for ln in lines:
        s = ln.strip()
        if "g_runnable_things[] = {" in s:
            in_table = True
        elif in_table and s.startswith("{ NULL"):
            out.append("\tLL_GPU_REFERENCE_ROWS")
            in_table = False
            done["rows"] = True
        elif in_table and re.match(r'\{\s*"', s):
            rows += 1
        if s == "#undef B" and done["rows"] and not done["create"]:
            out.append("#if B < 0")
            out.append(f"\tLL_GPU_REFERENCE_CREATE(run_task_class, {rows}, pagerank_iters, "
                       "root_node);")
            out.append("#endif")
            done["create"] = True
        out.append(ln)
        if s == '#include "benchmarks/triangle_counting.h"':
            out.append('#include "ll_gpu_reference.h"')
            done["include"] = True
    missing = [k for k, v in done.items() if not v]
    if missing:
        raise SystemExit(f"gpu_binding: anchor not found for {missing}")
    return "\n".join(out)


if __name__ == "__main__":
    src, dst = sys.argv[1], sys.argv[2]
    with open(src) as f:
        text = f.read()
    with open(dst, "w") as f:
        f.write(patch(text))
