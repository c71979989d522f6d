"""The drop-in boundary compiled against the reference: benchmark/benchmark.cc
of the reference with the GPU tasks registered by the three insertions of
oracle/gpu_binding.py (binding header llama_amd/csrc/host/ll_gpu_reference.h),
built by oracle/Makefile against the reference tree and libll_gpu.so.  On the
CPU: the driver lists every GPU task, its CPU tasks print the same results as
the unmodified reference build, and a GPU task without a GPU stops with the
library's error.  (tests/test_gpu_reference_binding.py runs the GPU tasks.)"""
import os
from pathlib import Path

import pytest

from binding_util import CPU_DRIVER, GPU_DRIVER, run, write_levels

REF = Path(os.environ.get("LLAMA_REFERENCE", "/root/reference"))
pytestmark = pytest.mark.skipif(not (REF / "benchmark" / "benchmark.cc").exists(),
                                reason="the reference tree is not in this container")

GPU_TASKS = ["gpu_pagerank", "gpu_pagerank_push", "gpu_sssp_unweighted", "gpu_bfs_count",
             "gpu_sssp_weighted", "gpu_bc_adj", "gpu_bc_random", "gpu_tc", "gpu_tc_i", "gpu_scc"]


def test_binding_patch_finds_its_anchors():
    import subprocess
    import sys
    root = Path(__file__).resolve().parent.parent
    r = subprocess.run([sys.executable, str(root / "oracle" / "gpu_binding.py"),
                        str(REF / "benchmark" / "benchmark.cc"), "/dev/stdout"],
                       capture_output=True, text=True, check=True)
    assert r.stdout.count('#include "ll_gpu_reference.h"') == 1
    assert "LL_GPU_REFERENCE_ROWS" in r.stdout and "LL_GPU_REFERENCE_CREATE" in r.stdout


def test_driver_lists_the_gpu_tasks():
    import subprocess
    assert GPU_DRIVER.exists(), "oracle/_ref/benchmark-memory-wd-gpu is not built"
    r = subprocess.run([str(GPU_DRIVER), "-h"], capture_output=True, text=True)
    for t in GPU_TASKS:
        assert f"  {t} " in r.stdout + r.stderr, t


def test_cpu_tasks_unchanged(tmp_path):
    files = write_levels(tmp_path, 8, 3, seed=5)
    for task, extra in (("pagerank_double", ("-n", "5")), ("sssp_unweighted", ("-R", "3")),
                        ("bc_random", ())):
        a = run(GPU_DRIVER, task, files, tmp_path / "a.txt", *extra)
        b = run(CPU_DRIVER, task, files, tmp_path / "b.txt", *extra)
        assert a == b, task


@pytest.mark.skipif(Path("/dev/kfd").exists(), reason="a GPU is present")
def test_gpu_task_without_gpu_reports_the_library_error(tmp_path):
    import subprocess
    files = write_levels(tmp_path, 6, 1, seed=2)
    r = subprocess.run([str(GPU_DRIVER), "-r", "gpu_pagerank", *files], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0
    assert "ll_gpu_create failed" in r.stderr
