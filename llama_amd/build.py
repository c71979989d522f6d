"""Build the native libraries in-tree.

  llama_amd/libll_host.so   host level builder, generator, .xs1 I/O (g++ -fopenmp)
  llama_amd/libll_gpu.so    HIP kernels + the ll_gpu_* C ABI (hipcc, gfx950)

Run ``python -m llama_amd.build`` (or ``__graft_entry__.build()``).  hipcc
cross-compiles for gfx950 without a GPU present.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"

ARCH = os.environ.get("LLAMA_AMD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # packed FP32 VALU forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fopenmp", "-fPIC", "-Wall", "-Wextra", "-DNDEBUG",
              f"-I{INCLUDE}"]

HIP_SOURCES = ["ll_gpu.hip", "prim.hip", "snapshot.hip", "rowblocks.hip", "pagerank.hip",
               "pagerank_push.hip", "pr_segments.hip", "pr_bins.hip",
               "bfs.hip", "bfs_batch.hip", "tc.hip", "scc.hip", "ppr.hip", "wpr.hip",
               "sssp.hip", "bc.hip", "noalloc.hip"]
HOST_SOURCES = ["host/ll_layout.cc", "host/ll_host.cc"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def _extra_hip_sources() -> list[str]:
    return [s for s in HIP_SOURCES if (CSRC / s).exists()]


def build_host() -> Path:
    BUILD.mkdir(exist_ok=True)
    out = PKG / "libll_host.so"
    tmp = BUILD / "libll_host.so.tmp"
    _run(["g++", *HOST_FLAGS, "-shared", "-o", str(tmp), *[str(CSRC / s) for s in HOST_SOURCES]])
    os.replace(tmp, out)
    return out


def build_gpu(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = _extra_hip_sources()
    objs = [BUILD / (Path(s).stem + ".o") for s in srcs]

    def one(i: int) -> None:
        _run([HIPCC, *HIP_FLAGS, "-c", str(CSRC / srcs[i]), "-o", str(objs[i])])

    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(srcs)))) as ex:
        list(ex.map(one, range(len(srcs))))
    out = PKG / "libll_gpu.so"
    tmp = BUILD / "libll_gpu.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *[str(o) for o in objs]])
    os.replace(tmp, out)
    return out


def build_cli() -> Path:
    """ll_gpu_bench: the C++ task driver (links both libraries, rpath $ORIGIN)."""
    out = PKG / "ll_gpu_bench"
    tmp = BUILD / "ll_gpu_bench.tmp"
    _run(["g++", *HOST_FLAGS, f"-I{CSRC / 'host'}", str(CSRC / "host" / "ll_gpu_bench.cc"),
          "-o", str(tmp), f"-L{PKG}", "-lll_host", "-lll_gpu", "-Wl,-rpath,$ORIGIN",
          "-Wl,--allow-shlib-undefined"])
    os.replace(tmp, out)
    return out


def build_all(jobs: int = 8) -> None:
    build_host()
    build_gpu(jobs)
    build_cli()


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build_all()
    print("built", PKG / "libll_host.so", PKG / "libll_gpu.so")
