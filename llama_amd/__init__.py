"""MI355X-native read-only analytics over LLAMA's multi-level snapshot CSR.

Native pieces (built in-tree by ``python -m llama_amd.build``):
  libll_host.so  level builder in LLAMA's layout, R-MAT/update generator, .xs1
  libll_gpu.so   HBM mirror + hand-written gfx950 kernels behind the ll_gpu_* C ABI

Python pieces:
  graph     HostGraph / DeviceGraph / GpuContext (ctypes over the C ABIs)
  tasks     benchmark-task registry mirroring benchmark/benchmark.cc
  autograd  torch.autograd.Function for weighted PageRank (forward + adjoint)
"""
from .graph import (DIR_IN, DIR_OUT, UNREACHED, DeviceGraph, Edges, GpuContext, GpuError,
                    HostGraph, build_multilevel)

__all__ = ["DIR_IN", "DIR_OUT", "UNREACHED", "DeviceGraph", "Edges", "GpuContext", "GpuError",
           "HostGraph", "build_multilevel"]
