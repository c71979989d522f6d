"""MI355X-native read-only analytics over LLAMA's multi-level snapshot CSR.

Native pieces (built in-tree by ``python -m llama_amd.build``):
  libll_host.so  level builder in LLAMA's layout, R-MAT/update generator, .xs1
  libll_gpu.so   HBM mirror + hand-written gfx950 kernels behind the ll_gpu_* C ABI

Python pieces:
  graph     HostGraph / DeviceGraph / GpuContext (ctypes over the C ABIs)
  autograd  torch.autograd.Function for weighted PageRank (forward + adjoint)

The benchmark-task registry (the reference's task interface) is C++:
csrc/host/ll_gpu_tasks.h, with the CLI csrc/host/ll_gpu_bench.cc and the
binding into the reference's own driver, csrc/host/ll_gpu_reference.h.
"""
from .graph import (DIR_IN, DIR_OUT, UNREACHED, DeviceGraph, Edges, GpuContext, GpuError,
                    HostGraph, build_multilevel)

__all__ = ["DIR_IN", "DIR_OUT", "UNREACHED", "DeviceGraph", "Edges", "GpuContext", "GpuError",
           "HostGraph", "build_multilevel"]
