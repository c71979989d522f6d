"""Weighted PageRank as a differentiable torch op over the HIP forward/adjoint.

    r = WeightedPageRank.apply(w, p, graph, iters, damping)

w: FP32 edge weights in the snapshot's out-edge order (DeviceGraph.adjacency
(DIR_OUT) order), p: FP32 teleport vector [V]; both CUDA tensors on the
graph's device.  r0 = p, r' = (1-d) p + d W^T r.  backward() runs the
transposed sweep lambda' = dL/dr + d W lambda for the same number of
iterations and returns dL/dw[e=(u,v)] = d r[u] lambda[v] and
dL/dp = (1-d) lambda (the fixed-point adjoint; exact once the forward
iteration has converged).  There is no non-native path: the graph must be a
DeviceGraph backed by libll_gpu.so.
"""
from __future__ import annotations

import torch

from .graph import DeviceGraph


class WeightedPageRank(torch.autograd.Function):
    @staticmethod
    def forward(ctx, w: torch.Tensor, p: torch.Tensor, graph: DeviceGraph, iters: int = 100,
                damping: float = 0.85, level: int = -1) -> torch.Tensor:
        if not isinstance(graph, DeviceGraph):
            raise TypeError("graph must be a llama_amd DeviceGraph")
        for t, name in ((w, "w"), (p, "p")):
            if not t.is_cuda or t.dtype != torch.float32:
                raise TypeError(f"{name} must be a float32 CUDA tensor")
        w = w.contiguous()
        p = p.contiguous()
        if p.numel() != graph.V:
            raise ValueError("teleport vector must have max_nodes entries")
        r = torch.empty_like(p)
        torch.cuda.current_stream(p.device).synchronize()
        graph.wpr_forward(w.data_ptr(), p.data_ptr(), r.data_ptr(), iters, damping, level)
        ctx.save_for_backward(w, r)
        ctx.graph, ctx.iters, ctx.damping, ctx.level = graph, iters, damping, level
        return r

    @staticmethod
    def backward(ctx, grad_r: torch.Tensor):
        w, r = ctx.saved_tensors
        g = grad_r.contiguous().to(torch.float32)
        dw = torch.empty_like(w)
        dp = torch.empty_like(r)
        torch.cuda.current_stream(r.device).synchronize()
        ctx.graph.wpr_adjoint(w.data_ptr(), r.data_ptr(), g.data_ptr(), ctx.iters, ctx.damping,
                              dw.data_ptr(), dp.data_ptr(), ctx.level)
        return dw, dp, None, None, None, None


def weighted_pagerank(w: torch.Tensor, p: torch.Tensor, graph: DeviceGraph, iters: int = 100,
                      damping: float = 0.85, level: int = -1) -> torch.Tensor:
    return WeightedPageRank.apply(w, p, graph, iters, damping, level)
