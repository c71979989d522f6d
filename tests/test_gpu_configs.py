"""Parity at the sizes of the single-level configurations: plain PageRank on
R-MAT scale 20 (10 iterations, the reference's own CPU case) against the FP64
oracle, and weighted PageRank with its adjoint on R-MAT scale 23 (row-normalised
U(0.1, 1) weights, uniform teleport, loss = sum of squared ranks over 1000
targets) against the FP64 oracle's forward and adjoint on every edge, per edge
on 200 sampled edges, and against FP64 central differences (h = 1e-6) of the
loss along a random +-1 direction on those 200 edges.

Bars: PageRank L1 relative error <= 1e-6 (FP32 compensated vs FP64); weighted
ranks <= 1e-5, gradients <= 1e-4 (L1 relative, FP32 vs FP64); the directional
derivative within 1e-4 relative of the central difference."""
import numpy as np
import pytest

import oracle
from conftest import gpu_case
from llama_amd.graph import DIR_OUT, DeviceGraph, build_multilevel

pytestmark = pytest.mark.gpu


def test_config1_pagerank_s20_vs_fp64(ctx):
    gpu_case("config 1: R-MAT s20 single level, PageRank 10 iterations vs FP64 oracle")
    g = build_multilevel(20, levels=1, seed=1, deletions=False)
    o = oracle.OracleGraph.from_host(g)
    dg = DeviceGraph(ctx, g)
    r, _ = dg.pagerank(10)
    ref, _ = o.pagerank(10)
    err = float(np.abs(r - ref).sum() / np.abs(ref).sum())
    print(f"config 1 PageRank L1 relative error vs FP64: {err:.3e}")
    assert err <= 1e-6
    # the task's finalize value (sum of ranks) to FP32 accuracy
    assert abs(r.sum() - ref.sum()) <= 1e-6 * ref.sum()
    dg.close()


def test_config4_weighted_pagerank_adjoint_s23(ctx):
    torch = pytest.importorskip("torch")
    from llama_amd.autograd import weighted_pagerank
    gpu_case("config 4: R-MAT s23 weighted PageRank + adjoint vs FP64 oracle and FD")
    g = build_multilevel(23, levels=1, seed=6, deletions=False)
    dg = DeviceGraph(ctx, g)
    ofs, adj = dg.adjacency(DIR_OUT)
    V, E = dg.V, adj.size
    deg = np.diff(ofs.astype(np.int64))
    src = np.repeat(np.arange(V, dtype=np.int64), deg)
    rng = np.random.default_rng(6)
    w = rng.uniform(0.1, 1.0, E)
    w /= np.bincount(src, weights=w, minlength=V)[src]  # row-normalised
    w32 = w.astype(np.float32)
    w = w32.astype(np.float64)  # the oracle sees the weights the device sees
    del src
    p = np.full(V, 1.0 / V)
    K, d = 100, 0.85
    tgt = np.random.default_rng(7).choice(V, 1000, replace=False)
    dev = torch.device("cuda:0")
    wt = torch.tensor(w32, device=dev, requires_grad=True)
    pt = torch.full((V,), 1.0 / V, dtype=torch.float32, device=dev, requires_grad=True)
    r = weighted_pagerank(wt, pt, dg, K, d)
    loss = (r[torch.tensor(tgt, device=dev)] ** 2).sum()
    loss.backward()
    rr = r.detach().cpu().numpy().astype(np.float64)
    dw = wt.grad.cpu().numpy().astype(np.float64)
    dp = pt.grad.cpu().numpy().astype(np.float64)
    del wt, pt, r

    gpu_case("config 4: FP64 oracle forward + adjoint")
    r_ref = oracle.wpr_forward(ofs, adj, w, p, K, d)
    err_r = np.abs(rr - r_ref).sum() / np.abs(r_ref).sum()
    gr = np.zeros(V)
    gr[tgt] = 2 * r_ref[tgt]
    dw_ref, dp_ref = oracle.wpr_adjoint(ofs, adj, w, r_ref, gr, K, d)
    err_w = np.abs(dw - dw_ref).sum() / np.abs(dw_ref).sum()
    err_p = np.abs(dp - dp_ref).sum() / np.abs(dp_ref).sum()
    samp = np.random.default_rng(8).choice(E, 200, replace=False)
    err_s = np.abs(dw[samp] - dw_ref[samp]).sum() / np.abs(dw_ref[samp]).sum()
    print(f"config 4 L1 rel: ranks {err_r:.2e}, dL/dw {err_w:.2e} (200 sampled {err_s:.2e}), "
          f"dL/dp {err_p:.2e}")
    assert err_r <= 1e-5 and err_w <= 1e-4 and err_p <= 1e-4 and err_s <= 1e-4

    gpu_case("config 4: FP64 central differences on 200 sampled edges")
    delta = np.random.default_rng(9).choice([-1.0, 1.0], samp.size)
    h = 1e-6

    def loss_at(sign):
        ws = w.copy()
        ws[samp] += sign * h * delta
        rs = oracle.wpr_forward(ofs, adj, ws, p, K, d)
        return float((rs[tgt] ** 2).sum())

    fd = (loss_at(1.0) - loss_at(-1.0)) / (2 * h)
    ad_ref = float((dw_ref[samp] * delta).sum())
    ad_gpu = float((dw[samp] * delta).sum())
    print(f"config 4 directional derivative: FD {fd:.6e}, FP64 adjoint {ad_ref:.6e}, "
          f"GPU adjoint {ad_gpu:.6e}")
    assert abs(ad_ref - fd) <= 1e-4 * abs(fd)
    assert abs(ad_gpu - fd) <= 1e-4 * abs(fd)
    dg.close()
