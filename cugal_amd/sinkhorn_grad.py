"""Differentiable log-domain Sinkhorn on the device.

``sinkhorn_plan(C, reg, iterations, checkpoint_every)`` runs `iterations`
unrolled iterations of the reference's log-domain update (v = -LSE_i(K + u),
u = -LSE_j(K + v), K = -C/reg; cugal/sinkhorn.py:107-111) from the given
duals and returns the plan exp(K + u + v).  Its backward is the hand-written
adjoint of the unrolled iterations (cg_sinkhorn_unrolled_backward): the
duals are kept every `checkpoint_every` iterations and each segment is
recomputed during the backward, so memory is O(n * iterations /
checkpoint_every) beyond the cost matrix.

``feature_sinkhorn_plan(fs, ft, w, lam, reg, ...)`` builds the cost
C = lam * sum_k w_k |fs_ik - ft_jk| on the device and returns the plan; its
backward yields dL/dw and dL/dlam by contracting dL/dC against the feature
distances inside the gradient pass (no n x n gradient is stored).
"""

from __future__ import annotations

import torch

from . import _native as nat
from .adjacency import Adjacency
from .feature_extraction import egonet_edges, node_features


def _checks(C: torch.Tensor, reg: float, iterations: int, checkpoint_every: int) -> int:
    nat.require_device_tensor(C, "cost", torch.float32)
    n = C.shape[0]
    if C.dim() != 2 or C.shape[1] != n:
        raise ValueError("cost must be square")
    if not reg > 0:
        raise ValueError("regularization must be > 0")
    if iterations < 0 or not 1 <= checkpoint_every <= 64:
        raise ValueError("iterations >= 0 and 1 <= checkpoint_every <= 64")
    return n


def _forward(C: torch.Tensor, reg: float, iterations: int, k: int, u0, v0):
    n = C.shape[0]
    dev = C.device
    u0 = torch.zeros(n, device=dev) if u0 is None else u0.detach().to(torch.float32).contiguous()
    v0 = torch.zeros(n, device=dev) if v0 is None else v0.detach().to(torch.float32).contiguous()
    ckpt = torch.empty(int(nat.query("cg_sinkhorn_checkpoint_floats", n, iterations, k)), device=dev)
    u = torch.empty(n, device=dev)
    v = torch.empty(n, device=dev)
    ws = nat.workspace(dev, nat.query("cg_sinkhorn_unrolled_workspace", n, k), "sinkhorn_grad")
    nat.call("cg_sinkhorn_unrolled_forward", nat.ptr(C), n, float(reg), int(iterations), int(k), nat.ptr(u0),
             nat.ptr(v0), nat.ptr(ckpt), nat.ptr(u), nat.ptr(v), nat.ptr(ws), ws.numel(), nat.stream_of(dev))
    plan = torch.empty_like(C)
    pws = nat.workspace(dev, 8 * n + 1024, "plan")
    nat.call("cg_sinkhorn_plan", nat.ptr(C), n, float(reg), nat.ptr(u), nat.ptr(v), nat.ptr(plan), nat.ptr(pws),
             pws.numel(), nat.stream_of(dev))
    return plan, ckpt, u, v


class SinkhornPlan(torch.autograd.Function):
    """plan = Sinkhorn(C); gradient with respect to C."""

    @staticmethod
    def forward(ctx, C, reg, iterations, checkpoint_every, u0=None, v0=None):
        C = C.detach().contiguous()
        _checks(C, reg, iterations, checkpoint_every)
        plan, ckpt, _, _ = _forward(C, reg, iterations, checkpoint_every, u0, v0)
        ctx.save_for_backward(C, ckpt)
        ctx.params = (float(reg), int(iterations), int(checkpoint_every))
        return plan

    @staticmethod
    def backward(ctx, grad_plan):
        C, ckpt = ctx.saved_tensors
        reg, iterations, k = ctx.params
        n = C.shape[0]
        grad_plan = grad_plan.to(torch.float32).contiguous()
        grad_cost = torch.empty_like(C)
        ws = nat.workspace(C.device, nat.query("cg_sinkhorn_unrolled_workspace", n, k), "sinkhorn_grad")
        nat.call("cg_sinkhorn_unrolled_backward", nat.ptr(C), n, reg, iterations, k, nat.ptr(ckpt),
                 nat.ptr(grad_plan), nat.ptr(grad_cost), None, None, None, 0.0, None, nat.ptr(ws), ws.numel(),
                 nat.stream_of(C.device))
        return grad_cost, None, None, None, None, None


def feature_cost(fs: torch.Tensor, ft: torch.Tensor, w: torch.Tensor, lam: float) -> torch.Tensor:
    """C = lam * sum_k w_k |fs_ik - ft_jk| for n x 4 feature matrices."""
    fs, ft = fs.to(torch.float32).contiguous(), ft.to(torch.float32).contiguous()
    nat.require_device_tensor(fs, "source features", torch.float32)
    n = fs.shape[0]
    if fs.shape != (n, 4) or ft.shape != (n, 4):
        raise ValueError("features must be n x 4")
    wd = w.detach().to(device=fs.device, dtype=torch.float32).contiguous()
    C = torch.empty((n, n), device=fs.device)
    nat.call("cg_feature_cost", nat.ptr(fs), nat.ptr(ft), nat.ptr(wd), float(lam), n, nat.ptr(C),
             nat.stream_of(fs.device))
    return C


class FeatureSinkhornPlan(torch.autograd.Function):
    """plan = Sinkhorn(lam * D(w)); gradients with respect to w and lam."""

    @staticmethod
    def forward(ctx, fs, ft, w, lam, reg, iterations, checkpoint_every):
        fs, ft = fs.detach().to(torch.float32).contiguous(), ft.detach().to(torch.float32).contiguous()
        lam_f = float(lam)
        C = feature_cost(fs, ft, w, lam_f)
        _checks(C, reg, iterations, checkpoint_every)
        plan, ckpt, _, _ = _forward(C, reg, iterations, checkpoint_every, None, None)
        wd = w.detach().to(device=fs.device, dtype=torch.float32).contiguous()
        ctx.save_for_backward(C, ckpt, fs, ft, wd)
        ctx.params = (float(reg), int(iterations), int(checkpoint_every), lam_f)
        ctx.w_meta = (w.dtype, w.device)
        ctx.lam_is_tensor = isinstance(lam, torch.Tensor)
        return plan

    @staticmethod
    def backward(ctx, grad_plan):
        C, ckpt, fs, ft, wd = ctx.saved_tensors
        reg, iterations, k, lam = ctx.params
        n = C.shape[0]
        grad_plan = grad_plan.to(torch.float32).contiguous()
        out = torch.empty(5, dtype=torch.float64, device=C.device)
        ws = nat.workspace(C.device, nat.query("cg_sinkhorn_unrolled_workspace", n, k), "sinkhorn_grad")
        nat.call("cg_sinkhorn_unrolled_backward", nat.ptr(C), n, reg, iterations, k, nat.ptr(ckpt),
                 nat.ptr(grad_plan), None, nat.ptr(fs), nat.ptr(ft), nat.ptr(wd), lam, nat.ptr(out), nat.ptr(ws),
                 ws.numel(), nat.stream_of(C.device))
        wdtype, wdev = ctx.w_meta
        grad_w = out[1:].to(device=wdev, dtype=wdtype)
        grad_lam = out[0].to(torch.float64) if ctx.lam_is_tensor else None
        return None, None, grad_w, grad_lam, None, None, None


def sinkhorn_plan(C: torch.Tensor, reg: float, iterations: int, checkpoint_every: int = 10, u0=None, v0=None):
    """Differentiable plan of `iterations` unrolled log-Sinkhorn iterations."""
    return SinkhornPlan.apply(C, reg, iterations, checkpoint_every, u0, v0)


def feature_sinkhorn_plan(fs, ft, w, lam, reg: float, iterations: int, checkpoint_every: int = 10):
    """Differentiable plan of the feature cost lam * D(w)."""
    return FeatureSinkhornPlan.apply(fs, ft, w, lam, reg, iterations, checkpoint_every)


def structural_features(adjacency: Adjacency) -> torch.Tensor:
    """n x 4 FP32: degree, clustering, mean neighbour degree, ego-net edges
    (FUGAL/pred.py:68-72 for the last)."""
    n = adjacency.number_of_nodes()
    tri = torch.empty(n, dtype=torch.int32, device=adjacency.row_pointers.device)
    f = node_features(adjacency, 1.0, tri)
    ego = egonet_edges(adjacency, tri).to(torch.float32)
    return torch.stack((f[:, 0], f[:, 1], f[:, 2], ego), dim=1).contiguous()
