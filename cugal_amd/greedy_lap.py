"""Greedy rounding of a quasi-permutation, the surface of
cugal/greedy_lap.py:7-85.

Every method runs in the HIP LAP kernels (cg_lap_greedy,
cg_lap_greedy_random).  Like the reference they mask the taken columns of
each visited row of the matrix in place.  DOUBLE_GREEDY and PARALLEL_GREEDY
order rows by a stable descending sort of the row maxima (the reference's
sorts are not stable, so inputs with tied row maxima may be visited in a
different order there).  RAND and MORE_RAND draw their randomness as the
reference does: the row order from torch.randperm on the matrix's device,
and one random.random() per visited row for MORE_RAND (the reference's
uniform(0, total) is total * random.random()); MORE_RAND's prefix sums are
taken in FP64.
"""

from __future__ import annotations

import random

import numpy as np
import torch

from . import _native as nat
from .config import Config, HungarianMethod

_METHODS = {HungarianMethod.GREEDY: 0, HungarianMethod.DOUBLE_GREEDY: 1, HungarianMethod.PARALLEL_GREEDY: 2}
_RANDOM_METHODS = {HungarianMethod.RAND: 0, HungarianMethod.MORE_RAND: 1}


def greedy_lap_device(cost_matrix: torch.Tensor, method: HungarianMethod, order: torch.Tensor | None = None,
                      uniforms: torch.Tensor | None = None) -> torch.Tensor:
    """Assignment (int32 device tensor, one column per row); masks
    `cost_matrix` in place as the reference does.  RAND and MORE_RAND take
    the row order (and MORE_RAND the per-row draws in [0, 1]); when omitted
    they are drawn as the reference draws them."""
    if method not in _METHODS and method not in _RANDOM_METHODS:
        raise NotImplementedError(f"{method} is not a greedy method")
    nat.require_device_tensor(cost_matrix, "cost_matrix", torch.float32)
    n = cost_matrix.shape[0]
    if cost_matrix.shape != (n, n):
        raise ValueError("cost_matrix must be square")
    dev = cost_matrix.device
    out = torch.empty(n, dtype=torch.int32, device=dev)
    ws = nat.workspace(dev, nat.query("cg_lap_workspace", n), "lap")
    if method in _METHODS:
        nat.call("cg_lap_greedy", nat.ptr(cost_matrix), n, _METHODS[method], nat.ptr(out), nat.ptr(ws), ws.numel(),
                 nat.stream_of(dev))
        return out
    if order is None:
        order = torch.randperm(n, device=dev)
    order = order.to(device=dev, dtype=torch.int32).contiguous()
    if order.numel() != n:
        raise ValueError("order must hold one entry per row")
    if method == HungarianMethod.MORE_RAND:
        if uniforms is None:
            uniforms = torch.tensor([random.random() for _ in range(n)], dtype=torch.float64)
        uniforms = uniforms.to(device=dev, dtype=torch.float64).contiguous()
        if uniforms.numel() != n:
            raise ValueError("uniforms must hold one draw per row")
    else:
        uniforms = None
    nat.call("cg_lap_greedy_random", nat.ptr(cost_matrix), n, _RANDOM_METHODS[method], nat.ptr(order),
             nat.ptr(uniforms), nat.ptr(out), nat.ptr(ws), ws.numel(), nat.stream_of(dev))
    if method == HungarianMethod.MORE_RAND and bool((out < 0).any()):
        # The reference's mask.nonzero()[0] raises here.
        raise IndexError("MORE_RAND: no prefix of a masked row reaches its draw (negative or NaN entries)")
    return out


def greedy_lap(cost_matrix: torch.Tensor, config: Config) -> list[int]:
    """Column chosen for each row, as a list (cugal/greedy_lap.py:7)."""
    assignment = greedy_lap_device(cost_matrix, config.hungarian_method)
    return np.asarray(assignment.cpu().numpy(), dtype=np.int64).tolist()
