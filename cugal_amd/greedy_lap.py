"""Greedy rounding of a quasi-permutation, the surface of
cugal/greedy_lap.py:7-85.

GREEDY and DOUBLE_GREEDY run in the HIP LAP kernels (cg_lap_greedy).  Like
the reference they mask the taken columns of each visited row of the matrix
in place; DOUBLE_GREEDY orders rows by a stable descending sort of the row
maxima (the reference's argsort is not stable, so inputs with tied row maxima
may be visited in a different order there).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from .config import Config, HungarianMethod

_METHODS = {HungarianMethod.GREEDY: 0, HungarianMethod.DOUBLE_GREEDY: 1}


def greedy_lap_device(cost_matrix: torch.Tensor, method: HungarianMethod) -> torch.Tensor:
    """Assignment (int32 device tensor, one column per row); masks
    `cost_matrix` in place as the reference does."""
    if method not in _METHODS:
        raise NotImplementedError(f"{method} is not implemented on the ROCm path")
    nat.require_device_tensor(cost_matrix, "cost_matrix", torch.float32)
    n = cost_matrix.shape[0]
    if cost_matrix.shape != (n, n):
        raise ValueError("cost_matrix must be square")
    out = torch.empty(n, dtype=torch.int32, device=cost_matrix.device)
    ws = nat.workspace(cost_matrix.device, nat.query("cg_lap_workspace", n), "lap")
    nat.call(
        "cg_lap_greedy", nat.ptr(cost_matrix), n, _METHODS[method], nat.ptr(out), nat.ptr(ws), ws.numel(),
        nat.stream_of(cost_matrix.device),
    )
    return out


def greedy_lap(cost_matrix: torch.Tensor, config: Config) -> list[int]:
    """Column chosen for each row, as a list (cugal/greedy_lap.py:7)."""
    assignment = greedy_lap_device(cost_matrix, config.hungarian_method)
    return np.asarray(assignment.cpu().numpy(), dtype=np.int64).tolist()
