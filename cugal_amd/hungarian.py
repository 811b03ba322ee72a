"""Exact linear assignment on the device, the replacement of the reference's
``cuda_kernels.dense_hungarian`` (cuda_kernels/hungarian/hungarian.cu:436-448)
and of the cugraph solver behind ``HungarianMethod.DENSE``
(cugal/pred.py:161-166).

The reference's DENSE branch hands ``1 - P`` to a minimiser; in FP32 that
rounds every entry of P below 2^-25 to the same cost.  ``dense_hungarian``
maximises P itself (``maximize=True``), which is exact; with
``maximize=False`` it minimises the given cost like the reference kernel.
"""

from __future__ import annotations

import torch

from . import _native as nat


def dense_hungarian(cost: torch.Tensor, maximize: bool = False, stats: bool = False):
    """Column of each row (int32 device tensor) of an optimal assignment of
    the square FP32 device matrix `cost`; with ``stats=True`` also returns
    (rows left free by the warm start, Dijkstra steps)."""
    nat.require_device_tensor(cost, "cost", torch.float32)
    n = cost.shape[0]
    if cost.dim() != 2 or cost.shape[1] != n:
        raise ValueError("cost must be square")
    dev = cost.device
    out = torch.empty(n, dtype=torch.int32, device=dev)
    st = torch.zeros(2, dtype=torch.int32, device=dev)
    ws = nat.workspace(dev, nat.query("cg_lap_hungarian_workspace", n), "hungarian")
    nat.call("cg_lap_hungarian", nat.ptr(cost), n, int(bool(maximize)), nat.ptr(out), nat.ptr(st), nat.ptr(ws),
             ws.numel(), nat.stream_of(dev))
    if stats:
        free_rows, steps = (int(x) for x in st.cpu().tolist())
        return out, (free_rows, steps)
    return out
