"""The reference's native module under its own name and signatures: the ten
functions cuda_kernels exports (cuda_kernels/kernels.cpp:32-70), each a
thin call of the libcugal_hip.so entry point that replaces it.

Like the reference's pybind11 functions they take caller-allocated torch
tensors and write their outputs in place, run on the current stream, and
return only the two scalars the reference returns (the marginal error and
the duality gap).  Unlike them they check dtypes, devices and shapes and
raise instead of printing.  Temporaries come from the library's
workspaces; the scalar-returning functions synchronise, as the reference's
do by returning a float.

A program written against the reference's kernels imports this module as
`cuda_kernels` (the repository root holds a one-line `cuda_kernels.py`
that re-exports it).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as nat


def _square(t: torch.Tensor, name: str) -> int:
    nat.require_device_tensor(t, name, torch.float32)
    if t.dim() != 2 or t.shape[0] != t.shape[1]:
        raise ValueError(f"{name} must be a square matrix, got shape {tuple(t.shape)}")
    return int(t.shape[0])


def _csr(col_indices: torch.Tensor, row_pointers: torch.Tensor) -> int:
    nat.require_device_tensor(row_pointers, "row_pointers", torch.int32)
    if col_indices.dtype != torch.int32 or col_indices.device != row_pointers.device:
        raise TypeError("col_indices must be int32 on the device of row_pointers")
    return int(row_pointers.numel()) - 1


def adjacency_matmul(col_indices: torch.Tensor, row_pointers: torch.Tensor, matrix: torch.Tensor,
                     out: torch.Tensor, negate: bool) -> None:
    """out = (negate ? -A : A) @ matrix (cuda_kernels/adjacency.cu:77-150)."""
    n = _csr(col_indices, row_pointers)
    matrix = matrix.contiguous()
    if _square(matrix, "matrix") != n or _square(out, "out") != n:
        raise ValueError("matrix and out must be n x n for the n-row adjacency")
    nat.call("cg_adjacency_matmul", nat.ptr(col_indices), nat.ptr(row_pointers), n, nat.ptr(matrix), nat.ptr(out),
             int(bool(negate)), nat.stream_of(out.device))


def create_adjacency(edges: torch.Tensor, col_indices: torch.Tensor, row_pointers: torch.Tensor) -> None:
    """CSR of a flat int32 edge list x0, y0, x1, y1, ... where the pair
    (x, y) is the entry A[y][x], columns sorted within each row
    (cuda_kernels/adjacency.cu:228-259).  The caller symmetrises, as the
    reference's Python does (cugal/adjacency.py:53-54)."""
    nat.require_device_tensor(edges, "edges", torch.int32)
    if edges.numel() % 2:
        raise ValueError("edges must hold an even number of values (pairs)")
    m = edges.numel() // 2
    nat.require_device_tensor(row_pointers, "row_pointers", torch.int32)
    if col_indices.dtype != torch.int32 or col_indices.numel() != m:
        raise ValueError(f"col_indices must be int32 with {m} entries")
    node_count = row_pointers.numel() - 1
    ws = nat.workspace(edges.device, nat.query("cg_csr_build_workspace", m, node_count, 0), "csr")
    nat.call("cg_csr_build", nat.ptr(edges) if m else None, m, node_count, 0, nat.ptr(col_indices) if m else None,
             nat.ptr(row_pointers), nat.ptr(ws), ws.numel(), nat.stream_of(edges.device))


def graph_features(col_indices: torch.Tensor, row_pointers: torch.Tensor, clustering: torch.Tensor,
                   degrees: torch.Tensor) -> None:
    """Per vertex the clustering coefficient and the degree
    (cuda_kernels/feature_extraction.cu:67-84)."""
    n = _csr(col_indices, row_pointers)
    nat.require_device_tensor(clustering, "clustering", torch.float32, n)
    nat.require_device_tensor(degrees, "degrees", torch.float32, n)
    nat.call("cg_graph_features", nat.ptr(col_indices), nat.ptr(row_pointers), n, nat.ptr(clustering),
             nat.ptr(degrees), None, nat.stream_of(degrees.device))


def average_neighbor_features(col_indices: torch.Tensor, row_pointers: torch.Tensor, features: torch.Tensor,
                              averages: torch.Tensor) -> None:
    """averages[v] = mean of features over the neighbours of v, 0 for an
    isolated vertex (cuda_kernels/feature_extraction.cu:108-125)."""
    n = _csr(col_indices, row_pointers)
    nat.require_device_tensor(features, "features", torch.float32, n)
    nat.require_device_tensor(averages, "averages", torch.float32, n)
    nat.call("cg_average_neighbor_features", nat.ptr(col_indices), nat.ptr(row_pointers), n, nat.ptr(features),
             nat.ptr(averages), nat.stream_of(averages.device))


def sinkhorn_log_step(K: torch.Tensor, add: torch.Tensor, out: torch.Tensor, momentum: float) -> None:
    """out[r] = (1 - momentum) out[r] - momentum LSE_c(K[r][c] + add[c])
    (cuda_kernels/sinkhorn_log.cu:100-114)."""
    n = _square(K, "K")
    nat.require_device_tensor(add, "add", torch.float32, n)
    nat.require_device_tensor(out, "out", torch.float32, n)
    nat.call("cg_sinkhorn_log_step", nat.ptr(K), n, nat.ptr(add), nat.ptr(out), float(momentum),
             nat.stream_of(K.device))


def sinkhorn_log_marginal(K: torch.Tensor, u: torch.Tensor, v: torch.Tensor) -> float:
    """sum_r |sum_c exp(K[r][c] + v[c] + u[r]) - 1| in FP64
    (cuda_kernels/sinkhorn_log.cu:68-98)."""
    n = _square(K, "K")
    nat.require_device_tensor(u, "u", torch.float32, n)
    nat.require_device_tensor(v, "v", torch.float32, n)
    ws = nat.workspace(K.device, nat.query("cg_sinkhorn_log_marginal_workspace", n), "marginal")
    result = np.zeros(1, dtype=np.float64)
    nat.call("cg_sinkhorn_log_marginal", nat.ptr(K), n, nat.ptr(u), nat.ptr(v), result.ctypes.data, nat.ptr(ws),
             ws.numel(), nat.stream_of(K.device))
    return float(result[0])


def add_distance(source: torch.Tensor, target: torch.Tensor, out: torch.Tensor) -> None:
    """out[x][y] += ||source[x] - target[y]||_2 over 4 features
    (cuda_kernels/distance.cu:11-53)."""
    n = _square(out, "out")
    for name, t in (("source", source), ("target", target)):
        nat.require_device_tensor(t, name, torch.float32, 4 * n)
    nat.call("cg_add_distance", nat.ptr(source), nat.ptr(target), n, nat.ptr(out), nat.stream_of(out.device))


def regularize(gradient: torch.Tensor, P: torch.Tensor, iteration: int) -> None:
    """gradient += iteration - 2 iteration P (cuda_kernels/regularize.cu:10-39)."""
    n = _square(gradient, "gradient")
    if _square(P, "P") != n:
        raise ValueError("P must match the gradient's shape")
    nat.call("cg_regularize", nat.ptr(gradient), nat.ptr(P), n, int(iteration), nat.stream_of(P.device))


def update_quasi_permutation_log(P: torch.Tensor, K: torch.Tensor, u: torch.Tensor, v: torch.Tensor, alpha: float,
                                 regularization: float) -> float:
    """P += alpha (exp(K + u + v) - P); returns |sum (exp(K + u + v) - P) C|
    with C = -regularization K (cuda_kernels/update_quasi_permutation.cu:
    7-56).  The library's kernel reads the cost C, so K is scaled into a
    temporary first."""
    n = _square(K, "K")
    if P.dim() != 2 or P.shape != K.shape:
        raise ValueError("P must match K's shape")
    nat.require_device_tensor(u, "u", torch.float32, n)
    nat.require_device_tensor(v, "v", torch.float32, n)
    C = K * (-float(regularization))
    gap = torch.empty(1, dtype=torch.float64, device=K.device)
    ws = nat.workspace(K.device, nat.query("cg_update_workspace", n), "update")
    nat.call("cg_update_quasi_permutation_log", nat.ptr(P), nat.row_stride(P, "P"), nat.ptr(C), n, nat.ptr(u),
             nat.ptr(v), float(alpha), float(regularization), nat.ptr(gap), nat.ptr(ws), ws.numel(),
             nat.stream_of(K.device))
    return abs(float(gap.item()))


def dense_hungarian(cost: torch.Tensor, output: torch.Tensor) -> None:
    """Minimum-cost assignment of a dense n x n cost; output (n int32, on
    the host as in the reference) receives the column of each row
    (cuda_kernels/hungarian/hungarian.cu:436-448).  The cost is left
    unchanged (the reference's solver overwrites it)."""
    n = _square(cost, "cost")
    if output.dtype != torch.int32 or output.numel() != n:
        raise ValueError(f"output must be an int32 tensor of {n} entries")
    ws = nat.workspace(cost.device, nat.query("cg_lap_hungarian_workspace", n), "hungarian")
    col = torch.empty(n, dtype=torch.int32, device=cost.device)
    nat.call("cg_lap_hungarian", nat.ptr(cost), n, 0, nat.ptr(col), None, nat.ptr(ws), ws.numel(),
             nat.stream_of(cost.device))
    output.copy_(col.to(output.device))


EXPORTS = ("adjacency_matmul", "create_adjacency", "graph_features", "average_neighbor_features",
           "sinkhorn_log_step", "sinkhorn_log_marginal", "add_distance", "regularize",
           "update_quasi_permutation_log", "dense_hungarian")
