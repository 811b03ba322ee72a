"""Node orders that make the gradient's CSR gathers local.

The gradient -2 A P B gathers, for each output row, the rows of P (then of
the intermediate) that the row's neighbour list names.  A thread block
gathers a 64-row tile of one column strip and the blocks of one XCD sweep
consecutive tiles, so the strip's lines are reused from that XCD's L2 when
neighbours have nearby labels.  A graph whose labels are scattered (a
relabelled copy, as every alignment target is) fetches each line many times.

Relabelling a graph changes no result of the alignment beyond the order of
floating-point sums: with permutations T (source) and S (target),
T A T^T (T P S^T) S B S^T = T (A P B) S^T, the distance and regulariser terms
are elementwise, and Sinkhorn and the Frank-Wolfe update commute with
relabelling rows and columns.  The loop therefore runs on relabelled graphs
and its quasi-permutation is copied back to the callers' labels
(cg_permute_copy) before the LAP, which sees the original labels.

The order is reverse Cuthill-McKee (breadth-first levels, neighbours of
earlier nodes first), computed on the host from the CSR, either of the whole
graph or of its edges that close at least two triangles: the shortcuts of a
small-world graph close none, and numbering without them follows the
graph's local structure (a rewired ring lattice gets its ring order back,
89 % of the edges within 64 labels, against 27 % for the plain order).
"auto" keeps the candidate (own labels, RCM, filtered RCM) that puts the
most edges within a small window (64 labels: an LDS tile's reach) plus a
large one (min(n/4, 8192) labels: an L2's reach of a gathered strip), and
relabels only where that gains clearly over the graph's own labels.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse
import scipy.sparse.csgraph
import torch

from . import _native as nat
from .adjacency import Adjacency, edges_to_csr

ORDERS = ("auto", "natural", "rcm")

# "auto" relabels only from this size: below it a column strip of the
# gathered matrix is a fraction of an XCD's L2 in any order.
AUTO_MIN_NODES = 8192
# ... and only when the order puts this many more of the edges within the
# window (the fraction of edges whose endpoints' labels differ by at most
# min(n / 4, 8192) rows: 1 MiB of a 128-byte-per-row strip).
AUTO_MIN_GAIN = 0.05


def _host_csr(adj: Adjacency):
    rp = adj.row_pointers.cpu().numpy().astype(np.int64)
    col = adj.col_indices.cpu().numpy().astype(np.int64)
    return rp, col


SMALL_WINDOW = 64
# edges closing fewer triangles than this are left out of the filtered order
FILTER_MIN_COMMON = 2


def window_locality(rp: np.ndarray, col: np.ndarray, position: np.ndarray | None = None,
                    window: int | None = None) -> float:
    """Fraction of CSR entries (r, c) with |pos(r) - pos(c)| <= window
    (default min(n/4, 8192)); `position` maps each node to its label
    (identity when None)."""
    n = len(rp) - 1
    if len(col) == 0:
        return 1.0
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    if position is not None:
        rows, cols = position[rows], position[col]
    else:
        cols = col
    if window is None:
        window = min(max(n // 4, 1), 8192)
    return float(np.mean(np.abs(rows - cols) <= window))


def locality_score(rp: np.ndarray, col: np.ndarray, position: np.ndarray | None = None) -> float:
    """Small-window plus large-window locality (0 to 2)."""
    return window_locality(rp, col, position, SMALL_WINDOW) + window_locality(rp, col, position)


def device_locality_scores(adj: Adjacency, positions: list) -> list[float]:
    """locality_score of each relabelling in `positions` (None = the graph's
    own labels), counted on the device from the CSR in HBM with one host
    synchronisation; equal to the host count (integer counts over nnz)."""
    n = adj.number_of_nodes()
    nnz = adj.col_indices.numel()
    if nnz == 0:
        return [2.0 for _ in positions]
    dev = adj.row_pointers.device
    counts = (adj.row_pointers[1:] - adj.row_pointers[:-1]).to(torch.int64)
    rows = torch.repeat_interleave(torch.arange(n, device=dev, dtype=torch.int64), counts, output_size=nnz)
    cols = adj.col_indices.to(torch.int64)
    large = min(max(n // 4, 1), 8192)
    hits = []
    for position in positions:
        if position is None:
            dist = (rows - cols).abs()
        else:
            pos = torch.as_tensor(position, device=dev, dtype=torch.int64)
            dist = (pos[rows] - pos[cols]).abs()
        hits.append((dist <= SMALL_WINDOW).sum())
        hits.append((dist <= large).sum())
    counted = torch.stack(hits).cpu().numpy()
    return [float(counted[2 * k]) / nnz + float(counted[2 * k + 1]) / nnz for k in range(len(positions))]


def rcm_position(rp: np.ndarray, col: np.ndarray) -> np.ndarray:
    """New label of each node under reverse Cuthill-McKee."""
    n = len(rp) - 1
    M = scipy.sparse.csr_matrix((np.ones(len(col), dtype=np.int8), col, rp), shape=(n, n))
    order = scipy.sparse.csgraph.reverse_cuthill_mckee(M, symmetric_mode=True)
    position = np.empty(n, dtype=np.int64)
    position[order] = np.arange(n, dtype=np.int64)
    return position


def filtered_rcm_position(adj: Adjacency, host_csr=None) -> np.ndarray:
    """Reverse Cuthill-McKee of the graph without its edges that have fewer
    than FILTER_MIN_COMMON common neighbours (counted on the device,
    cg_edge_common_neighbours); the edge filter is symmetric, so the
    filtered graph is undirected too.  `host_csr`: the CSR already copied to
    the host (_host_csr), if at hand."""
    n = adj.number_of_nodes()
    dev = adj.row_pointers.device
    nnz = adj.col_indices.numel()
    common = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)[:nnz]
    nat.call("cg_edge_common_neighbours", nat.ptr(adj.col_indices), nat.ptr(adj.row_pointers), n, nat.ptr(common),
             nat.stream_of(dev))
    keep = (common >= FILTER_MIN_COMMON).cpu().numpy()
    rp, col = host_csr if host_csr is not None else _host_csr(adj)
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(rp))
    M = scipy.sparse.csr_matrix((np.ones(int(keep.sum()), dtype=np.int8), (rows[keep], col[keep])), shape=(n, n))
    order = scipy.sparse.csgraph.reverse_cuthill_mckee(M, symmetric_mode=True)
    position = np.empty(n, dtype=np.int64)
    position[order] = np.arange(n, dtype=np.int64)
    return position


def choose_position(adj: Adjacency, mode: str) -> np.ndarray | None:
    """The relabelling of `adj` for the gathers (new label of each node), or
    None to keep its labels: the candidate (own labels, RCM, RCM of the
    triangle-closing edges) with the best locality score; "rcm" takes the
    better RCM even when the own labels score higher."""
    if mode not in ORDERS:
        raise ValueError(f"node_order must be one of {ORDERS}, got {mode!r}")
    n = adj.number_of_nodes()
    if mode == "natural" or n < 2 or adj.directed or (mode == "auto" and n < AUTO_MIN_NODES):
        return None
    rp, col = _host_csr(adj)
    candidates = [rcm_position(rp, col), filtered_rcm_position(adj, (rp, col))]
    own, *scores = device_locality_scores(adj, [None] + candidates)
    best = int(np.argmax(scores))
    if mode == "auto" and scores[best] < own + AUTO_MIN_GAIN:
        return None
    return candidates[best]


def relabel(adj: Adjacency, position: torch.Tensor) -> Adjacency:
    """The CSR of the graph with node v renamed position[v] (position an
    int64 device tensor, a permutation of 0..n-1); rows and the sorted
    column order are rebuilt on the device, repeated entries kept."""
    n = adj.number_of_nodes()
    dev = adj.row_pointers.device
    counts = (adj.row_pointers[1:] - adj.row_pointers[:-1]).to(torch.int64)
    rows = torch.repeat_interleave(torch.arange(n, device=dev), counts)
    cols = adj.col_indices.to(torch.int64)
    # a pair (x, y) is the entry [y][x]
    edges = torch.stack((position[cols], position[rows]), dim=1).to(torch.int32)
    col, rp = edges_to_csr(edges, n, symmetrize=False)
    return Adjacency(col, rp, loops_twice=adj.loops_twice)


def permute_rows(features: torch.Tensor, position: torch.Tensor) -> torch.Tensor:
    """Row position[v] of the result is row v of `features`."""
    out = torch.empty_like(features)
    out[position] = features
    return out


def copy_back(P: torch.Tensor, row_position: torch.Tensor | None, col_position: torch.Tensor | None) -> torch.Tensor:
    """Contiguous Q with Q[i][j] = P[row_position[i]][col_position[j]]: the
    quasi-permutation of relabelled graphs in the original labels."""
    n = P.shape[0]
    out = torch.empty((n, n), dtype=P.dtype, device=P.device)
    rmap = None if row_position is None else row_position.to(torch.int32).contiguous()
    cmap = None if col_position is None else col_position.to(torch.int32).contiguous()
    nat.call("cg_permute_copy", nat.ptr(P), nat.row_stride(P, "P"), n, nat.ptr(rmap), nat.ptr(cmap), nat.ptr(out), n,
             nat.stream_of(P.device))
    return out
