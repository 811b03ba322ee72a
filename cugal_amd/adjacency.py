"""CSR adjacency on the device, the surface of cugal/adjacency.py:18-156.

``Adjacency(col_indices, row_pointers)`` holds int32 CSR arrays on a ROCm
device with the columns of each row sorted.  It is built by the HIP CSR
kernels (histogram, scan, scatter, per-row rank sort) and multiplies through
the row-gather SpMM kernel; there is no host or torch fallback.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import networkx as nx
import numpy as np
import torch

from . import _native as nat


def _device(device) -> torch.device:
    dev = torch.device(device)
    if dev.type != "cuda":
        raise NotImplementedError(
            f"Adjacency lives on a ROCm device (torch device 'cuda'); got {device!r}. "
            "The CPU path of the reference is not part of this library."
        )
    return dev


def edges_to_csr(edges: torch.Tensor, node_count: int, symmetrize: bool) -> tuple[torch.Tensor, torch.Tensor]:
    """CSR of an (E, 2) int32 device tensor of pairs (x, y), each adding the
    entry A[y][x] (and A[x][y] when `symmetrize`)."""
    dev = edges.device
    edges = edges.to(torch.int32).contiguous()
    m = edges.shape[0]
    nnz = m * (2 if symmetrize else 1)
    col = torch.empty(max(nnz, 1), dtype=torch.int32, device=dev)[:nnz]
    rp = torch.empty(node_count + 1, dtype=torch.int32, device=dev)
    ws_bytes = nat.query("cg_csr_build_workspace", m, node_count, int(symmetrize))
    ws = nat.workspace(dev, ws_bytes, "csr")
    nat.call(
        "cg_csr_build", nat.ptr(edges) if m else None, m, node_count, int(symmetrize),
        nat.ptr(col) if nnz else None, nat.ptr(rp), nat.ptr(ws), ws.numel(), nat.stream_of(dev),
    )
    return col, rp


@dataclass
class Adjacency:
    col_indices: torch.Tensor
    """Column index of every non-zero, row by row (int32)."""

    row_pointers: torch.Tensor
    """Offset in col_indices where each row begins; length n + 1 (int32)."""

    loops_twice: bool = False
    """True when a self-loop is listed twice in its row, as the reference's
    symmetrised edge list gives it (cugal/adjacency.py:52-53): the loop
    contributes 2 to the SpMM; the 0/1 bit form cannot hold that."""

    directed: bool = False
    """True when the matrix need not be symmetric (a directed graph, or a
    dense matrix that is not symmetric): the gradient then takes the
    general form with the transpose (cugal/pred.py:38-51)."""

    _packed_bits: torch.Tensor | None = field(default=None, init=False, repr=False, compare=False)
    _transposed: "Adjacency | None" = field(default=None, init=False, repr=False, compare=False)

    @classmethod
    def from_dense(cls, dense: torch.Tensor) -> "Adjacency":
        """From a dense (n x n) matrix on the device; non-zeros become ones."""
        dev = _device(dense.device)
        rows, cols = torch.nonzero(dense, as_tuple=True)
        edges = torch.stack((cols, rows), dim=1).to(torch.int32)
        col, rp = edges_to_csr(edges, dense.shape[0], symmetrize=False)
        return cls(col, rp, directed=not bool(torch.equal(dense, dense.T)))

    @classmethod
    def from_graph(cls, graph: nx.Graph, device, loops_once: bool = False) -> "Adjacency":
        """From a networkx graph whose nodes are 0..n-1 (cugal/adjacency.py:
        46-69).  An undirected graph gives both directions of every edge, so
        a self-loop is listed twice, as in the reference's CSR; with
        `loops_once` it is listed once, the 1 on the diagonal of the
        reference's dense adjacency (nx.to_numpy_array, cugal/pred.py:
        106-110)."""
        dev = _device(device)
        node_count = graph.number_of_nodes()
        edge_list = list(nx.edges(graph))
        edges = np.asarray(edge_list, dtype=np.int64).reshape(-1, 2)
        if edges.size and (edges.min() < 0 or edges.max() >= node_count):
            raise ValueError("graph nodes must be labelled 0..n-1")
        undirected = not graph.is_directed()
        loops = int(np.count_nonzero(edges[:, 0] == edges[:, 1])) if edges.size else 0
        if undirected and loops and loops_once:
            # mirror the other edges on the host; the kernel then adds each
            # pair once
            other = edges[edges[:, 0] != edges[:, 1]]
            edges = np.concatenate((edges, other[:, ::-1]))
            t = torch.from_numpy(np.ascontiguousarray(edges).astype(np.int32)).to(dev)
            col, rp = edges_to_csr(t, node_count, symmetrize=False)
            return cls(col, rp)
        t = torch.from_numpy(edges.astype(np.int32)).to(dev)
        col, rp = edges_to_csr(t, node_count, symmetrize=undirected)
        # a directed graph: each edge (x, y) is the entry [y][x], so row y
        # lists the predecessors of y, as the reference's create_adjacency
        # gives them (cuda_kernels/adjacency.cu:228-259)
        return cls(col, rp, loops_twice=undirected and loops > 0, directed=not undirected)

    def transposed(self) -> "Adjacency":
        """The CSR of the transposed matrix (built once on the device and
        cached; the adjacency itself for a symmetric one)."""
        if not self.directed:
            return self
        if self._transposed is None:
            n = self.number_of_nodes()
            dev = self.row_pointers.device
            counts = (self.row_pointers[1:] - self.row_pointers[:-1]).to(torch.int64)
            rows = torch.repeat_interleave(torch.arange(n, device=dev), counts)
            # the pair (r, c) is the entry [c][r]
            edges = torch.stack((rows, self.col_indices.to(torch.int64)), dim=1).to(torch.int32)
            col, rp = edges_to_csr(edges, n, symmetrize=False)
            self._transposed = Adjacency(col, rp, directed=True)
            self._transposed._transposed = self
        return self._transposed

    def as_dense(self, dtype: torch.dtype) -> torch.Tensor:
        """Convert back to a dense matrix (format conversion)."""
        n = self.number_of_nodes()
        counts = (self.row_pointers[1:] - self.row_pointers[:-1]).to(torch.int64)
        rows = torch.repeat_interleave(torch.arange(n, device=self.col_indices.device), counts)
        dense = torch.zeros((n, n), dtype=dtype, device=self.col_indices.device)
        dense[rows, self.col_indices.to(torch.int64)] = 1
        return dense

    def packed_bits(self) -> torch.Tensor:
        """The 0/1 adjacency as bits, zero padded to np = n rounded up to
        256: word (k // 32) * np + i holds bit k % 32 of row i (int32
        storage of uint32 words).  The operand of the matrix-core gradient;
        built once on the device and cached on the object."""
        if self._packed_bits is None:
            n = self.number_of_nodes()
            dev = self.col_indices.device
            words = int(nat.query("cg_adjacency_bits_words", n))
            out = torch.empty(words, dtype=torch.int32, device=dev)
            nat.call("cg_adjacency_pack_bits", nat.ptr(self.col_indices), nat.ptr(self.row_pointers), n,
                     nat.ptr(out), nat.stream_of(dev))
            self._packed_bits = out
        return self._packed_bits

    def byte_size(self) -> int:
        return 4 * len(self.col_indices) + 4 * len(self.row_pointers)

    def number_of_nodes(self) -> int:
        return len(self.row_pointers) - 1

    def validate(self):
        col_indices, row_pointers = self.col_indices.cpu().numpy(), self.row_pointers.cpu().numpy()
        assert np.all(np.diff(row_pointers) >= 0), "row_pointers isn't sorted"
        assert row_pointers[0] == 0, "row_pointers doesn't start with 0"
        assert row_pointers[-1] == len(col_indices), "row_pointers doesn't end correctly"
        for row_index in range(len(row_pointers) - 1):
            cols = col_indices[row_pointers[row_index]:row_pointers[row_index + 1]]
            assert np.all(np.diff(cols) >= 0), "col_indices isn't sorted"
            assert np.all(cols <= self.number_of_nodes()), "invalid entries in col_indices"

    def mul(self, matrix: torch.Tensor, negate_lhs: bool = False) -> torch.Tensor:
        """self @ matrix for a dense n x n FP32 matrix (cugal/adjacency.py:
        128-143).  A transposed view is made contiguous first."""
        n = self.number_of_nodes()
        assert matrix.shape[0] == n and matrix.shape[1] == n, "matrix must match size"
        if matrix.dtype != torch.float32:
            raise NotImplementedError("Adjacency.mul: the SpMM kernel is FP32")
        matrix = matrix.contiguous()
        nat.require_device_tensor(matrix, "matrix", torch.float32)
        out = torch.empty_like(matrix)
        nat.call(
            "cg_adjacency_matmul", nat.ptr(self.col_indices), nat.ptr(self.row_pointers), n,
            nat.ptr(matrix), nat.ptr(out), int(negate_lhs), nat.stream_of(matrix.device),
        )
        return out
