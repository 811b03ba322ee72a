"""Structural node features and their distances, the surface of
cugal/feature_extraction.py:15-108.

Per node: degree, clustering coefficient, mean neighbour degree and mean
neighbour clustering (the order of cugal/feature_extraction.py:64 and :38),
computed by the HIP feature kernels from the CSR adjacency; with
Config.node_features = "fugal7" also the three counts of the original
FUGAL's 7-feature set (FUGAL/pred.py:24-110).  Integer counts
are exact and every ratio is the FP32 rounding of the FP64 value networkx
computes.
"""

from __future__ import annotations

from dataclasses import dataclass

import networkx as nx
import torch

from . import _native as nat
from .adjacency import Adjacency
from .config import Config


def _as_adjacency(graph: nx.Graph | Adjacency, config: Config) -> Adjacency:
    return graph if isinstance(graph, Adjacency) else Adjacency.from_graph(graph, config.device)


def node_features(adjacency: Adjacency, mu: float = 1.0, triangles: torch.Tensor | None = None) -> torch.Tensor:
    """n x 4 FP32 features scaled by `mu`; fills `triangles` (int32, n) with
    t(v) = 2 * (triangles through v) when given."""
    n = adjacency.number_of_nodes()
    dev = adjacency.row_pointers.device
    out = torch.empty((n, 4), dtype=torch.float32, device=dev)
    ws_bytes = nat.query("cg_node_features_workspace", n)
    ws = nat.workspace(dev, ws_bytes, "features")
    nat.call(
        "cg_node_features", nat.ptr(adjacency.col_indices), nat.ptr(adjacency.row_pointers), n, float(mu),
        nat.ptr(out), nat.ptr(triangles), nat.ptr(ws), ws.numel(), nat.stream_of(dev),
    )
    return out


def fugal_node_features(adjacency: Adjacency, mu: float = 1.0) -> torch.Tensor:
    """n x 8 FP32 rows scaled by `mu`: the 7 features of the original FUGAL
    (FUGAL/pred.py:24-110 with simple=False) -- degree, clustering, mean
    neighbour degree, mean neighbour clustering, ego-net edges, edges leaving
    the ego net, distinct nodes at distance 2 -- and a zero.  Simple graphs
    (no self-loops); the counts are exact."""
    n = adjacency.number_of_nodes()
    dev = adjacency.row_pointers.device
    out = torch.empty((n, 8), dtype=torch.float32, device=dev)
    ws = nat.workspace(dev, nat.query("cg_fugal_features_workspace", n), "features")
    nat.call(
        "cg_fugal_features", nat.ptr(adjacency.col_indices) if len(adjacency.col_indices) else None,
        nat.ptr(adjacency.row_pointers), n, float(mu), nat.ptr(out), nat.ptr(ws), ws.numel(), nat.stream_of(dev),
    )
    return out


def triangle_counts(adjacency: Adjacency) -> torch.Tensor:
    """t(v) = sum over neighbours u of |N(v) & N(u)| (twice the triangles
    through v), int32."""
    n = adjacency.number_of_nodes()
    dev = adjacency.row_pointers.device
    tri = torch.empty(n, dtype=torch.int32, device=dev)
    nat.call(
        "cg_graph_features", nat.ptr(adjacency.col_indices), nat.ptr(adjacency.row_pointers), n,
        None, None, nat.ptr(tri), nat.stream_of(dev),
    )
    return tri


def egonet_edges(adjacency: Adjacency, triangles: torch.Tensor | None = None) -> torch.Tensor:
    """Edges of each node's ego network, deg + triangles (FUGAL/pred.py:68-72)."""
    n = adjacency.number_of_nodes()
    dev = adjacency.row_pointers.device
    tri = triangle_counts(adjacency) if triangles is None else triangles
    out = torch.empty(n, dtype=torch.int32, device=dev)
    nat.call("cg_egonet_edges", nat.ptr(adjacency.row_pointers), nat.ptr(tri), n, nat.ptr(out), nat.stream_of(dev))
    return out


def extract_features_cuda(graph: nx.Graph | Adjacency, config: Config) -> torch.Tensor:
    """Unscaled n x 4 features on the device (cugal/feature_extraction.py:15-38)."""
    assert "cuda" in config.device
    features = node_features(_as_adjacency(graph, config), 1.0)
    if config.safe_mode:
        assert features[:, 1].isfinite().all(), "clustering tensor has NaN values"
        assert features[:, 0].isfinite().all(), "degrees tensor has NaN values"
    return features


def extract_features_nx(G: nx.Graph):
    """The reference's networkx feature path (cugal/feature_extraction.py:
    41-64) is its CPU path, which this library does not provide: the same
    four features come from the device kernels (extract_features_cuda,
    node_features), equal to networkx's values to FP32 rounding."""
    raise NotImplementedError(
        "extract_features_nx is the reference's CPU path; use extract_features_cuda / node_features on a ROCm device"
    )


@dataclass
class Features:
    """Features of source and target graph (already multiplied by mu)."""

    source: torch.Tensor
    target: torch.Tensor

    @classmethod
    def create(cls, source: nx.Graph | Adjacency, target: nx.Graph | Adjacency, config: Config) -> "Features":
        if "cuda" not in config.device:
            raise NotImplementedError("feature extraction runs on a ROCm device (config.device 'cuda')")
        if config.dtype != torch.float32:
            raise NotImplementedError("the feature kernels produce FP32 (config.dtype torch.float32)")
        if config.node_features == "cugal":
            extract = node_features
        elif config.node_features == "fugal7":
            extract = fugal_node_features
        else:
            raise ValueError(f"node_features must be 'cugal' or 'fugal7', got {config.node_features!r}")
        source_features = extract(_as_adjacency(source, config), config.mu)
        target_features = extract(_as_adjacency(target, config), config.mu)
        if config.safe_mode:
            assert source_features.isfinite().all(), "source feature tensor has NaN values"
            assert target_features.isfinite().all(), "target feature tensor has NaN values"
        return cls(source_features, target_features)

    def add_distance(self, out: torch.Tensor) -> torch.Tensor:
        """out += Euclidean distance matrix of the feature rows (the 4-feature
        rows of cuda_kernels/distance.cu; the gradient kernels recompute the
        distance of either feature set themselves)."""
        if self.source.shape[1] != 4:
            raise NotImplementedError("add_distance takes the 4-feature rows of the reference kernel")
        nat.require_device_tensor(out, "out", torch.float32)
        n = out.shape[0]
        # both contiguous copies stay referenced until the call returns (a
        # freed first temporary could be handed to the second)
        source, target = self.source.contiguous(), self.target.contiguous()
        nat.require_device_tensor(source, "source features", torch.float32, 4 * n)
        nat.require_device_tensor(target, "target features", torch.float32, 4 * n)
        nat.call("cg_add_distance", nat.ptr(source), nat.ptr(target), n, nat.ptr(out), nat.stream_of(out.device))
        return out

    def distance_matrix(self) -> torch.Tensor:
        """The n x n Euclidean distance matrix (cugal/feature_extraction.py:
        106-108 uses torch.cdist; this is the direct form)."""
        n = self.source.shape[0]
        out = torch.zeros((n, n), dtype=torch.float32, device=self.source.device)
        return self.add_distance(out)
