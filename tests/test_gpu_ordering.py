"""Relabelled graphs for the CSR gathers (cugal_amd/ordering.py): the
permuted copy, the gradient's equivariance under relabelling, and the
pipeline in either order against the oracle (test_gpu_pipeline.py)."""

import networkx as nx
import numpy as np
import pytest
import torch

from cugal_amd import _native as nat
from cugal_amd import ordering
from cugal_amd.adjacency import Adjacency
from cugal_amd.feature_extraction import Features, node_features
from cugal_amd.pred import _gradient_into

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1000, 1001])
def test_permute_copy_matches_indexing(device, n):
    g = torch.Generator(device=device).manual_seed(1)
    ld = n + 37
    src = torch.rand((n, ld), generator=g, device=device)[:, :n]
    rp = torch.randperm(n, generator=g, device=device)
    cp = torch.randperm(n, generator=g, device=device)
    for rmap, cmap in ((rp, cp), (None, cp), (rp, None), (None, None)):
        got = ordering.copy_back(src, rmap, cmap)
        want = src
        if rmap is not None:
            want = want[rmap]
        if cmap is not None:
            want = want[:, cmap]
        assert torch.equal(got, want.contiguous())


def test_gradient_is_equivariant_under_relabelling(device):
    # T A T^T (T P S^T) S B S^T = T (A P B) S^T, distance and regulariser
    # elementwise: the gradient of relabelled inputs is the relabelled
    # gradient, up to the order of the FP32 sums.
    n = 3000
    ga = nx.watts_strogatz_graph(n, 12, 0.1, seed=1)
    gb = nx.relabel_nodes(ga, dict(zip(range(n), np.random.default_rng(2).permutation(n).tolist())))
    A, B = Adjacency.from_graph(ga, device), Adjacency.from_graph(gb, device)
    feats = Features(node_features(A, 2.0), node_features(B, 2.0))
    P = torch.rand((n, n), generator=torch.Generator(device=device).manual_seed(3), device=device)
    P /= P.sum(1, keepdim=True)
    G = torch.empty((n, n), device=device)
    _gradient_into(G, A, B, P, feats, 4, 1.0)
    pa = torch.from_numpy(ordering.rcm_position(*ordering._host_csr(A))).to(device)
    pb = torch.from_numpy(ordering.rcm_position(*ordering._host_csr(B))).to(device)
    A2, B2 = ordering.relabel(A, pa), ordering.relabel(B, pb)
    f2 = Features(ordering.permute_rows(feats.source, pa), ordering.permute_rows(feats.target, pb))
    P2 = torch.empty_like(P)
    P2[pa[:, None], pb[None, :]] = P
    G2 = torch.empty_like(G)
    _gradient_into(G2, A2, B2, P2, f2, 4, 1.0)
    back = ordering.copy_back(G2, pa, pb)
    scale = float(G.abs().max())
    assert float((back - G).abs().max()) <= 2e-6 * scale
    # the relabelled CSR is the relabelled graph
    want = nx.relabel_nodes(gb, {v: int(pb[v]) for v in range(n)})
    ref = nx.to_scipy_sparse_array(want, nodelist=range(n), format="csr")
    ref.sort_indices()
    assert np.array_equal(B2.row_pointers.cpu().numpy(), ref.indptr)
    assert np.array_equal(B2.col_indices.cpu().numpy(), ref.indices)
