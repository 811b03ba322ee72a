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


def test_locality_order_choice(device):
    # A Watts-Strogatz graph (k = 20, 10 % rewired) under a random
    # relabelling: plain RCM keeps few edges within 64 labels (the
    # shortcuts spread its breadth-first levels); RCM of the edges that
    # close at least two triangles gets the ring order back.  "auto" picks
    # it for the shuffled copy and keeps the ring's own labels.
    n = 32768
    ring = nx.watts_strogatz_graph(n, 20, 0.1, seed=0)
    shuffled = nx.relabel_nodes(ring, dict(zip(range(n), np.random.default_rng(1).permutation(n).tolist())))
    a, b = Adjacency.from_graph(ring, device), Adjacency.from_graph(shuffled, device)
    rp, col = ordering._host_csr(b)
    plain = ordering.rcm_position(rp, col)
    filtered = ordering.filtered_rcm_position(b)
    assert np.array_equal(np.sort(filtered), np.arange(n))
    small = [ordering.window_locality(rp, col, p, 64) for p in (None, plain, filtered)]
    print(f"edges within 64 labels: own {small[0]:.3f}, RCM {small[1]:.3f}, filtered RCM {small[2]:.3f}")
    assert small[0] < 0.05 and small[1] < 0.5 < 0.85 < small[2]
    assert np.array_equal(ordering.choose_position(b, "auto"), filtered)
    assert ordering.choose_position(a, "auto") is None
    # the device's locality scores are the host counts exactly
    for adj in (a, b):
        arp, acol = ordering._host_csr(adj)
        positions = [None, ordering.rcm_position(arp, acol), filtered]
        got = ordering.device_locality_scores(adj, positions)
        assert got == [ordering.locality_score(arp, acol, p) for p in positions]
    # the common-neighbour counts against scipy, (M M) * M entrywise
    M = nx.to_scipy_sparse_array(shuffled, nodelist=range(n), format="csr")
    C = (M @ M).multiply(M).tocoo()
    lookup = {(int(r), int(c)): int(v) for r, c, v in zip(C.row, C.col, C.data)}
    common = torch.empty(b.col_indices.numel(), dtype=torch.int32, device=device)
    nat.call("cg_edge_common_neighbours", nat.ptr(b.col_indices), nat.ptr(b.row_pointers), n, nat.ptr(common),
             nat.stream_of(device))
    common = common.cpu().numpy()
    rows = np.repeat(np.arange(n), np.diff(rp))
    want = np.array([lookup.get((int(r), int(c)), 0) for r, c in zip(rows, col)])
    assert np.array_equal(common, want)
