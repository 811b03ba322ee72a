"""CSR build, node features and the gradient against the oracle
(cugal/adjacency.py, cugal/feature_extraction.py:41-64, cugal/pred.py:38-74)."""

import networkx as nx
import numpy as np
import pytest
import scipy.sparse
import torch

from cugal_amd import _native as nat
from cugal_amd.adjacency import Adjacency, edges_to_csr
from cugal_amd.feature_extraction import Features, egonet_edges, node_features, triangle_counts
from cugal_amd.config import Config
from cugal_amd.pred import _gradient_into
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu

GRAPHS = {
    "er": lambda: nx.erdos_renyi_graph(300, 0.05, seed=1),
    "ba": lambda: nx.barabasi_albert_graph(500, 5, seed=2),
    "ws": lambda: nx.watts_strogatz_graph(400, 10, 0.1, seed=3),
    "karate": nx.karate_club_graph,
    "petersen": nx.petersen_graph,
    "complete": lambda: nx.complete_graph(17),
    "star_isolated": lambda: nx.relabel_nodes(nx.star_graph(9), {i: i + 3 for i in range(10)}),
}


def _with_all_nodes(g):
    g = nx.Graph(g)
    n = max(g.nodes()) + 1
    g.add_nodes_from(range(n))
    return g


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_csr_matches_scipy(device, name):
    g = _with_all_nodes(GRAPHS[name]())
    adj = Adjacency.from_graph(g, device)
    ref = scipy.sparse.csr_matrix(nx.to_scipy_sparse_array(g, nodelist=range(g.number_of_nodes()), format="csr"))
    ref.sort_indices()
    assert np.array_equal(adj.row_pointers.cpu().numpy(), ref.indptr)
    assert np.array_equal(adj.col_indices.cpu().numpy(), ref.indices)


def test_csr_directed_and_empty_rows(device):
    edges = torch.tensor([[2, 5], [0, 5], [3, 1], [3, 1]], dtype=torch.int32, device=device)
    col, rp = edges_to_csr(edges, 8, symmetrize=False)
    # pair (x, y) gives A[y][x]; rows 0, 2-4, 6, 7 empty (incl. leading/trailing)
    assert rp.cpu().tolist() == [0, 0, 2, 2, 2, 2, 4, 4, 4]
    assert col.cpu().tolist() == [3, 3, 0, 2]


def test_csr_rejects_out_of_range(device):
    edges = torch.tensor([[0, 9]], dtype=torch.int32, device=device)
    with pytest.raises(nat.NativeError):
        edges_to_csr(edges, 4, symmetrize=True)


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_features_match_networkx(device, name):
    g = _with_all_nodes(GRAPHS[name]())
    adj = Adjacency.from_graph(g, device)
    tri = torch.empty(g.number_of_nodes(), dtype=torch.int32, device=device)
    got = node_features(adj, 1.0, tri).cpu().numpy()
    want = oracle.extract_features_nx(g)
    # degree exact; ratios are the FP32 rounding of the FP64 values
    assert np.array_equal(got[:, 0], want[:, 0].astype(np.float32))
    np.testing.assert_allclose(got[:, 1:], want[:, 1:].astype(np.float32), rtol=2 ** -23, atol=0)
    A = nx.to_numpy_array(g, nodelist=range(g.number_of_nodes()))
    assert np.array_equal(tri.cpu().numpy(), oracle.triangle_counts(A))
    ego = egonet_edges(adj, tri).cpu().numpy()
    assert np.array_equal(ego, oracle.egonet_edges(g))


def test_features_scaled_by_mu(device):
    g = _with_all_nodes(GRAPHS["ba"]())
    adj = Adjacency.from_graph(g, device)
    a = node_features(adj, 1.0)
    b = node_features(adj, 2.5)
    assert torch.equal(b, a * 2.5)


def test_distance_matches_direct_form(device):
    g = torch.Generator().manual_seed(4)
    fs = torch.rand((200, 4), generator=g) * 10
    ft = torch.rand((200, 4), generator=g) * 10
    f = Features(fs.to(device), ft.to(device))
    got = f.distance_matrix().double().cpu()
    want = torch.sqrt(((fs.double()[:, None, :] - ft.double()[None, :, :]) ** 2).sum(-1))
    np.testing.assert_allclose(got.numpy(), want.numpy(), rtol=1e-6, atol=1e-5)


@pytest.mark.parametrize("sparse", [True, False])
@pytest.mark.parametrize("n,seed,p,hub", [(256, 1, 0.02, False), (999, 2, 0.02, False), (1024, 3, 0.02, False),
                                          (512, 4, 0.25, False), (3000, 5, 0.002, True)])
def test_gradient_matches_oracle(device, sparse, n, seed, p, hub):
    # (512, p = 0.25): a 64-row tile lists ~8000 neighbours, several staged
    # index chunks; hub: node 0 is joined to every node, so its row alone
    # spans chunks.
    ga = nx.erdos_renyi_graph(n, p, seed=seed)
    gb = nx.erdos_renyi_graph(n, p, seed=seed + 100)
    if hub:
        ga.add_edges_from((0, j) for j in range(1, n))
        gb.add_edges_from((n - 1, j) for j in range(n - 1))
    A = Adjacency.from_graph(ga, device)
    B = Adjacency.from_graph(gb, device)
    gen = torch.Generator().manual_seed(seed)
    P = torch.rand((n, n), generator=gen)
    P /= P.sum(1, keepdim=True)
    cfg = Config(device="cuda", dtype=torch.float32, mu=0.5)
    feats = Features(node_features(A, cfg.mu), node_features(B, cfg.mu))
    lam = 7
    out = torch.empty((n, n), device=device)
    _gradient_into(out, A, B, P.to(device), feats, lam, 1.0 if sparse else 2.0)
    Ad = torch.tensor(nx.to_numpy_array(ga, nodelist=range(n)))
    Bd = torch.tensor(nx.to_numpy_array(gb, nodelist=range(n)))
    fs, ft = feats.source.double().cpu(), feats.target.double().cpu()
    D = torch.sqrt(((fs[:, None, :] - ft[None, :, :]) ** 2).sum(-1))
    Pd = P.double()
    ref = oracle.sparse_gradient(Ad, Bd, Pd, D, lam) if sparse else oracle.dense_gradient(Ad, Bd, Pd, D, lam)
    # FP32 sums of at most deg^2 terms of size <= 1 plus D and lambda terms:
    # relative 1e-5 of the magnitude, checked in absolute terms.
    scale = float(ref.abs().max())
    np.testing.assert_allclose(out.double().cpu().numpy(), ref.numpy(), rtol=0, atol=1e-5 * scale)


def _shuffled_rows(adj: Adjacency, seed: int) -> Adjacency:
    """The same CSR with every row's column list in a random order."""
    rp = adj.row_pointers.cpu().numpy()
    col = adj.col_indices.cpu().numpy().copy()
    rng = np.random.default_rng(seed)
    for r in range(len(rp) - 1):
        rng.shuffle(col[rp[r]:rp[r + 1]])
    return Adjacency(torch.from_numpy(col).to(adj.col_indices.device), adj.row_pointers.clone())


@pytest.mark.parametrize("n", [1024, 1000, 1001])
@pytest.mark.parametrize("variant", ["local", "unsorted", "long_row"])
def test_window_gradient_matches_oracle(device, n, variant):
    # Ring lattices with 10 % shortcuts in their own order: nearly every
    # neighbour lies within 16 rows of its 64-row tile, so both products
    # stage their windows in LDS (n = 1000: partial tiles; n = 1001, not a
    # multiple of 4: the plain gather).  "unsorted": every row's list
    # shuffled (its local run cannot be found by search, so the row is
    # gathered from global memory).  "long_row": one node joined to 100
    # others, which sends the graph to the plain gather.
    ga = nx.watts_strogatz_graph(n, 10, 0.1, seed=n)
    gb = nx.watts_strogatz_graph(n, 8, 0.1, seed=n + 1)
    if variant == "long_row":
        ga.add_edges_from((5, j) for j in range(300, 400))
    A = Adjacency.from_graph(ga, device)
    B = Adjacency.from_graph(gb, device)
    if variant == "unsorted":
        A, B = _shuffled_rows(A, 1), _shuffled_rows(B, 2)
    gen = torch.Generator().manual_seed(n)
    P = torch.rand((n, n), generator=gen)
    P /= P.sum(1, keepdim=True)
    feats = Features(node_features(Adjacency.from_graph(ga, device), 0.5),
                     node_features(Adjacency.from_graph(gb, device), 0.5))
    lam = 5
    out = torch.empty((n, n), device=device)
    _gradient_into(out, A, B, P.to(device), feats, lam, 1.0)
    Ad = torch.tensor(nx.to_numpy_array(ga, nodelist=range(n)))
    Bd = torch.tensor(nx.to_numpy_array(gb, nodelist=range(n)))
    fs, ft = feats.source.double().cpu(), feats.target.double().cpu()
    D = torch.sqrt(((fs[:, None, :] - ft[None, :, :]) ** 2).sum(-1))
    ref = oracle.sparse_gradient(Ad, Bd, P.double(), D, lam)
    scale = float(ref.abs().max())
    np.testing.assert_allclose(out.double().cpu().numpy(), ref.numpy(), rtol=0, atol=1e-5 * scale)


@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
@pytest.mark.parametrize("n", [1024, 1000])
def test_gradient_of_row_padded_P(device, precision, n):
    # The gradient of P given as a view with padded rows (gathered in place)
    # equals the gradient of the contiguous P bit for bit; a transposed view
    # is the matrix it shows.
    ga = nx.watts_strogatz_graph(n, 8, 0.2, seed=1)
    gb = nx.barabasi_albert_graph(n, 4, seed=2)
    A = Adjacency.from_graph(ga, device)
    B = Adjacency.from_graph(gb, device)
    cfg = Config(device="cuda", dtype=torch.float32, mu=0.5)
    feats = Features(node_features(A, cfg.mu), node_features(B, cfg.mu))
    P = torch.rand((n, n), generator=torch.Generator(device=device).manual_seed(3), device=device) / n
    ld = int(nat.query("cg_gradient_padded_ld", n))
    Pv = torch.empty((n, ld), device=device)[:, :n]
    Pv.copy_(P)
    want = torch.empty((n, n), device=device)
    got = torch.empty((n, n), device=device)
    _gradient_into(want, A, B, P, feats, 3, 1.0, precision)
    _gradient_into(got, A, B, Pv, feats, 3, 1.0, precision)
    assert torch.equal(got, want)
    _gradient_into(want, A, B, P.T.contiguous(), feats, 3, 1.0, precision)
    _gradient_into(got, A, B, P.T, feats, 3, 1.0, precision)
    assert torch.equal(got, want)


def test_adjacency_mul_matches_dense(device):
    n = 700
    g = nx.barabasi_albert_graph(n, 4, seed=9)
    A = Adjacency.from_graph(g, device)
    M = torch.randn((n, n), generator=torch.Generator().manual_seed(2))
    got = A.mul(M.to(device), negate_lhs=True).double().cpu()
    want = -torch.tensor(nx.to_numpy_array(g, nodelist=range(n))) @ M.double()
    np.testing.assert_allclose(got.numpy(), want.numpy(), atol=2e-4)
    # a transposed view is multiplied as the matrix it shows
    got_t = A.mul(M.to(device).T).double().cpu()
    want_t = torch.tensor(nx.to_numpy_array(g, nodelist=range(n))) @ M.double().T
    np.testing.assert_allclose(got_t.numpy(), want_t.numpy(), atol=2e-4)


def test_gradient_closed_form_uniform_P(device):
    # P = J/n  =>  A P B = d_A d_B^T / n  (no features, lambda 0).
    n = 512
    ga = nx.watts_strogatz_graph(n, 8, 0.2, seed=4)
    gb = nx.barabasi_albert_graph(n, 3, seed=5)
    A, B = Adjacency.from_graph(ga, device), Adjacency.from_graph(gb, device)
    P = torch.full((n, n), 1.0 / n, device=device)
    out = torch.empty_like(P)
    _gradient_into(out, A, B, P, None, 0, 1.0)
    da = torch.tensor([d for _, d in sorted(ga.degree())], dtype=torch.float64)
    db = torch.tensor([d for _, d in sorted(gb.degree())], dtype=torch.float64)
    np.testing.assert_allclose(out.double().cpu().numpy(), (-2 * da[:, None] * db[None, :] / n).numpy(), atol=1e-5)


def test_triangles_large_graph_invariant(device):
    # sum_v t(v) = 6 * #triangles (each triangle counted twice at each of
    # its 3 vertices); checked against networkx on a 32k-node graph.
    g = nx.watts_strogatz_graph(32768, 20, 0.1, seed=0)
    adj = Adjacency.from_graph(g, device)
    tri = triangle_counts(adj).cpu().numpy().astype(np.int64)
    assert tri.sum() == 6 * (sum(nx.triangles(g).values()) // 3)


def test_features_with_self_loops_follow_the_gpu_path(device):
    # Self-loops (one of the reference's datasets, in-arenas, has one): the
    # CSR lists a loop twice, as the reference's symmetrised edge list does,
    # and the features count CSR entries as the reference's GPU kernels do
    # (its networkx path ignores a loop for clustering and the means).
    # Node 40 has only a loop.
    g = nx.karate_club_graph()
    g.add_edges_from([(0, 0), (5, 5), (33, 33), (40, 40)])
    g = _with_all_nodes(g)
    adj = Adjacency.from_graph(g, device)
    assert adj.loops_twice
    rows = adj.row_pointers.cpu().numpy()
    cols = adj.col_indices.cpu().numpy()
    assert list(cols[rows[40]:rows[41]]) == [40, 40]
    once = Adjacency.from_graph(g, device, loops_once=True)
    assert not once.loops_twice
    assert once.col_indices.numel() == adj.col_indices.numel() - 4
    want_once = scipy.sparse.csr_matrix(nx.to_scipy_sparse_array(g, nodelist=range(g.number_of_nodes()), format="csr"))
    want_once.sort_indices()
    assert np.array_equal(once.row_pointers.cpu().numpy(), want_once.indptr)
    assert np.array_equal(once.col_indices.cpu().numpy(), want_once.indices)
    got = node_features(adj, 1.0).cpu().numpy()
    want = oracle.extract_features_csr(g)
    assert np.array_equal(got[:, 0], want[:, 0].astype(np.float32))
    np.testing.assert_allclose(got[:, 1:], want[:, 1:].astype(np.float32), rtol=2 ** -23, atol=0)
