"""The reference-named module (cuda_kernels.py -> cugal_amd/cuda_kernels.py):
each of the ten functions of cuda_kernels/kernels.cpp:32-70 called with the
reference's arguments on data, against networkx, scipy, numpy or an FP64
restatement of the kernel it replaces."""

import networkx as nx
import numpy as np
import pytest
import scipy.optimize
import scipy.sparse
import torch

import cuda_kernels

pytestmark = pytest.mark.gpu


def _graph_csr(g, device):
    """The reference's Adjacency.from_graph (cugal/adjacency.py:47-69):
    both directions listed in Python, flattened, then create_adjacency."""
    n = g.number_of_nodes()
    edges = list(nx.edges(g))
    edges += [(y, x) for x, y in edges]
    flat = torch.tensor([v for e in edges for v in e], dtype=torch.int32, device=device)
    col = torch.empty(len(edges), dtype=torch.int32, device=device)
    rp = torch.empty(n + 1, dtype=torch.int32, device=device)
    cuda_kernels.create_adjacency(flat, col, rp)
    return col, rp


def test_create_adjacency_symmetric_and_directed(device):
    g = nx.barabasi_albert_graph(300, 3, seed=1)
    g.add_node(300)  # an isolated last vertex
    col, rp = _graph_csr(g, device)
    ref = nx.to_scipy_sparse_array(g, nodelist=range(301), format="csr")
    ref.sort_indices()
    assert np.array_equal(rp.cpu().numpy(), ref.indptr)
    assert np.array_equal(col.cpu().numpy(), ref.indices)
    # directed pairs: (x, y) is the entry A[y][x]
    rng = np.random.default_rng(2)
    pairs = np.unique(rng.integers(0, 50, size=(400, 2)), axis=0)
    flat = torch.tensor(pairs.reshape(-1), dtype=torch.int32, device=device)
    col = torch.empty(len(pairs), dtype=torch.int32, device=device)
    rp = torch.empty(51, dtype=torch.int32, device=device)
    cuda_kernels.create_adjacency(flat, col, rp)
    M = scipy.sparse.csr_matrix((np.ones(len(pairs)), (pairs[:, 1], pairs[:, 0])), shape=(50, 50))
    M.sort_indices()
    assert np.array_equal(rp.cpu().numpy(), M.indptr)
    assert np.array_equal(col.cpu().numpy(), M.indices)


@pytest.mark.parametrize("negate", [False, True])
def test_adjacency_matmul(device, negate):
    g = nx.erdos_renyi_graph(257, 0.05, seed=3)
    col, rp = _graph_csr(g, device)
    M = torch.rand((257, 257), generator=torch.Generator().manual_seed(4))
    out = torch.empty((257, 257), device=device)
    cuda_kernels.adjacency_matmul(col, rp, M.to(device), out, negate)
    want = nx.to_numpy_array(g, nodelist=range(257)) @ M.double().numpy() * (-1 if negate else 1)
    np.testing.assert_allclose(out.double().cpu().numpy(), want, rtol=1e-6, atol=1e-6)


def test_graph_and_average_neighbor_features(device):
    g = nx.watts_strogatz_graph(400, 8, 0.2, seed=5)
    g.add_node(400)
    n = 401
    col, rp = _graph_csr(g, device)
    clustering = torch.zeros(n, device=device)
    degrees = torch.zeros(n, device=device)
    cuda_kernels.graph_features(col, rp, clustering, degrees)
    cl = nx.clustering(g)
    want_c = np.array([cl[v] for v in range(n)])
    want_d = np.array([g.degree(v) for v in range(n)], dtype=np.float64)
    assert np.array_equal(degrees.cpu().numpy(), want_d.astype(np.float32))
    np.testing.assert_allclose(clustering.cpu().numpy(), want_c.astype(np.float32), rtol=2 ** -23, atol=0)
    avg = torch.zeros(n, device=device)
    cuda_kernels.average_neighbor_features(col, rp, degrees, avg)
    want = np.array([np.mean([want_d[u] for u in g[v]]) if g.degree(v) else 0.0 for v in range(n)])
    np.testing.assert_allclose(avg.cpu().numpy(), want.astype(np.float32), rtol=2 ** -23, atol=0)


@pytest.mark.parametrize("momentum", [1.0, 0.7])
def test_sinkhorn_log_step_and_marginal(device, momentum):
    n = 300
    g = torch.Generator().manual_seed(6)
    K = -torch.rand((n, n), generator=g, dtype=torch.float64) * 8
    add = torch.randn(n, generator=g, dtype=torch.float64)
    out0 = torch.randn(n, generator=g, dtype=torch.float64)
    out = out0.float().to(device)
    cuda_kernels.sinkhorn_log_step(K.float().to(device), add.float().to(device), out, momentum)
    Kf, af = K.float().double(), add.float().double()
    want = (1 - momentum) * out0.float().double() - momentum * torch.logsumexp(Kf + af[None, :], dim=1)
    np.testing.assert_allclose(out.double().cpu().numpy(), want.numpy(), rtol=0, atol=2e-6)
    u = torch.randn(n, generator=g).float()
    v = torch.randn(n, generator=g).float()
    err = cuda_kernels.sinkhorn_log_marginal(K.float().to(device), u.to(device), v.to(device))
    want = float((torch.exp(Kf + v.double()[None, :] + u.double()[:, None]).sum(1) - 1).abs().sum())
    assert err == pytest.approx(want, rel=1e-5)


def test_add_distance_and_regularize(device):
    n = 129
    g = torch.Generator().manual_seed(7)
    fs = torch.rand((n, 4), generator=g) * 5
    ft = torch.rand((n, 4), generator=g) * 5
    base = torch.rand((n, n), generator=g)
    out = base.to(device)
    cuda_kernels.add_distance(fs.to(device), ft.to(device), out)
    want = base.double() + torch.sqrt(((fs.double()[:, None, :] - ft.double()[None, :, :]) ** 2).sum(-1))
    np.testing.assert_allclose(out.double().cpu().numpy(), want.numpy(), rtol=1e-6, atol=1e-6)
    P = torch.rand((n, n), generator=g)
    G = base.clone().to(device)
    cuda_kernels.regularize(G, P.to(device), 5)
    want = base.double() + 5 - 10 * P.double()
    np.testing.assert_allclose(G.double().cpu().numpy(), want.numpy(), rtol=0, atol=4e-6)


def test_update_quasi_permutation_log(device):
    n, reg, alpha = 200, 0.4, 0.25
    g = torch.Generator().manual_seed(8)
    C = torch.rand((n, n), generator=g) * 3
    K = (C / -reg).float()
    u = -torch.rand(n, generator=g)
    v = -torch.rand(n, generator=g)
    P0 = torch.rand((n, n), generator=g) / n
    P = P0.to(device)
    gap = cuda_kernels.update_quasi_permutation_log(P, K.to(device), u.to(device), v.to(device), alpha, reg)
    plan = torch.exp(K.double() + u.double()[:, None] + v.double()[None, :])
    want = P0.double() + alpha * (plan - P0.double())
    np.testing.assert_allclose(P.double().cpu().numpy(), want.numpy(), rtol=1e-5, atol=1e-9)
    want_gap = abs(float(((plan - P0.double()) * (K.double() * -reg)).sum()))
    assert gap == pytest.approx(want_gap, rel=1e-5)


def test_dense_hungarian(device):
    n = 120
    cost = torch.rand((n, n), generator=torch.Generator().manual_seed(9))
    out = torch.empty(n, dtype=torch.int32)
    dev_cost = cost.to(device)
    cuda_kernels.dense_hungarian(dev_cost, out)
    _, want = scipy.optimize.linear_sum_assignment(cost.double().numpy())
    assert np.array_equal(out.numpy(), want)
    assert torch.equal(dev_cost.cpu(), cost)
