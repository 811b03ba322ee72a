"""Directed graphs on the dense path (cugal/pred.py:38-51, :106-110): the
general gradient -A^T P B - A P B^T + D + lambda (1 - 2P) (twice lambda on
the dense path) in cg_gradient_directed against the FP64 oracle, the
sparse_gradient branch with explicit transposes (cugal/pred.py:65-67), and
cugal() on directed pairs against the oracle pipeline, whose features
follow the reference's GPU path (a directed edge x -> y in row y of the
CSR)."""

import copy

import networkx as nx
import numpy as np
import pytest
import scipy.sparse
import torch

from cugal_amd import Config, HungarianMethod, SinkhornMethod, cugal
from cugal_amd.adjacency import Adjacency
from cugal_amd.config import Config as Cfg
from cugal_amd.feature_extraction import Features, node_features
from cugal_amd.pred import dense_gradient, sparse_gradient
from cugal_amd.profile import Profile
from cugal_amd.synthetic import alignment_accuracy, answer_columns
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu


def _pair(n, p, seed):
    ga = nx.gnp_random_graph(n, p, seed=seed, directed=True)
    gb = nx.gnp_random_graph(n, p, seed=seed + 50, directed=True)
    return ga, gb


def test_transposed_csr(device):
    g = nx.gnp_random_graph(300, 0.03, seed=1, directed=True)
    g.add_edge(5, 5)
    adj = Adjacency.from_graph(g, device)
    assert adj.directed
    M = nx.to_scipy_sparse_array(g, nodelist=range(300), format="csr")  # M[x][y] = 1 for x -> y
    for got, want in ((adj, scipy.sparse.csr_matrix(M.T)), (adj.transposed(), scipy.sparse.csr_matrix(M))):
        want.sort_indices()
        assert np.array_equal(got.row_pointers.cpu().numpy(), want.indptr)
        assert np.array_equal(got.col_indices.cpu().numpy(), want.indices)
    assert adj.transposed().transposed() is adj


@pytest.mark.parametrize("n,seed", [(256, 1), (999, 2), (1024, 3)])
def test_dense_gradient_of_directed_graphs_matches_oracle(device, n, seed):
    ga, gb = _pair(n, 0.03, seed)
    Ad = torch.tensor(nx.to_numpy_array(ga, nodelist=range(n)))
    Bd = torch.tensor(nx.to_numpy_array(gb, nodelist=range(n)))
    # features of the GPU path: the CSR rows of the directed graphs
    feats = Features(node_features(Adjacency.from_graph(ga, device), 0.5),
                     node_features(Adjacency.from_graph(gb, device), 0.5))
    P = torch.rand((n, n), generator=torch.Generator().manual_seed(seed))
    P /= P.sum(1, keepdim=True)
    lam = 5
    got = dense_gradient(Ad.float().to(device), Bd.float().to(device), P.to(device), feats, lam)
    fs, ft = feats.source.double().cpu(), feats.target.double().cpu()
    D = torch.sqrt(((fs[:, None, :] - ft[None, :, :]) ** 2).sum(-1))
    want = oracle.dense_gradient(Ad, Bd, P.double(), D, lam)
    scale = float(want.abs().max())
    np.testing.assert_allclose(got.double().cpu().numpy(), want.numpy(), rtol=0, atol=1e-5 * scale)
    # the sparse function's general branch: one lambda term
    A = Adjacency.from_graph(ga, device).transposed()
    B = Adjacency.from_graph(gb, device).transposed()
    got = sparse_gradient(A, B, A.transposed(), B.transposed(), P.to(device), feats, lam)
    want = -Ad.T @ P.double() @ Bd - Ad @ P.double() @ Bd.T + D + lam * (1 - 2 * P.double())
    np.testing.assert_allclose(got.double().cpu().numpy(), want.numpy(), rtol=0, atol=1e-5 * scale)


def test_symmetric_input_to_the_general_form_equals_the_symmetric_form(device):
    # A symmetric pair through the directed kernel gives -2 A P B: the two
    # kernels agree to FP32 rounding (the general one adds two products).
    n = 500
    ga = nx.watts_strogatz_graph(n, 6, 0.2, seed=4)
    gb = nx.barabasi_albert_graph(n, 3, seed=5)
    A, B = Adjacency.from_graph(ga, device), Adjacency.from_graph(gb, device)
    P = torch.rand((n, n), generator=torch.Generator(device=device).manual_seed(6), device=device)
    sym = sparse_gradient(A, B, A, B, P, None, 2)
    At = Adjacency(A.col_indices, A.row_pointers, directed=True)
    Bt = Adjacency(B.col_indices, B.row_pointers, directed=True)
    gen = sparse_gradient(A, B, At, Bt, P, None, 2)
    assert float((gen - sym).abs().max()) <= 2e-6 * float(sym.abs().max())


def _directed_instance(n, p, seed, drop=0.03):
    g = nx.gnp_random_graph(n, p, seed=seed, directed=True)
    perm = np.random.default_rng(seed).permutation(n)
    edges = [(int(perm[x]), int(perm[y])) for x, y in g.edges()]
    keep = np.random.default_rng(seed + 1).random(len(edges)) >= drop
    t = nx.DiGraph()
    t.add_nodes_from(range(n))
    t.add_edges_from(e for e, k in zip(edges, keep) if k)
    return g, t, perm


@pytest.mark.parametrize("method", ["LOG", "STANDARD"])
def test_directed_pipeline_matches_oracle(device, method):
    src, tgt, perm = _directed_instance(300, 0.03, 7)
    reg = 0.5 if method == "LOG" else 1.0
    gpu = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod[method],
                 sinkhorn_regularization=reg, mu=1.0, iter_count=3, frank_wolfe_iter_count=4,
                 use_sparse_adjacency=False, hungarian_method=HungarianMethod.DOUBLE_GREEDY)
    ref = oracle.OracleConfig(sinkhorn_regularization=reg, sinkhorn_method=method, mu=1.0, iter_count=3,
                              frank_wolfe_iter_count=4, use_sparse_adjacency=False)
    _, mapping = cugal(copy.deepcopy(src), copy.deepcopy(tgt), gpu, Profile())
    run = oracle.align(copy.deepcopy(src), copy.deepcopy(tgt), ref)
    acc = alignment_accuracy(perm, answer_columns(mapping))
    ref_acc = alignment_accuracy(perm, answer_columns(run.mapping))
    print(f"directed {method}: accuracy {acc:.4f} vs oracle {ref_acc:.4f}")
    assert abs(acc - ref_acc) <= 0.005 + 1e-12


def test_directed_fixed_iteration_quasi_permutation_matches_oracle(device):
    src, tgt, _ = _directed_instance(200, 0.05, 9)
    gpu = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
                 sinkhorn_regularization=1.0, sinkhorn_threshold=0.0, sinkhorn_iterations=30, mu=1.0,
                 iter_count=2, frank_wolfe_iter_count=3, use_sparse_adjacency=False,
                 hungarian_method=HungarianMethod.SCIPY)
    ref = oracle.OracleConfig(sinkhorn_regularization=1.0, sinkhorn_method="LOG", sinkhorn_threshold=0.0,
                              sinkhorn_iterations=30, mu=1.0, iter_count=2, frank_wolfe_iter_count=3,
                              use_sparse_adjacency=False, hungarian_method="SCIPY")
    P, mapping = cugal(copy.deepcopy(src), copy.deepcopy(tgt), gpu, Profile())
    run = oracle.align(copy.deepcopy(src), copy.deepcopy(tgt), ref)
    R = run.P.numpy()
    np.testing.assert_allclose(P, R, rtol=1e-5, atol=1e-11 * float(R.max()))
    assert [m for _, m in mapping] == [m for _, m in run.mapping]


def test_directed_graphs_refuse_the_sparse_path_and_the_matrix_cores(device):
    src, tgt, _ = _directed_instance(50, 0.1, 11)
    with pytest.raises(NotImplementedError, match="undirected"):
        cugal(copy.deepcopy(src), copy.deepcopy(tgt), Cfg(device="cuda:0", dtype=torch.float32,
                                                           use_sparse_adjacency=True), Profile())
    with pytest.raises(NotImplementedError):
        cugal(copy.deepcopy(src), copy.deepcopy(tgt), Cfg(device="cuda:0", dtype=torch.float32,
                                                           gradient_precision="bf16x2"), Profile())
