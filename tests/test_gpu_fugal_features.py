"""The 7-feature set of the original FUGAL (FUGAL/pred.py:24-110, simple=False)
on the device (cg_fugal_features), the gradient with 8-float feature rows in
both forms, and the alignment with Config.node_features = "fugal7", against
the oracle."""

import copy

import networkx as nx
import numpy as np
import pytest
import torch

from cugal_amd import Config, HungarianMethod, SinkhornMethod, cugal
from cugal_amd.adjacency import Adjacency
from cugal_amd.feature_extraction import Features, fugal_node_features, node_features
from cugal_amd.pred import _gradient_into
from cugal_amd.synthetic import alignment_accuracy, answer_columns, erdos_renyi
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu

GRAPHS = {
    "er": lambda: nx.erdos_renyi_graph(300, 0.05, seed=1),
    "ba_hubs": lambda: nx.barabasi_albert_graph(600, 3, seed=2),
    "ws": lambda: nx.watts_strogatz_graph(400, 10, 0.1, seed=3),
    "petersen": nx.petersen_graph,
    "complete": lambda: nx.complete_graph(17),
    "star_isolated": lambda: nx.relabel_nodes(nx.star_graph(9), {i: i + 3 for i in range(10)}),
    "dense_hub": lambda: nx.compose(nx.erdos_renyi_graph(700, 0.3, seed=5), nx.star_graph(699)),
}


def _with_all_nodes(g):
    g = nx.Graph(g)
    g.add_nodes_from(range(max(g.nodes()) + 1))
    return g


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_fugal_features_match_oracle(device, name):
    g = _with_all_nodes(GRAPHS[name]())
    A = Adjacency.from_graph(g, device)
    mu = 0.75
    got = fugal_node_features(A, mu).cpu().numpy()
    want = oracle.extract_features_fugal(g)
    assert got.shape == (g.number_of_nodes(), 8) and np.all(got[:, 7] == 0)
    # The first four are cg_node_features' rows; the counts are exact.
    np.testing.assert_array_equal(got[:, :4], node_features(A, mu).cpu().numpy())
    np.testing.assert_array_equal(got[:, 4:7], (want[:, 4:7] * mu).astype(np.float32))


@pytest.mark.parametrize("precision", ["fp32", "bf16x2"])
@pytest.mark.parametrize("n,p", [(300, 0.05), (1024, 0.02)])
def test_gradient_with_fugal_features(device, precision, n, p):
    ga = nx.erdos_renyi_graph(n, p, seed=n)
    gb = nx.erdos_renyi_graph(n, p, seed=n + 1)
    A, B = Adjacency.from_graph(ga, device), Adjacency.from_graph(gb, device)
    feats = Features(fugal_node_features(A, 0.5), fugal_node_features(B, 0.5))
    P = torch.rand((n, n), generator=torch.Generator().manual_seed(n))
    P /= P.sum(1, keepdim=True)
    lam = 5
    out = torch.empty((n, n), device=device)
    _gradient_into(out, A, B, P.to(device), feats, lam, 1.0, precision=precision)
    Ad = torch.tensor(nx.to_numpy_array(ga, nodelist=range(n)))
    Bd = torch.tensor(nx.to_numpy_array(gb, nodelist=range(n)))
    fs = torch.tensor(oracle.extract_features_fugal(ga)) * 0.5
    ft = torch.tensor(oracle.extract_features_fugal(gb)) * 0.5
    D = torch.sqrt(((fs[:, None, :] - ft[None, :, :]) ** 2).sum(-1))
    ref = oracle.sparse_gradient(Ad, Bd, P.double(), D, lam)
    # The distance's seven squares in FP32 (relative 2e-6 of D) plus the
    # product's error (FP32 sums, or the hi+lo split on the matrix cores).
    apb = 2.0 * (Ad @ P.double() @ Bd)
    err = (out.double().cpu() - ref).abs()
    rel = 1e-6 if precision == "fp32" else 4e-5
    assert bool((err <= rel * apb + 4e-6 * (D + lam * (1 + 2 * P.double())) + 1e-6).all()), float(err.max())


def test_alignment_with_fugal_features_matches_oracle(device):
    inst = erdos_renyi(400, 0.02, 0.05, seed=0)
    gpu = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
                 sinkhorn_regularization=0.3, mu=1.0, iter_count=5, frank_wolfe_iter_count=5,
                 use_sparse_adjacency=True, hungarian_method=HungarianMethod.DOUBLE_GREEDY, node_features="fugal7")
    ref = oracle.OracleConfig(sinkhorn_regularization=0.3, sinkhorn_method="LOG", mu=1.0, iter_count=5,
                              frank_wolfe_iter_count=5, use_sparse_adjacency=True, node_features="fugal7")
    _, mapping = cugal(copy.deepcopy(inst.source), copy.deepcopy(inst.target), gpu)
    run = oracle.align(copy.deepcopy(inst.source), copy.deepcopy(inst.target), ref)
    acc = alignment_accuracy(inst.source_mapping, answer_columns(mapping))
    ref_acc = alignment_accuracy(inst.source_mapping, answer_columns(run.mapping))
    print(f"\n7-feature alignment: accuracy {acc:.4f} vs oracle {ref_acc:.4f}")
    assert abs(acc - ref_acc) <= 0.005 + 1e-12
    g = copy.deepcopy(inst.source)
    g.add_edge(0, 0)
    with pytest.raises(ValueError):
        cugal(g, copy.deepcopy(inst.target), gpu)
