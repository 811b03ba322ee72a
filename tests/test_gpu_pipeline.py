"""End-to-end cugal() on the device against the FP64 oracle pipeline
(cugal/pred.py:196-249) on seeded synthetic instances."""

import copy

import networkx as nx
import numpy as np
import pytest
import torch

from cugal_amd import Config, HungarianMethod, SinkhornMethod, cugal
from cugal_amd.profile import Phase, Profile
from cugal_amd.synthetic import alignment_accuracy, answer_columns, erdos_renyi, watts_strogatz
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu


def _configs(sparse, reg, mu, iters, fw, lap="DOUBLE_GREEDY", fused=True, thr=1e-3):
    gpu = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
                 sinkhorn_regularization=reg, sinkhorn_threshold=thr, mu=mu, iter_count=iters,
                 frank_wolfe_iter_count=fw, use_sparse_adjacency=sparse, sinkhorn_fused=fused,
                 hungarian_method=HungarianMethod[lap])
    ref = oracle.OracleConfig(sinkhorn_regularization=reg, sinkhorn_method="LOG", sinkhorn_threshold=thr, mu=mu,
                              iter_count=iters, frank_wolfe_iter_count=fw, use_sparse_adjacency=sparse,
                              hungarian_method=lap)
    return gpu, ref


@pytest.mark.parametrize("sparse", [True, False])
@pytest.mark.parametrize("fused", [True, False])
def test_pipeline_accuracy_matches_oracle(device, sparse, fused):
    inst = erdos_renyi(400, 0.02, 0.05, seed=0)
    gpu, ref = _configs(sparse, 0.3, 2.0, 5, 5, fused=fused)
    profile = Profile()
    P, mapping = cugal(copy.deepcopy(inst.source), copy.deepcopy(inst.target), gpu, profile)
    run = oracle.align(copy.deepcopy(inst.source), copy.deepcopy(inst.target), ref)
    acc = alignment_accuracy(inst.source_mapping, answer_columns(mapping))
    ref_acc = alignment_accuracy(inst.source_mapping, answer_columns(run.mapping))
    assert abs(acc - ref_acc) <= 0.005 + 1e-12, (acc, ref_acc)
    assert profile.frank_wolfe_iterations == 25
    assert len(profile.sinkhorn_profiles) == 25
    assert set(profile.phase_times) == {Phase.FEATURE_EXTRACTION, Phase.GRADIENT, Phase.SINKHORN, Phase.HUNGARIAN}
    got_iters = [p.iteration_count for p in profile.sinkhorn_profiles]
    # FP32 vs FP64 can move a threshold crossing by one check period
    assert all(abs(a - b) <= 10 for a, b in zip(got_iters, run.sinkhorn_iterations))


def test_reference_default_config_er1000(device):
    # Erdos-Renyi G(1000, 0.01), seed 0, 5% of edges removed and as many
    # non-edges added, seeded permutation; the reference's default Config
    # (cugal/config.py:59-120: STANDARD, reg 1.0, <= 500 iterations, thr
    # 1e-3, mu 0.5, 15 x 10 FW, dense adjacency with the doubled lambda term,
    # DOUBLE_GREEDY) on the device in FP32 against the FP64 oracle: node
    # accuracy within 0.5 pp.
    inst = erdos_renyi(1000, 0.01, 0.05, seed=0)
    gpu = Config(device="cuda:0", dtype=torch.float32)
    assert gpu.sinkhorn_method == SinkhornMethod.STANDARD and not gpu.use_sparse_adjacency
    profile = Profile()
    _, mapping = cugal(copy.deepcopy(inst.source), copy.deepcopy(inst.target), gpu, profile)
    run = oracle.align(copy.deepcopy(inst.source), copy.deepcopy(inst.target), oracle.OracleConfig())
    acc = alignment_accuracy(inst.source_mapping, answer_columns(mapping))
    ref_acc = alignment_accuracy(inst.source_mapping, answer_columns(run.mapping))
    print(f"\nER(1000) default config: accuracy {acc:.4f} vs oracle {ref_acc:.4f}; Sinkhorn iterations "
          f"{sum(p.iteration_count for p in profile.sinkhorn_profiles)} vs {sum(run.sinkhorn_iterations)}")
    assert abs(acc - ref_acc) <= 0.005 + 1e-12, (acc, ref_acc)
    assert profile.frank_wolfe_iterations == 150


def test_isomorphic_pair_recovered(device):
    inst = watts_strogatz(600, 10, 0.2, 0.0, seed=2)
    gpu, _ = _configs(True, 0.3, 2.0, 6, 6)
    _, mapping = cugal(copy.deepcopy(inst.source), copy.deepcopy(inst.target), gpu, Profile())
    assert alignment_accuracy(inst.source_mapping, answer_columns(mapping)) == 1.0


@pytest.mark.parametrize("order", ["natural", "rcm"])
def test_fixed_iteration_quasi_permutation_matches_oracle(device, order):
    # With a zero threshold every solve runs its full iteration count, so the
    # FP32 trajectory can be compared entry by entry with the FP64 one; with
    # "rcm" the loop runs on relabelled graphs and P comes back in the
    # graphs' own labels.
    inst = erdos_renyi(200, 0.05, 0.05, seed=4)
    gpu, ref = _configs(True, 1.0, 1.0, 2, 3, lap="SCIPY", thr=0.0)
    gpu = Config(**{**gpu.__dict__, "sinkhorn_iterations": 30, "node_order": order})
    ref.sinkhorn_iterations = 30
    P, mapping = cugal(copy.deepcopy(inst.source), copy.deepcopy(inst.target), gpu, Profile())
    run = oracle.align(copy.deepcopy(inst.source), copy.deepcopy(inst.target), ref)
    # Q within 1e-5 relative of the FP64 oracle (measured 1.4e-6), entries
    # below 1e-6 of the largest to 1e-5 of that floor.
    R = run.P.numpy()
    np.testing.assert_allclose(P, R, rtol=1e-5, atol=1e-11 * float(R.max()))
    assert [m for _, m in mapping] == [m for _, m in run.mapping]


def test_safe_mode_and_config_errors(device):
    inst = erdos_renyi(100, 0.05, 0.0, seed=1)
    gpu, _ = _configs(True, 0.5, 1.0, 1, 1)
    cfg = Config(**{**gpu.__dict__, "safe_mode": True})
    cugal(copy.deepcopy(inst.source), copy.deepcopy(inst.target), cfg, Profile())
    with pytest.raises(NotImplementedError):
        cugal(copy.deepcopy(inst.source), copy.deepcopy(inst.target),
              Config(**{**gpu.__dict__, "sinkhorn_method": SinkhornMethod.MIX}), Profile())


@pytest.mark.parametrize("form", ["log_fused_csr", "log_plain_csr", "standard_fused", "log_bf16x2_dense_lap"])
def test_alignment_is_deterministic(device, form):
    # Every reduction runs in a fixed order (no floating-point atomics), so
    # two runs on the same inputs give bitwise-equal P and the same mapping.
    from cugal_amd.pred import align_device
    from cugal_amd.pred import _adjacency_of

    inst = erdos_renyi(500, 0.02, 0.05, seed=8)
    gpu, _ = _configs(True, 0.3, 2.0, 3, 4, fused=form != "log_plain_csr")
    if form == "standard_fused":
        gpu = Config(**{**gpu.__dict__, "sinkhorn_method": SinkhornMethod.STANDARD, "sinkhorn_regularization": 1.0})
    if form == "log_bf16x2_dense_lap":
        gpu = Config(**{**gpu.__dict__, "gradient_precision": "bf16x2", "hungarian_method": HungarianMethod.DENSE})
    src, tgt = copy.deepcopy(inst.source), copy.deepcopy(inst.target)
    node_count = max(max(src.nodes()), max(tgt.nodes()))
    src.add_nodes_from(range(node_count))
    tgt.add_nodes_from(range(node_count))
    A, B = _adjacency_of(src, gpu), _adjacency_of(tgt, gpu)
    P1, c1 = align_device(A, B, gpu)
    P1 = P1.clone()
    P2, c2 = align_device(A, B, gpu)
    assert torch.equal(P1, P2)
    assert torch.equal(torch.as_tensor(c1).cpu(), torch.as_tensor(c2).cpu())


def test_safe_mode_reports_a_non_finite_plan(device):
    # Forced from the host: complete graphs give a cost near -2n, so at
    # reg 0.01 the STANDARD kernel exp(-C/reg) overflows FP32 to inf and the
    # plan becomes NaN (0 * inf), as in the reference; safe_mode stops with
    # the reference's message before the LAP.
    import networkx as nx

    g = nx.complete_graph(30)
    perm = np.random.default_rng(0).permutation(30)
    h = nx.relabel_nodes(g, {i: int(perm[i]) for i in range(30)})
    cfg = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.STANDARD,
                 sinkhorn_regularization=0.01, sinkhorn_iterations=20, iter_count=1, frank_wolfe_iter_count=1,
                 use_sparse_adjacency=False, safe_mode=True)
    with pytest.raises(AssertionError, match="quasi_permutation tensor has NaN values"):
        cugal(g, h, cfg, Profile())


@pytest.mark.parametrize("method", ["LOG", "STANDARD"])
def test_pipeline_at_an_odd_size(device, method):
    # n = 333: the fused Sinkhorn sweep, the float4 gathers and the float4
    # update all fall back to their scalar forms; same accuracy as the oracle.
    import networkx as nx

    from cugal_amd.synthetic import NoiseLevel, make_instance

    inst = make_instance(nx.erdos_renyi_graph(333, 0.03, seed=6), NoiseLevel(0.0, 0.05, True), seed=6)
    src, tgt = copy.deepcopy(inst.source), copy.deepcopy(inst.target)
    reg = 0.3 if method == "LOG" else 1.0
    gpu = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod[method],
                 sinkhorn_regularization=reg, mu=2.0, iter_count=4, frank_wolfe_iter_count=4,
                 use_sparse_adjacency=True)
    ref = oracle.OracleConfig(sinkhorn_regularization=reg, sinkhorn_method=method, mu=2.0, iter_count=4,
                              frank_wolfe_iter_count=4, use_sparse_adjacency=True)
    P, mapping = cugal(src, tgt, gpu, Profile())
    assert P.shape[0] % 4 != 0, P.shape
    run = oracle.align(copy.deepcopy(inst.source), copy.deepcopy(inst.target), ref)
    acc = alignment_accuracy(inst.source_mapping, answer_columns(mapping))
    ref_acc = alignment_accuracy(inst.source_mapping, answer_columns(run.mapping))
    assert abs(acc - ref_acc) <= 0.005 + 1e-12, (acc, ref_acc)



_TINY = {
    # one edge: the two nodes are interchangeable, so P stays uniform and the
    # LAP's lowest-index rule decides
    "edge": (lambda: nx.Graph([(0, 1)]), [1, 0]),
    # a tree with branches of lengths 1, 2 and 3 at node 2: no automorphism
    "asymmetric_tree": (lambda: nx.Graph([(0, 1), (1, 2), (2, 3), (3, 4), (4, 5), (2, 6)]), [4, 0, 6, 2, 5, 3, 1]),
    # one node: every kernel at n = 1
    "single_node": (lambda: nx.empty_graph(1), [0]),
    # no edges: zero features, P stays uniform, every LAP breaks exact ties
    "edgeless": (lambda: nx.empty_graph(5), [3, 0, 4, 1, 2]),
    # the same tree on labels 0-2, 4-7: cugal() pads the source with node 3
    # and the target with node 2 (n = 8, the fused Sinkhorn width)
    "padded_tree": (lambda: nx.Graph([(0, 1), (1, 2), (2, 4), (4, 5), (5, 6), (2, 7)]), [5, 7, 0, 2, 6, 1, 4, 3]),
}


@pytest.mark.parametrize("name", sorted(_TINY))
@pytest.mark.parametrize("method,lap", [("LOG", "DOUBLE_GREEDY"), ("STANDARD", "GREEDY"), ("LOG", "SCIPY"),
                                        ("STANDARD", "DENSE")])
def test_tiny_graphs_match_oracle(device, name, method, lap):
    # Graphs of 2 to 7 nodes: every kernel runs far below its tile size
    # (scalar fallbacks, partial wavefronts).  With a zero threshold each
    # solve runs its full count, so Q follows the FP64 oracle entry by entry
    # and the rounded mapping is the oracle's; DENSE (the exact device LAP)
    # is held to scipy's optimum, since a tie may have several optimal
    # assignments.
    make, perm = _TINY[name]
    source = make()
    source.remove_edges_from(list(nx.selfloop_edges(source)))
    target = nx.relabel_nodes(source, {i: perm[i] for i in source.nodes()})
    reg = 0.3 if method == "LOG" else 1.0
    gpu = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod[method],
                 sinkhorn_regularization=reg, sinkhorn_threshold=0.0, sinkhorn_iterations=40, mu=1.0, iter_count=3,
                 frank_wolfe_iter_count=3, use_sparse_adjacency=True, hungarian_method=HungarianMethod[lap])
    ref = oracle.OracleConfig(sinkhorn_regularization=reg, sinkhorn_method=method, sinkhorn_threshold=0.0, mu=1.0,
                              iter_count=3, frank_wolfe_iter_count=3, use_sparse_adjacency=True,
                              hungarian_method="SCIPY" if lap == "DENSE" else lap)
    ref.sinkhorn_iterations = 40
    P, mapping = cugal(copy.deepcopy(source), copy.deepcopy(target), gpu, Profile())
    run = oracle.align(copy.deepcopy(source), copy.deepcopy(target), ref)
    R = run.P.numpy()
    assert P.shape == R.shape
    np.testing.assert_allclose(P, R, rtol=1e-5, atol=1e-6 * float(R.max()))
    cols = [m for _, m in mapping]
    ref_cols = [m for _, m in run.mapping]
    if lap == "DENSE":
        assert sorted(cols) == list(range(len(cols)))
        rows = np.arange(len(cols))
        assert abs(R[rows, cols].sum() - R[rows, ref_cols].sum()) <= 1e-6 * len(cols)
    else:
        assert cols == ref_cols


@pytest.mark.parametrize("method,sparse,precision", [("LOG", True, "auto"), ("LOG", False, "fp32"),
                                                     ("STANDARD", True, "fp32"), ("STANDARD", False, "fp32"),
                                                     ("LOG", False, "bf16x2")])
def test_self_loops_match_oracle(device, method, sparse, precision):
    # A looped graph (G(40, 0.12), seed 3, no automorphism, with loops at 0,
    # 5, 33 and a lone looped node 40) against its relabelled copy: the
    # sparse path counts a loop twice (the reference's CSR), the dense path
    # once (nx.to_numpy_array); features are the GPU path's in both.  Fixed
    # iteration counts, so Q follows the FP64 oracle entry by entry.
    g = nx.Graph(nx.erdos_renyi_graph(40, 0.12, seed=3))
    g.add_edges_from([(0, 0), (5, 5), (33, 33), (40, 40)])
    g.add_nodes_from(range(41))
    perm = np.random.default_rng(3).permutation(41)
    h = nx.relabel_nodes(g, {i: int(perm[i]) for i in range(41)})
    reg = 0.3 if method == "LOG" else 1.0
    gpu = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod[method],
                 sinkhorn_regularization=reg, sinkhorn_threshold=0.0, sinkhorn_iterations=40, mu=1.0, iter_count=3,
                 frank_wolfe_iter_count=3, use_sparse_adjacency=sparse, gradient_precision=precision,
                 hungarian_method=HungarianMethod.SCIPY)
    ref = oracle.OracleConfig(sinkhorn_regularization=reg, sinkhorn_method=method, sinkhorn_threshold=0.0, mu=1.0,
                              iter_count=3, frank_wolfe_iter_count=3, use_sparse_adjacency=sparse,
                              hungarian_method="SCIPY")
    ref.sinkhorn_iterations = 40
    P, mapping = cugal(copy.deepcopy(g), copy.deepcopy(h), gpu, Profile())
    run = oracle.align(copy.deepcopy(g), copy.deepcopy(h), ref)
    R = run.P.numpy()
    # bf16x2: the hi + lo split of P carries 16 significant bits
    rtol = 1e-5 if precision != "bf16x2" else 1e-4
    np.testing.assert_allclose(P, R, rtol=rtol, atol=1e-6 * float(R.max()))
    assert [m for _, m in mapping] == [m for _, m in run.mapping]


def test_self_loops_refuse_the_matrix_core_form_on_the_sparse_path(device):
    g = nx.Graph(nx.karate_club_graph())
    g.add_edge(0, 0)
    cfg = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG, iter_count=1,
                 frank_wolfe_iter_count=1, use_sparse_adjacency=True, gradient_precision="bf16x2")
    with pytest.raises(NotImplementedError, match="self-loop counts twice"):
        cugal(copy.deepcopy(g), copy.deepcopy(g), cfg, Profile())
