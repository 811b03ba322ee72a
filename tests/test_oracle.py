"""Pins of the FP64 oracle (oracle/fugal.py) against facts that do not come
from the oracle itself: closed forms, brute force, networkx and scipy."""

import itertools

import networkx as nx
import numpy as np
import pytest
import torch

from oracle import fugal as oracle


def _adj(g):
    return nx.to_numpy_array(g, nodelist=sorted(g.nodes()), weight=None)


# ---------------------------------------------------------------- features --

@pytest.mark.parametrize("n", [3, 5, 9])
def test_features_complete_graph(n):
    f = oracle.extract_features_nx(nx.complete_graph(n))
    np.testing.assert_array_equal(f, np.tile([n - 1, 1.0, n - 1, 1.0], (n, 1)))


def test_features_cycle_and_star():
    f = oracle.extract_features_nx(nx.cycle_graph(7))
    np.testing.assert_array_equal(f, np.tile([2, 0.0, 2, 0.0], (7, 1)))
    s = oracle.extract_features_nx(nx.star_graph(5))  # centre 0, leaves 1..5
    np.testing.assert_array_equal(s[0], [5, 0.0, 1.0, 0.0])
    np.testing.assert_array_equal(s[1:], np.tile([1, 0.0, 5.0, 0.0], (5, 1)))


def test_features_isolated_node_is_zero():
    g = nx.path_graph(4)
    g.add_node(4)
    f = oracle.extract_features_nx(g)
    np.testing.assert_array_equal(f[4], [0, 0, 0, 0])


def _brute_triangles(A):
    n = A.shape[0]
    t = np.zeros(n, dtype=np.int64)
    for a, b, c in itertools.combinations(range(n), 3):
        if A[a, b] and A[b, c] and A[a, c]:
            t[[a, b, c]] += 1
    return t


@pytest.mark.parametrize("g", [nx.karate_club_graph(), nx.petersen_graph(),
                               nx.erdos_renyi_graph(40, 0.2, seed=1), nx.watts_strogatz_graph(30, 6, 0.3, seed=2)])
def test_triangles_and_clustering_brute_force(g):
    A = _adj(g)
    tri = _brute_triangles(A)
    assert np.array_equal(oracle.triangle_counts(A), 2 * tri)
    d = A.sum(1)
    with np.errstate(invalid="ignore", divide="ignore"):
        c = np.where(d >= 2, 2 * tri / (d * (d - 1)), 0.0)
    f = oracle.extract_features_nx(g)
    np.testing.assert_allclose(f[:, 1], c, rtol=1e-15, atol=0)
    # neighbour means from adjacency products
    with np.errstate(invalid="ignore", divide="ignore"):
        nd = np.where(d > 0, (A @ d) / d, 0.0)
        nc = np.where(d > 0, (A @ c) / d, 0.0)
    np.testing.assert_allclose(f[:, 2], nd, rtol=1e-14)
    np.testing.assert_allclose(f[:, 3], nc, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("g", [nx.karate_club_graph(), nx.petersen_graph(), nx.star_graph(7),
                               nx.erdos_renyi_graph(60, 0.1, seed=3), nx.watts_strogatz_graph(30, 6, 0.3, seed=2)])
def test_gpu_path_features_equal_networkx_on_simple_graphs(g):
    g = nx.Graph(g)
    g.add_nodes_from(range(max(g.nodes()) + 2))  # one isolated node
    np.testing.assert_allclose(oracle.extract_features_csr(g), oracle.extract_features_nx(g), rtol=1e-14, atol=0)


def _loop_graph():
    g = nx.Graph(nx.karate_club_graph())
    g.add_edges_from([(0, 0), (5, 5), (33, 33), (40, 40)])
    g.add_nodes_from(range(41))
    return g


# networkx's path takes the mean of an empty list for the lone looped node
# and maps the NaN to 0, as the reference does
@pytest.mark.filterwarnings("ignore:Mean of empty slice:RuntimeWarning")
@pytest.mark.filterwarnings("ignore:invalid value encountered:RuntimeWarning")
def test_gpu_path_features_with_self_loops():
    # Entry-count matrix M (a loop is two entries of its row); the sorted
    # merge of two rows counts min(M[v,w], M[k,w]) matches per w, and row v
    # visits k M[v,k] times.
    g = _loop_graph()
    n = g.number_of_nodes()
    M = np.zeros((n, n), dtype=np.int64)
    for x, y in g.edges():
        M[x, y] += 1
        M[y, x] += 1
    t = np.array([sum(M[v, k] * np.minimum(M[v], M[k]).sum() for k in range(n)) for v in range(n)])
    d = M.sum(1).astype(np.float64)
    clus = np.where((t == 0) | (d < 2), 0.0, t / np.maximum(d * (d - 1), 1))
    nd = np.where(d > 0, (M @ d) / np.maximum(d, 1), 0.0)
    nc = np.where(d > 0, (M @ clus) / np.maximum(d, 1), 0.0)
    f = oracle.extract_features_csr(g)
    np.testing.assert_array_equal(f[:, 0], d)
    np.testing.assert_allclose(f[:, 1], clus, rtol=1e-15, atol=0)
    np.testing.assert_allclose(f[:, 2], nd, rtol=1e-14, atol=0)
    np.testing.assert_allclose(f[:, 3], nc, rtol=1e-14, atol=1e-15)
    # a lone looped vertex: row [v, v], t = 2 + 2, clustering 4 / (2 * 1)
    np.testing.assert_array_equal(f[40], [2.0, 2.0, 2.0, 2.0])
    # and the CPU path disagrees there: networkx ignores the loop
    np.testing.assert_array_equal(oracle.extract_features_nx(g)[40], [2.0, 0.0, 0.0, 0.0])


@pytest.mark.filterwarnings("ignore:Mean of empty slice:RuntimeWarning")
@pytest.mark.filterwarnings("ignore:invalid value encountered:RuntimeWarning")
def test_gpu_path_features_of_a_directed_graph():
    # The reference's CSR of a directed graph puts edge x -> y in row y
    # (create_adjacency reads the pair (x, y) as the entry [y][x]), so the
    # GPU path's features are those of M = A^T (A = nx.to_numpy_array):
    # degree = in-degree, t(v) = sum_k M[v,k] sum_w min(M[v,w], M[k,w]), and
    # means over predecessors.
    g = nx.gnp_random_graph(60, 0.08, seed=5, directed=True)
    g.add_edge(7, 7)
    n = g.number_of_nodes()
    A = nx.to_numpy_array(g, nodelist=range(n)).astype(np.int64)
    M = A.T
    t = np.array([sum(M[v, k] * np.minimum(M[v], M[k]).sum() for k in range(n)) for v in range(n)])
    d = M.sum(1).astype(np.float64)
    assert np.array_equal(d, [g.in_degree(v) for v in range(n)])
    clus = np.where((t == 0) | (d < 2), 0.0, t / np.maximum(d * (d - 1), 1))
    nd = np.where(d > 0, (M @ d) / np.maximum(d, 1), 0.0)
    nc = np.where(d > 0, (M @ clus) / np.maximum(d, 1), 0.0)
    f = oracle.extract_features_csr(g)
    np.testing.assert_array_equal(f[:, 0], d)
    np.testing.assert_allclose(f[:, 1], clus, rtol=1e-15, atol=0)
    np.testing.assert_allclose(f[:, 2], nd, rtol=1e-14, atol=0)
    np.testing.assert_allclose(f[:, 3], nc, rtol=1e-14, atol=1e-15)


def test_dense_gradient_of_a_directed_pair_brute_force():
    # -A^T P B - A P B^T entry by entry from the edge lists (x -> y is
    # A[x][y]); with P the permutation k -> perm[k] both terms count edge
    # pairs: (A^T P B)_ij = #{k : k -> i in ga, perm[k] -> j in gb} and
    # (A P B^T)_ij = #{k : i -> k in ga, j -> perm[k] in gb}.
    ga = nx.gnp_random_graph(12, 0.3, seed=1, directed=True)
    gb = nx.gnp_random_graph(12, 0.3, seed=2, directed=True)
    n = 12
    A = torch.tensor(nx.to_numpy_array(ga, nodelist=range(n)), dtype=torch.float64)
    B = torch.tensor(nx.to_numpy_array(gb, nodelist=range(n)), dtype=torch.float64)
    perm = np.random.default_rng(3).permutation(n)
    P = torch.zeros((n, n), dtype=torch.float64)
    P[torch.arange(n), torch.as_tensor(perm)] = 1.0
    G = oracle.dense_gradient(A, B, P, torch.zeros((n, n), dtype=torch.float64), 0)
    want = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            t1 = sum(1 for k in range(n) if ga.has_edge(k, i) and gb.has_edge(perm[k], j))
            t2 = sum(1 for k in range(n) if ga.has_edge(i, k) and gb.has_edge(j, perm[k]))
            want[i, j] = -(t1 + t2)
    np.testing.assert_array_equal(G.numpy(), want)


def test_adjacency_counts_a_loop_once_dense_twice_sparse():
    g = _loop_graph()
    dense = oracle.adjacency_matrix(g, sparse=False).numpy()
    sparse = oracle.adjacency_matrix(g, sparse=True).numpy()
    off = ~np.eye(g.number_of_nodes(), dtype=bool)
    np.testing.assert_array_equal(dense[off], sparse[off])
    for v in (0, 5, 33, 40):
        assert dense[v, v] == 1.0 and sparse[v, v] == 2.0
    assert np.count_nonzero(np.diag(sparse)) == 4


def test_karate_and_petersen_totals():
    assert oracle.triangle_counts(_adj(nx.karate_club_graph())).sum() == 6 * 45
    assert oracle.triangle_counts(_adj(nx.petersen_graph())).sum() == 0


@pytest.mark.parametrize("g", [nx.karate_club_graph(), nx.erdos_renyi_graph(50, 0.15, seed=4)])
def test_egonet_edges_is_degree_plus_triangles(g):
    A = _adj(g)
    assert np.array_equal(oracle.egonet_edges(g), A.sum(1).astype(np.int64) + _brute_triangles(A))


def test_distance_matrix_direct_form():
    g = torch.Generator().manual_seed(0)
    a = torch.rand((60, 4), generator=g, dtype=torch.float64) * 7
    b = torch.rand((60, 4), generator=g, dtype=torch.float64) * 7
    direct = np.sqrt(((a.numpy()[:, None, :] - b.numpy()[None, :, :]) ** 2).sum(-1))
    np.testing.assert_allclose(oracle.distance_matrix(a, b).numpy(), direct, atol=1e-9)


# ---------------------------------------------------------------- gradient --

def _pair(seed, n=30):
    ga = nx.erdos_renyi_graph(n, 0.2, seed=seed)
    gb = nx.erdos_renyi_graph(n, 0.2, seed=seed + 1)
    return torch.tensor(_adj(ga)), torch.tensor(_adj(gb))


def test_gradient_closed_form_uniform():
    A, B = _pair(1)
    n = A.shape[0]
    P = torch.full((n, n), 1.0 / n, dtype=torch.float64)
    G = oracle.sparse_gradient(A, B, P, torch.zeros_like(P), 0)
    da, db = A.sum(1), B.sum(1)
    np.testing.assert_allclose(G.numpy(), (-2 * torch.outer(da, db) / n).numpy(), atol=1e-12)


def test_gradient_at_permutation_brute_force():
    # (A Pi B)_ij = sum_k A_ik B_{pi(k) j}: count of neighbours k of i whose
    # image pi(k) is a neighbour of j.
    A, B = _pair(3, n=12)
    n = A.shape[0]
    perm = np.random.RandomState(0).permutation(n)
    Pi = torch.zeros((n, n), dtype=torch.float64)
    Pi[torch.arange(n), torch.tensor(perm)] = 1
    G = oracle.sparse_gradient(A, B, Pi, torch.zeros_like(Pi), 0)
    want = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            want[i, j] = -2 * sum(1 for k in range(n) if A[i, k] and B[perm[k], j])
    np.testing.assert_array_equal(G.numpy(), want)


def test_dense_minus_sparse_is_one_regulariser():
    A, B = _pair(5)
    n = A.shape[0]
    P = torch.rand((n, n), dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    D = torch.rand((n, n), dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    lam = 3
    diff = oracle.dense_gradient(A, B, P, D, lam) - oracle.sparse_gradient(A, B, P, D, lam)
    np.testing.assert_allclose(diff.numpy(), (lam * (1 - 2 * P)).numpy(), atol=1e-12)


# ---------------------------------------------------------------- sinkhorn --

def test_sinkhorn_separable_cost_uniform_plan():
    n = 40
    g = torch.Generator().manual_seed(0)
    C = torch.rand(n, 1, generator=g, dtype=torch.float64) * 3 + torch.rand(1, n, generator=g, dtype=torch.float64) * 3
    for method in ("LOG", "STANDARD"):
        st = oracle.SinkhornState(n, method)
        if method == "LOG":
            K, u, v, _ = st.solve_log(C, 0.5, 3, 0.0, 10)
            plan = oracle.plan_log(K, u, v)
        else:
            K, u, v, _ = st.solve_standard(C, 0.5, 3, 0.0, 10)
            plan = oracle.plan_standard(K, u, v)
        np.testing.assert_allclose(plan.numpy(), np.full((n, n), 1.0 / n), rtol=1e-12)


def test_sinkhorn_two_by_two_closed_form():
    # C = [[0, c], [c, 0]]: K = [[1, e], [e, 1]], e = exp(-c/reg), whose
    # doubly stochastic scaling is K / (1 + e).
    c, reg = 1.3, 0.7
    C = torch.tensor([[0.0, c], [c, 0.0]], dtype=torch.float64)
    st = oracle.SinkhornState(2, "LOG")
    K, u, v, _ = st.solve_log(C, reg, 50, 0.0, 10)
    e = np.exp(-c / reg)
    np.testing.assert_allclose(oracle.plan_log(K, u, v).numpy(), np.array([[1, e], [e, 1]]) / (1 + e), rtol=1e-12)


def test_log_and_standard_agree_and_are_doubly_stochastic():
    n = 64
    C = torch.rand((n, n), dtype=torch.float64, generator=torch.Generator().manual_seed(3)) * 5
    a = oracle.SinkhornState(n, "LOG")
    K1, u1, v1, r1 = a.solve_log(C, 1.0, 500, 1e-9, 10)
    b = oracle.SinkhornState(n, "STANDARD")
    K2, u2, v2, r2 = b.solve_standard(C, 1.0, 500, 1e-9, 10)
    p1, p2 = oracle.plan_log(K1, u1, v1), oracle.plan_standard(K2, u2, v2)
    np.testing.assert_allclose(p1.numpy(), p2.numpy(), rtol=1e-7, atol=1e-12)
    for p in (p1, p2):
        np.testing.assert_allclose(p.sum(0).numpy(), np.ones(n), atol=1e-8)
        np.testing.assert_allclose(p.sum(1).numpy(), np.ones(n), atol=1e-8)
    assert r1.errors[-1] < 1e-9 and r2.errors[-1] < 1e-9


def test_small_regularisation_approaches_lap():
    import scipy.optimize

    n = 12
    C = torch.rand((n, n), dtype=torch.float64, generator=torch.Generator().manual_seed(7))
    _, col = scipy.optimize.linear_sum_assignment(C.numpy())
    st = oracle.SinkhornState(n, "LOG")
    K, u, v, _ = st.solve_log(C, 0.005, 5000, 1e-10, 10)
    plan = oracle.plan_log(K, u, v)
    assert np.array_equal(plan.argmax(1).numpy(), col)


def test_log_marginal_quirk_corrupts_kernel():
    # The reference's in-place marginal check (cugal/sinkhorn.py:67-74) leaves
    # K changed; the default path does not.
    n = 10
    C = torch.rand((n, n), dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    st = oracle.SinkhornState(n, "LOG")
    K = -C / 1.0
    before = K.clone()
    st.marginal_error_log(K, quirk=False)
    assert torch.equal(K, before)
    st.marginal_error_log(K, quirk=True)
    assert not torch.equal(K, before)


def test_lehmann_momentum_rule():
    st = oracle.SinkhornState(4, "LOG")
    assert st.lehmann_momentum(None, 10, [1.0, 0.5]) == 1.0
    assert st.lehmann_momentum(40, 10, [1.0, 0.5]) == 1.0  # 2*10 < 40
    w = st.lehmann_momentum(20, 10, [1.0, 0.5])            # ratio 0.5 / 1.0 = 0.5
    assert w == pytest.approx(2.0 / (1.0 + np.sqrt(1 - 0.5 ** 0.1)))
    assert 1.0 < w < 2.0


# ------------------------------------------------------------- update, LAP --

def test_update_alpha_limits():
    n = 8
    P = torch.rand((n, n), dtype=torch.float64)
    plan = torch.rand((n, n), dtype=torch.float64)
    Q = P.clone()
    oracle.update_quasi_permutation(Q, plan, 1.0)
    assert torch.equal(Q, plan)
    Q = P.clone()
    oracle.update_quasi_permutation(Q, plan, 0.0)
    assert torch.equal(Q, P)


def test_update_keeps_doubly_stochastic():
    n = 16
    st = oracle.SinkhornState(n, "LOG")
    C = torch.rand((n, n), dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    K, u, v, _ = st.solve_log(C, 0.5, 1000, 1e-12, 10)
    P = torch.full((n, n), 1.0 / n, dtype=torch.float64)
    oracle.update_quasi_permutation(P, oracle.plan_log(K, u, v), 2.0 / 3.0)
    np.testing.assert_allclose(P.sum(0).numpy(), np.ones(n), atol=1e-9)
    np.testing.assert_allclose(P.sum(1).numpy(), np.ones(n), atol=1e-9)


@pytest.mark.parametrize("alpha", [1.0, 0.5, 2.0 / 7.0])
def test_duality_gap_closed_forms(alpha):
    # From P = J/n toward a permutation plan Pi, step = alpha (Pi - J/n), so
    # the gap |sum G (step / alpha)| (cugal/pred.py:140-141) is
    # |sum_i G[i, pi(i)] - (1/n) sum_ij G_ij|, whatever alpha is.
    n = 9
    g = torch.Generator().manual_seed(11)
    G = torch.randn((n, n), dtype=torch.float64, generator=g)
    perm = torch.randperm(n, generator=g)
    Pi = torch.zeros((n, n), dtype=torch.float64)
    Pi[torch.arange(n), perm] = 1.0
    P = torch.full((n, n), 1.0 / n, dtype=torch.float64)
    step = oracle.update_quasi_permutation(P, Pi, alpha)
    want = abs(float(G[torch.arange(n), perm].sum()) - float(G.sum()) / n)
    assert oracle.duality_gap(G, step, alpha) == pytest.approx(want, rel=1e-12)
    # a hand case: G = [[1, 2], [3, 4]], P = I, plan = the swap: the gap is
    # |(2 + 3) - (1 + 4)| = 0; with G = [[1, 5], [3, 4]] it is |8 - 5| = 3
    P2 = torch.eye(2, dtype=torch.float64)
    swap = torch.tensor([[0.0, 1.0], [1.0, 0.0]], dtype=torch.float64)
    step2 = oracle.update_quasi_permutation(P2.clone(), swap, 0.25)
    assert oracle.duality_gap(torch.tensor([[1.0, 2.0], [3.0, 4.0]], dtype=torch.float64), step2, 0.25) == 0.0
    assert oracle.duality_gap(torch.tensor([[1.0, 5.0], [3.0, 4.0]], dtype=torch.float64), step2,
                              0.25) == pytest.approx(3.0, rel=1e-15)


def _brute_lap(M):
    n = M.shape[0]
    best, arg = -np.inf, None
    for perm in itertools.permutations(range(n)):
        s = M[np.arange(n), perm].sum()
        if s > best:
            best, arg = s, perm
    return np.array(arg)


@pytest.mark.parametrize("seed", range(5))
def test_scipy_lap_brute_force(seed):
    M = torch.rand((7, 7), dtype=torch.float64, generator=torch.Generator().manual_seed(seed))
    assert np.array_equal(oracle.scipy_lap(M), _brute_lap(M.numpy()))


@pytest.mark.parametrize("method", ["GREEDY", "DOUBLE_GREEDY", "RAND", "PARALLEL_GREEDY"])
def test_greedy_optimal_on_dominant_matrix(method):
    # With one entry per row larger than the rest of its row and column,
    # greedy is optimal in any row order, so it equals the brute-force LAP.
    n = 7
    g = torch.Generator().manual_seed(1)
    perm = torch.randperm(n, generator=g)
    M = torch.rand((n, n), generator=g, dtype=torch.float64) * 0.1
    M[torch.arange(n), perm] = 1.0 + torch.rand(n, generator=g, dtype=torch.float64)
    want = _brute_lap(M.numpy())
    order = torch.randperm(n, generator=g).tolist()
    assert oracle.greedy_lap(M.clone(), method, order=order) == list(want)


@pytest.mark.parametrize("seed", range(4))
def test_rand_with_fixed_orders_is_greedy_and_double_greedy(seed):
    g = torch.Generator().manual_seed(seed)
    M = torch.rand((9, 9), generator=g, dtype=torch.float64)
    assert oracle.greedy_lap(M.clone(), "RAND", order=list(range(9))) == oracle.greedy_lap(M.clone(), "GREEDY")
    sorted_rows = sorted(range(9), key=lambda r: -float(M[r].max()))
    assert oracle.greedy_lap(M.clone(), "RAND", order=sorted_rows) == oracle.greedy_lap(M.clone(), "DOUBLE_GREEDY")


def test_parallel_greedy_regimes():
    g = torch.Generator().manual_seed(3)
    # No row maximum above 0.5: the walk alone, i.e. DOUBLE_GREEDY.
    M = torch.rand((8, 8), generator=g, dtype=torch.float64) * 0.5
    assert oracle.greedy_lap(M.clone(), "PARALLEL_GREEDY") == oracle.greedy_lap(M.clone(), "DOUBLE_GREEDY")
    # Every maximum above 0.5 on distinct columns: the plain row argmax, and
    # P is left unmasked.
    perm = torch.randperm(8, generator=g)
    M = torch.rand((8, 8), generator=g, dtype=torch.float64) * 0.5
    M[torch.arange(8), perm] = 0.6
    Q = M.clone()
    assert oracle.greedy_lap(Q, "PARALLEL_GREEDY") == perm.tolist()
    assert torch.equal(Q, M)
    # Mixed: row 0 (0.9 at column 0) is served in parallel; row 1, whose
    # maximum is column 0 as well, is walked with column 0 masked.
    M = torch.tensor([[0.9, 0.05, 0.05], [0.45, 0.4, 0.15], [0.1, 0.2, 0.3]], dtype=torch.float64)
    Q = M.clone()
    assert oracle.greedy_lap(Q, "PARALLEL_GREEDY") == [0, 1, 2]
    assert Q[1, 0] == 0 and Q[0, 0] == 0.9 and Q[2, 0] == 0


def test_more_rand_inverse_cdf_and_masking():
    M = torch.tensor([[0.2, 0.5, 0.3], [0.6, 0.3, 0.1], [0.0, 0.0, 1.0]], dtype=torch.float64)
    # Row 0 first: U in (0, 0.2] -> 0, (0.2, 0.7] -> 1, (0.7, 1) -> 2.  Row 1
    # then has the taken column masked to 0.
    for u0, want0 in ((0.1, 0), (0.3, 1), (0.69, 1), (0.71, 2), (0.999, 2)):
        Q = M.clone()
        res = oracle.greedy_lap(Q, "MORE_RAND", order=[0, 1, 2], uniforms=[u0, 0.5, 0.5])
        assert res[0] == want0
        assert Q[1, want0] == 0
    # A draw of 0 takes column 0 even where the entry is 0 (r <= cumsum).
    Q = M.clone()
    assert oracle.greedy_lap(Q, "MORE_RAND", order=[2, 0, 1], uniforms=[0.0, 0.5, 0.5])[2] == 0
    # A permutation matrix: every positive draw takes its one column.
    perm = [2, 0, 1]
    Pm = torch.zeros((3, 3), dtype=torch.float64)
    Pm[[0, 1, 2], perm] = 1.0
    assert oracle.greedy_lap(Pm, "MORE_RAND", order=[1, 2, 0], uniforms=[0.3, 0.9, 1e-9]) == perm


def test_more_rand_frequencies():
    # The first visited row is sampled in proportion to its entries.
    rng = np.random.default_rng(0)
    M = torch.tensor([[0.1, 0.4, 0.2, 0.3], [1.0, 0, 0, 0], [0, 1.0, 0, 0], [0, 0, 1.0, 0]], dtype=torch.float64)
    draws = 4000
    counts = np.zeros(4)
    for u in rng.random(draws):
        counts[oracle.greedy_lap(M.clone(), "MORE_RAND", order=[0, 1, 2, 3], uniforms=[u, 0.5, 0.5, 0.5])[0]] += 1
    p = np.array([0.1, 0.4, 0.2, 0.3])
    sigma = np.sqrt(draws * p * (1 - p))
    assert np.all(np.abs(counts - draws * p) < 4 * sigma)


def test_double_greedy_order_and_masking():
    # Row 1 holds the largest maximum, so it is served first and takes
    # column 0; row 0 then sees column 0 masked to 0 and takes column 1.
    M = torch.tensor([[0.5, 0.4, 0.1], [0.9, 0.05, 0.05], [0.2, 0.3, 0.45]], dtype=torch.float64)
    Q = M.clone()
    assert oracle.greedy_lap(Q, "DOUBLE_GREEDY") == [1, 0, 2]
    assert Q[0, 0] == 0 and Q[1, 0] == 0.9          # masking is in place, taken cols only
    assert oracle.greedy_lap(M.clone(), "GREEDY") == [0, 1, 2]


# ------------------------------------------------------------- pipeline --

def test_pipeline_recovers_isomorphic_pair():
    from cugal_amd.synthetic import alignment_accuracy, answer_columns, watts_strogatz

    inst = watts_strogatz(80, 6, 0.3, 0.0, seed=5)
    run = oracle.align(inst.source, inst.target, oracle.OracleConfig(
        sinkhorn_regularization=0.3, sinkhorn_method="LOG", mu=2.0, iter_count=5, frank_wolfe_iter_count=5,
        use_sparse_adjacency=True))
    assert alignment_accuracy(inst.source_mapping, answer_columns(run.mapping)) == 1.0


def test_forced_iterations_follow_a_run():
    # The oracle given a run's own per-solve Sinkhorn counts (threshold
    # replaced by the counts) reproduces that run bit for bit; a count
    # changed in one solve changes the result.
    import copy

    from cugal_amd.synthetic import erdos_renyi

    inst = erdos_renyi(50, 0.1, 0.05, seed=3)
    cfg = dict(sinkhorn_regularization=0.3, sinkhorn_method="LOG", sinkhorn_threshold=1e-3, mu=2.0, iter_count=3,
               frank_wolfe_iter_count=3, use_sparse_adjacency=True)
    free = oracle.align(copy.deepcopy(inst.source), copy.deepcopy(inst.target), oracle.OracleConfig(**cfg))
    assert len(set(free.sinkhorn_iterations)) > 1  # stops at different counts
    forced = oracle.align(copy.deepcopy(inst.source), copy.deepcopy(inst.target),
                          oracle.OracleConfig(**cfg, forced_iterations=list(free.sinkhorn_iterations)))
    assert forced.sinkhorn_iterations == free.sinkhorn_iterations
    assert torch.equal(forced.P, free.P)
    counts = list(free.sinkhorn_iterations)
    counts[2] += 1
    other = oracle.align(copy.deepcopy(inst.source), copy.deepcopy(inst.target),
                         oracle.OracleConfig(**cfg, forced_iterations=counts))
    assert other.sinkhorn_iterations == counts and not torch.equal(other.P, free.P)


def test_log_marginal_quirk_forms():
    # The in-place LOG marginal check (cugal/sinkhorn.py:67-74) first runs
    # at iteration 10, so through 10 iterations every form is the clean
    # solve.  After it, the CUDA branch's form ("gpu") still takes the
    # v-update from the untouched K (the copy of :99-100): v after 12
    # iterations is -LSE_i(K0 + u after 11); the CPU branch's form (True)
    # reads the rewritten K there and departs from that.
    gen = torch.Generator().manual_seed(5)
    C = torch.rand(12, 12, generator=gen, dtype=torch.float64)
    K0 = -C / 0.3

    def run(quirk, its):
        st = oracle.SinkhornState(12, "LOG")
        _, u, v, res = st.solve_log(C.clone(), 0.3, its, 0.0, 10, quirk_log_marginal=quirk)
        return u.clone(), v.clone(), res

    clean = run(False, 10)
    for quirk in (True, "gpu"):
        u, v, _ = run(quirk, 10)
        assert torch.equal(u, clean[0]) and torch.equal(v, clean[1])
    for quirk, follows in ((True, False), ("gpu", True)):
        u11, _, res = run(quirk, 11)
        _, v12, _ = run(quirk, 12)
        assert len(res.errors) == 1 and res.errors[0] > 1.0  # the check's error carries -u - v
        want = -torch.logsumexp(K0 + u11[:, None], 0)
        assert torch.allclose(v12, want, rtol=0, atol=1e-12) is follows


def test_pipeline_default_config_small():
    # Reference default Config (STANDARD, FP64, dense quirk, DOUBLE_GREEDY).
    from cugal_amd.synthetic import alignment_accuracy, answer_columns, erdos_renyi

    inst = erdos_renyi(60, 0.1, 0.0, seed=2)
    run = oracle.align(inst.source, inst.target, oracle.OracleConfig(iter_count=4, frank_wolfe_iter_count=4))
    assert alignment_accuracy(inst.source_mapping, answer_columns(run.mapping)) >= 0.9


def test_unrolled_plan_is_solve_log_and_feature_cost_brute_force():
    n = 30
    C = torch.rand((n, n), dtype=torch.float64, generator=torch.Generator().manual_seed(4)) * 3
    st = oracle.SinkhornState(n, "LOG")
    K, u, v, _ = st.solve_log(C, 0.4, 17, 0.0, 10)
    np.testing.assert_allclose(oracle.unrolled_plan(C, 0.4, 17).numpy(), oracle.plan_log(K, u, v).numpy(), rtol=1e-13)
    fs = torch.rand((5, 4), dtype=torch.float64, generator=torch.Generator().manual_seed(1))
    ft = torch.rand((5, 4), dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    w = torch.tensor([0.5, 1.0, 1.5, 2.0], dtype=torch.float64)
    want = np.array([[0.3 * sum(float(w[k]) * abs(float(fs[i, k] - ft[j, k])) for k in range(4)) for j in range(5)]
                     for i in range(5)])
    np.testing.assert_allclose(oracle.feature_cost(fs, ft, w, 0.3).numpy(), want, rtol=1e-14)


# ---- alignment quality (benchmark/experiment.py:115-141) ----------------------

def _brute_counts(S, T, ans, gt):
    """Conserved ordered node pairs by explicit enumeration."""
    m = len(ans)
    inter = sum(1 for a in range(m) for b in range(m) if S[ans[a], ans[b]] == 1 and T[gt[a], gt[b]] == 1)
    ones_t = sum(1 for a in range(m) for b in range(m) if T[gt[a], gt[b]] == 1)
    return inter, int((S == 1).sum()), ones_t


def _scores(S, T, ans, gt):
    return (oracle.induced_conserved_structure(S, T, ans, gt), oracle.edge_correctness(S, T, ans, gt),
            oracle.symmetric_substructure(S, T, ans, gt))


def test_structure_scores_hand_cases():
    # S: path 0-1-2; T: star around 0 (edges 0-1, 0-2); ground truth identity.
    S = nx.to_numpy_array(nx.Graph([(0, 1), (1, 2)]), nodelist=[0, 1, 2])
    T = nx.to_numpy_array(nx.Graph([(0, 1), (0, 2)]), nodelist=[0, 1, 2])
    gt = np.array([0, 1, 2])
    # Answer (1, 0, 2) maps S's edges onto T's: everything is conserved.
    assert _scores(S, T, np.array([1, 0, 2]), gt) == (1.0, 1.0, 1.0)
    # The identity answer conserves one of two edges: 2 of 4 ordered pairs.
    ics, ec, s3 = _scores(S, T, gt, gt)
    assert (ics, ec) == (0.5, 0.5) and abs(s3 - 1 / 3) < 1e-15
    # An answer outside the graph scores 0 (the reference catches the error).
    assert _scores(S, T, np.array([0, 1, 3]), gt) == (0, 0, 0)


def test_structure_scores_brute_force_and_noise_free_pairs():
    from cugal_amd.synthetic import erdos_renyi

    rng = np.random.default_rng(3)
    for seed in range(3):
        inst = erdos_renyi(40, 0.15, 0.1, seed=seed)
        # cugal() pads both graphs before the harness scores them (isolated
        # nodes are missing from an edge list).
        inst.source.add_nodes_from(range(40))
        inst.target.add_nodes_from(range(40))
        S = nx.to_numpy_array(inst.source)
        T = nx.to_numpy_array(inst.target)
        gt = inst.source_mapping
        ans = rng.permutation(len(gt))
        inter, ones_s, ones_t = _brute_counts(S, T, ans, gt)
        ics, ec, s3 = _scores(S, T, ans, gt)
        assert ec == inter / ones_s and ics == inter / ones_t and s3 == inter / (ones_s + ones_t - inter)
    # A noise-free pair, graphs in sorted node order.  The reference applies
    # the answer to S and the truth to T, so T[truth][:, truth] = S and the
    # identity answer scores 1, while the true alignment in general does not
    # (the standard definition, S against T[answer][:, answer], is this
    # formula with (identity, answer), and scores the truth 1).
    inst = erdos_renyi(50, 0.1, 0.0, seed=7)
    n = len(inst.source_mapping)
    S = nx.to_numpy_array(inst.source, nodelist=range(n))
    T = nx.to_numpy_array(inst.target, nodelist=range(n))
    assert _scores(S, T, np.arange(n), inst.source_mapping) == (1.0, 1.0, 1.0)
    assert _scores(S, T, inst.source_mapping, inst.source_mapping)[1] < 0.5


# ---- FUGAL's 7-feature set (FUGAL/pred.py:24-110) -----------------------------

def test_fugal_features_closed_forms():
    F = oracle.extract_features_fugal(nx.complete_graph(9))
    assert np.all(F[:, 4] == 36) and np.all(F[:, 5] == 0) and np.all(F[:, 6] == 0)
    F = oracle.extract_features_fugal(nx.star_graph(6))  # centre 0, leaves 1..6
    assert list(F[0, 4:]) == [6, 0, 0]
    assert np.all(F[1:, 4:] == [1, 5, 5])
    F = oracle.extract_features_fugal(nx.path_graph(6))
    assert F[:, 4:].tolist() == [[1, 1, 1], [2, 1, 1], [2, 2, 2], [2, 2, 2], [2, 1, 1], [1, 1, 1]]
    F = oracle.extract_features_fugal(nx.petersen_graph())  # 3-regular, girth 5
    assert np.all(F[:, 4:] == [3, 6, 6])
    g = nx.Graph([(0, 1)])
    g.add_node(2)
    assert oracle.extract_features_fugal(g)[2].tolist() == [0.0] * 7
    simple = oracle.extract_features_fugal(nx.karate_club_graph(), simple=True)
    np.testing.assert_array_equal(simple[:, :4], oracle.extract_features_nx(nx.karate_club_graph()))
    assert np.all(simple[:, 4:] == 0)


def test_fugal_features_against_networkx_routines():
    for g in (nx.erdos_renyi_graph(80, 0.08, seed=1), nx.barabasi_albert_graph(90, 3, seed=2),
              nx.watts_strogatz_graph(70, 6, 0.2, seed=3)):
        F = oracle.extract_features_fugal(g)
        for v in sorted(g.nodes()):
            closed = set(g.neighbors(v)) | {v}
            assert F[v, 4] == g.subgraph(closed).number_of_edges()
            assert F[v, 5] == len(list(nx.edge_boundary(g, closed)))
            dist = nx.single_source_shortest_path_length(g, v, cutoff=2)
            assert F[v, 6] == sum(1 for d in dist.values() if d == 2)


@pytest.mark.parametrize("simple", [True, False])
def test_cugal_oracle_agrees_with_the_original_fugal(simple):
    # The original FUGAL's loop (fresh Sinkhorn scalings per step, no
    # epsilon) and this oracle's cugal loop in the same setting (STANDARD,
    # warm-started, reg 1, no threshold, lambda term once, SCIPY rounding)
    # reach the same fixed points once each Sinkhorn solve has converged
    # (at the reference's 500 iterations the warm and fresh starts still
    # differ by ~3e-4, at 2000 by ~2e-6): the same assignment and P within
    # 1e-5.  The 7 features widen the cost range, and Sinkhorn at reg 1
    # converges far more slowly (2.6e-4 apart at 2000 iterations, 1.5e-4 at
    # 8000): the same assignment, P within 5e-4.
    from cugal_amd.synthetic import watts_strogatz

    inst = watts_strogatz(60, 6, 0.2, 0.05, seed=4)
    src, tgt = inst.source, inst.target
    assert min(d for _, d in src.degree()) > 0 and min(d for _, d in tgt.degree()) > 0
    P0, ans0 = oracle.fugal_original(src, tgt, niter=3, simple=simple, mu=0.5, sinkhorn_iterations=2000)
    run = oracle.align(src.copy(), tgt.copy(), oracle.OracleConfig(
        sinkhorn_regularization=1.0, sinkhorn_method="STANDARD", sinkhorn_threshold=0.0, sinkhorn_iterations=2000,
        mu=0.5, iter_count=3,
        frank_wolfe_iter_count=10, use_sparse_adjacency=True, hungarian_method="SCIPY",
        node_features="cugal" if simple else "fugal7"))
    np.testing.assert_allclose(run.P.numpy(), P0.numpy(), rtol=0, atol=1e-5 if simple else 5e-4)
    assert [c for _, c in run.mapping] == [c for _, c in ans0]
