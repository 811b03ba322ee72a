"""FP64 CPU restatement of cuGAL's alignment path (test infrastructure).

Each function cites the reference lines it restates and names its pin (the
independent check in tests/test_oracle.py).  The reference runs its CPU path
with torch in FP64; so does this module, with the same op order where the
reference's rounding is part of the result.

Reference quirks (see the `quirks` arguments):
  * dense_gradient adds lambda (1 - 2P) twice (cugal/pred.py:46-50): kept,
    since it is the reference's result for use_sparse_adjacency=False;
  * the LOG solver's marginal check overwrites K with exp(K+u+v)-u-v
    (cugal/sinkhorn.py:67-74): off by default (quirk_log_marginal=True
    reproduces it in the CPU branch, "gpu" in the CUDA branch, whose
    v-update reads a copy of K taken before the loop);
  * DOUBLE_GREEDY's argsort is not stable (cugal/greedy_lap.py:52): a stable
    sort is used, identical whenever the row maxima are distinct;
  * a self-loop counts once in the dense adjacency and twice in the sparse
    one, and the GPU path's features count its CSR entries where networkx
    ignores it (adjacency_matrix, extract_features_csr).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import networkx as nx
import numpy as np
import scipy.optimize
import torch

M_EPS = 1e-16


# ---------------------------------------------------------------------------
# Features
# ---------------------------------------------------------------------------

def extract_features_nx(G: nx.Graph) -> np.ndarray:
    """[degree, clustering, mean neighbour degree, mean neighbour clustering]
    per node in sorted node order, cugal/feature_extraction.py:41-64.
    Pin: closed forms on K_n, C_n, stars, the Petersen graph; clustering
    against brute-force triangle enumeration; neighbour means against
    numpy adjacency products."""
    node_list = sorted(G.nodes())
    node_degree_dict = dict(G.degree())
    node_clustering_dict = dict(nx.clustering(G))
    degs = [node_degree_dict[n] for n in node_list]
    clusts = [node_clustering_dict[n] for n in node_list]
    neighbor_degs = [
        np.mean([node_degree_dict[m] for m in G.neighbors(n) if m != n]) if node_degree_dict[n] > 0 else 0
        for n in node_list
    ]
    neighbor_clusts = [
        np.mean([node_clustering_dict[m] for m in G.neighbors(n) if m != n]) if node_degree_dict[n] > 0 else 0
        for n in node_list
    ]
    return np.nan_to_num(np.stack((degs, clusts, neighbor_degs, neighbor_clusts)).T)


def csr_lists(G: nx.Graph) -> list[list[int]]:
    """The sorted neighbour lists of the reference's CSR for a graph
    labelled 0..n-1 (cugal/adjacency.py:46-69 over create_adjacency,
    cuda_kernels/adjacency.cu:228-259, which reads each pair (x, y) as the
    entry [y][x]): an undirected graph lists every edge of nx.edges(G) in
    both directions, so a self-loop twice; a directed graph lists each edge
    x -> y once, in row y (the predecessors of y).
    Pin: through extract_features_csr (equal to networkx on simple graphs;
    the entry-count matrix built from G.edges() on looped or directed
    graphs)."""
    n = G.number_of_nodes()
    rows = [[] for _ in range(n)]
    for x, y in G.edges():
        rows[y].append(x)
        if not G.is_directed():
            rows[x].append(y)
    return [sorted(r) for r in rows]


def extract_features_csr(G: nx.Graph) -> np.ndarray:
    """The four features as the reference's GPU path computes them
    (cugal/feature_extraction.py:15-38 over cuda_kernels/
    feature_extraction.cu:13-106): degree = CSR row length; t = sum over
    the row's entries k of the sorted-merge intersection size of rows v and
    k; clustering = t / (d (d - 1)), 0 when t = 0 or d < 2; neighbour means
    over the row's entries, 0 for an empty row.  On simple graphs this is
    extract_features_nx; with self-loops the two reference paths differ
    (networkx ignores a loop for clustering and the means).
    Pin: equal to extract_features_nx on simple graphs; on graphs with loops
    the multiset formula t(v) = sum_k M[v,k] sum_w min(M[v,w], M[k,w]) over
    the entry-count matrix M, and closed forms (a lone looped vertex: 2, 2,
    2, 2)."""
    rows = csr_lists(G)
    n = len(rows)

    def merge_count(a, b):
        i = j = c = 0
        while i < len(a) and j < len(b):
            if a[i] == b[j]:
                c += 1
                i += 1
                j += 1
            elif a[i] > b[j]:
                j += 1
            else:
                i += 1
        return c

    deg = np.array([len(r) for r in rows], dtype=np.float64)
    t = np.array([sum(merge_count(rows[v], rows[k]) for k in rows[v]) for v in range(n)], dtype=np.float64)
    clus = np.where((t == 0) | (deg < 2), 0.0, t / np.maximum(deg * (deg - 1), 1.0))
    nd = np.array([np.mean([deg[k] for k in r]) if r else 0.0 for r in rows])
    nc = np.array([np.mean([clus[k] for k in r]) if r else 0.0 for r in rows])
    return np.stack((deg, clus, nd, nc)).T


def adjacency_matrix(G: nx.Graph, sparse: bool, dtype=torch.float64) -> torch.Tensor:
    """The adjacency the reference's gradient multiplies by: the dense path
    takes nx.to_numpy_array (a self-loop is 1 on the diagonal,
    cugal/pred.py:106-110); the sparse path its CSR, where a loop is listed
    twice (cugal/adjacency.py:52-53), so the diagonal holds 2.
    Pin: equal to the scipy CSR of the graph on simple graphs; the loop
    diagonal by construction of csr_lists."""
    A = torch.tensor(nx.to_numpy_array(G, nodelist=sorted(G.nodes())), dtype=dtype)
    if sparse:
        for x, y in G.edges():
            if x == y:
                A[x, x] = 2.0
    return A


def triangle_counts(adjacency: np.ndarray) -> np.ndarray:
    """t(v) = sum_{u in N(v)} |N(v) & N(u)| = 2 * triangles through v =
    (A^3)_vv, the quantity of cuda_kernels/feature_extraction.cu:41-57.
    Pin: brute-force enumeration of vertex triples; nx.triangles."""
    A = (adjacency != 0).astype(np.int64)
    return np.einsum("ij,ji->i", A @ A, A)


def egonet_edges(G: nx.Graph) -> np.ndarray:
    """Edges of each node's ego network (FUGAL/pred.py:68-72).  Pin:
    deg + triangles on simple graphs."""
    return np.array([nx.ego_graph(G, v).number_of_edges() for v in sorted(G.nodes())], dtype=np.int64)


def extract_features_fugal(G: nx.Graph, simple: bool = False) -> np.ndarray:
    """The node features of the original FUGAL, FUGAL/pred.py:24-110, in
    sorted node order: the 4 of extract_features_nx and, with simple=False,
    ego-net edges, edges leaving the ego net (edge tuples of the ego net's
    nodes not in the ego graph) and distinct nodes at distance 2; n x 7
    (columns 4-6 zero when simple).
    Pin: closed forms on K_n, stars, paths and the Petersen graph;
    nx.edge_boundary of the closed neighbourhood and BFS distances
    (single_source_shortest_path_length) for the two new counts."""
    node_list = sorted(G.nodes())
    deg = dict(G.degree())
    egonets = {v: nx.ego_graph(G, v) for v in node_list}
    F = np.zeros((len(node_list), 7))
    F[:, :4] = extract_features_nx(G)
    if not simple:
        F[:, 4] = [egonets[v].number_of_edges() if deg[v] > 0 else 0 for v in node_list]
        F[:, 5] = [
            len([edge for edge in set.union(*[set(G.edges(j)) for j in egonets[v].nodes])
                 if not egonets[v].has_edge(*edge)])
            for v in node_list
        ]
        F[:, 6] = [
            len(set(p for m in G.neighbors(v) for p in G.neighbors(m)) - set(G.neighbors(v)) - {v})
            if deg[v] > 0 else 0
            for v in node_list
        ]
    return np.nan_to_num(F)


def distance_matrix(source: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """torch.cdist of the scaled features, cugal/feature_extraction.py:106-108.
    Pin: direct FP64 Euclidean form (abs tolerance: cdist's mm path
    rounds differently for n > 25)."""
    return torch.cdist(source, target)


# ---------------------------------------------------------------------------
# Gradient
# ---------------------------------------------------------------------------

def dense_gradient(A: torch.Tensor, B: torch.Tensor, P: torch.Tensor, D: torch.Tensor, iteration: int,
                   double_lambda: bool = True) -> torch.Tensor:
    """-A^T P B - A P B^T + D + lambda(1 - 2P) [+ lambda - 2 lambda P again],
    cugal/pred.py:38-51.  Pin: closed forms at P = J/n (d_A d_B^T / n) and
    at a permutation (common-neighbour counts by brute force)."""
    gradient = -A.T @ P @ B - A @ P @ B.T
    gradient = gradient + D + iteration * (1 - 2 * P)
    if double_lambda:
        gradient += iteration - iteration * 2 * P
    return gradient


def sparse_gradient(A: torch.Tensor, B: torch.Tensor, P: torch.Tensor, D: torch.Tensor, iteration: int):
    """Symmetric sparse form -2 (B (A P)^T)^T + D + lambda(1 - 2P),
    cugal/pred.py:54-74 (A, B symmetric).  Pin: equals dense_gradient minus
    one regulariser term on undirected graphs."""
    gradient = (B @ (A @ P).T).T
    gradient *= -2
    gradient = gradient + D
    gradient += iteration - iteration * 2 * P
    return gradient


# ---------------------------------------------------------------------------
# Sinkhorn
# ---------------------------------------------------------------------------

def scale_kernel_matrix(K, u, v):
    """cugal/sinkhorn.py:20-23 (in place)."""
    K *= v.reshape(1, -1)
    K *= u.reshape(-1, 1)
    return K


def scale_kernel_matrix_log(K, log_u, log_v):
    """cugal/sinkhorn.py:26-30 (in place)."""
    K += log_u[:, None]
    K += log_v[None, :]
    torch.exp(K, out=K)
    return K


def marginal_error(K, u, v) -> float:
    """Column-marginal L1 error with in-place scale/unscale,
    cugal/sinkhorn.py:33-40."""
    K *= v.reshape(1, -1)
    K *= u.reshape(-1, 1)
    dS = torch.sum(abs(torch.sum(K, dim=0) - 1)).item()
    K /= v.reshape(1, -1)
    K /= u.reshape(-1, 1)
    return dS


@dataclass
class SinkhornResult:
    errors: list = field(default_factory=list)
    iteration_count: int = 0


class SinkhornState:
    """Warm-started duals and the two solvers of cugal/sinkhorn.py:51-144.
    Pin: separable costs (plan = J/n), the 2x2 closed form, doubly
    stochastic invariants, LOG == STANDARD fixed point, the eps -> 0 limit
    against scipy's LAP."""

    def __init__(self, n: int, method: str, dtype=torch.float64):
        if method == "LOG":
            vector = torch.zeros(n, dtype=dtype)
        else:
            vector = torch.full((n,), 1 / n, dtype=dtype)
        self.u = vector.clone()
        self.v = vector

    def lehmann_momentum(self, momentum_start, eval_freq, errors) -> float:
        """cugal/sinkhorn.py:78-88."""
        if momentum_start is None or len(errors) * eval_freq < momentum_start:
            return 1.0
        index = momentum_start // eval_freq
        if len(errors) < index - 1:
            return 1.0
        ratio = min(errors[index - 1] / errors[index - 2], 0.99)
        if math.isnan(ratio):
            ratio = 0.99
        return 2.0 / (1.0 + math.sqrt(1 - ratio ** (1.0 / eval_freq)))

    def marginal_error_log(self, K: torch.Tensor, quirk: bool) -> float:
        """Column-marginal L1 error of exp(K + u + v).  With `quirk`, the
        reference's in-place version (cugal/sinkhorn.py:67-74), which leaves K
        = exp(K+u+v) - u - v; otherwise on a copy."""
        if quirk:
            K += self.u[:, None]
            K += self.v[None, :]
            torch.exp(K, out=K)
            K -= self.u[:, None]
            K -= self.v[None, :]
            return torch.sum(abs(torch.sum(K, dim=0) - 1)).item()
        plan = torch.exp(K + self.u[:, None] + self.v[None, :])
        return torch.sum(abs(torch.sum(plan, dim=0) - 1)).item()

    def solve_log(self, C, reg, iterations, threshold, eval_freq, momentum_start=None, quirk_log_marginal=False):
        """cugal/sinkhorn.py:90-120 (CPU branch: torch.logsumexp).
        quirk_log_marginal: False, True (the CPU branch with the in-place
        marginal check), or "gpu": the CUDA branch's version of it, where the
        v-update reads the transposed copy of K taken before the loop
        (:99-100, :104), which the check never touches, and only the
        u-update and the returned K see the rewritten K."""
        K = -C / reg
        K_v = K.clone() if quirk_log_marginal == "gpu" else K
        errors, momentum = [], 1.0
        iteration = -1
        for iteration in range(iterations):
            self.v = self.v * (1 - momentum) - momentum * torch.logsumexp(K_v + self.u[:, None], 0)
            self.u = self.u * (1 - momentum) - momentum * torch.logsumexp(K + self.v[None, :], 1)
            if iteration % eval_freq == 0 and iteration != 0:
                errors.append(self.marginal_error_log(K, bool(quirk_log_marginal)))
                if errors[-1] < threshold:
                    break
                momentum = self.lehmann_momentum(momentum_start, eval_freq, errors)
        return K, self.u, self.v, SinkhornResult(errors, iteration + 1)

    def solve_standard(self, C, reg, iterations, threshold, eval_freq, momentum_start=None):
        """cugal/sinkhorn.py:122-144."""
        K = torch.exp(C / -reg)
        errors, momentum = [], 1
        iteration = -1
        for iteration in range(iterations):
            self.v = self.v ** (1 - momentum) * (1 / (self.u @ K + M_EPS)) ** momentum
            self.u = self.u ** (1 - momentum) * (1 / (K @ self.v + M_EPS)) ** momentum
            if iteration % eval_freq == 0:
                errors.append(marginal_error(K, self.u, self.v))
                if errors[-1] < threshold:
                    break
                momentum = self.lehmann_momentum(momentum_start, eval_freq, errors)
        return K, self.u, self.v, SinkhornResult(errors, iteration + 1)


def plan_log(K, u, v) -> torch.Tensor:
    return torch.exp(K + u[:, None] + v[None, :])


def plan_standard(K, u, v) -> torch.Tensor:
    return K * v.reshape(1, -1) * u.reshape(-1, 1)


# ---------------------------------------------------------------------------
# Frank-Wolfe update
# ---------------------------------------------------------------------------

def update_quasi_permutation(P, plan, alpha) -> torch.Tensor:
    """P += alpha (plan - P); returns the step (cugal/pred.py:77-86)."""
    q = plan.clone()
    q -= P
    q *= alpha
    P += q
    return q


def duality_gap(gradient, step, alpha) -> float:
    """|sum G (step / alpha)| (cugal/pred.py:140-141).  Pin: from P = J/n
    toward a permutation plan the gap is |sum_i G[i, pi(i)] - sum G / n|
    for every alpha; a 2 x 2 hand case."""
    return abs(torch.sum(gradient * (step / alpha)).item())


# ---------------------------------------------------------------------------
# LAP
# ---------------------------------------------------------------------------

def greedy_lap(cost_matrix: torch.Tensor, method: str, order=None, uniforms=None) -> list[int]:
    """GREEDY (cugal/greedy_lap.py:10-19), RAND (:21-32), MORE_RAND
    (:34-48), DOUBLE_GREEDY (:50-63) and PARALLEL_GREEDY (:65-85), in place
    on cost_matrix like the reference, with stable sorts.

    RAND and MORE_RAND take the row order the reference draws with
    torch.randperm (`order`); MORE_RAND also takes the k-th visited row's
    draw U_k of random.random() (`uniforms`): the reference's
    uniform(0, total) is 0 + total * U_k.  Its prefix sums are taken in FP64
    (the reference's device cumsum rounds in FP32 in a parallel order that no
    restatement reproduces); the column is the first j with r <= cumsum_j.

    Pins: brute force over permutations where greedy is optimal; RAND with
    the identity / the sorted order equals GREEDY / DOUBLE_GREEDY; hand-built
    masking cases; PARALLEL_GREEDY's two regimes (all maxima above 0.5: plain
    argmax; none: DOUBLE_GREEDY); MORE_RAND's selection frequencies against
    the masked row's proportions."""
    n = cost_matrix.size(1)
    taken = torch.ones(n, dtype=cost_matrix.dtype)
    res = [0] * n
    if method == "GREEDY":
        order = range(cost_matrix.size(0))
    elif method in ("RAND", "MORE_RAND"):
        order = [int(r) for r in order]
        if sorted(order) != list(range(cost_matrix.size(0))):
            raise ValueError("order must be a permutation of the rows")
    elif method == "DOUBLE_GREEDY":
        max_values_in_rows = cost_matrix.max(dim=1).values
        order = torch.sort(max_values_in_rows, descending=True, stable=True).indices.tolist()
    elif method == "PARALLEL_GREEDY":
        max_values_in_rows, max_col_in_rows = cost_matrix.max(dim=1)
        order_values, order_t = torch.sort(max_values_in_rows, descending=True, stable=True)
        half_mask = order_values > 0.5
        for row in order_t[half_mask].tolist():
            res[row] = int(max_col_in_rows[row])
            taken[res[row]] = 0
        order = order_t[~half_mask].tolist()
    else:
        raise ValueError(method)
    for k, row in enumerate(order):
        cost_matrix[row] *= taken
        if method == "MORE_RAND":
            cumsum = np.cumsum(cost_matrix[row].to(torch.float64).numpy())
            r = cumsum[-1] * float(uniforms[k])
            hits = np.nonzero(r <= cumsum)[0]
            if len(hits) == 0:
                raise IndexError("no prefix reaches the draw")
            m = int(hits[0])
        else:
            m = int(cost_matrix[row].argmax())
        taken[m] = 0
        res[row] = m
    return res


def scipy_lap(quasi_permutation: torch.Tensor) -> np.ndarray:
    """SCIPY branch, cugal/pred.py:156-158.  Pin: brute force for n <= 7."""
    _, col = scipy.optimize.linear_sum_assignment(quasi_permutation.numpy(), maximize=True)
    return col


# ---------------------------------------------------------------------------
# The whole path
# ---------------------------------------------------------------------------

@dataclass
class OracleConfig:
    """The reference Config fields the path reads (cugal/config.py:59-120)."""

    sinkhorn_regularization: float = 1.0
    sinkhorn_method: str = "STANDARD"
    sinkhorn_iterations: int = 500
    sinkhorn_threshold: float = 1e-3
    sinkhorn_eval_freq: int = 10
    mu: float = 0.5
    iter_count: int = 15
    frank_wolfe_iter_count: int = 10
    frank_wolfe_threshold: float | None = None
    use_sparse_adjacency: bool = False
    sinkhorn_momentum_start: int | None = None
    hungarian_method: str = "DOUBLE_GREEDY"
    quirk_log_marginal: bool | str = False  # False, True or "gpu" (SinkhornState.solve_log)
    node_features: str = "cugal"  # or "fugal7": FUGAL/pred.py:24-110, simple=False
    # Test tool, not a reference field: per-solve Sinkhorn iteration counts
    # that replace the threshold stop (solve k runs exactly
    # forced_iterations[k] iterations), to follow another run's stop
    # decisions.
    forced_iterations: list | None = None


@dataclass
class OracleRun:
    P: torch.Tensor
    mapping: list
    sinkhorn_iterations: list
    duality_gaps: list
    sinkhorn_errors: list = field(default_factory=list)  # per solve, the checked marginal errors
    cost_absmax: list = field(default_factory=list)  # per solve, max |C|


def pad_graphs(source: nx.Graph, target: nx.Graph):
    """cugal/pred.py:209-212 (in place; range excludes the maximum label)."""
    node_count = max(max(source.nodes()), max(target.nodes()))
    source.add_nodes_from(set(range(node_count)) - set(source.nodes()))
    target.add_nodes_from(set(range(node_count)) - set(target.nodes()))


def align(source: nx.Graph, target: nx.Graph, config: OracleConfig, dtype=torch.float64) -> OracleRun:
    """cugal() with the CPU path (cugal/pred.py:196-249, :89-148, :151-193).
    Pin: exact recovery of the permutation on noise-free isomorphic pairs
    without automorphisms."""
    pad_graphs(source, target)

    def extract(G):
        if config.node_features != "cugal":
            return extract_features_fugal(G)
        # Graphs with self-loops or directed edges: the features of the
        # reference's GPU path (the path a device run stands in for); the
        # CPU path's networkx features differ there and agree elsewhere.
        if nx.number_of_selfloops(G) or G.is_directed():
            return extract_features_csr(G)
        return extract_features_nx(G)

    fs = torch.tensor(extract(source), dtype=dtype) * config.mu
    ft = torch.tensor(extract(target), dtype=dtype) * config.mu
    D = distance_matrix(fs, ft)
    A = adjacency_matrix(source, config.use_sparse_adjacency, dtype)
    B = adjacency_matrix(target, config.use_sparse_adjacency, dtype)
    n = A.shape[0]
    state = SinkhornState(n, config.sinkhorn_method, dtype)
    P = torch.full((n, n), 1 / n, dtype=dtype)
    iterations, gaps, errors, cost_absmax = [], [], [], []
    for lam in range(config.iter_count):
        for it in range(1, config.frank_wolfe_iter_count + 1):
            if config.use_sparse_adjacency:
                G = sparse_gradient(A, B, P, D, lam)
            else:
                G = dense_gradient(A, B, P, D, lam)
            forced = config.forced_iterations
            its = config.sinkhorn_iterations if forced is None else forced[len(iterations)]
            threshold = config.sinkhorn_threshold if forced is None else 0.0
            if config.sinkhorn_method == "LOG":
                K, u, v, res = state.solve_log(G, config.sinkhorn_regularization, its, threshold,
                                               config.sinkhorn_eval_freq, config.sinkhorn_momentum_start,
                                               config.quirk_log_marginal)
                plan = plan_log(K, u, v)
            else:
                K, u, v, res = state.solve_standard(G, config.sinkhorn_regularization, its, threshold,
                                                    config.sinkhorn_eval_freq, config.sinkhorn_momentum_start)
                plan = plan_standard(K, u, v)
            iterations.append(res.iteration_count)
            errors.append(list(res.errors))
            cost_absmax.append(float(G.abs().max()))
            alpha = 2.0 / float(2.0 + it)
            step = update_quasi_permutation(P, plan, alpha)
            gap = duality_gap(G, step, alpha)
            gaps.append(gap)
            if config.frank_wolfe_threshold is not None and gap < config.frank_wolfe_threshold:
                break
    Q = P.clone()
    if config.hungarian_method in ("GREEDY", "DOUBLE_GREEDY"):
        col = greedy_lap(Q, config.hungarian_method)
    else:
        col = list(scipy_lap(Q))
    mapping = [(i, int(col[i])) for i in range(n)]
    return OracleRun(Q, mapping, iterations, gaps, errors, cost_absmax)


# ---------------------------------------------------------------------------
# Differentiable references: unrolled log-domain Sinkhorn and its loss
# ---------------------------------------------------------------------------

def unrolled_plan(C: torch.Tensor, reg: float, iterations: int, u0=None, v0=None) -> torch.Tensor:
    """`iterations` plain log-domain iterations of cugal/sinkhorn.py:107-111
    (momentum 1, no stopping test) from (u0, v0) = 0 by default; returns
    exp(K + u + v).  Differentiable with torch autograd; any device.
    Pin: the separable-cost and 2x2 closed forms via solve_log, whose loop
    this is, and FP64 central differences of its own loss."""
    n = C.shape[0]
    K = -C / reg
    u = torch.zeros(n, dtype=C.dtype, device=C.device) if u0 is None else u0
    v = torch.zeros(n, dtype=C.dtype, device=C.device) if v0 is None else v0
    for _ in range(iterations):
        v = -torch.logsumexp(K + u[:, None], 0)
        u = -torch.logsumexp(K + v[None, :], 1)
    return torch.exp(K + u[:, None] + v[None, :])


def feature_cost(fs: torch.Tensor, ft: torch.Tensor, w: torch.Tensor, lam) -> torch.Tensor:
    """lam * sum_k w_k |fs_ik - ft_jk| (weighted L1 over the features)."""
    return lam * (w[None, None, :] * (fs[:, None, :] - ft[None, :, :]).abs()).sum(-1)


# ---------------------------------------------------------------------------
# Alignment quality (benchmark/experiment.py:115-141)
# ---------------------------------------------------------------------------

def _aligned_adjacencies(source_adjacency, target_adjacency, source_mapping, target_mapping):
    adj1 = source_adjacency[source_mapping][:, source_mapping]
    adj2 = target_adjacency[target_mapping][:, target_mapping]
    return adj1, adj2


def edge_correctness(source_adjacency, target_adjacency, source_mapping, target_mapping) -> float:
    """benchmark/experiment.py:115-121, dense numpy (0 on any error, as there).
    Called as Result.calculate does (:365-376): source_mapping = the computed
    answer, target_mapping = the ground truth, adjacencies from
    nx.to_numpy_array (rows in the graphs' node order).
    Pin: brute force over ordered node pairs on small graphs, 1 for the
    ground-truth answer on a noise-free pair, and the identities
    EC * #S = ICS * #T = conserved pairs."""
    try:
        adj1, adj2 = _aligned_adjacencies(source_adjacency, target_adjacency, source_mapping, target_mapping)
        return np.sum(adj1 + adj2 == 2) / np.sum(source_adjacency == 1)
    except Exception:
        return 0


def induced_conserved_structure(source_adjacency, target_adjacency, source_mapping, target_mapping) -> float:
    """benchmark/experiment.py:124-130.  Pin: as edge_correctness."""
    try:
        adj1, adj2 = _aligned_adjacencies(source_adjacency, target_adjacency, source_mapping, target_mapping)
        return np.sum(adj1 + adj2 == 2) / np.sum(adj2 == 1)
    except Exception:
        return 0


def symmetric_substructure(source_adjacency, target_adjacency, source_mapping, target_mapping) -> float:
    """benchmark/experiment.py:133-140.  Pin: as edge_correctness, and
    S3 = inter / (#S + #T - inter) from the brute-force counts."""
    try:
        adj1, adj2 = _aligned_adjacencies(source_adjacency, target_adjacency, source_mapping, target_mapping)
        intersection = np.sum(adj1 + adj2 == 2)
        return intersection / (np.sum(source_adjacency == 1) + np.sum(adj2 == 1) - intersection)
    except Exception:
        return 0


# ---------------------------------------------------------------------------
# The original FUGAL (second oracle of the Frank-Wolfe loop)
# ---------------------------------------------------------------------------

def sinkhorn_knopp_fugal(C: torch.Tensor, reg: float, max_iter: int) -> torch.Tensor:
    """FUGAL/sinkhorn.py:89-218 as convex_init calls it (marginals of ones,
    no log, so exactly max_iter iterations from u = v = 1/n): v = 1 / (u^T K),
    u = 1 / (K v), plan = diag(u) K diag(v).
    Pin: its plan agrees with solve_standard's after convergence (same fixed
    point, tests/test_oracle.py)."""
    n = C.shape[0]
    K = torch.exp(C / -reg)
    u = torch.ones(n, dtype=C.dtype) / n
    v = torch.ones(n, dtype=C.dtype) / n
    ones = torch.ones(n, dtype=C.dtype)
    for _ in range(max_iter):
        v = ones / (u @ K)
        u = ones / (K @ v)
    return u.reshape(-1, 1) * K * v.reshape(1, -1)


def fugal_original(source: nx.Graph, target: nx.Graph, niter: int, simple: bool, mu: float,
                   sinkhorn_iterations: int = 500):
    """Fugal.main + convex_init + convertToPermHungarian (FUGAL/Fugal.py:33-65,
    FUGAL/pred.py:122-143, :150-164) on graphs labelled 0..n-1: self-loops on
    isolated rows, features of FUGAL/pred.py:24-110, D = Euclidean distances,
    for lambda in 0..niter-1 and it in 1..10:
      G = -A^T P B - A P B^T + mu D + lambda (1 - 2P), q = Sinkhorn(G, reg 1,
      500 iterations (the reference's maxIter), fresh scalings),
      P += 2 / (2 + it) (q - P);
    then scipy's maximising assignment.  Returns (P, list of (row, col))."""
    n = max(source.number_of_nodes(), target.number_of_nodes())
    Src = nx.to_numpy_array(source, nodelist=range(n))
    Tar = nx.to_numpy_array(target, nodelist=range(n))
    for M in (Src, Tar):
        for i in range(n):
            if M[i].sum() == 0:
                M[i, i] = 1
    A = torch.tensor(Src, dtype=torch.float64)
    B = torch.tensor(Tar, dtype=torch.float64)
    F1 = extract_features_fugal(nx.from_numpy_array(Src), simple)
    F2 = extract_features_fugal(nx.from_numpy_array(Tar), simple)
    D = torch.tensor(np.sqrt(np.maximum(((F1[:, None, :] - F2[None, :, :]) ** 2).sum(-1), 0.0)), dtype=torch.float64)
    P = torch.ones((n, n), dtype=torch.float64) / n
    ones = torch.ones((n, n), dtype=torch.float64)
    K = mu * D
    for i in range(niter):
        for it in range(1, 11):
            G = -torch.mm(torch.mm(A.T, P), B) - torch.mm(torch.mm(A, P), B.T) + K + i * (ones - 2 * P)
            q = sinkhorn_knopp_fugal(G, 1.0, sinkhorn_iterations)
            alpha = 2.0 / float(2.0 + it)
            P = P + alpha * (q - P)
    row, col = scipy.optimize.linear_sum_assignment(P.numpy(), maximize=True)
    return P, [(int(r), int(c)) for r, c in zip(row, col)]
