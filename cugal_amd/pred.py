"""The FUGAL alignment pipeline, the surface of cugal/pred.py:30-249.

``cugal(source, target, config, profile)`` pads the two graphs to one node
range, builds their CSR adjacency and structural features on the device,
runs ``iter_count`` lambda stages of ``frank_wolfe_iter_count`` Frank-Wolfe
steps (gradient -> Sinkhorn -> quasi-permutation update) and rounds the
result with the configured LAP.  Every n x n pass runs in the HIP kernels of
libcugal_hip.so; the host only sequences them (the SCIPY and SPARSE LAPs run
scipy on the host, as the reference does).  Supported: config.device a ROCm
device ("cuda"), dtype torch.float32, SinkhornMethod.LOG or STANDARD,
undirected graphs on either path and directed graphs on the dense path
(use_sparse_adjacency=False, as the reference).  Anything else raises rather
than falling back.
"""

from __future__ import annotations

import networkx as nx
import numpy as np
import scipy.optimize
import scipy.sparse
import scipy.sparse.csgraph
import torch

from . import _native as nat
from . import ordering
from .adjacency import Adjacency
from .config import Config, HungarianMethod, SinkhornMethod
from .feature_extraction import Features
from .greedy_lap import greedy_lap
from .hungarian import dense_hungarian
from .profile import Phase, Profile, SinkhornProfile, TimeStamp
from .sinkhorn import LogKernel, SinkhornState, StandardKernel


def _check_config(config: Config) -> None:
    if "cuda" not in config.device:
        raise NotImplementedError(
            "this library runs the alignment on a ROCm device: set Config(device='cuda', ...). "
            "The CPU path of the reference is not provided."
        )
    if config.dtype != torch.float32:
        raise NotImplementedError("the HIP kernels compute in FP32: set Config(dtype=torch.float32)")
    if config.sinkhorn_method not in (SinkhornMethod.LOG, SinkhornMethod.STANDARD):
        raise NotImplementedError(
            "SinkhornMethod.MIX has no solver in the reference (cugal/sinkhorn.py:146-156): use LOG or STANDARD"
        )
    if config.gradient_precision not in ("fp32", "bf16x2", "auto"):
        raise NotImplementedError(f"gradient_precision {config.gradient_precision!r} is not available")
    if config.node_order not in ordering.ORDERS:
        raise ValueError(f"node_order must be one of {ordering.ORDERS}, got {config.node_order!r}")


# Density above which the matrix-core gradient beats the CSR gathers on one
# MI355X, per n.  Measured per gradient on Erdos-Renyi pairs (random
# labels, so the gathers get no window locality): n = 16384: CSR 7.2 / 14.2 /
# 21.5 ms at densities 0.4 / 0.8 / 1.2 %, matrix cores 20.7 ms throughout
# (crossing at 1.17 %); n = 32768: CSR 40 / 80 / 123 / 171 ms at 0.2 / 0.4 /
# 0.6 / 0.8 %, matrix cores 159-160 ms (crossing at 0.75 %).  The gathers'
# cost per entry grows with n as the column strips outgrow the caches, so
# the crossing falls with n; the rule interpolates the two measurements as a
# power of n.
MFMA_DENSITY_16K = 0.0117
MFMA_DENSITY_EXPONENT = 0.64


def mfma_density(n: int) -> float:
    """Density from which "auto" takes the matrix-core gradient."""
    return MFMA_DENSITY_16K * (16384.0 / max(n, 1)) ** MFMA_DENSITY_EXPONENT


def gradient_precision_for(A: Adjacency, B: Adjacency, precision: str) -> str:
    """Resolve "auto": the matrix cores when the adjacencies are dense
    enough, else the CSR gathers.  An adjacency that lists a self-loop twice
    (the sparse form of a graph with loops) needs the CSR gathers: the
    matrix-core operand holds 0/1 entries."""
    if A.directed or B.directed:
        if precision == "bf16x2":
            raise NotImplementedError(
                "gradient_precision='bf16x2' is the symmetric form -2 A P B; a directed graph takes the CSR form"
            )
        return "fp32"
    if A.loops_twice or B.loops_twice:
        if precision == "bf16x2":
            raise NotImplementedError(
                "gradient_precision='bf16x2' holds 0/1 adjacency entries; with use_sparse_adjacency=True a "
                "self-loop counts twice (the reference's CSR lists it twice): use 'fp32' or 'auto'"
            )
        return "fp32" if precision == "auto" else precision
    if precision != "auto":
        return precision
    n = A.number_of_nodes()
    nnz = 0.5 * (A.col_indices.numel() + B.col_indices.numel())
    return "bf16x2" if nnz > mfma_density(n) * n * n else "fp32"


def _gradient_into(
    out: torch.Tensor, A: Adjacency, B: Adjacency, P: torch.Tensor, features, iteration: int, factor: float,
    precision: str = "fp32", A_transpose: Adjacency | None = None, B_transpose: Adjacency | None = None,
) -> torch.Tensor:
    """out = -2 A P B + D + factor * iteration * (1 - 2P) for symmetric A, B;
    for directed ones (or given transposes) -(A P B^T + A^T P B) + D +
    factor * iteration * (1 - 2P).  `features` is a Features (distance
    recomputed in the kernel), an n x n distance tensor (added afterwards)
    or None.  precision "fp32": CSR adjacency times FP32 P (cg_gradient,
    cg_gradient_directed); "bf16x2": dense 0/1 BF16 adjacency times P split
    into BF16 hi + lo on the matrix cores, FP32 accumulation
    (cg_gradient_mfma, symmetric graphs)."""
    n = P.shape[0]
    dev = P.device
    if P.dim() != 2 or P.stride(1) != 1:
        P = P.contiguous()
    ldp = nat.row_stride(P, "P")
    fs = ft = None
    fdim = 4
    if isinstance(features, Features):
        fs, ft = features.source.contiguous(), features.target.contiguous()
        fdim = int(fs.shape[1])
    general = A.directed or B.directed or (A_transpose is not None and A_transpose is not A) or (
        B_transpose is not None and B_transpose is not B)
    precision = gradient_precision_for(A, B, precision)
    if general:
        if precision == "bf16x2":
            raise NotImplementedError("the matrix-core gradient is the symmetric form -2 A P B")
        At = A_transpose if A_transpose is not None else A.transposed()
        Bt = B_transpose if B_transpose is not None else B.transposed()
        ws = nat.workspace(dev, nat.query("cg_gradient_directed_workspace", n), "gradient")
        nat.call(
            "cg_gradient_directed", nat.ptr(A.col_indices), nat.ptr(A.row_pointers), nat.ptr(At.col_indices),
            nat.ptr(At.row_pointers), nat.ptr(B.col_indices), nat.ptr(B.row_pointers), nat.ptr(Bt.col_indices),
            nat.ptr(Bt.row_pointers), n, nat.ptr(P), ldp, nat.ptr(fs), nat.ptr(ft), fdim, int(iteration),
            float(factor), nat.ptr(out), nat.ptr(ws), ws.numel(), nat.stream_of(dev),
        )
    elif precision == "bf16x2":
        ws = nat.workspace(dev, nat.query("cg_gradient_mfma_workspace", n), "gradient_mfma")
        nat.call(
            "cg_gradient_mfma", nat.ptr(A.packed_bits()), nat.ptr(B.packed_bits()), n, nat.ptr(P), ldp, nat.ptr(fs),
            nat.ptr(ft), fdim, int(iteration), float(factor), nat.ptr(out), nat.ptr(ws), ws.numel(), nat.stream_of(dev),
        )
    elif precision == "fp32":
        ws = nat.workspace(dev, nat.query("cg_gradient_workspace", n), "gradient")
        nat.call(
            "cg_gradient", nat.ptr(A.col_indices), nat.ptr(A.row_pointers), nat.ptr(B.col_indices),
            nat.ptr(B.row_pointers), n, nat.ptr(P), ldp, nat.ptr(fs), nat.ptr(ft), fdim, int(iteration),
            float(factor), nat.ptr(out), nat.ptr(ws), ws.numel(), nat.stream_of(dev),
        )
    else:
        raise NotImplementedError(f"gradient precision {precision!r}")
    if isinstance(features, torch.Tensor):
        out += features
    return out


def add_feature_distance(gradient: torch.Tensor, features: torch.Tensor | Features) -> torch.Tensor:
    """gradient += D (cugal/pred.py:30-35)."""
    if isinstance(features, Features):
        return features.add_distance(gradient)
    gradient += features
    return gradient


def dense_gradient(A: torch.Tensor, B: torch.Tensor, P: torch.Tensor, features, iteration: int) -> torch.Tensor:
    """-A^T P B - A P B^T + D + 2 iteration (1 - 2P) for dense adjacency
    tensors, symmetric or not (cugal/pred.py:38-51, whose regulariser is
    added twice); the non-zeros of A and B are taken as CSR on the device."""
    out = torch.empty_like(P)
    return _gradient_into(out, Adjacency.from_dense(A), Adjacency.from_dense(B), P, features, iteration, 2.0)


def sparse_gradient(A: Adjacency, B: Adjacency, A_transpose: Adjacency, B_transpose: Adjacency, P: torch.Tensor,
                    features, iteration: int) -> torch.Tensor:
    """-2 A P B + D + iteration (1 - 2P) when A is A_transpose and B is
    B_transpose, else the general -(A^T P B + A P B^T) + D + iteration
    (1 - 2P) (cugal/pred.py:54-74)."""
    out = torch.empty_like(P)
    if A is A_transpose and B is B_transpose:
        return _gradient_into(out, A, B, P, features, iteration, 1.0)
    return _gradient_into(out, A, B, P, features, iteration, 1.0, A_transpose=A_transpose, B_transpose=B_transpose)


def update_quasi_permutation_log(P: torch.Tensor, K: LogKernel, u: torch.Tensor, v: torch.Tensor, alpha: float,
                                 sinkhorn_regularization: float | None = None, gap: torch.Tensor | None = None):
    """P += alpha (exp(K + u + v) - P) in one pass; returns |sum C (plan - P)|
    as a float, or leaves the signed sum in `gap` (device FP64) without
    synchronising when `gap` is given (cuda_kernels/update_quasi_permutation.cu:40-56)."""
    if not isinstance(K, LogKernel):
        K = LogKernel(-K * sinkhorn_regularization, sinkhorn_regularization)
    n = P.shape[0]
    dev = P.device
    out = torch.empty(1, dtype=torch.float64, device=dev) if gap is None else gap
    ws = nat.workspace(dev, nat.query("cg_update_workspace", n), "update")
    nat.call(
        "cg_update_quasi_permutation_log", nat.ptr(P), nat.row_stride(P, "P"), nat.ptr(K.cost), n, nat.ptr(u),
        nat.ptr(v), float(alpha), float(K.regularization), nat.ptr(out), nat.ptr(ws), ws.numel(),
        nat.stream_of(dev),
    )
    return abs(float(out.item())) if gap is None else None


def update_quasi_permutation_standard(P: torch.Tensor, K: StandardKernel, u: torch.Tensor, v: torch.Tensor,
                                      alpha: float, gap: torch.Tensor | None = None):
    """P += alpha (diag(u) K diag(v) - P) in one pass with K = exp(-C/reg)
    (cugal/pred.py:77-86 with scale_kernel_matrix); the gap as in
    update_quasi_permutation_log."""
    n = P.shape[0]
    dev = P.device
    out = torch.empty(1, dtype=torch.float64, device=dev) if gap is None else gap
    ws = nat.workspace(dev, nat.query("cg_update_workspace", n), "update")
    nat.call(
        "cg_update_quasi_permutation_standard", nat.ptr(P), nat.row_stride(P, "P"), nat.ptr(K.cost), n,
        nat.ptr(u), nat.ptr(v), float(alpha), float(K.regularization), nat.ptr(out), nat.ptr(ws), ws.numel(),
        nat.stream_of(dev),
    )
    return abs(float(out.item())) if gap is None else None


def update_quasi_permutation(P: torch.Tensor, K, u: torch.Tensor, v: torch.Tensor, alpha: float,
                             sinkhorn_method: SinkhornMethod) -> torch.Tensor:
    """Reference-compatible form (cugal/pred.py:77-86): updates P and returns
    the step alpha (plan - P)."""
    before = P.clone()
    gap = torch.empty(1, dtype=torch.float64, device=P.device)
    match sinkhorn_method:
        case SinkhornMethod.LOG:
            update_quasi_permutation_log(P, K, u, v, alpha, gap=gap)
        case SinkhornMethod.STANDARD:
            update_quasi_permutation_standard(P, K, u, v, alpha, gap=gap)
        case _:
            raise NotImplementedError("SinkhornMethod.MIX has no solver in the reference")
    return P - before


def _adjacency_of(graph, config: Config, loops_once: bool = False) -> Adjacency:
    """The CSR the features are computed from (for a directed graph, row y
    lists the predecessors of y, as the reference's create_adjacency)."""
    if isinstance(graph, Adjacency):
        return graph
    if isinstance(graph, torch.Tensor):
        return Adjacency.from_dense(graph.to(config.device))
    return Adjacency.from_graph(graph, config.device, loops_once=loops_once)


def _gradient_adjacency_of(graph, config: Config) -> Adjacency:
    """The matrix the gradient multiplies by: nx.to_numpy_array of the graph
    on the dense path (a loop once; a directed edge x -> y the entry [x][y],
    cugal/pred.py:106-110), its CSR on the sparse path (undirected only, as
    the reference asserts at cugal/pred.py:98-105)."""
    if isinstance(graph, (Adjacency, torch.Tensor)):
        return _adjacency_of(graph, config)
    if nx.is_directed(graph):
        if config.use_sparse_adjacency:
            raise NotImplementedError(
                "directed graphs take the dense path (use_sparse_adjacency=False); the reference's sparse path "
                "asserts undirected graphs (cugal/pred.py:98-105)"
            )
        return Adjacency.from_graph(graph, config.device).transposed()
    return Adjacency.from_graph(graph, config.device, loops_once=not config.use_sparse_adjacency)


def find_quasi_permutation_matrix(A, B, features, config: Config, profile: Profile) -> torch.Tensor:
    """The Frank-Wolfe loop of cugal/pred.py:89-148: lambda = 0..iter_count-1,
    step it = 1..frank_wolfe_iter_count with alpha = 2 / (2 + it), Sinkhorn
    duals warm-started across all steps."""
    _check_config(config)
    # Graphs given as networkx: the dense path's adjacency holds a self-loop
    # once (nx.to_numpy_array), the sparse path's CSR twice.
    A = _gradient_adjacency_of(A, config)
    B = _gradient_adjacency_of(B, config)
    n = A.number_of_nodes()
    if B.number_of_nodes() != n:
        raise ValueError(f"graphs differ in size ({n} vs {B.number_of_nodes()})")
    dev = A.row_pointers.device
    precision = gradient_precision_for(A, B, config.gradient_precision)
    # The CSR gathers run on graphs relabelled for locality where that
    # pays; P is copied back to the original labels at the end.
    row_position = col_position = None
    if precision == "fp32" and (features is None or isinstance(features, Features)):
        pa = ordering.choose_position(A, config.node_order)
        pb = ordering.choose_position(B, config.node_order)
        if pa is not None:
            row_position = torch.from_numpy(pa).to(dev)
            A = ordering.relabel(A, row_position)
        if pb is not None:
            col_position = torch.from_numpy(pb).to(dev)
            B = ordering.relabel(B, col_position)
        if features is not None and (pa is not None or pb is not None):
            features = Features(
                features.source if pa is None else ordering.permute_rows(features.source, row_position),
                features.target if pb is None else ordering.permute_rows(features.target, col_position),
            )
    sinkhorn_state = SinkhornState(n, config)
    # P lives in a row-padded buffer during the loop: the gradient gathers
    # its rows in place (a 2^k-float row stride would alias L2 sets and cost
    # a copy per step); the caller gets a contiguous copy.
    ldp = int(nat.query("cg_gradient_padded_ld", n))
    P = torch.empty((n, ldp), device=dev, dtype=torch.float32)[:, :n]
    P.fill_(1.0 / n)
    gradient = torch.empty((n, n), device=dev, dtype=torch.float32)
    gap = torch.empty(1, dtype=torch.float64, device=dev)
    factor = 1.0 if config.use_sparse_adjacency else 2.0
    for lam in range(config.iter_count):
        for it in range(1, config.frank_wolfe_iter_count + 1):
            start_time = TimeStamp(config.device)
            _gradient_into(gradient, A, B, P, features, lam, factor, precision)
            profile.log_time(start_time, Phase.GRADIENT)

            start_time = TimeStamp(config.device)
            sinkhorn_profile = SinkhornProfile()
            K, u, v = sinkhorn_state.solve(gradient, config, sinkhorn_profile)
            profile.sinkhorn_profiles.append(sinkhorn_profile)
            profile.log_time(start_time, Phase.SINKHORN)

            alpha = 2.0 / float(2.0 + it)
            if config.sinkhorn_method == SinkhornMethod.STANDARD:
                update_quasi_permutation_standard(P, K, u, v, alpha, gap=gap)
            else:
                update_quasi_permutation_log(P, K, u, v, alpha, gap=gap)

            profile.frank_wolfe_iterations += 1
            if config.frank_wolfe_threshold is not None and abs(float(gap.item())) < config.frank_wolfe_threshold:
                break
    if row_position is not None or col_position is not None:
        return ordering.copy_back(P, row_position, col_position)
    return P.contiguous()


def hungarian(quasi_permutation: torch.Tensor, config: Config, profile: Profile):
    """Round to a permutation (cugal/pred.py:151-178); returns the column
    of each row."""
    profile.sparsity = (
        torch.sum(torch.le(quasi_permutation, torch.finfo(quasi_permutation.dtype).eps)).item()
        / quasi_permutation.numel()
    )
    start_time = TimeStamp(config.device)
    match config.hungarian_method:
        case HungarianMethod.SCIPY:
            _, column_indices = scipy.optimize.linear_sum_assignment(quasi_permutation.cpu().numpy(), maximize=True)
        case (HungarianMethod.GREEDY | HungarianMethod.RAND | HungarianMethod.MORE_RAND | HungarianMethod.DOUBLE_GREEDY
              | HungarianMethod.PARALLEL_GREEDY):
            column_indices = greedy_lap(quasi_permutation, config)
        case HungarianMethod.DENSE:
            column_indices = dense_hungarian(quasi_permutation, maximize=True).cpu().numpy().astype(np.int64)
        case HungarianMethod.SPARSE:
            sparse = scipy.sparse.csr_matrix(quasi_permutation.cpu().numpy())
            _, column_indices = scipy.sparse.csgraph.min_weight_full_bipartite_matching(sparse, maximize=True)
        case _:
            raise NotImplementedError(f"{config.hungarian_method} is not implemented on the ROCm path")
    profile.log_time(start_time, Phase.HUNGARIAN)
    return column_indices


def convert_to_permutation_matrix(quasi_permutation: torch.Tensor, config: Config, profile: Profile):
    """(quasi-permutation as numpy, [(i, col_i)]) (cugal/pred.py:181-193).
    The reference also builds a dense permutation matrix it never returns;
    that allocation is skipped."""
    col_ind = hungarian(quasi_permutation, config, profile)
    mapping = [(i, int(col_ind[i])) for i in range(len(quasi_permutation))]
    return to_host(quasi_permutation), mapping


def align_device(source: Adjacency, target: Adjacency, config: Config, profile: Profile | None = None,
                 gradient_adjacencies: tuple[Adjacency, Adjacency] | None = None):
    """The alignment from device-resident CSR adjacencies: features, the
    Frank-Wolfe loop and the LAP.  Returns (P on the device, column of each
    row).  ``cugal`` is this plus graph conversion and the host copy of P.
    `gradient_adjacencies` replaces (source, target) in the gradient: the
    dense path of a graph with self-loops (loops listed once there, twice in
    the CSR the features are computed from)."""
    _check_config(config)
    profile = Profile() if profile is None else profile
    start_time = TimeStamp(config.device)
    features = Features.create(source, target, config)
    if config.safe_mode:
        assert features.source.isfinite().all(), "source feature tensor has NaN values"
        assert features.target.isfinite().all(), "target feature tensor has NaN values"
    profile.log_time(start_time, Phase.FEATURE_EXTRACTION)

    grad_source, grad_target = gradient_adjacencies or (source, target)
    quasi_permutation = find_quasi_permutation_matrix(grad_source, grad_target, features, config, profile)
    if config.safe_mode:
        assert quasi_permutation.isfinite().all(), "quasi_permutation tensor has NaN values"
    col_ind = hungarian(quasi_permutation, config, profile)
    return quasi_permutation, col_ind


def to_host(tensor: torch.Tensor) -> np.ndarray:
    """A device tensor as a numpy array, copied through page-locked host
    memory (about 13x faster than a pageable copy for a 4 GiB P; the array
    keeps its pinned buffer alive)."""
    host = torch.empty(tensor.shape, dtype=tensor.dtype, pin_memory=True)
    host.copy_(tensor)
    return host.numpy()


def cugal(source: nx.Graph, target: nx.Graph, config: Config, profile: Profile = Profile()):
    """Run the alignment (cugal/pred.py:196-249).  Returns the quasi-
    permutation (after rounding, as numpy) and the mapping from source to
    target nodes.  Like the reference, pads both graphs in place with the
    nodes of range(max label) they lack, and `profile` defaults to one shared
    instance."""
    _check_config(config)
    if config.use_sparse_adjacency and (nx.is_directed(source) or nx.is_directed(target)):
        raise NotImplementedError(
            "directed graphs take the dense path (use_sparse_adjacency=False); the reference's sparse path asserts "
            "undirected graphs (cugal/pred.py:98-105)"
        )
    before = TimeStamp("cpu")

    node_count = max(max(source.nodes()), max(target.nodes()))
    source.add_nodes_from(set(range(node_count)) - set(source.nodes()))
    target.add_nodes_from(set(range(node_count)) - set(target.nodes()))
    if source.number_of_nodes() != target.number_of_nodes():
        # range(max label) leaves out the largest label, so one graph can
        # keep a node the other lacks (the reference then fails on a shape
        # mismatch inside the gradient)
        raise ValueError(f"graphs differ in size after padding ({source.number_of_nodes()} vs "
                         f"{target.number_of_nodes()} nodes)")
    if config.node_features == "fugal7" and (nx.number_of_selfloops(source) or nx.number_of_selfloops(target)
                                             or nx.is_directed(source) or nx.is_directed(target)):
        raise ValueError("the 7-feature set is defined here for simple undirected graphs")

    source_adj = _adjacency_of(source, config)
    target_adj = _adjacency_of(target, config)
    gradient_adjacencies = None
    if source_adj.directed or target_adj.directed or (
            not config.use_sparse_adjacency and (source_adj.loops_twice or target_adj.loops_twice)):
        # The reference's dense gradient multiplies by nx.to_numpy_array (a
        # loop is 1 on the diagonal, a directed edge x -> y the entry
        # [x][y]); its features come from the CSR (a loop twice, a directed
        # edge in the row of its head).
        gradient_adjacencies = (_gradient_adjacency_of(source, config), _gradient_adjacency_of(target, config))

    quasi_permutation, col_ind = align_device(source_adj, target_adj, config, profile, gradient_adjacencies)
    mapping = [(i, int(col_ind[i])) for i in range(len(quasi_permutation))]
    output = to_host(quasi_permutation), mapping
    if config.safe_mode:
        assert np.isfinite(output[0]).all(), "output permutation matrix has NaN values"

    profile.time = TimeStamp("cpu").elapsed_seconds(before)
    profile.max_memory = torch.cuda.max_memory_allocated(config.device)
    torch.cuda.reset_peak_memory_stats(config.device)
    return output
