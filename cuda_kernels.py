"""The reference's native module name (cuda_kernels/kernels.cpp:32-70) on
top of libcugal_hip.so; the functions live in cugal_amd/cuda_kernels.py."""

from cugal_amd.cuda_kernels import (  # noqa: F401
    EXPORTS, adjacency_matmul, add_distance, average_neighbor_features, create_adjacency, dense_hungarian,
    graph_features, regularize, sinkhorn_log_marginal, sinkhorn_log_step, update_quasi_permutation_log,
)
