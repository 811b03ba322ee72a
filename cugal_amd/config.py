"""Configuration of the alignment, field-for-field the surface of
cugal/config.py:11-141 (same enum names and values, same fields, defaults,
order and dict round trip), plus four extension fields at the end: the
Sinkhorn sweep structure, the gradient's arithmetic, the node-feature set
and the labels the loop runs on.  Their defaults give the reference's
results up to the rounding of FP32 sums.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass
from enum import Enum
from typing import Optional

import torch


class SinkhornMethod(str, Enum):
    """The method used for Sinkhorn-Knopp (cugal/config.py:11-27)."""

    STANDARD = "STANDARD"
    """The standard Sinkhorn-Knopp method (needs double precision)."""

    LOG = "LOG"
    """Sinkhorn-Knopp in logarithmic space: the MI355X kernel path."""

    MIX = "MIX"
    """Declared by the reference but never implemented there
    (cugal/sinkhorn.py:146-156 has no branch for it); rejected here."""


class HungarianMethod(Enum):
    """The method used to round the quasi-permutation (cugal/config.py:30-52)."""

    SCIPY = "SCIPY"
    GREEDY = "GREEDY"
    RAND = "RAND"
    MORE_RAND = "MORE_RAND"
    DOUBLE_GREEDY = "DOUBLE_GREEDY"
    PARALLEL_GREEDY = "PARALLEL_GREEDY"
    DENSE = "DENSE"
    SPARSE = "SPARSE"


@dataclass(frozen=True)
class Config:
    """Configuration of the cuGAL algorithm (cugal/config.py:55-141)."""

    safe_mode: bool = False
    """Check features, the quasi-permutation and the output for NaN/Inf."""

    device: str = "cpu"
    """The torch device used for computations.  The kernels here run on a
    ROCm device, spelled "cuda" / "cuda:N" by torch."""

    dtype: torch.dtype = torch.float64
    """The data type used for computations (the kernels here are FP32)."""

    torch_sinkhorn: bool = False
    """Reference flag selecting its torch update over the CUDA kernel
    (cugal/pred.py:134); the update here is always the HIP kernel."""

    sinkhorn_regularization: float = 1.0
    """Regularization of the cost matrix when running Sinkhorn."""

    sinkhorn_method: SinkhornMethod = SinkhornMethod.STANDARD
    """The version of Sinkhorn used."""

    sinkhorn_iterations: int = 500
    """The maximum number of Sinkhorn iterations performed."""

    dynamic_sinkhorn_regularization: bool = False
    """Unused by the reference (cugal/config.py:82); kept for the surface."""

    sinkhorn_scaling: float = 1.0
    """Unused by the reference (cugal/config.py:85); kept for the surface."""

    sinkhorn_threshold: float = 1e-3
    """The column-marginal L1 error at which Sinkhorn stops."""

    sinkhorn_eval_freq: int = 10
    """Iterations between marginal-error checks."""

    sinkhorn_momentum: bool = False
    """Unused by the reference (cugal/config.py:94); kept for the surface."""

    mu: float = 0.5
    """The contribution of node features in finding the alignment."""

    iter_count: int = 15
    """The number of lambda stages."""

    frank_wolfe_iter_count: int = 10
    """The number of Frank-Wolfe iterations per stage."""

    frank_wolfe_threshold: float | None = None
    """Duality gap below which a stage stops early (None: never)."""

    use_sparse_adjacency: bool = False
    """Use the sparse gradient form (lambda term once).  When False the dense
    form of the reference is followed, which adds the lambda term twice
    (cugal/pred.py:46-50)."""

    sinkhorn_cache_size: int = 0
    """Unused by the reference (cugal/config.py:112); kept for the surface."""

    sinkhorn_momentum_start: Optional[int] = None
    """Iteration from which Lehmann momentum is applied (None: never)."""

    recompute_distance: bool = False
    """Reference switch between a stored and a recomputed distance matrix.
    The kernels here always recompute the distance in the gradient epilogue
    (no n x n matrix); the flag is accepted and has no further effect."""

    hungarian_method: HungarianMethod = HungarianMethod.DOUBLE_GREEDY
    """The LAP used to round the quasi-permutation."""

    # ---- extension fields (not in the reference) --------------------------

    sinkhorn_fused: bool = True
    """Read the cost once per Sinkhorn iteration (row and column sweeps in
    one pass, n <= 32768) instead of twice.  Same iterates either way."""

    gradient_precision: str = "fp32"
    """Arithmetic of the A P B contraction: "fp32" (CSR adjacency times
    dense FP32 P), "bf16x2" (the 0/1 adjacency, exact in BF16, times P split
    into BF16 hi + lo on the matrix cores, FP32 accumulation) or "auto" (the
    matrix cores when the adjacencies hold more non-zeros than the measured
    crossing -- 1.17 % of n^2 at n = 16384, 0.75 % at n = 32768 --, where
    they become the faster form, else "fp32")."""

    node_features: str = "cugal"
    """The node features behind the distance term: "cugal" (the 4 of
    cugal/feature_extraction.py:41-64) or "fugal7" (the 7 of the original
    FUGAL, FUGAL/pred.py:24-110 with simple=False: the same 4 plus ego-net
    edges, edges leaving the ego net and nodes at distance 2)."""

    node_order: str = "auto"
    """Labels the Frank-Wolfe loop runs on: "natural" (the graphs' own),
    "rcm" (each graph relabelled by reverse Cuthill-McKee, so the CSR
    gathers of the gradient read nearby rows) or "auto" (RCM for a graph
    where it keeps clearly more edges local, from n = 8192, on the CSR
    gradient form).  The quasi-permutation is returned in the graphs' own
    labels; a relabelling changes only the order of floating-point sums."""

    def convert_tensor(self, input: torch.Tensor) -> torch.Tensor:
        """Convert tensor to correct type and dtype."""
        return input.to(dtype=self.dtype, device=self.device)

    def to_dict(self) -> dict:
        config_dict = dataclasses.asdict(self)
        config_dict["dtype"] = str(self.dtype).removeprefix("torch.")
        config_dict["sinkhorn_method"] = self.sinkhorn_method.value
        config_dict["hungarian_method"] = self.hungarian_method.value
        return config_dict

    @classmethod
    def from_dict(cls, config_dict: dict):
        config_dict = dict(config_dict)
        config_dict["dtype"] = getattr(torch, config_dict["dtype"])
        config_dict["sinkhorn_method"] = SinkhornMethod(config_dict["sinkhorn_method"])
        config_dict["hungarian_method"] = HungarianMethod[config_dict["hungarian_method"]]
        return cls(**config_dict)
