"""FP64 checkers for the full-size workloads, computed on the device with
torch (row chunks, so no n x n FP64 matrix is held at n = 65536).

They restate, in FP64, the formulas the reference's path computes:
  gradient (sparse form, cugal/pred.py:54-74):
      G = -2 A P B + D + factor * lam * (1 - 2P)
  log-Sinkhorn (cugal/sinkhorn.py:90-120): K = -C/reg,
      v = -LSE_i(K + u), u = -LSE_j(K + v)
  plan exp(K + u + v), its row sums and column-marginal L1 error
      (cugal/sinkhorn.py:26-40)
  Frank-Wolfe update and gap (cugal/pred.py:77-86, :140-141):
      P' = P + alpha (plan - P), gap = sum G (plan - P)
"""

from __future__ import annotations

import numpy as np
import scipy.sparse
import torch

LOG2E = 1.4426950408889634


def exponent_floor(C_absmax: float, reg: float, u=None, v=None) -> float:
    """Relative error of an FP32 log-domain plan entry: its exponent
    K2 = C * (-log2 e / reg) + u2 + v2 carries about 2^-24 (4 + |K2|)
    absolute rounding, which is a relative error of the plan value.  With
    the duals given, their magnitudes count too (the sum is rounded at the
    size of its largest term)."""
    mag = C_absmax / reg
    if u is not None:
        mag += float(u.abs().max()) + float(v.abs().max())
    return 2.0 ** -24 * (4.0 + mag * LOG2E)


def update_bound(plan, old, want, alpha: float, floor: float):
    """Elementwise bound of the FP32 update P + alpha (plan - P): the plan's
    exponent error, then a few roundings at the size of the operands."""
    return alpha * plan * 4 * floor + 2.0 ** -22 * (abs(want) + alpha * (plan + abs(old))) + 1e-30


def max_relative(got, want, rtol: float, what: str, tiny: float = 1e-200) -> float:
    """Assert |got - want| <= rtol * want + tiny elementwise (want >= 0, as
    plan entries are); return the largest relative error over the entries
    above `tiny` (entries below it, in both, underflow in FP32)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    err = np.abs(got - want)
    assert np.all(err <= rtol * want + tiny), what
    big = want > tiny
    return float(np.max(err[big] / want[big])) if big.any() else 0.0


def csr_host(adj) -> scipy.sparse.csr_matrix:
    n = adj.number_of_nodes()
    rp = adj.row_pointers.cpu().numpy()
    col = adj.col_indices.cpu().numpy()
    return scipy.sparse.csr_matrix((np.ones(len(col)), col, rp), shape=(n, n))


def gradient_rows(A, B, P: torch.Tensor, features, lam: float, factor: float, rows):
    """FP64 rows of G = -2 A P B + D + factor lam (1 - 2P) for symmetric
    A, B, and the elementwise magnitudes (|2APB|, |D| + |lam term|) that
    bound its rounding.  Row i of A P B is (sum_{k in N_A(i)} P[k, :]) B;
    B symmetric, so y B = B y."""
    Bs = csr_host(B)
    arp = A.row_pointers.cpu().numpy()
    acol = A.col_indices.cpu().numpy()
    refs, prods, others = [], [], []
    for i in rows:
        nbrs = acol[arp[i]:arp[i + 1]]
        if len(nbrs):
            y = P[torch.as_tensor(nbrs, device=P.device, dtype=torch.long)].double().sum(0).cpu().numpy()
        else:
            y = np.zeros(P.shape[1])
        prod = 2.0 * (Bs @ y)
        p_i = P[int(i)].double().cpu().numpy()
        reg_term = factor * lam * (1.0 - 2.0 * p_i)
        if features is not None:
            fs = features.source[int(i)].double().cpu()
            ft = features.target.double().cpu()
            dist = torch.sqrt(((fs[None, :] - ft) ** 2).sum(1)).numpy()
        else:
            dist = np.zeros_like(p_i)
        refs.append(-prod + dist + reg_term)
        prods.append(np.abs(prod))
        others.append(np.abs(dist) + np.abs(reg_term))
    return np.array(refs), np.array(prods), np.array(others)


def _chunks(n: int, rows_per_chunk: int):
    for r0 in range(0, n, rows_per_chunk):
        yield r0, min(n, r0 + rows_per_chunk)


def log_sinkhorn(C: torch.Tensor, reg: float, u0: torch.Tensor, v0: torch.Tensor, iterations: int,
                 rows_per_chunk: int = 2048):
    """`iterations` FP64 log-domain iterations from (u0, v0) (FP32 values
    taken exactly), streaming C in row chunks; returns FP64 (u, v)."""
    n = C.shape[0]
    u = u0.double().clone()
    v = v0.double().clone()
    for _ in range(iterations):
        # v_j = -LSE_i(K_ij + u_i): running column (max, sum) over chunks
        m = torch.full((n,), -float("inf"), dtype=torch.float64, device=C.device)
        s = torch.zeros(n, dtype=torch.float64, device=C.device)
        for r0, r1 in _chunks(n, rows_per_chunk):
            X = C[r0:r1].double() / -reg + u[r0:r1, None]
            cm = X.max(0).values
            mn = torch.maximum(m, cm)
            s = s * torch.exp(m - mn) + torch.exp(X - mn[None, :]).sum(0)
            m = mn
        v = -(m + torch.log(s))
        for r0, r1 in _chunks(n, rows_per_chunk):
            X = C[r0:r1].double() / -reg + v[None, :]
            u[r0:r1] = -torch.logsumexp(X, dim=1)
    return u, v


def plan_rows(C: torch.Tensor, reg: float, u: torch.Tensor, v: torch.Tensor, rows) -> np.ndarray:
    """FP64 plan exp(-C/reg + u + v) on the given rows."""
    idx = torch.as_tensor(list(rows), device=C.device, dtype=torch.long)
    X = C[idx].double() / -reg + u.double()[idx, None] + v.double()[None, :]
    return torch.exp(X).cpu().numpy()


def marginals(C: torch.Tensor, reg: float, u: torch.Tensor, v: torch.Tensor, rows_per_chunk: int = 2048):
    """FP64 row sums and column sums of exp(-C/reg + u + v)."""
    n = C.shape[0]
    u, v = u.double(), v.double()
    col = torch.zeros(n, dtype=torch.float64, device=C.device)
    row = torch.empty(n, dtype=torch.float64, device=C.device)
    for r0, r1 in _chunks(n, rows_per_chunk):
        E = torch.exp(C[r0:r1].double() / -reg + u[r0:r1, None] + v[None, :])
        col += E.sum(0)
        row[r0:r1] = E.sum(1)
    return row, col


def inner(C: torch.Tensor, P: torch.Tensor, rows_per_chunk: int = 2048) -> tuple[float, float]:
    """sum C * P and sum |C * P| in FP64."""
    n = C.shape[0]
    s = a = 0.0
    for r0, r1 in _chunks(n, rows_per_chunk):
        prod = C[r0:r1].double() * P[r0:r1].double()
        s += float(prod.sum())
        a += float(prod.abs().sum())
    return s, a


def plan_inner(C: torch.Tensor, reg: float, u: torch.Tensor, v: torch.Tensor, rows_per_chunk: int = 2048):
    """sum C * plan and sum |C * plan| in FP64, plan = exp(-C/reg + u + v)."""
    n = C.shape[0]
    u, v = u.double(), v.double()
    s = a = 0.0
    for r0, r1 in _chunks(n, rows_per_chunk):
        Cd = C[r0:r1].double()
        prod = Cd * torch.exp(Cd / -reg + u[r0:r1, None] + v[None, :])
        s += float(prod.sum())
        a += float(prod.abs().sum())
    return s, a


def random_quasi_permutation(n: int, ld: int, device, seed: int, rows_per_chunk: int = 4096) -> torch.Tensor:
    """A positive row-stochastic n x n FP32 matrix in a buffer of row
    stride `ld` (the layout the Frank-Wolfe loop keeps P in), filled in row
    chunks from a seeded device generator."""
    P = torch.empty((n, ld), device=device, dtype=torch.float32)[:, :n]
    g = torch.Generator(device=device).manual_seed(seed)
    for r0, r1 in _chunks(n, rows_per_chunk):
        X = torch.rand((r1 - r0, n), generator=g, device=device) + 0.05
        P[r0:r1] = X / X.sum(1, keepdim=True)
    return P
