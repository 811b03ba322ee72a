"""Exact LAP on the device (cg_lap_hungarian) against scipy's
linear_sum_assignment, which the oracle pins by brute force
(oracle.fugal.scipy_lap; tests/test_oracle.py)."""

import itertools
import time

import numpy as np
import pytest
import scipy.optimize
import torch

from cugal_amd._native import NativeError
from cugal_amd.hungarian import dense_hungarian
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu


def _objective(M: np.ndarray, cols) -> float:
    return float(M.astype(np.float64)[np.arange(len(cols)), np.asarray(cols)].sum())


def _is_perm(cols, n):
    return sorted(int(c) for c in cols) == list(range(n))


@pytest.mark.parametrize("maximize", [False, True])
@pytest.mark.parametrize("seed", range(6))
def test_brute_force_small(device, seed, maximize):
    g = torch.Generator().manual_seed(seed)
    n = 1 + seed
    M = torch.rand((n, n), generator=g)
    best = None
    for perm in itertools.permutations(range(n)):
        s = float(M.double()[torch.arange(n), torch.tensor(perm)].sum())
        if best is None or (s > best[0] if maximize else s < best[0]):
            best = (s, list(perm))
    got = dense_hungarian(M.to(device), maximize=maximize).cpu().tolist()
    assert got == best[1]


@pytest.mark.parametrize("maximize", [False, True])
@pytest.mark.parametrize("n", [100, 1000, 1025, 3000])
def test_random_matches_scipy(device, n, maximize):
    g = torch.Generator().manual_seed(n)
    M = torch.rand((n, n), generator=g)
    got = dense_hungarian(M.to(device), maximize=maximize).cpu().numpy()
    _, want = scipy.optimize.linear_sum_assignment(M.numpy().astype(np.float64), maximize=maximize)
    assert _is_perm(got, n)
    assert np.array_equal(got, want)


def test_integer_costs_with_ties_reach_the_optimum(device):
    # Many optimal assignments: the objective must equal scipy's.
    rng = np.random.default_rng(3)
    M = rng.integers(0, 10, size=(300, 300)).astype(np.float32)
    got = dense_hungarian(torch.from_numpy(M).to(device)).cpu().numpy()
    _, want = scipy.optimize.linear_sum_assignment(M)
    assert _is_perm(got, 300)
    assert _objective(M, got) == _objective(M, want)


def test_constant_matrix_gives_a_permutation(device):
    M = torch.full((2000, 2000), 0.25, device=device)
    got = dense_hungarian(M, maximize=True).cpu().tolist()
    assert _is_perm(got, 2000)


@pytest.mark.parametrize("n", [4096, 8192, 12000])
def test_quasi_permutation_matches_scipy(device, n):
    # The shape HungarianMethod.DENSE sees: a doubly stochastic matrix
    # concentrated near a permutation, with competing rows.
    g = torch.Generator().manual_seed(n)
    perm = torch.randperm(n, generator=g)
    Q = torch.rand((n, n), generator=g) * (0.2 / n)
    Q[torch.arange(n), perm] += 0.5 + 0.3 * torch.rand(n, generator=g)
    swap = torch.randperm(n, generator=g)[: n // 10]
    Q[swap, perm[torch.roll(swap, 1)]] += 0.6  # 10 % of rows with a rival column
    Qd = Q.to(device)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    got, (free_rows, steps) = dense_hungarian(Qd, maximize=True, stats=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    got = got.cpu().numpy()
    t1 = time.perf_counter()
    want = oracle.scipy_lap(Q.double())
    dts = time.perf_counter() - t1
    print(f"\nhungarian n={n}: device {dt * 1e3:.1f} ms ({free_rows} free rows after the warm start, "
          f"{steps} Dijkstra steps); scipy {dts * 1e3:.1f} ms")
    assert np.array_equal(got, want)


def test_full_size_quasi_permutation(device):
    n = 32768
    g = torch.Generator(device=device).manual_seed(0)
    perm = torch.randperm(n, generator=g, device=device)
    Q = torch.rand((n, n), generator=g, device=device) * (0.2 / n)
    Q[torch.arange(n, device=device), perm] += 0.5
    swap = torch.randperm(n, generator=g, device=device)[: n // 20]
    Q[swap, perm[torch.roll(swap, 1)]] += 0.6
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    got, (free_rows, steps) = dense_hungarian(Q, maximize=True, stats=True)
    torch.cuda.synchronize()
    print(f"\nhungarian n={n}: {(time.perf_counter() - t0) * 1e3:.1f} ms, {free_rows} free rows, {steps} steps")
    got = got.long()
    assert _is_perm(got.cpu().tolist(), n)
    # Optimality certificate on the device: the objective is at least that
    # of the planted assignment and of the greedy row argmax where it is a
    # permutation; and no 2-swap improves it.
    obj = Q[torch.arange(n, device=device), got].double().sum()
    planted = Q[torch.arange(n, device=device), perm].double().sum()
    assert float(obj) >= float(planted) - 1e-9
    a, b = torch.randint(0, n, (2, 200000), device=device)
    gain = (Q[a, got[b]] + Q[b, got[a]]).double() - (Q[a, got[a]] + Q[b, got[b]]).double()
    assert float(gain.max()) <= 1e-6


def test_bad_costs_raise(device):
    M = torch.rand((64, 64), device=device)
    M[3, 5] = float("nan")
    with pytest.raises(NativeError):
        dense_hungarian(M)
    M = torch.rand((64, 64), device=device)
    M[:, 7] = float("inf")
    with pytest.raises(NativeError):
        dense_hungarian(M)
