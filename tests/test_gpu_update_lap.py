"""FW update (cugal/pred.py:77-86, :133-141) and greedy LAP
(cugal/greedy_lap.py:10-63) against the oracle."""

import numpy as np
import pytest
import torch

from cugal_amd.config import HungarianMethod
from cugal_amd.greedy_lap import greedy_lap_device
from cugal_amd.pred import update_quasi_permutation_log
from cugal_amd.sinkhorn import LogKernel
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [128, 1000, 777])
def test_update_matches_oracle(device, n):
    g = torch.Generator().manual_seed(n)
    C = torch.rand((n, n), generator=g) * 5
    u = torch.randn(n, generator=g) * 0.5
    v = torch.randn(n, generator=g) * 0.5 - 3.0
    P = torch.rand((n, n), generator=g) / n
    reg, alpha = 0.7, 2.0 / 3.0
    Pd = P.to(device).contiguous()
    gap = torch.empty(1, dtype=torch.float64, device=device)
    update_quasi_permutation_log(Pd, LogKernel(C.to(device), reg), u.to(device), v.to(device), alpha, gap=gap)
    K = -C.double() / reg
    plan = oracle.plan_log(K, u.double(), v.double())
    Pr = P.double().clone()
    step = oracle.update_quasi_permutation(Pr, plan, alpha)
    want_gap = oracle.duality_gap(C.double(), step, alpha)
    np.testing.assert_allclose(Pd.double().cpu().numpy(), Pr.numpy(), rtol=1e-5, atol=1e-9)
    assert abs(abs(float(gap.item())) - want_gap) <= 1e-4 * want_gap + 1e-6


def test_update_alpha_one_gives_plan(device):
    n = 256
    g = torch.Generator().manual_seed(1)
    C = torch.rand((n, n), generator=g)
    u, v = torch.randn(n, generator=g), torch.randn(n, generator=g)
    P = torch.rand((n, n), generator=g).to(device)
    update_quasi_permutation_log(P, LogKernel(C.to(device), 1.0), u.to(device), v.to(device), 1.0,
                                 gap=torch.empty(1, dtype=torch.float64, device=device))
    plan = torch.exp(-C.double() + u.double()[:, None] + v.double()[None, :])
    # P + (plan - P) in FP32 cancels: the error is an ulp of max(P, plan)
    # (P ~ U(0, 1)), not of the plan entry.
    np.testing.assert_allclose(P.double().cpu().numpy(), plan.numpy(), rtol=2e-6, atol=1.2e-7)


def _distinct_rowmax_matrix(n, seed):
    g = torch.Generator().manual_seed(seed)
    M = torch.rand((n, n), generator=g) / n
    perm = torch.randperm(n, generator=g)
    M[torch.arange(n), perm] += torch.rand(n, generator=g) * 0.5 + 0.1
    return M


@pytest.mark.parametrize("method", ["GREEDY", "DOUBLE_GREEDY"])
@pytest.mark.parametrize("n", [50, 1000, 1537])
def test_greedy_matches_oracle(device, method, n):
    M = _distinct_rowmax_matrix(n, seed=n)
    Md = M.to(device)
    got = greedy_lap_device(Md, HungarianMethod[method]).cpu().numpy()
    Mr = M.clone()
    want = oracle.greedy_lap(Mr, method)
    assert got.tolist() == want
    # the matrix is masked in place exactly like the reference's
    assert torch.equal(Md.cpu(), Mr)


def test_double_greedy_zero_row_takes_first_index(device):
    # A row that is all zero after masking picks column 0 even though it is
    # taken (argmax of an all-equal row is its first index), as the reference
    # does: rows 1 and 2 both end up on column 0.
    M = torch.tensor([[0.9, 0.1, 0.0], [0.8, 0.0, 0.0], [0.0, 0.0, 0.0]])
    got = greedy_lap_device(M.clone().to(device), HungarianMethod.DOUBLE_GREEDY).cpu().tolist()
    assert got == oracle.greedy_lap(M.clone(), "DOUBLE_GREEDY") == [0, 0, 0]


@pytest.mark.parametrize("method", ["GREEDY", "DOUBLE_GREEDY"])
def test_greedy_fallback_paths_match_oracle(device, method):
    # Identical rows: the row maxima tie (stable order) and after 16 rows every
    # row's candidate list is taken, so the walk scans whole masked rows.  A
    # zero block exercises the rule that a best value of 0 loses to the first
    # (possibly taken) column.
    n = 300
    g = torch.Generator().manual_seed(8)
    row = torch.rand(n, generator=g)
    M = row.repeat(n, 1)
    M[200:, :] = 0.0
    M[250:, 5] = 1e-3
    got = greedy_lap_device(M.clone().to(device), HungarianMethod[method])
    Mr = M.clone()
    want = oracle.greedy_lap(Mr, method)
    assert got.cpu().tolist() == want
    Md = M.clone().to(device)
    greedy_lap_device(Md, HungarianMethod[method])
    assert torch.equal(Md.cpu(), Mr)


def test_greedy_full_size_is_permutation(device):
    n = 32768
    g = torch.Generator(device=device).manual_seed(0)
    M = torch.rand((n, n), generator=g, device=device) * 1e-3
    perm = torch.randperm(n, generator=g, device=device)
    M[torch.arange(n, device=device), perm] += 1.0
    got = greedy_lap_device(M, HungarianMethod.DOUBLE_GREEDY)
    assert torch.equal(got.long(), perm)
