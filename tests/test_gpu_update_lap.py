"""FW update (cugal/pred.py:77-86, :133-141) and greedy LAP
(cugal/greedy_lap.py:10-85) against the oracle."""

import numpy as np
import pytest
import torch

from cugal_amd._native import NativeError
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


@pytest.mark.parametrize("n,pad", [(1000, 64), (1024, 96), (777, 3)])
def test_update_of_row_padded_P(device, n, pad):
    # P as the n x n view of an n x (n + pad) buffer: the same P and gap as a
    # contiguous P, and the padding columns are left alone.
    g = torch.Generator().manual_seed(n)
    C = torch.rand((n, n), generator=g).to(device) * 5
    u = (torch.randn(n, generator=g) * 0.5).to(device)
    v = (torch.randn(n, generator=g) * 0.5 - 3.0).to(device)
    P = (torch.rand((n, n), generator=g) / n).to(device)
    buf = torch.full((n, n + pad), -7.0, device=device)
    Pv = buf[:, :n]
    Pv.copy_(P)
    gaps = torch.empty(2, dtype=torch.float64, device=device)
    update_quasi_permutation_log(P, LogKernel(C, 0.7), u, v, 0.4, gap=gaps[0:1])
    update_quasi_permutation_log(Pv, LogKernel(C, 0.7), u, v, 0.4, gap=gaps[1:2])
    assert torch.equal(P, Pv)
    assert float(gaps[0]) == float(gaps[1])
    assert bool((buf[:, n:] == -7.0).all())
    with pytest.raises(ValueError):
        update_quasi_permutation_log(P.T, LogKernel(C, 0.7), u, v, 0.4, gap=gaps[0:1])


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


ALL_GREEDY = ["GREEDY", "DOUBLE_GREEDY", "PARALLEL_GREEDY", "RAND", "MORE_RAND"]


def _run_both(device, M, method, seed):
    """The device method and the oracle on the same matrix and draws;
    returns (device assignment, oracle assignment, device P, oracle P)."""
    n = M.shape[0]
    rng = np.random.default_rng(seed)
    order = rng.permutation(n)
    uniforms = rng.random(n)
    Md = M.clone().to(device)
    got = greedy_lap_device(Md, HungarianMethod[method], order=torch.from_numpy(order),
                            uniforms=torch.from_numpy(uniforms)).cpu().tolist()
    Mr = M.clone()
    want = oracle.greedy_lap(Mr, method, order=order.tolist(), uniforms=uniforms.tolist())
    return got, want, Md.cpu(), Mr


@pytest.mark.parametrize("method", ALL_GREEDY)
@pytest.mark.parametrize("n", [50, 1000, 1537])
def test_greedy_matches_oracle(device, method, n):
    M = _distinct_rowmax_matrix(n, seed=n)
    got, want, Md, Mr = _run_both(device, M, method, seed=n)
    assert got == want
    # the matrix is masked in place exactly like the reference's
    assert torch.equal(Md, Mr)


def test_parallel_greedy_shared_column(device):
    # Two rows above 0.5 on the same column both take it (the reference
    # assigns the parallel set without checking), and neither is masked.
    M = torch.tensor([[0.6, 0.4, 0.0], [0.7, 0.3, 0.0], [0.2, 0.5, 0.3]])
    got, want, Md, Mr = _run_both(device, M, "PARALLEL_GREEDY", seed=0)
    assert got == want == [0, 0, 1]
    assert torch.equal(Md, Mr) and Md[0, 0] == 0.6 and Md[2, 0] == 0.0


def test_random_methods_reject_bad_inputs(device):
    M = torch.rand((64, 64), device=device)
    with pytest.raises(NativeError):
        greedy_lap_device(M, HungarianMethod.RAND, order=torch.zeros(64, dtype=torch.int64))
    with pytest.raises(NativeError):
        greedy_lap_device(M, HungarianMethod.RAND, order=torch.arange(1, 65))
    with pytest.raises(NativeError):
        greedy_lap_device(M, HungarianMethod.MORE_RAND, order=torch.arange(64),
                          uniforms=torch.full((64,), 1.5, dtype=torch.float64))


def test_more_rand_unreachable_draw_raises(device):
    # Prefix sums [-1, -0.5] never reach 0.1 * (-0.5): the reference's
    # mask.nonzero()[0] raises, and so do both restatements.
    M = torch.tensor([[-1.0, 0.5], [0.2, 0.8]])
    with pytest.raises(IndexError):
        greedy_lap_device(M.to(device), HungarianMethod.MORE_RAND, order=torch.tensor([0, 1]),
                          uniforms=torch.tensor([0.1, 0.5], dtype=torch.float64))
    with pytest.raises(IndexError):
        oracle.greedy_lap(M.clone(), "MORE_RAND", order=[0, 1], uniforms=[0.1, 0.5])


def test_random_methods_draw_like_the_reference(device):
    # With no order given, RAND visits rows in torch.randperm(n, device)
    # order: same seed, same order, same result as passing that order.
    M = _distinct_rowmax_matrix(500, seed=4).to(device)
    torch.manual_seed(11)
    got = greedy_lap_device(M.clone(), HungarianMethod.RAND)
    torch.manual_seed(11)
    order = torch.randperm(500, device=device)
    assert torch.equal(got, greedy_lap_device(M.clone(), HungarianMethod.RAND, order=order))


def test_double_greedy_zero_row_takes_first_index(device):
    # A row that is all zero after masking picks column 0 even though it is
    # taken (argmax of an all-equal row is its first index), as the reference
    # does: rows 1 and 2 both end up on column 0.
    M = torch.tensor([[0.9, 0.1, 0.0], [0.8, 0.0, 0.0], [0.0, 0.0, 0.0]])
    got = greedy_lap_device(M.clone().to(device), HungarianMethod.DOUBLE_GREEDY).cpu().tolist()
    assert got == oracle.greedy_lap(M.clone(), "DOUBLE_GREEDY") == [0, 0, 0]


@pytest.mark.parametrize("method", ALL_GREEDY)
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
    got, want, Md, Mr = _run_both(device, M, method, seed=8)
    assert got == want
    assert torch.equal(Md, Mr)


def test_greedy_full_size_is_permutation(device):
    n = 32768
    g = torch.Generator(device=device).manual_seed(0)
    M = torch.rand((n, n), generator=g, device=device) * 1e-3
    perm = torch.randperm(n, generator=g, device=device)
    M[torch.arange(n, device=device), perm] += 1.0
    got = greedy_lap_device(M, HungarianMethod.DOUBLE_GREEDY)
    assert torch.equal(got.long(), perm)


@pytest.mark.parametrize("method", ["PARALLEL_GREEDY", "RAND", "MORE_RAND"])
def test_greedy_variants_at_8192_match_oracle(device, method):
    M = _distinct_rowmax_matrix(8192, seed=5)
    got, want, Md, Mr = _run_both(device, M, method, seed=5)
    assert got == want
    assert torch.equal(Md, Mr)
