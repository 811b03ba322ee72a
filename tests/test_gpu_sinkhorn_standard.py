"""HIP STANDARD Sinkhorn (cg_sinkhorn_standard_solve) and its Frank-Wolfe
update against the FP64 oracle of cugal/sinkhorn.py:122-144 and
cugal/pred.py:77-86, through the C ABI."""

import copy

import numpy as np
import pytest
import torch

from cugal_amd import Config, HungarianMethod, SinkhornMethod, cugal
from cugal_amd.pred import update_quasi_permutation_standard
from cugal_amd.profile import Profile
from cugal_amd.sinkhorn import StandardKernel, solve_standard_native, solve_log_native
from cugal_amd.synthetic import alignment_accuracy, answer_columns, erdos_renyi
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu


def _cost(n, seed, scale=5.0):
    g = torch.Generator().manual_seed(seed)
    return torch.rand((n, n), generator=g, dtype=torch.float64) * scale


def _run_gpu(C64, reg, iters, thr, freq, fused, device, momentum_start=None):
    n = C64.shape[0]
    C = C64.to(torch.float32).to(device).contiguous()
    u = torch.full((n,), 1.0 / n, device=device)
    v = torch.full((n,), 1.0 / n, device=device)
    errors, count = solve_standard_native(C, reg, u, v, iters, thr, freq, momentum_start, fused=fused)
    torch.cuda.synchronize()
    return u.double().cpu(), v.double().cpu(), errors, count


def _run_oracle(C64, reg, iters, thr, freq, momentum_start=None):
    st = oracle.SinkhornState(C64.shape[0], "STANDARD")
    return st.solve_standard(C64.to(torch.float32).double(), reg, iters, thr, freq, momentum_start)


# FP32 scalings against FP64: each u_i (v_j) is one over an FP32 sum of n
# positive terms (relative rounding ~ n^(1/2) 2^-24 per step, damped by the
# iteration).  Measured at n <= 4096 after 40 iterations: scalings within
# 2.2e-7 relative, plans within 2.3e-7; held to 5e-6 and 1e-5.
S_RTOL = 5e-6
PLAN_RTOL = 1e-5


def std_error_floor(n):
    """FP32 floor of sum_j |v_j sum_i u_i K_ij - 1|: n columns, each a
    product of FP32 quantities carrying ~8 ulps."""
    return n * 8 * 2.0 ** -24


@pytest.mark.parametrize("n", [2600, 5000, 7000, 12288, 20480, 2601, 5002, 8191, 12289, 20483])
def test_fused_matches_two_sweep_at_every_block_shape(device, n):
    # Fused STANDARD sweeps of 3, 5, 7 (partial) and 3, 5 (exact) float4
    # groups per thread, with caching loads up to n = 8192, against the
    # two-sweep path: same scalings, same error trace.  2601 ... 20483 are
    # widths that are not a multiple of 4 (rows at 4-byte boundaries, the
    # last group cut by n), on both block sizes and load policies.
    g = torch.Generator(device=device).manual_seed(n)
    C = torch.rand((n, n), generator=g, device=device) * 5.0
    out = {}
    for fused in (True, False):
        u = torch.full((n,), 1.0 / n, device=device)
        v = torch.full((n,), 1.0 / n, device=device)
        errors, count = solve_standard_native(C, 1.0, u, v, 31, 0.0, 10, None, fused=fused)
        assert count == 31
        out[fused] = (u.double().cpu(), v.double().cpu(), errors)
    (uf, vf, ef), (ut, vt, et) = out[True], out[False]
    np.testing.assert_allclose(uf.numpy(), ut.numpy(), rtol=S_RTOL)
    np.testing.assert_allclose(vf.numpy(), vt.numpy(), rtol=S_RTOL)
    assert len(ef) == len(et) == 4
    np.testing.assert_allclose(ef, et, rtol=2e-3, atol=std_error_floor(n))


@pytest.mark.parametrize("n", [64, 256, 1000, 333, 1001, 4096])
@pytest.mark.parametrize("fused", [True, False])
def test_fixed_iterations_match_oracle(device, n, fused):
    C = _cost(n, seed=n)
    reg = 1.0
    u, v, errors, count = _run_gpu(C, reg, 40, 0.0, 10, fused, device)
    K, ru, rv, res = _run_oracle(C, reg, 40, 0.0, 10)
    assert count == res.iteration_count == 40
    np.testing.assert_allclose(u.numpy(), ru.numpy(), rtol=S_RTOL)
    np.testing.assert_allclose(v.numpy(), rv.numpy(), rtol=S_RTOL)
    np.testing.assert_allclose(oracle.plan_standard(K, u, v).numpy(), oracle.plan_standard(K, ru, rv).numpy(),
                               rtol=PLAN_RTOL)
    # errors at iterations 0, 10, 20, 30 (the reference checks iteration 0)
    assert len(errors) == len(res.errors) == 4
    np.testing.assert_allclose(errors, res.errors, rtol=2e-3, atol=std_error_floor(n))


@pytest.mark.parametrize("fused", [True, False])
def test_threshold_stop_matches_oracle(device, fused):
    n = 512
    C = _cost(n, seed=3, scale=8.0)
    reg, thr = 0.7, 1e-2
    u, v, errors, count = _run_gpu(C, reg, 500, thr, 10, fused, device)
    _, ru, rv, res = _run_oracle(C, reg, 500, thr, 10)
    assert res.iteration_count < 500
    assert count == res.iteration_count
    assert errors[-1] < thr <= errors[-2]
    np.testing.assert_allclose(u.numpy(), ru.numpy(), rtol=1e-3)
    np.testing.assert_allclose(v.numpy(), rv.numpy(), rtol=1e-3)


@pytest.mark.parametrize("fused", [True, False])
def test_momentum_matches_oracle(device, fused):
    n = 300 if not fused else 320
    C = _cost(n, seed=11, scale=6.0)
    reg = 0.8
    u, v, errors, count = _run_gpu(C, reg, 60, 0.0, 10, fused, device, momentum_start=20)
    _, ru, rv, res = _run_oracle(C, reg, 60, 0.0, 10, momentum_start=20)
    np.testing.assert_allclose(u.numpy(), ru.numpy(), rtol=1e-3)
    np.testing.assert_allclose(v.numpy(), rv.numpy(), rtol=1e-3)
    np.testing.assert_allclose(errors, res.errors, rtol=1e-2, atol=std_error_floor(n))


def test_standard_and_log_reach_the_same_plan(device):
    # Both methods iterate towards the same doubly stochastic scaling.
    n = 256
    C = _cost(n, seed=5).to(torch.float32).to(device)
    reg = 1.0
    us = torch.full((n,), 1.0 / n, device=device)
    vs = torch.full((n,), 1.0 / n, device=device)
    solve_standard_native(C, reg, us, vs, 300, 0.0, 10, None)
    ul = torch.zeros(n, device=device)
    vl = torch.zeros(n, device=device)
    solve_log_native(C, reg, ul, vl, 300, 0.0, 10, None)
    K = torch.exp(-C.double() / reg)
    plan_s = K * us.double()[:, None] * vs.double()[None, :]
    plan_l = torch.exp(-C.double() / reg + ul.double()[:, None] + vl.double()[None, :])
    np.testing.assert_allclose(plan_s.cpu().numpy(), plan_l.cpu().numpy(), rtol=1e-3, atol=1e-9)
    np.testing.assert_allclose(plan_s.sum(0).cpu().numpy(), 1.0, atol=1e-4)


def test_standard_update_matches_oracle(device):
    n = 777
    g = torch.Generator().manual_seed(2)
    C = torch.rand((n, n), generator=g) * 5
    u = torch.rand(n, generator=g) / n + 0.5 / n
    v = torch.rand(n, generator=g) + 0.5
    P = torch.rand((n, n), generator=g) / n
    reg, alpha = 0.9, 0.4
    Pd = P.to(device)
    gap = torch.empty(1, dtype=torch.float64, device=device)
    update_quasi_permutation_standard(Pd, StandardKernel(C.to(device), reg), u.to(device), v.to(device), alpha,
                                      gap=gap)
    K = torch.exp(C.double() / -reg)
    Pr = P.double().clone()
    step = oracle.update_quasi_permutation(Pr, oracle.plan_standard(K, u.double(), v.double()), alpha)
    want_gap = oracle.duality_gap(C.double(), step, alpha)
    np.testing.assert_allclose(Pd.double().cpu().numpy(), Pr.numpy(), rtol=1e-5, atol=1e-9)
    assert abs(abs(float(gap.item())) - want_gap) <= 1e-4 * want_gap + 1e-6


@pytest.mark.parametrize("sparse", [True, False])
def test_pipeline_standard_matches_oracle(device, sparse):
    inst = erdos_renyi(400, 0.02, 0.05, seed=0)
    gpu = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.STANDARD,
                 sinkhorn_regularization=1.0, mu=0.5, iter_count=5, frank_wolfe_iter_count=5,
                 use_sparse_adjacency=sparse, hungarian_method=HungarianMethod.DOUBLE_GREEDY)
    ref = oracle.OracleConfig(sinkhorn_regularization=1.0, sinkhorn_method="STANDARD", mu=0.5, iter_count=5,
                              frank_wolfe_iter_count=5, use_sparse_adjacency=sparse)
    profile = Profile()
    _, mapping = cugal(copy.deepcopy(inst.source), copy.deepcopy(inst.target), gpu, profile)
    run = oracle.align(copy.deepcopy(inst.source), copy.deepcopy(inst.target), ref)
    acc = alignment_accuracy(inst.source_mapping, answer_columns(mapping))
    ref_acc = alignment_accuracy(inst.source_mapping, answer_columns(run.mapping))
    print(f"\nSTANDARD pipeline sparse={sparse}: accuracy {acc:.4f} vs oracle {ref_acc:.4f}")
    assert abs(acc - ref_acc) <= 0.005 + 1e-12, (acc, ref_acc)
    got_iters = [p.iteration_count for p in profile.sinkhorn_profiles]
    assert all(abs(a - b) <= 10 for a, b in zip(got_iters, run.sinkhorn_iterations))


# ---- FP64 (the reference's default dtype) ---------------------------------

def _run_gpu64(C64, reg, iters, thr, freq, device, momentum_start=None):
    from cugal_amd.sinkhorn import solve_standard_native_f64

    n = C64.shape[0]
    C = C64.to(device).contiguous()
    u = torch.full((n,), 1.0 / n, dtype=torch.float64, device=device)
    v = torch.full((n,), 1.0 / n, dtype=torch.float64, device=device)
    errors, count = solve_standard_native_f64(C, reg, u, v, iters, thr, freq, momentum_start)
    torch.cuda.synchronize()
    return u.cpu(), v.cpu(), errors, count


@pytest.mark.parametrize("n", [64, 333, 1000, 4096])
def test_fp64_fixed_iterations_match_oracle(device, n):
    # FP64 against FP64: the same iteration, sums in a different order, so
    # the scalings agree to a few ulps per iteration (held to 1e-11).
    C = _cost(n, seed=100 + n)
    u, v, errors, count = _run_gpu64(C, 1.0, 40, 0.0, 10, device)
    _, uo, vo, res = oracle.SinkhornState(n, "STANDARD").solve_standard(C.clone(), 1.0, 40, 0.0, 10)
    assert count == 40 == res.iteration_count
    np.testing.assert_allclose(u.numpy(), uo.numpy(), rtol=1e-11)
    np.testing.assert_allclose(v.numpy(), vo.numpy(), rtol=1e-11)
    np.testing.assert_allclose(errors, res.errors, rtol=1e-8, atol=1e-13 * n)


def test_fp64_threshold_stop_and_momentum_match_oracle(device):
    n = 2000
    C = _cost(n, seed=7, scale=8.0)
    for momentum_start in (None, 40):
        u, v, errors, count = _run_gpu64(C, 0.5, 500, 1e-9, 10, device, momentum_start)
        st = oracle.SinkhornState(n, "STANDARD")
        _, uo, vo, res = st.solve_standard(C.clone(), 0.5, 500, 1e-9, 10, momentum_start)
        print(f"FP64 STANDARD momentum_start={momentum_start}: {count} iterations (oracle {res.iteration_count}), "
              f"last error {errors[-1]:.3e} (oracle {res.errors[-1]:.3e})")
        assert count == res.iteration_count < 500
        assert len(errors) == len(res.errors)
        np.testing.assert_allclose(errors, res.errors, rtol=1e-6, atol=1e-14 * n)
        np.testing.assert_allclose(u.numpy(), uo.numpy(), rtol=1e-9)
        np.testing.assert_allclose(v.numpy(), vo.numpy(), rtol=1e-9)


def test_fp64_small_regularization_beyond_fp32_range(device):
    # reg 0.01 with costs up to 5: exp(-C/reg) reaches e^-500 ~ 1e-217, far
    # below FP32's range (every FP32 entry of K past C/reg ~ 104 is 0) and
    # inside FP64's; the FP64 solve follows the oracle there.
    n = 300
    C = _cost(n, seed=9, scale=5.0)
    u, v, errors, count = _run_gpu64(C, 0.01, 60, 0.0, 10, device)
    _, uo, vo, res = oracle.SinkhornState(n, "STANDARD").solve_standard(C.clone(), 0.01, 60, 0.0, 10)
    plan = u[:, None] * torch.exp(C / -0.01) * v[None, :]
    plan_o = uo[:, None] * torch.exp(C / -0.01) * vo[None, :]
    assert torch.isfinite(plan).all()
    np.testing.assert_allclose(plan.numpy(), plan_o.numpy(), rtol=1e-8, atol=1e-300)
    np.testing.assert_allclose(errors, res.errors, rtol=1e-6)


def test_fp64_update_matches_oracle(device):
    from cugal_amd import _native as nat

    n, reg, alpha = 500, 0.7, 0.4
    C = _cost(n, seed=11)
    g = torch.Generator().manual_seed(12)
    u = torch.rand(n, generator=g, dtype=torch.float64) / n
    v = torch.rand(n, generator=g, dtype=torch.float64) * 3
    P0 = torch.rand((n, n), generator=g, dtype=torch.float64) / n
    Cd, ud, vd = C.to(device), u.to(device), v.to(device)  # held across the call
    for standard in (1, 0):
        P = P0.to(device)
        gap = torch.empty(1, dtype=torch.float64, device=device)
        ws = nat.workspace(device, nat.query("cg_update_workspace", n), "update")
        nat.call("cg_update_quasi_permutation_f64", standard, nat.ptr(P), n, nat.ptr(Cd), n, nat.ptr(ud),
                 nat.ptr(vd), alpha, reg, nat.ptr(gap), nat.ptr(ws), ws.numel(), nat.stream_of(device))
        K = torch.exp(C / -reg)
        plan = oracle.plan_standard(K, u, v) if standard else torch.exp(C / -reg + u[:, None] + v[None, :])
        Q = P0.clone()
        step = oracle.update_quasi_permutation(Q, plan, alpha)
        np.testing.assert_allclose(P.cpu().numpy(), Q.numpy(), rtol=1e-13, atol=1e-18)
        want_gap = float((C * (step / alpha)).sum())
        assert float(gap) == pytest.approx(want_gap, rel=1e-10)
