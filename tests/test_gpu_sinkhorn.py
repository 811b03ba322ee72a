"""HIP log-Sinkhorn (cg_sinkhorn_log_solve) against the FP64 oracle of
cugal/sinkhorn.py:90-120, through the C ABI."""

import numpy as np
import pytest
import torch

from cugal_amd import _native as nat
from cugal_amd.sinkhorn import solve_log_native
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu


def _cost(n, seed, scale=10.0):
    g = torch.Generator().manual_seed(seed)
    return torch.rand((n, n), generator=g, dtype=torch.float64) * scale


def _run_gpu(C64, reg, iters, thr, freq, fused, device, u0=None, v0=None, momentum_start=None):
    n = C64.shape[0]
    C = C64.to(torch.float32).to(device).contiguous()
    u = (torch.zeros(n) if u0 is None else u0).to(torch.float32).to(device).contiguous()
    v = (torch.zeros(n) if v0 is None else v0).to(torch.float32).to(device).contiguous()
    errors, count, _ = solve_log_native(C, reg, u, v, iters, thr, freq, momentum_start, fused=fused)
    torch.cuda.synchronize()
    return u.double().cpu(), v.double().cpu(), errors, count


def _run_oracle(C64, reg, iters, thr, freq, u0=None, v0=None, momentum_start=None):
    n = C64.shape[0]
    st = oracle.SinkhornState(n, "LOG")
    if u0 is not None:
        st.u, st.v = u0.double().clone(), v0.double().clone()
    K, u, v, res = st.solve_log(C64.to(torch.float32).double(), reg, iters, thr, freq, momentum_start)
    return K, u, v, res


# FP32 log-domain iterates against FP64: |C/reg| <= 33 here, so potentials
# carry about 33 * 2^-24 ~ 2e-6 absolute rounding per step; 5e-5 absolute on
# u, v bounds the accumulated drift (measured 6e-7, or 7e-6 on the two-sweep
# path where it is a gauge shift that cancels in the plan).
U_ATOL = 5e-5


def error_floor(n, C, reg):
    """FP32 floor of the column-marginal L1 error: each column sum carries
    the rounding of exponents of magnitude up to |C|/reg * log2(e) (unit
    roundoff 2^-24 each), so sum_j |c_j - 1| cannot be resolved below about
    n * 2^-24 * (4 + max|K2|).  The FP64 oracle reaches ~1e-13; the FP32
    kernels stop improving at this floor (the reference's FP32 kernels do
    too)."""
    kmax = float(C.abs().max()) / reg * 1.4426950408889634
    return n * 2.0 ** -24 * (4.0 + kmax)


def assert_errors_close(got, want, n, C, reg, rtol=2e-3):
    assert len(got) == len(want)
    np.testing.assert_allclose(got, want, rtol=rtol, atol=error_floor(n, C, reg))


@pytest.mark.parametrize("n", [64, 256, 1000, 333, 379, 1001, 4096])
@pytest.mark.parametrize("fused", [True, False])
def test_fixed_iterations_match_oracle(device, n, fused):
    C = _cost(n, seed=n)
    reg = 0.3 if n != 333 else 1.0
    u, v, errors, count = _run_gpu(C, reg, 40, 0.0, 10, fused, device)
    K, ru, rv, res = _run_oracle(C, reg, 40, 0.0, 10)
    assert count == res.iteration_count == 40
    np.testing.assert_allclose(u.numpy(), ru.numpy(), atol=U_ATOL * max(1.0, 1.0 / reg))
    np.testing.assert_allclose(v.numpy(), rv.numpy(), atol=U_ATOL * max(1.0, 1.0 / reg))
    plan = oracle.plan_log(K, u, v)
    ref_plan = oracle.plan_log(K, ru, rv)
    # The plan within 1e-5 relative of FP64 (measured 5e-7 to 8e-7; the
    # two-sweep path's potentials carry a 7e-6 gauge shift u + c, v - c that
    # cancels in the plan).  Entries below 1e-6 of the largest are held to
    # 1e-5 of that floor in absolute terms.
    floor = 1e-6 * float(ref_plan.max())
    np.testing.assert_allclose(plan.numpy(), ref_plan.numpy(), rtol=1e-5, atol=1e-5 * floor)
    # errors recorded at iterations 10, 20, 30 (not 0)
    assert len(errors) == len(res.errors) == 3
    assert_errors_close(errors, res.errors, n, C, reg)


def test_fused_equals_two_sweep(device):
    n = 2048
    C = _cost(n, seed=7, scale=3.0)
    a = _run_gpu(C, 0.2, 60, 0.0, 10, True, device)
    b = _run_gpu(C, 0.2, 60, 0.0, 10, False, device)
    np.testing.assert_allclose(a[0].numpy(), b[0].numpy(), atol=2e-5)
    np.testing.assert_allclose(a[1].numpy(), b[1].numpy(), atol=2e-5)


def test_fused_rows_out_of_shifted_range(device):
    # Warm start from one cost, then rows 5 and 900 shift by +-200 (C/reg =
    # -+667): their mass under the previous duals under/overflows FP32, so
    # the fused sweep redoes them with the exact max shift.  Same iterates
    # as the two-sweep path and as the FP64 oracle.
    n, reg = 1024, 0.3
    C1 = _cost(n, seed=21, scale=3.0)
    u0, v0, _, _ = _run_gpu(C1, reg, 40, 0.0, 10, True, device)
    C2 = C1.clone()
    C2[5] += 200.0
    C2[900] -= 200.0
    a = _run_gpu(C2, reg, 30, 0.0, 10, True, device, u0=u0, v0=v0)
    b = _run_gpu(C2, reg, 30, 0.0, 10, False, device, u0=u0, v0=v0)
    _, ru, rv, _ = _run_oracle(C2, reg, 30, 0.0, 10, u0=u0, v0=v0)
    assert np.isfinite(a[0].numpy()).all() and np.isfinite(a[1].numpy()).all()
    np.testing.assert_allclose(a[0].numpy(), b[0].numpy(), atol=1e-3)
    np.testing.assert_allclose(a[1].numpy(), b[1].numpy(), atol=1e-3)
    np.testing.assert_allclose(a[0].numpy(), ru.numpy(), atol=2e-3)
    np.testing.assert_allclose(a[1].numpy(), rv.numpy(), atol=2e-3)


def test_threshold_stop_and_warm_start(device):
    n = 512
    C1 = _cost(n, seed=11, scale=2.0)
    C2 = C1 + 0.05 * _cost(n, seed=12, scale=1.0)
    u, v, e1, c1 = _run_gpu(C1, 0.5, 500, 1e-3, 10, True, device)
    K, ru, rv, r1 = _run_oracle(C1, 0.5, 500, 1e-3, 10)
    assert error_floor(n, C1, 0.5) < 1e-3  # the threshold is resolvable in FP32 here
    assert c1 == r1.iteration_count
    assert e1[-1] < 1e-3 and len(e1) == len(r1.errors)
    # second solve warm-starts from the first one's duals
    u2, v2, e2, c2 = _run_gpu(C2, 0.5, 500, 1e-3, 10, True, device, u0=u, v0=v)
    st = oracle.SinkhornState(n, "LOG")
    st.u, st.v = ru.clone(), rv.clone()
    _, ru2, rv2, r2 = st.solve_log(C2.to(torch.float32).double(), 0.5, 500, 1e-3, 10)
    assert c2 == r2.iteration_count
    np.testing.assert_allclose(u2.numpy(), ru2.numpy(), atol=1e-4)


@pytest.mark.parametrize("fused", [True, False])
def test_momentum_matches_oracle(device, fused):
    # Errors at iterations 10 and 20 (0.28 and 4.7e-3 in FP64) sit above the
    # FP32 floor, so the Lehmann factor computed from their ratio is the
    # oracle's; later errors reach the floor.
    n, reg = 300, 0.05
    C = _cost(n, seed=5, scale=10.0)
    u, v, e, c = _run_gpu(C, reg, 80, 0.0, 10, fused, device, momentum_start=20)
    K, ru, rv, res = _run_oracle(C, reg, 80, 0.0, 10, momentum_start=20)
    assert e[0] > 0.1 and e[1] > 1e-3
    assert_errors_close(e, res.errors, n, C, reg, rtol=2e-2)
    plan = oracle.plan_log(K, u, v)
    ref_plan = oracle.plan_log(K, ru, rv)
    np.testing.assert_allclose(plan.numpy(), ref_plan.numpy(), rtol=2e-3, atol=1e-7)


def test_separable_cost_gives_uniform_plan(device):
    # C_ij = f_i + g_j: the scaled kernel is rank one, so one iteration
    # reaches the plan with every entry 1/n (row and column sums 1).
    n = 1024
    g = torch.Generator().manual_seed(3)
    f = torch.rand(n, generator=g, dtype=torch.float64) * 5
    h = torch.rand(n, generator=g, dtype=torch.float64) * 5
    C = f[:, None] + h[None, :]
    u, v, errors, _ = _run_gpu(C, 0.7, 11, 0.0, 10, True, device)
    plan = torch.exp(-C / 0.7 + u[:, None] + v[None, :])
    np.testing.assert_allclose(plan.numpy(), np.full((n, n), 1.0 / n), rtol=2e-5)
    assert errors[0] < 1e-3


@pytest.mark.parametrize("n", [8192, 32768])
def test_large_marginals(device, n):
    # Size-independent properties at full size: after a solve the rows of
    # the plan sum to 1 (the last half-step normalises rows) and the reported
    # column error equals the column sums' L1 deviation.
    g = torch.Generator(device=device).manual_seed(1)
    C = torch.rand((n, n), generator=g, device=device) * 4.0
    u = torch.zeros(n, device=device)
    v = torch.zeros(n, device=device)
    errors, count, _ = solve_log_native(C, 0.3, u, v, 21, 0.0, 10, None, fused=True)
    plan = torch.exp(-C.double() / 0.3 + u.double()[:, None] + v.double()[None, :])
    rows = plan.sum(1)
    cols = plan.sum(0)
    assert count == 21
    torch.testing.assert_close(rows, torch.ones_like(rows), rtol=0, atol=5e-4)
    col_err = float((cols - 1).abs().sum())
    assert abs(errors[-1] - col_err) <= 1e-2 * col_err + 1e-3


@pytest.mark.parametrize("n", [2600, 5000, 7000, 12288, 16384, 20480, 28000,
                               2601, 5002, 7393, 8191, 12287, 12289, 20483, 32767])
def test_fused_matches_two_sweep_at_every_block_shape(device, n):
    # The fused sweep's block holds 1 to 8 float4 groups of a row per thread
    # (256 threads up to n = 8192, 1024 above; 2600, 5000, 7000 take 3, 5
    # and 7 groups with partial last groups, 12288 and 20480 exactly 3 and 5,
    # 28000 seven), reads matrices up to n = 12288 with caching loads and
    # alternating row order and wider ones with non-temporal loads in a
    # fixed order.  Widths that are not a multiple of 4 (2601 ... 32767: every
    # residue, both block sizes, both load policies, one group short of 8 x
    # 1024 x 4) start their rows at 4-byte boundaries and end them inside a
    # group.  Each against the two-sweep path (pinned to the oracle at
    # small n), from a warm start so the shifted row step runs: same
    # potentials within the drift bound, same error trace.
    g = torch.Generator(device=device).manual_seed(n)
    C = torch.rand((n, n), generator=g, device=device) * 4.0
    u0 = torch.rand(n, generator=g, device=device) - 0.5
    v0 = torch.rand(n, generator=g, device=device) - 0.5
    out = {}
    for fused in (True, False):
        u, v = u0.clone(), v0.clone()
        errors, count, _ = solve_log_native(C, 0.3, u, v, 31, 0.0, 10, None, fused=fused)
        assert count == 31
        out[fused] = (u.double().cpu(), v.double().cpu(), errors)
    (uf, vf, ef), (ut, vt, et) = out[True], out[False]
    np.testing.assert_allclose(uf.numpy(), ut.numpy(), atol=U_ATOL / 0.3)
    np.testing.assert_allclose(vf.numpy(), vt.numpy(), atol=U_ATOL / 0.3)
    assert len(ef) == len(et) == 3
    np.testing.assert_allclose(ef, et, rtol=2e-3, atol=n * 2.0 ** -24 * (4.0 + 4.0 / 0.3 * 1.4427))


def test_reference_step_and_marginal(device):
    # cg_sinkhorn_log_step / cg_sinkhorn_log_marginal keep the reference
    # kernels' contracts (cuda_kernels/sinkhorn_log.cu:8-114).
    n = 300
    K = -_cost(n, seed=9, scale=3.0)
    add = torch.randn(n, dtype=torch.float64)
    out = torch.randn(n, dtype=torch.float64)
    Kd, addd = K.float().to(device), add.float().to(device)
    outd = out.float().to(device)
    nat.call("cg_sinkhorn_log_step", Kd.data_ptr(), n, addd.data_ptr(), outd.data_ptr(), 0.7,
             nat.stream_of(device))
    ref = 0.3 * out - 0.7 * torch.logsumexp(K.float().double() + add.float().double()[None, :], 1)
    np.testing.assert_allclose(outd.double().cpu().numpy(), ref.numpy(), atol=2e-5)
    from cugal_amd.sinkhorn import marginal_error_log
    u = torch.randn(n, dtype=torch.float64) * 0.1
    v = torch.randn(n, dtype=torch.float64) * 0.1
    got = marginal_error_log(Kd.clone(), u.float().to(device), v.float().to(device))
    want = float((torch.exp(K.float().double() + v.float().double()[None, :] + u.float().double()[:, None]).sum(1) - 1).abs().sum())
    assert abs(got - want) <= 1e-4 * want


@pytest.mark.parametrize("fused", [True, False])
def test_sweep_timer_samples_and_changes_nothing(device, fused):
    # The sweep timer records events around the sweeps of iterations 5, 15,
    # 25, ... only; it must not change the iterates, and it counts only the
    # sampled sweeps that ran before the stop.
    from cugal_amd.sinkhorn import SweepTiming, solve_log_native

    n = 512
    C = (torch.rand((n, n), generator=torch.Generator().manual_seed(9)) * 4).to(device)
    runs = []
    for timing in (False, True):
        u = torch.zeros(n, device=device)
        v = torch.zeros(n, device=device)
        SweepTiming.reset(timing)
        errors, iters, _ = solve_log_native(C, 0.3, u, v, 500, 1e-3, 10, None, fused=fused)
        runs.append((u.cpu(), v.cpu(), iters, errors, SweepTiming.total_ms, SweepTiming.launches))
        SweepTiming.reset(False)
    (u0, v0, it0, e0, _, _), (u1, v1, it1, e1, ms, launches) = runs
    assert torch.equal(u0, u1) and torch.equal(v0, v1) and it0 == it1 and e0 == e1
    assert 10 < it0 < 500, it0
    sampled = (it0 - 5 + 9) // 10
    assert launches == sampled * (1 if fused else 2), (launches, it0)
    assert ms > 0.0


def test_concurrent_solves_from_two_threads(device):
    # Two host threads, each with its own stream and workspace, run
    # threshold-stopped solves through the C ABI at the same time; the stop
    # flags and control mirrors are per thread, so each gets exactly what it
    # gets alone (iterates bitwise, same iteration count).
    import threading

    import numpy as np

    def solve(C, reg, stream):
        n = C.shape[0]
        with torch.cuda.stream(stream):
            u = torch.zeros(n, device=device)
            v = torch.zeros(n, device=device)
            ws = torch.empty(int(nat.query("cg_sinkhorn_workspace", n, 1, 500, 10)), dtype=torch.uint8,
                             device=device)
            cap = 500 // 10 + 4
            errors = np.zeros(cap)
            count = np.zeros(1, dtype=np.int32)
            iters = np.zeros(1, dtype=np.int32)
            for _ in range(3):  # several solves, warm-started, to overlap many checks
                nat.call("cg_sinkhorn_log_solve", nat.ptr(C), n, float(reg), nat.ptr(u), nat.ptr(v), 500, 1e-3, 10,
                         -1, 1, errors.ctypes.data, cap, count.ctypes.data, iters.ctypes.data, None, None,
                         nat.ptr(ws), ws.numel(), nat.stream_of(C.device))
            stream.synchronize()
            return u.cpu(), v.cpu(), int(iters[0])

    costs = [(_cost(2048, seed=31, scale=2.0).float().to(device), 0.5),
             (_cost(2048, seed=32, scale=6.0).float().to(device), 0.3)]
    alone = [solve(C, reg, torch.cuda.Stream(device)) for C, reg in costs]
    results = [None, None]

    def worker(k):
        results[k] = solve(costs[k][0], costs[k][1], torch.cuda.Stream(device))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    for k in range(2):
        assert results[k] is not None
        assert torch.equal(results[k][0], alone[k][0]) and torch.equal(results[k][1], alone[k][1])
        assert results[k][2] == alone[k][2] and 10 <= alone[k][2] < 500


def test_concurrent_solves_through_the_python_api(device):
    # The same through SinkhornState.solve: the library's own workspace is
    # kept per (device, thread, stream), so two threads on their own streams
    # solve at once and each gets what it gets alone.
    import threading

    from cugal_amd.config import Config, SinkhornMethod
    from cugal_amd.sinkhorn import SinkhornState

    cfg = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
                 sinkhorn_regularization=0.4, sinkhorn_threshold=1e-3)

    def solve(C, stream):
        with torch.cuda.stream(stream):
            state = SinkhornState(C.shape[0], cfg)
            for _ in range(3):
                state.solve(C, cfg)
            stream.synchronize()
            return state.u.cpu(), state.v.cpu()

    costs = [_cost(2048, seed=33, scale=2.0).float().to(device), _cost(3000, seed=34, scale=5.0).float().to(device)]
    alone = [solve(C, torch.cuda.Stream(device)) for C in costs]
    results = [None, None]

    def worker(k):
        results[k] = solve(costs[k], torch.cuda.Stream(device))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    for k in range(2):
        assert results[k] is not None
        assert torch.equal(results[k][0], alone[k][0]) and torch.equal(results[k][1], alone[k][1])
