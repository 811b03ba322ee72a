"""The hot path at the full sizes of the benchmark workloads, checked stage by
stage against FP64 (tests/fp64_checks.py) and against the oracle where it
finishes in seconds:

* Watts-Strogatz n = 32768, k = 20, p = 0.1, 1 % noise: both gradient forms
  (CSR gathers and the matrix-core BF16 hi + lo form) on sampled rows and
  against each other over the whole matrix; the fused Sinkhorn solve, the
  marginals and the Frank-Wolfe update and gap.
* Stochastic block model 64 x 1024 (n = 65536, degrees 48 / 16, 1 % noise):
  every n x n offset above 2^31; gradient rows, the two-sweep Sinkhorn, the
  update and gap, and DOUBLE_GREEDY (a permutation that satisfies the
  greedy's defining conditions, and the in-place masking of P).
* Barabasi-Albert n = 8192, m = 5, 2 % noise: one lambda stage of ten
  Frank-Wolfe steps through the library's own stages, each stage against
  FP64 on the device, then DOUBLE_GREEDY against the oracle.

Reference formulas: cugal/pred.py:54-86, :133-141, cugal/sinkhorn.py:90-120,
cugal/greedy_lap.py:50-63.
"""

import networkx as nx
import numpy as np
import pytest
import torch

import fp64_checks as f64
from cugal_amd import _native as nat
from cugal_amd.config import Config, HungarianMethod, SinkhornMethod
from cugal_amd.feature_extraction import Features, node_features
from cugal_amd.greedy_lap import greedy_lap_device
from cugal_amd.pred import _gradient_into, update_quasi_permutation_log
from cugal_amd.profile import SinkhornProfile
from cugal_amd.sinkhorn import LogKernel, SinkhornState, solve_log_native
from cugal_amd.synthetic import NoiseLevel, barabasi_albert, make_edge_instance, stochastic_block
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _ws32768():
    g = nx.watts_strogatz_graph(32768, 20, 0.1, seed=0)
    return make_edge_instance(np.array(g.edges), NoiseLevel(0.0, 0.01, True), 0)


def _sample_rows(n, count, seed):
    rng = np.random.default_rng(seed)
    rows = {0, n - 1, n - 2, n // 2}
    rows.update(int(r) for r in rng.choice(n, size=count, replace=False))
    return sorted(rows)


def _assert_gradient_rows(got, ref, prod, other, rel_prod, rel_other, what):
    err = np.abs(got - ref)
    bound = rel_prod * prod + rel_other * other + 1e-30
    worst = float(np.max(err / bound))
    _say(f"{what}: max |G - G64| = {float(err.max()):.3e}, max error / bound = {worst:.3f}")
    assert worst <= 1.0, what


def test_ws32768_gradients_against_fp64(device):
    inst = _ws32768()
    n = inst.n
    A, B = inst.adjacencies(device)
    feats = Features(node_features(A, 2.0), node_features(B, 2.0))
    ld = int(nat.query("cg_gradient_padded_ld", n))
    P = f64.random_quasi_permutation(n, ld, device, seed=1)
    lam = 7
    rows = _sample_rows(n, 12, 5)
    ref, prod, other = f64.gradient_rows(A, B, P, feats, lam, 1.0, rows)
    G = torch.empty((n, n), device=device)
    idx = torch.as_tensor(rows, device=device)
    # CSR gathers, FP32: sums of at most deg_A * deg_B = ~400 positive terms
    _gradient_into(G, A, B, P, feats, lam, 1.0, precision="fp32")
    _assert_gradient_rows(G[idx].double().cpu().numpy(), ref, prod, other, 2e-5, 2e-6, "CSR gradient n=32768")
    # matrix cores: P and A P split into BF16 hi + lo (2^-18 relative each)
    _gradient_into(G, A, B, P, feats, lam, 1.0, precision="bf16x2")
    _assert_gradient_rows(G[idx].double().cpu().numpy(), ref, prod, other, 4e-5, 2e-6, "bf16x2 gradient n=32768")
    # the two forms over the whole matrix, on -2 A P B alone
    _gradient_into(G, A, B, P, None, 0, 1.0, precision="bf16x2")
    G32 = torch.empty_like(G)
    _gradient_into(G32, A, B, P, None, 0, 1.0, precision="fp32")
    rel = float(((G - G32).abs() / (G32.abs() + 1e-30)).max())
    _say(f"bf16x2 vs CSR over the whole matrix: max relative difference {rel:.3e}")
    assert rel <= 6e-5


def test_ws32768_sinkhorn_and_update_against_fp64(device):
    inst = _ws32768()
    n = inst.n
    A, B = inst.adjacencies(device)
    feats = Features(node_features(A, 2.0), node_features(B, 2.0))
    ld = int(nat.query("cg_gradient_padded_ld", n))
    P = f64.random_quasi_permutation(n, ld, device, seed=2)
    C = torch.empty((n, n), device=device)
    _gradient_into(C, A, B, P, feats, 3, 1.0)
    reg = 0.3
    floor = f64.exponent_floor(float(C.abs().max()), reg)
    rows = _sample_rows(n, 12, 6)
    u = torch.zeros(n, device=device)
    v = torch.zeros(n, device=device)
    # two iterations of the fused solve against two FP64 iterations
    solve_log_native(C, reg, u, v, 2, 0.0, 10, None, fused=True)
    u64, v64 = f64.log_sinkhorn(C, reg, torch.zeros(n, device=device), torch.zeros(n, device=device), 2)
    got = f64.plan_rows(C, reg, u, v, rows)
    want = f64.plan_rows(C, reg, u64, v64, rows)
    rel = f64.max_relative(got, want, 8 * floor, "fused Sinkhorn plan, 2 iterations")
    _say(f"fused Sinkhorn n=32768, 2 iterations: plan max relative error {rel:.3e} (exponent floor {floor:.2e})")
    # 39 more (checks at 10, 20, 30; the last one checks the returned state)
    errors, count, _ = solve_log_native(C, reg, u, v, 41, 0.0, 10, None, fused=True)
    assert count == 41 and len(errors) == 4
    row, col = f64.marginals(C, reg, u, v)
    err64 = float((col - 1).abs().sum())
    _say(f"column-marginal error: device {errors[-1]:.6e}, FP64 of the same duals {err64:.6e}")
    assert abs(errors[-1] - err64) <= 2e-3 * err64 + n * floor
    assert float((row - 1).abs().max()) <= 4 * floor
    # Frank-Wolfe update with alpha = 2 / (2 + 1)
    alpha = 2.0 / 3.0
    P_old = P.clone()
    gap = torch.empty(1, dtype=torch.float64, device=device)
    update_quasi_permutation_log(P, LogKernel(C, reg), u, v, alpha, gap=gap)
    plan = f64.plan_rows(C, reg, u, v, rows)
    idx = torch.as_tensor(rows, device=device)
    old = P_old[idx].double().cpu().numpy()
    want = old + alpha * (plan - old)
    got = P[idx].double().cpu().numpy()
    ufloor = f64.exponent_floor(float(C.abs().max()), reg, u, v)
    assert np.all(np.abs(got - want) <= f64.update_bound(plan, old, want, alpha, ufloor))
    s_plan, a_plan = f64.plan_inner(C, reg, u, v)
    s_old, a_old = f64.inner(C, P_old)
    gap64 = s_plan - s_old
    _say(f"gap: device {float(gap):.9e}, FP64 {gap64:.9e}")
    assert abs(float(gap) - gap64) <= 4 * ufloor * a_plan + 1e-6 * (a_plan + a_old)


def _greedy_conditions(P0: torch.Tensor, masked: torch.Tensor, col: torch.Tensor, rows_per_chunk: int = 2048):
    """DOUBLE_GREEDY (cugal/greedy_lap.py:50-63) visits rows by descending
    row maximum (stable) and gives each the first column of the largest
    entry among the columns no earlier row took, zeroing the taken columns
    of that row in place.  An assignment is that greedy's exactly when it
    satisfies these conditions row by row, checked here in row chunks:
    a permutation; col[r] = first argmax of P0[r] over the columns taken at
    or after r's visit; the matrix left behind equals P0 with those earlier
    columns zeroed."""
    n = P0.shape[0]
    dev = P0.device
    col = col.to(torch.int64)
    assert torch.equal(torch.sort(col).values, torch.arange(n, device=dev)), "not a permutation"
    order = torch.sort(P0.max(1).values, descending=True, stable=True).indices
    visit = torch.empty(n, dtype=torch.int64, device=dev)
    visit[order] = torch.arange(n, device=dev)
    taken_at = torch.empty(n, dtype=torch.int64, device=dev)
    taken_at[col] = visit
    for r0 in range(0, n, rows_per_chunk):
        r1 = min(n, r0 + rows_per_chunk)
        earlier = taken_at[None, :] < visit[r0:r1, None]
        M = torch.where(earlier, torch.zeros((), device=dev), P0[r0:r1])
        assert torch.equal(M.argmax(1), col[r0:r1]), f"greedy choice differs in rows {r0}..{r1}"
        assert torch.equal(masked[r0:r1], M), f"masked rows differ in {r0}..{r1}"


def test_sbm65536_stages_beyond_int32_offsets(device):
    _say("sbm n=65536: generating")
    inst = stochastic_block(64, 1024, 48.0, 16.0, 0.01, seed=0)
    n = inst.n
    assert n * n >= 2 ** 32
    A, B = inst.adjacencies(device)
    feats = Features(node_features(A, 2.0), node_features(B, 2.0))
    ld = int(nat.query("cg_gradient_padded_ld", n))
    P = f64.random_quasi_permutation(n, ld, device, seed=3)
    rows = _sample_rows(n, 10, 7) + [40000, 50000, 65000]
    rows = sorted(set(rows))
    _say("sbm n=65536: gradient")
    C = torch.empty((n, n), device=device)
    _gradient_into(C, A, B, P, feats, 5, 1.0)
    ref, prod, other = f64.gradient_rows(A, B, P, feats, 5, 1.0, rows)
    idx = torch.as_tensor(rows, device=device)
    _assert_gradient_rows(C[idx].double().cpu().numpy(), ref, prod, other, 2e-5, 2e-6, "CSR gradient n=65536")

    _say("sbm n=65536: two-sweep Sinkhorn")
    reg = 0.3
    floor = f64.exponent_floor(float(C.abs().max()), reg)
    u = torch.zeros(n, device=device)
    v = torch.zeros(n, device=device)
    solve_log_native(C, reg, u, v, 2, 0.0, 10, None, fused=True)  # above the fused width: two sweeps
    u64, v64 = f64.log_sinkhorn(C, reg, torch.zeros(n, device=device), torch.zeros(n, device=device), 2)
    got = f64.plan_rows(C, reg, u, v, rows)
    want = f64.plan_rows(C, reg, u64, v64, rows)
    rel = f64.max_relative(got, want, 8 * floor, "two-sweep Sinkhorn plan, 2 iterations")
    _say(f"two-sweep Sinkhorn n=65536, 2 iterations: plan max relative error {rel:.3e} (floor {floor:.2e})")
    errors, count, _ = solve_log_native(C, reg, u, v, 11, 0.0, 10, None, fused=True)
    assert count == 11 and len(errors) == 1
    row, col = f64.marginals(C, reg, u, v)
    err64 = float((col - 1).abs().sum())
    _say(f"column-marginal error: device {errors[-1]:.6e}, FP64 of the same duals {err64:.6e}")
    assert abs(errors[-1] - err64) <= 2e-3 * err64 + n * floor
    assert float((row - 1).abs().max()) <= 4 * floor

    _say("sbm n=65536: update")
    alpha = 0.5
    P_old = P.clone()
    gap = torch.empty(1, dtype=torch.float64, device=device)
    update_quasi_permutation_log(P, LogKernel(C, reg), u, v, alpha, gap=gap)
    plan = f64.plan_rows(C, reg, u, v, rows)
    old = P_old[idx].double().cpu().numpy()
    want = old + alpha * (plan - old)
    got = P[idx].double().cpu().numpy()
    ufloor = f64.exponent_floor(float(C.abs().max()), reg, u, v)
    assert np.all(np.abs(got - want) <= f64.update_bound(plan, old, want, alpha, ufloor))
    s_plan, a_plan = f64.plan_inner(C, reg, u, v)
    s_old, a_old = f64.inner(C, P_old)
    gap64 = s_plan - s_old
    _say(f"gap: device {float(gap):.9e}, FP64 {gap64:.9e}")
    assert abs(float(gap) - gap64) <= 4 * ufloor * a_plan + 1e-6 * (a_plan + a_old)
    C = None

    _say("sbm n=65536: DOUBLE_GREEDY")
    P_old.copy_(P)
    Q = P.contiguous() if P.stride(0) != n else P
    assignment = greedy_lap_device(Q, HungarianMethod.DOUBLE_GREEDY)
    _greedy_conditions(P_old, Q, assignment)


def test_ba8192_lambda_stage_per_stage_fp64(device):
    # One lambda stage (lambda = 0, ten Frank-Wolfe steps) of the bench's
    # configuration; every stage is checked against FP64 on the device
    # from the same inputs, then the LAP against the oracle.
    inst = barabasi_albert(8192, 5, 0.02, seed=0)
    n = max(max(inst.source.nodes()), max(inst.target.nodes())) + 1
    inst.source.add_nodes_from(range(n))
    inst.target.add_nodes_from(range(n))
    from cugal_amd.adjacency import Adjacency

    A = Adjacency.from_graph(inst.source, device)
    B = Adjacency.from_graph(inst.target, device)
    config = Config(device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
                    sinkhorn_regularization=0.3, sinkhorn_threshold=1e-3, mu=2.0, use_sparse_adjacency=True,
                    hungarian_method=HungarianMethod.DOUBLE_GREEDY)
    feats = Features.create(A, B, config)
    Ad = A.as_dense(torch.float64)
    Bd = B.as_dense(torch.float64)
    fs, ft = feats.source.double(), feats.target.double()
    D = torch.sqrt(((fs[:, None, :] - ft[None, :, :]) ** 2).sum(-1))
    fs = ft = None
    ld = int(nat.query("cg_gradient_padded_ld", n))
    P = torch.empty((n, ld), device=device)[:, :n]
    P.fill_(1.0 / n)
    G = torch.empty((n, n), device=device)
    gap = torch.empty(1, dtype=torch.float64, device=device)
    state = SinkhornState(n, config)
    reg = config.sinkhorn_regularization
    worst = {"gradient": 0.0, "update": 0.0, "gap": 0.0}
    for it in range(1, 11):
        lam = 0
        _gradient_into(G, A, B, P, feats, lam, 1.0)
        P64 = P.double()
        prod = 2.0 * ((Ad @ P64) @ Bd)
        err = (G.double() - (-prod + D + lam * (1 - 2 * P64))).abs()
        bound = 2e-5 * prod + 2e-6 * (D.abs() + abs(lam)) + 1e-30
        worst["gradient"] = max(worst["gradient"], float((err / bound).max()))
        assert float((err / bound).max()) <= 1.0, f"gradient, step {it}"
        P64 = prod = err = bound = None

        u0, v0 = state.u.clone(), state.v.clone()
        floor = f64.exponent_floor(float(G.abs().max()), reg)
        K64 = G.double() / -reg
        # (a) the iteration's arithmetic: ten iterations from the warm start
        u1, v1 = u0.clone(), v0.clone()
        solve_log_native(G, reg, u1, v1, 10, 0.0, 10, None, fused=True)
        u64, v64 = f64.log_sinkhorn(G, reg, u0, v0, 10)
        plan64 = torch.exp(K64 + u64[:, None] + v64[None, :])
        plan = torch.exp(K64 + u1.double()[:, None] + v1.double()[None, :])
        big = plan64 > 1e-200
        rel = float(((plan - plan64).abs()[big] / plan64[big]).max())
        worst["plan10"] = max(worst.get("plan10", 0.0), rel / floor)
        assert bool(((plan - plan64).abs() <= 16 * floor * plan64 + 1e-200).all()), f"plan, 10 its, step {it}"
        # (b) the whole solve (threshold stop or the 500 cap): at |C|/reg of
        # thousands the problem contracts slowly and FP32 rounding moves the
        # trajectory along its slow directions, so entries drift well above
        # the one-step floor; the plan's mass per row and the transport cost
        # stay within 1e-3 and 1e-5 of FP64
        sp = SinkhornProfile()
        _, u, v = state.solve(G, config, sp)
        count = sp.iteration_count
        u64, v64 = f64.log_sinkhorn(G, reg, u0, v0, count)
        plan64 = torch.exp(K64 + u64[:, None] + v64[None, :])
        plan = torch.exp(K64 + u.double()[:, None] + v.double()[None, :])
        rel = float(((plan - plan64).abs()[big] / plan64[big]).max())
        tv = float((plan - plan64).abs().sum(1).max())
        Gd = G.double()
        cost, cost64 = float((Gd * plan).sum()), float((Gd * plan64).sum())
        worst["plan_rel"] = max(worst.get("plan_rel", 0.0), rel)
        worst["plan_tv"] = max(worst.get("plan_tv", 0.0), tv)
        worst["cost_rel"] = max(worst.get("cost_rel", 0.0), abs(cost - cost64) / abs(cost64))
        assert tv <= 1e-3, f"plan row mass, step {it}: {tv:.3e}"
        assert abs(cost - cost64) <= 1e-5 * abs(cost64), f"transport cost, step {it}"
        plan64 = u64 = v64 = big = Gd = None
        ufloor = f64.exponent_floor(float(G.abs().max()), reg, u, v)

        alpha = 2.0 / (2.0 + it)
        P_old = P.double()
        update_quasi_permutation_log(P, LogKernel(G, reg), u, v, alpha, gap=gap)
        want = P_old + alpha * (plan - P_old)
        err = (P.double() - want).abs()
        bound = f64.update_bound(plan, P_old, want, alpha, ufloor)
        worst["update"] = max(worst["update"], float((err / bound).max()))
        assert float((err / bound).max()) <= 1.0, f"update, step {it}"
        Gd = G.double()
        gap64 = float((Gd * (plan - P_old)).sum())
        tol = 4 * ufloor * float((Gd * plan).abs().sum()) + 1e-6 * float((Gd * (plan + P_old)).abs().sum())
        worst["gap"] = max(worst["gap"], abs(float(gap) - gap64) / tol)
        assert abs(float(gap) - gap64) <= tol, f"gap, step {it}"
        plan = K64 = P_old = want = err = bound = Gd = None
        _say(f"ba8192 step {it}: {count} Sinkhorn iterations; worst error / bound so far {worst}")
    # the LAP of the stage's P: bit-exact against the oracle, masking included
    Q = P.contiguous()
    Qh = Q.cpu()
    cols = greedy_lap_device(Q, HungarianMethod.DOUBLE_GREEDY).cpu().numpy()
    want = oracle.greedy_lap(Qh, "DOUBLE_GREEDY")
    assert np.array_equal(cols, np.array(want))
    assert torch.equal(Q.cpu(), Qh)
