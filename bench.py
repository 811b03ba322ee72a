"""Benchmark of the FUGAL alignment path on one MI355X.

Workload (default): Watts-Strogatz n=32768, k=20, rewire p=0.1, 1% edge noise
(remove + refill) and a seeded permutation; cuGAL's published GPU settings
(LOG Sinkhorn FP32, regularization 0.3, threshold 1e-3, <= 500 iterations,
mu 2, 15 lambda stages x 10 Frank-Wolfe steps, sparse adjacency) with the
DOUBLE_GREEDY LAP.  One step = one whole alignment (features, gradient,
Sinkhorn, update, LAP) from CSR adjacencies already resident in HBM to P
and the assignment in HBM (`cugal_amd.pred.align_device`).  The API-level
extras of `cugal()` -- building the CSR from networkx graphs and copying P
to the host -- are timed separately and reported as `api_extra_s`.

Reported: `value` = seconds per alignment on the "mixed" path -- the fused
Sinkhorn (the cost read once per iteration; FP32 potentials, FP64 column
sums, marginal errors and Frank-Wolfe gap) with the gradient in the form
the adjacency density calls for (`gradient_precision="auto"`: the CSR
gathers here, the matrix cores for dense graphs) on graphs relabelled for
gather locality where that pays (`node_order="auto"`) -- beside the plain
FP32 path (two-sweep Sinkhorn, CSR gradient) on the same input, node
accuracies, the Sinkhorn sweep roofline (bytes = 4 n^2 per sweep over the
HIP-event time of the sweep launches of every 10th iteration, recorded on
the launch stream inside the timed run) and the FP64 CPU oracle timed on a
bounded sample of the same Sinkhorn work.  `precision_form` gives the
matrix-core gradient's time and rate, the alignment time it implies (and,
when the budget allows, the measured whole alignment), and every stage's
error against FP64 at this size from the alignment's own P: gradient in
both forms, Sinkhorn plan and marginal error, update and duality gap.
Further legs: `differentiation_form` (Erdos-Renyi n=4096, 100 unrolled
log-Sinkhorn iterations with checkpoints every 10; forward and hand-written
backward times and the gradients' error against torch autograd in FP64),
`reference_case` (the reference's own CPU-runnable case, Erdos-Renyi
n=1000 with its default Config, cugal() on the device and the FP64 oracle
on the host cores), `ba8192`, `sbm65536` and `published` (below).  The run keeps to a
wall-clock budget (`--budget-s`, counted from the interpreter's start):
after the timed steps each leg runs only if its estimate fits and is
otherwise recorded as skipped, so the JSON line always prints.

Other workloads (`--workload`, also run as legs of the default run):
`ba8192` (Barabasi-Albert n=8192, m=5, 2% edge noise; the LAP as greedy and
as the exact device Hungarian, checked against scipy on the same P; alone,
also the FP64 oracle timed per stage and extrapolated) and `sbm65536`
(stochastic block model, 64 blocks of 1024, intra/inter degrees 48/16, 1%
noise, 15 lambda stages of one Frank-Wolfe step, greedy LAP; P is 17 GB);
`er4096grad` (the differentiation form alone); `published` (the
reference's published benchmark experiment on the seven graphs it ships,
with its seed, noise and Config, beside its published times and
accuracies; see published_experiment_leg).
"""

from __future__ import annotations

import argparse
import copy
import json
import os
import sys
import time

T_START = time.time()
ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_config(fused: bool, args, precision: str = "auto", lap: str = "DOUBLE_GREEDY"):
    import torch

    from cugal_amd import Config, HungarianMethod, SinkhornMethod

    return Config(
        device="cuda:0", dtype=torch.float32, sinkhorn_method=SinkhornMethod.LOG,
        sinkhorn_regularization=args.reg, sinkhorn_iterations=args.sinkhorn_iterations,
        sinkhorn_threshold=args.threshold, sinkhorn_eval_freq=10, mu=args.mu, iter_count=args.iter_count,
        frank_wolfe_iter_count=args.fw_iter_count, use_sparse_adjacency=True,
        hungarian_method=HungarianMethod[lap], sinkhorn_fused=fused, gradient_precision=precision,
        node_order=getattr(args, "node_order", "auto"),
    )


def device_inputs(inst, device):
    """Padded graphs (as cugal() pads them) and their CSR in HBM."""
    from cugal_amd.adjacency import Adjacency

    source, target = copy.deepcopy(inst.source), copy.deepcopy(inst.target)
    node_count = max(max(source.nodes()), max(target.nodes()))
    source.add_nodes_from(set(range(node_count)) - set(source.nodes()))
    target.add_nodes_from(set(range(node_count)) - set(target.nodes()))
    return Adjacency.from_graph(source, device), Adjacency.from_graph(target, device)


def run_alignment(inst, config, timing: bool, inputs=None):
    import torch

    from cugal_amd.pred import align_device
    from cugal_amd.profile import Profile
    from cugal_amd.sinkhorn import SweepTiming
    from cugal_amd.synthetic import alignment_accuracy

    A, B = inputs if inputs is not None else device_inputs(inst, config.device)
    SweepTiming.reset(timing)
    profile = Profile()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P, col_ind = align_device(A, B, config, profile)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    acc = alignment_accuracy(inst.source_mapping, col_ind)
    sweeps = (SweepTiming.total_ms, SweepTiming.launches)
    SweepTiming.reset(False)
    return elapsed, acc, profile, sweeps, P


def api_extra_seconds(inst, P, device) -> float:
    """What cugal() adds around align_device: CSR from the networkx graphs
    and the host copy of P."""
    import torch

    from cugal_amd.pred import to_host

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    device_inputs(inst, device)
    torch.cuda.synchronize()
    to_host(P)
    return time.perf_counter() - t0


def measured_traffic(n: int):
    """HBM bytes per launch of the fused sweep from the committed counter
    passes (FETCH_SIZE + WRITE_SIZE, calibrated by a device copy), when they
    were taken at this n; else None."""
    path = os.path.join(ROOT, "profiles", "traffic_ws32768.json")
    try:
        data = json.load(open(path))
    except (OSError, ValueError):
        return None
    kernels = data.get("kernels", {})
    entry = kernels.get("sk_fused_sweep<1024, 8, true, false>")
    if data.get("n") != n or entry is None:
        return None
    return entry["read_bytes"] + entry["write_bytes"]


def gradient_error_sample(A, B, P, rows: int = 16, seed: int = 0):
    """The precision-form gradient -2 A P B (no distance or regulariser
    term) on the device against FP64 on `rows` sampled rows, and against the
    FP32 CSR gradient on the whole matrix.  FP64 row i is
    -2 (sum_{k in N_A(i)} P[k, :]) B, from the rows of P it needs."""
    import numpy as np
    import scipy.sparse
    import torch

    from cugal_amd.pred import _gradient_into

    n = P.shape[0]
    g16 = torch.empty_like(P)
    g32 = torch.empty_like(P)
    _gradient_into(g16, A, B, P, None, 0, 1.0, precision="bf16x2")
    _gradient_into(g32, A, B, P, None, 0, 1.0, precision="fp32")
    scale = float(g32.abs().max())
    vs_fp32 = float((g16 - g32).abs().max()) / scale
    g32 = None
    brp = B.row_pointers.cpu().numpy()
    bcol = B.col_indices.cpu().numpy()
    Bs = scipy.sparse.csr_matrix((np.ones(len(bcol)), bcol, brp), shape=(n, n))
    arp = A.row_pointers.cpu().numpy()
    acol = A.col_indices.cpu().numpy()
    rng = np.random.default_rng(seed)
    worst = 0.0
    for i in rng.choice(n, size=min(rows, n), replace=False):
        nbrs = acol[arp[i]:arp[i + 1]]
        if len(nbrs) == 0:
            continue
        y = P[torch.as_tensor(nbrs, device=P.device, dtype=torch.long)].double().sum(0).cpu().numpy()
        ref = -2.0 * (Bs.T @ y)  # B symmetric: y B = B^T y
        got = g16[int(i)].double().cpu().numpy()
        worst = max(worst, float(np.max(np.abs(got - ref))) / max(float(np.max(np.abs(ref))), 1e-300))
    return {"max_err_vs_fp64_sampled_rows": worst, "rows": int(min(rows, n)),
            "max_err_vs_fp32_csr": vs_fp32, "norm": "max |error| / max |reference| of the row or matrix"}


def _torch_unrolled_plan(C, reg: float, iterations: int):
    """Plain-PyTorch FP64 unrolled log-Sinkhorn (v = -LSE_i(K + u), then
    u = -LSE_j(K + v), K = -C/reg, from zero duals), differentiated by torch
    autograd: the reference the hand-written adjoint is checked against."""
    import torch

    K = -C / reg
    u = torch.zeros(C.shape[0], dtype=C.dtype, device=C.device)
    v = torch.zeros_like(u)
    for _ in range(iterations):
        v = -torch.logsumexp(K + u[:, None], dim=0)
        u = -torch.logsumexp(K + v[None, :], dim=1)
    return torch.exp(K + u[:, None] + v[None, :])


def differentiation_form(seed: int = 0, repeats: int = 5):
    """Erdos-Renyi n=4096, p=0.004, 3% edge noise; cost lam * D(w) over
    degree, clustering, mean neighbour degree and ego-net edges with
    w ~ U(0.5, 1.5); 100 unrolled log-Sinkhorn iterations at eps = 0.01 with
    the duals checkpointed every 10; loss = minus the plan mass on the
    ground-truth mapping.  Times the forward (cost build + iterations + plan)
    and the hand-written backward with HIP events on the launch stream
    (median of `repeats`), and states the error of dL/dw, dL/dlam and the
    plan against torch autograd through an FP64 unrolled Sinkhorn."""
    import torch

    from cugal_amd.adjacency import Adjacency
    from cugal_amd.sinkhorn_grad import feature_sinkhorn_plan, structural_features
    from cugal_amd.synthetic import erdos_renyi

    dev = torch.device("cuda:0")
    inst = erdos_renyi(4096, 0.004, 0.03, seed=seed)
    src, tgt = inst.source, inst.target
    n = max(max(src.nodes()), max(tgt.nodes())) + 1
    src.add_nodes_from(range(n))
    tgt.add_nodes_from(range(n))
    fs = structural_features(Adjacency.from_graph(src, dev))
    ft = structural_features(Adjacency.from_graph(tgt, dev))
    w = 0.5 + torch.rand(4, generator=torch.Generator().manual_seed(seed + 11))
    lam, reg, T, k = 0.05, 0.01, 100, 10
    G = torch.zeros((n, n), dtype=torch.float32, device=dev)
    G[torch.arange(n, device=dev), torch.as_tensor(inst.source_mapping, device=dev)] = -1.0
    fwd, bwd = [], []
    for r in range(repeats + 1):
        wg = w.to(dev).requires_grad_(True)
        lg = torch.tensor(lam, dtype=torch.float64, device=dev, requires_grad=True)
        start, mid, end = (torch.cuda.Event(enable_timing=True) for _ in range(3))
        torch.cuda.synchronize()
        start.record()
        plan = feature_sinkhorn_plan(fs, ft, wg, lg, reg, T, k)
        mid.record()
        (plan * G).sum().backward()
        end.record()
        torch.cuda.synchronize()
        if r > 0:  # the first run loads the kernels
            fwd.append(start.elapsed_time(mid))
            bwd.append(mid.elapsed_time(end))
    fwd.sort()
    bwd.sort()
    fwd_ms, bwd_ms = fwd[len(fwd) // 2], bwd[len(bwd) // 2]
    w64 = w.double().to(dev).requires_grad_(True)
    l64 = torch.tensor(lam, dtype=torch.float64, device=dev, requires_grad=True)
    d = (fs.double()[:, None, :] - ft.double()[None, :, :]).abs()
    ref = _torch_unrolled_plan(l64 * (d * w64).sum(-1), reg, T)
    d = None
    (ref * G.double()).sum().backward()
    plan_err = float((plan.detach().double() - ref.detach()).abs().max() / ref.detach().abs().max())
    ref = None
    rel_w = float((wg.grad.double() - w64.grad).abs().max() / w64.grad.abs().max())
    rel_l = abs(float(lg.grad) - float(l64.grad)) / abs(float(l64.grad))
    return {
        "workload": f"erdos_renyi n={n} p=0.004, 3% edge noise; cost lam*D(w) over 4 structural features, "
                    f"lam {lam}, eps {reg}, {T} unrolled iterations, checkpoints every {k}",
        "n": n, "forward_ms": fwd_ms, "backward_ms": bwd_ms, "backward_over_forward": bwd_ms / fwd_ms,
        "repeats": repeats, "grad_w": wg.grad.tolist(), "grad_w_fp64": w64.grad.tolist(),
        "grad_lam": float(lg.grad), "grad_lam_fp64": float(l64.grad),
        "max_rel_err_grad_w": rel_w, "rel_err_grad_lam": rel_l, "max_rel_err_plan": plan_err,
        "reference": "torch autograd through an FP64 unrolled log-Sinkhorn on the device",
    }


def cpu_baseline(n: int, reg: float, seconds_budget: float, total_iterations: int):
    """FP64 oracle log-Sinkhorn (torch.logsumexp, cugal/sinkhorn.py:107-111)
    on a cost of the bench's size; times whole iterations until the budget
    is spent and scales to the GPU run's Sinkhorn iteration count."""
    import torch

    from oracle import fugal as oracle

    g = torch.Generator().manual_seed(0)
    C = torch.rand((n, n), generator=g, dtype=torch.float64) * 4.0
    state = oracle.SinkhornState(n, "LOG")
    done, t0 = 0, time.perf_counter()
    while True:
        state.solve_log(C, reg, 1, 0.0, 10)
        done += 1
        if time.perf_counter() - t0 > seconds_budget:
            break
    per_iter = (time.perf_counter() - t0) / done
    return {
        "value": per_iter * total_iterations,
        "unit": "s (Sinkhorn only, extrapolated)",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": f"FP64 oracle log-Sinkhorn at n={n}: {done} iteration(s) timed ({per_iter:.2f} s each), "
                  f"scaled to the {total_iterations} iterations of the GPU alignment",
    }


def cpu_stage_baseline(source, target, reg: float, fw_steps: int, iterations: int, seconds_budget: float):
    """Per-stage FP64 oracle timing at the workload's size, extrapolated to
    the GPU run: one sparse-form gradient (cugal/pred.py:54-74, dense torch
    products on the host) times the Frank-Wolfe step count, plus the
    log-Sinkhorn iteration time (cpu_baseline) times the GPU run's
    iteration count."""
    import networkx as nx
    import torch

    from oracle import fugal as oracle

    n = source.number_of_nodes()
    A = torch.tensor(nx.to_numpy_array(source, nodelist=range(n)), dtype=torch.float64)
    B = torch.tensor(nx.to_numpy_array(target, nodelist=range(n)), dtype=torch.float64)
    g = torch.Generator().manual_seed(0)
    P = torch.rand((n, n), generator=g, dtype=torch.float64)
    P /= P.sum(1, keepdim=True)
    D = torch.rand((n, n), generator=g, dtype=torch.float64)
    t0 = time.perf_counter()
    oracle.sparse_gradient(A, B, P, D, 3)
    t_grad = time.perf_counter() - t0
    del A, B, P, D
    sk = cpu_baseline(n, reg, seconds_budget, iterations)
    return {
        "value": t_grad * fw_steps + sk["value"],
        "unit": "s (gradient + Sinkhorn, extrapolated per stage)",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": f"FP64 oracle at n={n}: one gradient {t_grad:.2f} s x {fw_steps} Frank-Wolfe steps; "
                  + sk["sample"],
    }


def reference_case(seed: int = 0):
    """The reference's own CPU-runnable case: Erdos-Renyi G(1000, 0.01),
    seed 0, 5 % of edges removed and as many non-edges added, seeded
    permutation; the reference's default Config (cugal/config.py:59-120:
    STANDARD, reg 1.0, <= 500 iterations, thr 1e-3, mu 0.5, 15 x 10
    Frank-Wolfe steps, dense adjacency with the doubled lambda term,
    DOUBLE_GREEDY).  cugal() on the device in FP32 and the FP64 oracle
    pipeline on the host cores, each timed end to end (graphs in, mapping
    out), with both node accuracies."""
    import torch

    from cugal_amd import Config, cugal
    from cugal_amd.profile import Profile
    from cugal_amd.synthetic import alignment_accuracy, answer_columns, erdos_renyi
    from oracle import fugal as oracle

    inst = erdos_renyi(1000, 0.01, 0.05, seed=seed)
    cfg = Config(device="cuda:0", dtype=torch.float32)
    cugal(copy.deepcopy(inst.source), copy.deepcopy(inst.target), cfg, Profile())  # warm-up
    torch.cuda.synchronize()
    profile = Profile()
    t0 = time.perf_counter()
    _, mapping = cugal(copy.deepcopy(inst.source), copy.deepcopy(inst.target), cfg, profile)
    torch.cuda.synchronize()
    gpu_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    run = oracle.align(copy.deepcopy(inst.source), copy.deepcopy(inst.target), oracle.OracleConfig())
    cpu_s = time.perf_counter() - t0
    return {
        "workload": "erdos_renyi n=1000 p=0.01 seed 0, 5% edge noise + refill, seeded permutation; the reference's "
                    "default Config (STANDARD, reg 1.0, <=500 its, thr 1e-3, mu 0.5, 15x10 FW, dense, DOUBLE_GREEDY)",
        "gpu_seconds": gpu_s, "gpu_accuracy": alignment_accuracy(inst.source_mapping, answer_columns(mapping)),
        "gpu_sinkhorn_iterations": sum(p.iteration_count for p in profile.sinkhorn_profiles),
        "cpu_seconds": cpu_s, "cpu_accuracy": alignment_accuracy(inst.source_mapping, answer_columns(run.mapping)),
        "cpu_sinkhorn_iterations": sum(run.sinkhorn_iterations), "cpu_threads": torch.get_num_threads(),
        "cpu": "FP64 oracle (restatement of the reference's CPU path), end to end on the host cores",
    }


def probe(args) -> int:
    """A few launches of every hot kernel on the bench workload, plus a
    device copy of one n x n FP32 matrix (4 n^2 bytes read, 4 n^2 written)
    that calibrates the FETCH_SIZE / WRITE_SIZE counters.  For profilers."""
    import torch

    from cugal_amd import _native as nat
    from cugal_amd.adjacency import Adjacency
    from cugal_amd.config import HungarianMethod
    from cugal_amd.feature_extraction import Features, node_features
    from cugal_amd.greedy_lap import greedy_lap_device
    from cugal_amd.hungarian import dense_hungarian
    from cugal_amd import _native as nat
    from cugal_amd.pred import _gradient_into, update_quasi_permutation_log, update_quasi_permutation_standard
    from cugal_amd.sinkhorn import LogKernel, StandardKernel, solve_log_native, solve_standard_native
    from cugal_amd.synthetic import watts_strogatz

    from cugal_amd import ordering
    from cugal_amd.sinkhorn import solve_standard_native_f64

    n = args.n
    inst = watts_strogatz(n, args.k, args.p, args.noise, seed=args.seed)
    dev = torch.device("cuda:0")
    A = Adjacency.from_graph(inst.source, dev)
    B = Adjacency.from_graph(inst.target, dev)
    feats = Features(node_features(A, args.mu), node_features(B, args.mu))
    # the CSR gradient on the labels the Frank-Wolfe loop uses
    # (node_order "auto": the target relabelled by reverse Cuthill-McKee)
    pb = ordering.choose_position(B, "auto")
    if pb is not None:
        pos = torch.from_numpy(pb).to(dev)
        Bo = ordering.relabel(B, pos)
        feats_o = Features(feats.source, ordering.permute_rows(feats.target, pos))
    else:
        Bo, feats_o = B, feats
    g = torch.Generator(device=dev).manual_seed(0)
    # P row-padded, as the alignment keeps it (the gradient gathers it in place)
    P = torch.empty((n, int(nat.query("cg_gradient_padded_ld", n))), device=dev)[:, :n]
    P.copy_(torch.rand((n, n), generator=g, device=dev))
    P /= P.sum(1, keepdim=True)
    G = torch.empty((n, n), device=dev)
    for lam in (3, 4):
        _gradient_into(G, A, B, P, feats, lam, 1.0, precision="bf16x2")
    for lam in (3, 4):
        _gradient_into(G, A, Bo, P, feats_o, lam, 1.0)
    # the general (directed) form on the same matrices, as the dense path of
    # a directed pair runs it
    Ad = Adjacency(A.col_indices, A.row_pointers, directed=True)
    Bd = Adjacency(B.col_indices, B.row_pointers, directed=True)
    _gradient_into(G, Ad, Bd, P, feats, 3, 2.0)
    us = torch.full((n,), 1.0 / n, device=dev)
    vs = torch.full((n,), 1.0 / n, device=dev)
    solve_standard_native(G, 4.0 * args.reg, us, vs, 10, 0.0, 10, None, fused=True)
    u = torch.zeros(n, device=dev)
    v = torch.zeros(n, device=dev)
    solve_log_native(G, args.reg, u, v, 20, 0.0, 10, None, fused=True)
    # FP64 STANDARD (the reference's default dtype) on an n/2 slice
    h = n // 2
    G64 = G[:h, :h].double().contiguous()
    u64 = torch.full((h,), 1.0 / h, dtype=torch.float64, device=dev)
    v64 = torch.full((h,), 1.0 / h, dtype=torch.float64, device=dev)
    solve_standard_native_f64(G64, 4.0 * args.reg, u64, v64, 5, 0.0, 10, None)
    G64 = u64 = v64 = None
    solve_log_native(G, args.reg, u, v, 10, 0.0, 10, None, fused=False)
    gap = torch.empty(1, dtype=torch.float64, device=dev)
    Q = P.clone()
    update_quasi_permutation_standard(Q, StandardKernel(G, 4.0 * args.reg), us, vs, 0.5, gap=gap)
    update_quasi_permutation_log(P, LogKernel(G, args.reg), u, v, 0.5, gap=gap)
    Q.copy_(P)
    greedy_lap_device(Q, HungarianMethod.PARALLEL_GREEDY)
    Q.copy_(P)
    greedy_lap_device(Q, HungarianMethod.MORE_RAND)
    # The exact LAP on a quasi-permutation with 5% rival entries (the shape
    # the alignment produces; a uniform random matrix would take far longer).
    perm = torch.randperm(n, generator=g, device=dev)
    Q = torch.rand((n, n), generator=g, device=dev) * (0.2 / n)
    Q[torch.arange(n, device=dev), perm] += 0.5
    swap = torch.randperm(n, generator=g, device=dev)[: n // 20]
    Q[swap, perm[torch.roll(swap, 1)]] += 0.6
    dense_hungarian(Q, maximize=True)
    Q = None
    greedy_lap_device(P.contiguous(), HungarianMethod.DOUBLE_GREEDY)
    copy = torch.empty_like(G)
    for _ in range(3):
        copy.copy_(G)
    torch.cuda.synchronize()
    print(json.dumps({"probe": "ok", "n": n, "copy_bytes": 4 * n * n}), flush=True)
    return 0


class Budget:
    """Wall clock of the whole run, counted from the interpreter's start
    (torch's first import included), against `seconds`: the optional legs
    after the timed steps run only while their estimate fits, so the JSON
    line always prints in time."""

    def __init__(self, seconds: float):
        self.seconds = seconds

    def left(self) -> float:
        return self.seconds - (time.time() - T_START)

    def allows(self, name: str, needs: float) -> bool:
        if self.left() >= needs:
            return True
        log(f"skipping {name}: {self.left():.0f} s of the wall-clock budget left, it needs ~{needs:.0f} s")
        return False

    def skipped(self, needs: float) -> dict:
        return {"skipped": f"wall-clock budget: {max(0.0, self.left()):.0f} s left of {self.seconds:.0f}, "
                           f"leg needs ~{needs:.0f} s"}


def _fp64_checks():
    """The FP64 device checkers shared with the tests (tests/fp64_checks.py)."""
    path = os.path.join(ROOT, "tests")
    if path not in sys.path:
        sys.path.insert(0, path)
    import fp64_checks

    return fp64_checks


def stage_errors(inputs, P, args, rows: int = 16, seed: int = 0):
    """Error of each stage against FP64 at the workload's size, from the
    alignment's own quasi-permutation P: the gradient in both forms (with
    distance and regulariser terms, lambda 7) on sampled rows; ten
    iterations of the fused Sinkhorn solve from zero duals, the plan on
    sampled rows and the column-marginal error against FP64 iterations on
    the device; the Frank-Wolfe update (alpha 2/3) and its duality gap
    against FP64 over the whole matrix."""
    import numpy as np
    import torch

    from cugal_amd import _native as nat
    from cugal_amd.feature_extraction import Features, node_features
    from cugal_amd.pred import _gradient_into, update_quasi_permutation_log
    from cugal_amd.sinkhorn import LogKernel, solve_log_native

    f64 = _fp64_checks()
    A, B = inputs
    n = P.shape[0]
    dev = P.device
    feats = Features(node_features(A, args.mu), node_features(B, args.mu))
    ld = int(nat.query("cg_gradient_padded_ld", n))
    Pp = torch.empty((n, ld), device=dev)[:, :n]
    Pp.copy_(P)
    rng = np.random.default_rng(seed)
    sample = sorted({0, n - 1, *[int(r) for r in rng.choice(n, size=rows, replace=False)]})
    idx = torch.as_tensor(sample, device=dev)
    lam = 7
    ref, prod, other = f64.gradient_rows(A, B, Pp, feats, lam, 1.0, sample)
    G = torch.empty((n, n), device=dev)
    out = {"rows": len(sample)}
    for form in ("bf16x2", "fp32"):
        _gradient_into(G, A, B, Pp, feats, lam, 1.0, precision=form)
        got = G[idx].double().cpu().numpy()
        err = np.abs(got - ref)
        out[f"gradient_{form}"] = {
            "max_abs_err": float(err.max()),
            "max_rel_err_of_row_max": float(np.max(err.max(1) / np.abs(ref).max(1))),
            "max_err_over_rounding_bound": float(np.max(err / (1e-5 * prod + 1e-6 * other + 1e-30))),
        }
    reg = args.reg
    u = torch.zeros(n, device=dev)
    v = torch.zeros(n, device=dev)
    iters = 10
    errors, _, _ = solve_log_native(G, reg, u, v, iters + 1, 0.0, 10, None, fused=True)
    u64, v64 = f64.log_sinkhorn(G, reg, torch.zeros(n, device=dev), torch.zeros(n, device=dev), iters + 1)
    got = f64.plan_rows(G, reg, u, v, sample)
    want = f64.plan_rows(G, reg, u64, v64, sample)
    big = want > 1e-200
    _, col = f64.marginals(G, reg, u, v)
    err64 = float((col - 1).abs().sum())
    out["sinkhorn"] = {
        "iterations": iters + 1, "regularization": reg,
        "plan_max_rel_err": float(np.max(np.abs(got - want)[big] / want[big])),
        "plan_exponent_floor": f64.exponent_floor(float(G.abs().max()), reg),
        "marginal_error_device": errors[-1], "marginal_error_fp64_of_same_duals": err64,
    }
    alpha = 2.0 / 3.0
    P_old = Pp.clone()
    gap = torch.empty(1, dtype=torch.float64, device=dev)
    update_quasi_permutation_log(Pp, LogKernel(G, reg), u, v, alpha, gap=gap)
    plan = f64.plan_rows(G, reg, u, v, sample)
    old = P_old[idx].double().cpu().numpy()
    want_p = old + alpha * (plan - old)
    s_plan, a_plan = f64.plan_inner(G, reg, u, v)
    s_old, a_old = f64.inner(G, P_old)
    gap64 = s_plan - s_old
    err_p = np.abs(Pp[idx].double().cpu().numpy() - want_p)
    out["update"] = {
        "alpha": alpha,
        "P_max_err_of_row_max": float(np.max(err_p.max(1) / np.abs(want_p).max(1))),
        "gap_device": float(gap), "gap_fp64": gap64,
        # the gap is a difference of two sums of size sum|G plan|; each plan
        # entry carries the exponent floor, so its error scales with that sum
        "gap_abs_err_over_sum_abs_G_plan": abs(float(gap) - gap64) / a_plan,
        "sum_abs_G_plan": a_plan, "sum_abs_G_P": a_old,
    }
    out["product_only"] = gradient_error_sample(A, B, P_old, rows=rows, seed=seed)
    out["reference"] = "FP64 on the device from the same FP32 inputs (gradient rows by CSR products, " \
                       "log-Sinkhorn by chunked logsumexp, update and gap by chunked sums)"
    return out


def mfma_gradient_seconds(inputs, n: int, repeats: int = 3) -> float:
    """Seconds per matrix-core gradient (HIP events on the launch stream)."""
    import torch

    from cugal_amd import _native as nat
    from cugal_amd.pred import _gradient_into

    A, B = inputs
    dev = torch.device("cuda:0")
    ld = int(nat.query("cg_gradient_padded_ld", n))
    P = torch.empty((n, ld), device=dev)[:, :n]
    P.fill_(1.0 / n)
    G = torch.empty((n, n), device=dev)
    _gradient_into(G, A, B, P, None, 3, 1.0, precision="bf16x2")
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(repeats):
        _gradient_into(G, A, B, P, None, 3, 1.0, precision="bf16x2")
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / repeats / 1e3


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--probe", action="store_true", help="launch each hot kernel a few times (for profilers)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--k", type=int, default=20)
    ap.add_argument("--p", type=float, default=0.1)
    ap.add_argument("--noise", type=float, default=0.01)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--reg", type=float, default=0.3)
    ap.add_argument("--mu", type=float, default=2.0)
    ap.add_argument("--threshold", type=float, default=1e-3)
    ap.add_argument("--sinkhorn-iterations", type=int, default=500)
    ap.add_argument("--iter-count", type=int, default=15)
    ap.add_argument("--fw-iter-count", type=int, default=10)
    ap.add_argument("--node-order", choices=("auto", "natural", "rcm"), default="auto")
    ap.add_argument("--workload", choices=("ws32768", "ba8192", "sbm65536", "er4096grad", "published"),
                    default="ws32768")
    ap.add_argument("--graphs", default=None,
                    help="with --workload published: comma-separated graph kinds (e.g. BIO_DMELA) instead of all seven")
    ap.add_argument("--runs", type=int, default=4, help="with --workload published: runs per graph")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiation-form (backward) run")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain two-sweep comparison run")
    ap.add_argument("--no-precision-form", action="store_true", help="skip the matrix-core gradient run")
    ap.add_argument("--no-extra-workloads", action="store_true", help="skip the BA n=8192 and SBM n=65536 legs")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--budget-s", type=float, default=560.0,
                    help="wall-clock seconds the whole run may take; optional legs are skipped past it")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        log("only --gpus 1 is supported: the alignment runs on a single MI355X")
        return 2

    import torch

    from cugal_amd.synthetic import watts_strogatz

    assert torch.cuda.is_available(), "bench.py needs a ROCm device"
    if args.probe:
        return probe(args)
    if args.workload == "ba8192":
        extra = ba8192_leg(args, cpu=not args.no_cpu)
        print(_contract_line(extra.pop("seconds"), args, extra.pop("workload"), 8192, extra), flush=True)
        return 0
    if args.workload == "sbm65536":
        extra = sbm65536_leg(args, precision_form=not args.no_precision_form)
        print(_contract_line(extra.pop("seconds"), args, extra.pop("workload"), extra.pop("n"), extra), flush=True)
        return 0
    if args.workload == "published":
        pub = published_experiment_leg(runs=args.runs, only=args.graphs.split(",") if args.graphs else None)
        print(_contract_line(pub["total_seconds"], args, pub.pop("workload") + "; value = the seven graphs' mean "
                             "cugal() seconds, summed", max(g["n"] for g in pub["graphs"].values()),
                             {"data": "the reference's shipped graphs with the published seeded noise",
                              "published_experiment": pub}), flush=True)
        return 0
    if args.workload == "er4096grad":
        diff = differentiation_form(args.seed)
        log(f"differentiation form: forward {diff['forward_ms']:.1f} ms, backward {diff['backward_ms']:.1f} ms")
        print(_contract_line(diff["backward_ms"] * 1e-3, args, diff["workload"] + " (value = backward seconds)",
                             diff["n"], {"differentiation_form": diff}), flush=True)
        return 0
    budget = Budget(args.budget_s)
    log(f"device {torch.cuda.get_device_name(0)}; generating WS(n={args.n}, k={args.k}, p={args.p})")
    t = time.perf_counter()
    inst = watts_strogatz(args.n, args.k, args.p, args.noise, seed=args.seed)
    log(f"instance ready in {time.perf_counter() - t:.1f}s "
        f"({inst.source.number_of_edges()} / {inst.target.number_of_edges()} edges)")

    # Kernel load / first-launch costs, outside every timed region.
    small = watts_strogatz(512, 10, 0.1, 0.01, seed=1)
    for fused in (True, False):
        run_alignment(small, make_config(fused, argparse.Namespace(**{**vars(args), "iter_count": 1,
                                                                      "fw_iter_count": 2})), False)
        run_alignment(small, make_config(fused, argparse.Namespace(**{**vars(args), "iter_count": 1,
                                                                      "fw_iter_count": 2})), False)
    inputs = device_inputs(inst, "cuda:0")

    def plain_leg(step_s: float, alignments_left: int):
        # The two-sweep comparison the metric is a ratio against; it runs
        # after the first warm-up when there is one (outside every timed
        # step), so the K timed steps cannot crowd it out of the budget.
        if args.no_plain:
            return None
        # The metric is a ratio to this leg: it may use the 30 s the optional
        # legs leave under the wall-clock limit (they are skipped instead).
        needs = 2.05 * step_s + 5 + alignments_left * step_s
        if not budget.allows("plain", needs - 30):
            return budget.skipped(needs)
        e, a, pprof, (pms, pcnt), _ = run_alignment(inst, make_config(False, args, precision="fp32"), True, inputs)
        log(f"plain two-sweep: {e:.2f}s acc {a:.4f}")
        return {"seconds": e, "accuracy": a,
                "sinkhorn_iterations": sum(p.iteration_count for p in pprof.sinkhorn_profiles),
                "sweep_ms": pms, "sweep_launches": pcnt,
                "path": "two-sweep Sinkhorn (two reads of C per iteration), FP32 CSR gradient"}

    plain = None
    plain_done = False
    for w in range(args.warmup):
        e, a, _, _, _ = run_alignment(inst, make_config(True, args), False, inputs)
        log(f"warmup {w}: {e:.2f}s acc {a:.4f}")
        if w == 0:
            plain = plain_leg(e, args.warmup - 1 + args.steps)
            plain_done = True

    times, accs, sweep_ms, sweep_n, iters_total, profile, P = [], [], 0.0, 0, 0, None, None
    for s in range(args.steps):
        P = None
        e, a, profile, (ms, cnt), P = run_alignment(inst, make_config(True, args), True, inputs)
        times.append(e)
        accs.append(a)
        sweep_ms += ms
        sweep_n += cnt
        iters_total = sum(p.iteration_count for p in profile.sinkhorn_profiles)
        log(f"step {s}: fused {e:.2f}s acc {a:.4f} sinkhorn iterations {iters_total}")
    value = sum(times) / len(times)
    step_s = max(times)
    api_extra = api_extra_seconds(inst, P, "cuda:0")

    from cugal_amd.pred import gradient_precision_for, mfma_density
    from cugal_amd.profile import Phase

    form = gradient_precision_for(inputs[0], inputs[1], "auto")
    grad_s = profile.phase_times.get(Phase.GRADIENT, 0.0)
    fw_steps = profile.frank_wolfe_iterations

    if not plain_done:
        plain = plain_leg(step_s, 0)

    cpu = None
    if not args.no_cpu:
        # FP64 oracle Sinkhorn iterations at n = 16384 take ~0.7 s each
        # the contract's CPU baseline comes first among the legs after the
        # timed steps; a few iterations (0.7 s each) are enough to time
        secs = min(args.cpu_seconds, budget.left() - 15)
        if secs >= 2:
            log("timing the FP64 CPU oracle sample")
            cpu = cpu_baseline(min(args.n, 16384), args.reg, secs, iters_total)
            if args.n > 16384:
                scale = (args.n / 16384) ** 2
                cpu["value"] *= scale
                cpu["sample"] += f"; measured at n=16384 and scaled by (n/16384)^2 = {scale:.1f}"
        else:
            cpu = budget.skipped(17)

    precision = None
    if not args.no_precision_form and budget.allows("precision-form stage errors", 20):
        t_mfma = mfma_gradient_seconds(inputs, args.n)
        est = value - grad_s + fw_steps * t_mfma
        npad = -(-args.n // 256) * 256
        precision = {
            "form": "gradient on the matrix cores: exact 0/1 BF16 adjacency times P split into BF16 hi + lo, "
                    "FP32 accumulation; FP32 log-Sinkhorn; FP64 marginal errors and duality gap",
            "gradient_seconds_each": t_mfma,
            "mfma_tflops_issued": 8.0 * npad ** 3 / t_mfma / 1e12,
            "mfma_tflops_algorithmic": 4.0 * args.n ** 3 / t_mfma / 1e12,
            "mfma_peak_tflops": 2500.0, "mfma_frac_algorithmic": 4.0 * args.n ** 3 / t_mfma / 1e12 / 2500.0,
            "estimated_seconds": est,
            "estimate": f"the headline alignment with its {fw_steps} CSR gradients ({grad_s:.2f} s) replaced by "
                        f"{fw_steps} timed matrix-core gradients",
            "estimated_ratio_to_plain": (plain["seconds"] / est) if plain and "seconds" in plain else None,
            "why_not_faster_here": f"at this graph's density the two CSR gathers need ~8.6e10 reads per gradient "
                                   f"against 4 n^3 = {4.0 * args.n ** 3:.2e} MFMA flops; the matrix cores win "
                                   f"above ~{100 * mfma_density(args.n):.2f} % density at this n",
            "stage_errors_vs_fp64": stage_errors(inputs, P, args),
        }
        log(f"precision form: MFMA gradient {t_mfma * 1e3:.1f} ms; per-stage errors measured")
    P = None

    legs = {}
    if not args.no_cpu:
        if budget.allows("reference case", 12):
            legs["reference_case"] = reference_case()
            rc = legs["reference_case"]
            log(f"reference case ER(1000), default Config: device {rc['gpu_seconds']:.2f}s acc "
                f"{rc['gpu_accuracy']:.4f}; FP64 oracle on {rc['cpu_threads']} threads "
                f"{rc['cpu_seconds']:.2f}s acc {rc['cpu_accuracy']:.4f}")
        else:
            legs["reference_case"] = budget.skipped(12)
    if not args.no_diff:
        if budget.allows("differentiation form", 10):
            diff = differentiation_form(args.seed)
            legs["differentiation_form"] = diff
            log(f"differentiation form: forward {diff['forward_ms']:.1f} ms, backward {diff['backward_ms']:.1f} ms "
                f"(x{diff['backward_over_forward']:.2f}); dL/dw rel err {diff['max_rel_err_grad_w']:.2e}")
        else:
            legs["differentiation_form"] = budget.skipped(10)
    if not args.no_extra_workloads:
        if budget.allows("BA n=8192", 25):
            legs["ba8192"] = ba8192_leg(args, cpu=False)
        else:
            legs["ba8192"] = budget.skipped(25)
        if budget.allows("SBM n=65536", 60):
            legs["sbm65536"] = sbm65536_leg(args, precision_form=False)
        else:
            legs["sbm65536"] = budget.skipped(60)
        if budget.allows("published experiment", 110):
            legs["published_experiment"] = published_experiment_leg(budget=budget)
        else:
            legs["published_experiment"] = budget.skipped(110)
    if precision is not None and not args.no_precision_form:
        if budget.allows("precision form, whole alignment", 2.8 * step_s + 10):
            e, a, xprof, _, _ = run_alignment(inst, make_config(True, args, precision="bf16x2"), False, inputs)
            precision["seconds"] = e
            precision["accuracy"] = a
            precision["gradient_seconds"] = xprof.phase_times.get(Phase.GRADIENT, 0.0)
            precision["ratio_to_plain"] = (plain["seconds"] / e) if plain and "seconds" in plain else None
            log(f"precision form (MFMA gradient): {e:.2f}s acc {a:.4f}")
        else:
            precision["seconds"] = budget.skipped(2.8 * step_s)["skipped"]

    n = args.n
    avg_sweep_ms = sweep_ms / max(1, sweep_n)
    sweep_bytes = 4.0 * n * n
    achieved = sweep_bytes / (avg_sweep_ms * 1e-3) / 1e9 if sweep_n else None
    roofline = {
        "kernel": "sk_fused_sweep (one read of C per Sinkhorn iteration)",
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": measured_traffic(n),
        "bytes_per_launch": sweep_bytes, "avg_launch_ms": avg_sweep_ms, "launches": sweep_n,
    }

    phase = {k.value: v for k, v in profile.phase_times.items()} if profile else {}
    order = "natural labels" if args.node_order == "natural" else f"node_order {args.node_order}"
    out = {
        "metric": "align time (s) n=32k mixed vs FP32 plain @ equal node acc; Sinkhorn HBM GB/s",
        "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": value * 1e3, "higher_is_better": False, "scaling": "strong", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic",
        "config": {"workload": f"watts_strogatz n={n} k={args.k} p={args.p}, {args.noise:.0%} edge noise, "
                               f"seeded permutation; LOG reg {args.reg}, thr {args.threshold}, mu {args.mu}, "
                               f"{args.iter_count}x{args.fw_iter_count} FW, DOUBLE_GREEDY; value = fused Sinkhorn "
                               f"(one read of C per iteration) with the {form} gradient "
                               f"({'CSR gathers' if form == 'fp32' else 'matrix cores'}, chosen by "
                               f"gradient_precision auto at this density), {order}",
                   "n": n, "parallelism": "single GPU"},
        "accuracy": sum(accs) / len(accs), "sinkhorn_iterations": iters_total,
        "sinkhorn_gbs": achieved, "phase_times": phase, "api_extra_s": api_extra, "plain": plain,
        "speedup_vs_plain": (plain["seconds"] / value) if plain and "seconds" in plain else None,
        "precision_form": precision,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    out.update(legs)
    out["wall_s"] = time.time() - T_START
    print(json.dumps(out), flush=True)
    return 0


def _contract_line(value, args, workload: str, n: int, extra: dict) -> str:
    out = {
        "metric": "align time (s) n=32k mixed vs FP32 plain @ equal node acc; Sinkhorn HBM GB/s",
        "value": value, "unit": "s", "n_gpus": 1, "steps": 1, "warmup": args.warmup, "ms_per_step": value * 1e3,
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
        "config": {"workload": workload, "n": n, "parallelism": "single GPU"},
    }
    out.update(extra)
    return json.dumps(out)


def ba8192_leg(args, cpu: bool) -> dict:
    """Barabasi-Albert n=8192, m=5, 2% edge noise: the alignment with the
    exact device Hungarian (DENSE); the same P rounded by scipy's
    linear_sum_assignment on the host and by the greedy walk; with `cpu`,
    the FP64 oracle timed per stage and extrapolated."""
    import numpy as np
    import scipy.optimize
    import torch

    from cugal_amd.config import HungarianMethod
    from cugal_amd.greedy_lap import greedy_lap_device
    from cugal_amd.hungarian import dense_hungarian
    from cugal_amd.synthetic import alignment_accuracy, barabasi_albert

    n = 8192
    ba_args = argparse.Namespace(**{**vars(args), "iter_count": 15, "fw_iter_count": 10})
    inst = barabasi_albert(n, 5, 0.02, seed=args.seed)
    small = barabasi_albert(512, 5, 0.02, seed=1)
    warm = argparse.Namespace(**{**vars(ba_args), "iter_count": 1, "fw_iter_count": 2})
    run_alignment(small, make_config(True, warm, lap="DENSE"), False)
    inputs = device_inputs(inst, "cuda:0")
    e, acc, prof, (ms, cnt), P = run_alignment(inst, make_config(True, ba_args, lap="DENSE"), True, inputs)
    iters = sum(p.iteration_count for p in prof.sinkhorn_profiles)
    col_dense = dense_hungarian(P, maximize=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    col_dense, (free_rows, dsteps) = dense_hungarian(P, maximize=True, stats=True)
    torch.cuda.synchronize()
    t_dense = time.perf_counter() - t0
    Ph = P.cpu().numpy().astype(np.float64)
    t0 = time.perf_counter()
    _, col_scipy = scipy.optimize.linear_sum_assignment(Ph, maximize=True)
    t_scipy = time.perf_counter() - t0
    Q = P.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    col_greedy = greedy_lap_device(Q, HungarianMethod.DOUBLE_GREEDY)
    torch.cuda.synchronize()
    t_greedy = time.perf_counter() - t0
    col_dense = col_dense.cpu().numpy()
    col_greedy = col_greedy.cpu().numpy()
    avg = ms / max(1, cnt)
    out = {
        "workload": f"barabasi_albert n={n} m=5, 2% edge noise, seeded permutation; LOG reg {args.reg}, "
                    f"15x10 FW, DENSE (device Hungarian)",
        "seconds": e, "accuracy": acc, "sinkhorn_iterations": iters,
        "phase_times": {k.value: v for k, v in prof.phase_times.items()},
        "lap": {
            "hungarian_device_s": t_dense, "hungarian_free_rows_after_warm_start": free_rows,
            "hungarian_dijkstra_steps": dsteps, "scipy_s": t_scipy,
            "hungarian_equals_scipy": bool(np.array_equal(col_dense, col_scipy)),
            "objective_hungarian": float(Ph[np.arange(n), col_dense].sum()),
            "objective_scipy": float(Ph[np.arange(n), col_scipy].sum()),
            "greedy_s": t_greedy, "accuracy_greedy": alignment_accuracy(inst.source_mapping, col_greedy),
            "accuracy_hungarian": alignment_accuracy(inst.source_mapping, col_dense),
        },
        "sinkhorn_gbs": 4.0 * n * n / (avg * 1e-3) / 1e9 if cnt else None,
    }
    from cugal_amd.profile import Phase

    out["sinkhorn_us_per_iteration"] = prof.phase_times.get(Phase.SINKHORN, 0.0) / max(1, iters) * 1e6
    log(f"ba8192: {e:.2f}s acc {acc:.4f}; hungarian {t_dense * 1e3:.1f} ms vs scipy {t_scipy * 1e3:.0f} ms, "
        f"equal={out['lap']['hungarian_equals_scipy']}")
    if cpu:
        log("timing the FP64 CPU oracle per stage")
        out["cpu_baseline"] = cpu_stage_baseline(inst.source, inst.target, args.reg, prof.frank_wolfe_iterations,
                                                 iters, args.cpu_seconds)
        log(f"FP64 oracle estimate: {out['cpu_baseline']['value']:.0f}s ({out['cpu_baseline']['sample']})")
    return out


def sbm65536_leg(args, precision_form: bool) -> dict:
    """Stochastic block model: 64 blocks of 1024 nodes, intra/inter degrees
    48/16, 1% edge noise; 15 lambda stages of one Frank-Wolfe step, greedy
    LAP.  P (17 GB) and the cost stay in HBM; the Sinkhorn uses the
    two-sweep kernels (the fused sweep keeps v in LDS up to n = 32768)."""
    import torch

    from cugal_amd.profile import Phase
    from cugal_amd.synthetic import alignment_accuracy, stochastic_block

    t = time.perf_counter()
    inst = stochastic_block(64, 1024, 48.0, 16.0, 0.01, seed=args.seed)
    n = inst.n
    log(f"sbm instance n={n} in {time.perf_counter() - t:.1f}s ({len(inst.source_edges)} edges)")
    cfg_args = argparse.Namespace(**{**vars(args), "iter_count": 15, "fw_iter_count": 1})
    small = watts_small()
    run_alignment(small, make_config(True, argparse.Namespace(**{**vars(args), "iter_count": 1,
                                                                 "fw_iter_count": 2})), False)
    inputs = inst.adjacencies("cuda:0")
    e, acc, prof, (ms, cnt), P = run_alignment(inst, make_config(True, cfg_args), True, inputs)
    P = None
    avg = ms / max(1, cnt)
    gbs = 4.0 * n * n / (avg * 1e-3) / 1e9 if cnt else None
    out = {
        "workload": f"stochastic_block 64 x 1024, degrees 48/16, 1% edge noise, seeded permutation; LOG reg "
                    f"{args.reg}, 15 lambda stages x 1 FW step, DOUBLE_GREEDY",
        "n": n, "seconds": e, "accuracy": acc,
        "sinkhorn_iterations": sum(p.iteration_count for p in prof.sinkhorn_profiles),
        "phase_times": {k.value: v for k, v in prof.phase_times.items()},
        "sinkhorn_gbs": gbs, "sinkhorn_sweep_avg_ms": avg, "sinkhorn_sweeps": cnt,
        "roofline": {"kernel": "sk_col_sweep / sk_row_sweep (two reads of C per iteration)", "bound": "hbm",
                     "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS if gbs else None, "traffic": None},
    }
    log(f"sbm65536: {e:.2f}s acc {acc:.4f}, sinkhorn sweeps {gbs:.0f} GB/s, "
        f"phases {[(k.value, round(v, 2)) for k, v in prof.phase_times.items()]}")
    if precision_form:
        # The precision form at this scale: the gradient on the matrix cores
        # (0.1 % density, so "auto" above chose the CSR gathers).
        ex, ax, xprof, _, Px = run_alignment(inst, make_config(True, cfg_args, precision="bf16x2"), False, inputs)
        grad_s = xprof.phase_times.get(Phase.GRADIENT, 0.0)
        steps = xprof.frank_wolfe_iterations
        npad = -(-n // 256) * 256
        issued = steps * 8.0 * npad ** 3 / grad_s / 1e12 if grad_s else None
        out["precision_form"] = {
            "seconds": ex, "accuracy": ax, "gradient_seconds": grad_s, "frank_wolfe_steps": steps,
            "phase_times": {k.value: v for k, v in xprof.phase_times.items()},
            "mfma_tflops_issued": issued,
            "mfma_tflops_algorithmic": steps * 4.0 * n ** 3 / grad_s / 1e12 if grad_s else None,
            "mfma_peak_tflops": 2500.0, "mfma_frac": (issued / 2500.0) if issued else None,
            "gradient_error": gradient_error_sample(inputs[0], inputs[1], Px),
        }
        Px = None
        log(f"sbm65536 precision form: {ex:.2f}s acc {ax:.4f}, gradient {grad_s:.2f}s ({issued:.0f} TFLOP/s issued)")
    return out


PUBLISHED_RESULTS = os.path.join(ROOT, "tests", "golden", "published_results.json")
PUBLISHED_DATA = os.path.join(ROOT, "tests", "golden")


def published_experiment_leg(runs: int = 4, forced: bool = True, budget=None, only=None) -> dict:
    """The reference's published benchmark experiment (its settings are in
    benchmark/results/CUGAL_res.json:2-81): the seven graphs it ships under
    benchmark/data/, its three noise levels, seed and four runs, and its
    cuGAL Config (FP32 LOG, reg 0.3, <= 500 iterations, thr 1e-3, mu 2,
    15 x 10 FW, sparse adjacency, SCIPY LAP).  The noisy pairs are drawn from
    the seed in the reference's order (benchmark/experiment.py:440-460: every
    graph, every level, every run), so they are the published runs' pairs.
    cugal() runs on the first level's pairs (source noise 0.1), each timed as
    Profile.time, the wall time of one call (cugal/pred.py:243), beside the
    published cuGAL mean time and accuracy.  At this threshold the solves
    of the five larger graphs run to the 500-iteration cap, the published
    ones and this build's alike; the two small graphs stop earlier.  The
    published accuracies and iteration counts come from a source version
    other than the current one (the FP64 oracle of the current source lands
    where this build does on these pairs, and with the current source's
    in-place LOG marginal check, cugal/sinkhorn.py:67-74, no solve
    converges), so they are context; parity on these pairs is held against
    the oracle's alignment of them (tests/golden/published_pairs_oracle.json).
    With `forced`, one run per graph is
    repeated with the threshold at 0, so that every solve runs 500
    iterations, and the Sinkhorn time per iteration compares directly with
    the published Sinkhorn time over its iteration count.  `only` limits the
    run to the named graph kinds (the pairs are still drawn for all seven, so
    they stay the published ones)."""
    import numpy as np
    import torch

    from cugal_amd import Config, cugal
    from cugal_amd.experiment import DATA_FILES, GraphKind, NoiseLevel, generate_graph, load_edge_list
    from cugal_amd.profile import Phase, Profile
    from cugal_amd.synthetic import alignment_accuracy, answer_columns

    published = json.load(open(PUBLISHED_RESULTS))["cugal"]
    exp = published["experiment"]
    cfg = Config.from_dict({**exp["algorithms"][0]["config"], "device": "cuda:0"})
    forced_cfg = Config.from_dict({**exp["algorithms"][0]["config"], "device": "cuda:0", "sinkhorn_threshold": 0.0})
    levels = [NoiseLevel(**lv) for lv in exp["noise_levels"]]
    first = levels[0]
    generator = np.random.default_rng(seed=exp["seed"])
    pairs = {}
    for g in exp["graphs"]:
        kind = GraphKind[g["kind"]]
        graph = load_edge_list(os.path.join(PUBLISHED_DATA, DATA_FILES[kind]))
        drawn = [[generate_graph(graph, None, generator, level) for _ in range(exp["num_runs"])] for level in levels]
        pairs[kind.value] = drawn[0]
    pub = {(e["graph"], e["noise"]["source_noise"]): e for e in published["entries"]}

    # first-use kernel loading outside every timed call
    src0, tgt0, _, _ = pairs["IN_ARENAS"][0]
    cugal(copy.deepcopy(src0), copy.deepcopy(tgt0), cfg, Profile())
    torch.cuda.synchronize()

    graphs = {}
    total, total_pub = 0.0, 0.0
    for g in exp["graphs"]:
        name = g["kind"]
        if only is not None and name not in only:
            continue
        if budget is not None and budget.left() < 30:
            graphs[name] = budget.skipped(30)
            continue
        times, accs, its = [], [], []
        for source, target, mapping, _ in pairs[name][:runs]:
            profile = Profile()
            _, answer = cugal(copy.deepcopy(source), copy.deepcopy(target), cfg, profile)
            times.append(profile.time)
            accs.append(alignment_accuracy(mapping, answer_columns(answer)))
            its.append(sum(p.iteration_count for p in profile.sinkhorn_profiles))
        ref = pub[(name, first.source_noise)]
        entry = {
            "n": max(pairs[name][0][0].number_of_nodes(), pairs[name][0][1].number_of_nodes()),
            "runs": len(times), "seconds": sum(times) / len(times), "accuracy": sum(accs) / len(accs),
            "sinkhorn_iterations": sum(its) / len(its),
            "published_seconds": ref["time"], "published_accuracy": ref["accuracy"],
            "published_sinkhorn_iterations": ref["sinkhorn_iterations"],
        }
        entry["time_ratio_published_over_this"] = ref["time"] / entry["seconds"]
        total += entry["seconds"]
        total_pub += ref["time"]
        if forced:
            source, target, mapping, _ = pairs[name][0]
            profile = Profile()
            _, answer = cugal(copy.deepcopy(source), copy.deepcopy(target), forced_cfg, profile)
            iters = sum(p.iteration_count for p in profile.sinkhorn_profiles)
            sk = profile.phase_times.get(Phase.SINKHORN, 0.0)
            entry["forced_500"] = {
                "seconds": profile.time, "sinkhorn_seconds": sk, "sinkhorn_iterations": iters,
                "accuracy": alignment_accuracy(mapping, answer_columns(answer)),
                "ms_per_iteration": sk / max(1, iters) * 1e3,
                "published_ms_per_iteration": (ref["sinkhorn_time"] / ref["sinkhorn_iterations"] * 1e3
                                               if ref["sinkhorn_iterations"] else None),
            }
        graphs[name] = entry
        log(f"published experiment {name} (n={entry['n']}): {entry['seconds']:.2f}s acc {entry['accuracy']:.4f} "
            f"({entry['sinkhorn_iterations']:.0f} its) against published {ref['time']:.2f}s acc {ref['accuracy']:.4f}"
            + (f"; 500-iteration solves {entry['forced_500']['ms_per_iteration']:.3f} ms/it against "
               f"{entry['forced_500']['published_ms_per_iteration']:.3f}" if forced else ""))
    return {
        "workload": "the reference's published experiment (benchmark/results/CUGAL_res.json): bio-dmela, ca-Erdos992, "
                    "ca-GrQc, ca-netscience, in-arenas, inf-power, inf-euroroad at source noise 0.1, seed 283841238, "
                    "its cuGAL Config (FP32 LOG reg 0.3, <=500 its, thr 1e-3, mu 2, 15x10 FW, sparse, SCIPY)",
        "published_gpu": "not stated in the reference", "graphs": graphs,
        "total_seconds": total, "published_total_seconds": total_pub,
    }


def watts_small():
    from cugal_amd.synthetic import watts_strogatz

    return watts_strogatz(512, 10, 0.1, 0.01, seed=1)


if __name__ == "__main__":
    sys.exit(main())
