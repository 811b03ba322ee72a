"""Hand-written adjoint of the unrolled log-Sinkhorn (cugal_amd.sinkhorn_grad)
against torch autograd through the FP64 restatement of the reference's
log-domain iteration and against FP64 central differences."""

import numpy as np
import pytest
import torch

from cugal_amd.sinkhorn_grad import feature_cost, feature_sinkhorn_plan, sinkhorn_plan, structural_features
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu


def _rand(shape, seed, scale=1.0, dtype=torch.float64):
    return torch.rand(shape, generator=torch.Generator().manual_seed(seed), dtype=dtype) * scale


@pytest.mark.parametrize("k", [1, 3, 7, 20])
def test_grad_cost_matches_torch_autograd(device, k):
    n, reg, T = 200, 0.5, 20
    C = _rand((n, n), 1, 3.0).float()
    M = _rand((n, n), 2) - 0.5
    C_gpu = C.to(device).requires_grad_(True)
    plan = sinkhorn_plan(C_gpu, reg, T, checkpoint_every=k)
    (plan.double() * M.to(device)).sum().backward()
    C64 = C.double().requires_grad_(True)
    ref_plan = oracle.unrolled_plan(C64, reg, T)
    (ref_plan * M).sum().backward()
    np.testing.assert_allclose(plan.detach().double().cpu().numpy(), ref_plan.detach().numpy(), rtol=1e-4, atol=1e-8)
    g, r = C_gpu.grad.double().cpu(), C64.grad
    scale = float(r.abs().max())
    np.testing.assert_allclose(g.numpy(), r.numpy(), rtol=0, atol=2e-3 * scale)


def _grad_vs_autograd(device, C, reg, T, k, offset=0, seed=2):
    n = C.shape[0]
    M = _rand((n, n), seed) - 0.5
    if offset:
        # A cost whose data pointer is not 16-byte aligned takes the separate
        # row and column passes instead of the fused adjoint sweep.
        store = torch.empty(n * n + offset, device=device)
        C_gpu = store[offset:].view(n, n)
        C_gpu.copy_(C.to(device))
        assert C_gpu.data_ptr() % 16 != 0
        C_gpu.requires_grad_(True)
    else:
        C_gpu = C.to(device).requires_grad_(True)
    plan = sinkhorn_plan(C_gpu, reg, T, checkpoint_every=k)
    (plan.double() * M.to(device)).sum().backward()
    C64 = C.double().to(device).requires_grad_(True)
    ref_plan = oracle.unrolled_plan(C64, reg, T)
    (ref_plan * M.to(device)).sum().backward()
    np.testing.assert_allclose(plan.detach().double().cpu().numpy(), ref_plan.detach().cpu().numpy(), rtol=1e-4,
                               atol=1e-8)
    g, r = C_gpu.grad.double().cpu(), C64.grad.cpu()
    scale = float(r.abs().max())
    np.testing.assert_allclose(g.numpy(), r.numpy(), rtol=0, atol=2e-3 * scale)


@pytest.mark.parametrize("offset", [0, 1])
@pytest.mark.parametrize("n,k,T", [(256, 4, 9), (1024, 10, 20)])
def test_grad_fused_and_separate_passes(device, n, k, T, offset):
    C = _rand((n, n), 21 + n, 3.0).float()
    _grad_vs_autograd(device, C, 0.4, T, k, offset)


@pytest.mark.parametrize("n", [2, 3, 4, 8, 17, 333])
def test_grad_small_and_odd_sizes(device, n):
    # Widths far below a block (4 and 8 take the fused adjoint sweep with
    # one live thread group; the odd ones the separate passes).
    C = _rand((n, n), 51 + n, 3.0).float()
    _grad_vs_autograd(device, C, 0.5, 7, 3)


def test_grad_longest_segment(device):
    # 64 iterations per segment: the gradient pass holds (4 * 64 + 1) * 64
    # floats of LDS (above the 64 KiB a launch gets without asking).
    C = _rand((128, 128), 31, 2.0).float()
    _grad_vs_autograd(device, C, 0.5, 70, 64)


def test_grad_above_fused_size(device):
    # n > 8192: the column vectors no longer fit in LDS; separate passes.
    C = _rand((8200, 8200), 41, 3.0).float()
    _grad_vs_autograd(device, C, 0.5, 4, 3)


def test_grad_zero_iterations_is_plan_derivative(device):
    # plan = exp(-C/reg) with no iterations: dL/dC = -G plan / reg.
    n, reg = 64, 0.8
    C = _rand((n, n), 3, 2.0).float()
    G = _rand((n, n), 4) - 0.5
    C_gpu = C.to(device).requires_grad_(True)
    plan = sinkhorn_plan(C_gpu, reg, 0, checkpoint_every=5)
    plan.backward(G.float().to(device))
    want = -G * torch.exp(-C.double() / reg) / reg
    np.testing.assert_allclose(C_gpu.grad.double().cpu().numpy(), want.numpy(), rtol=2e-5, atol=1e-9)


def _oracle_loss(fs, ft, w, lam, reg, T, M):
    return (oracle.unrolled_plan(oracle.feature_cost(fs, ft, w, lam), reg, T) * M).sum()


def test_grad_weights_and_lambda_match_central_differences(device):
    n, reg, T, k = 150, 0.3, 30, 10
    fs = _rand((n, 4), 5, 5.0)
    ft = _rand((n, 4), 6, 5.0)
    w = 0.5 + _rand(4, 7)
    lam = 0.7
    M = _rand((n, n), 8) - 0.5
    wg = w.float().to(device).requires_grad_(True)
    lg = torch.tensor(lam, dtype=torch.float64, device=device, requires_grad=True)
    plan = feature_sinkhorn_plan(fs.float().to(device), ft.float().to(device), wg, lg, reg, T, k)
    (plan.double() * M.to(device)).sum().backward()
    # FP64 central differences, h = 1e-6
    h = 1e-6
    fd_w = []
    for m in range(4):
        e = torch.zeros(4, dtype=torch.float64)
        e[m] = h
        fd_w.append((_oracle_loss(fs, ft, w + e, lam, reg, T, M) - _oracle_loss(fs, ft, w - e, lam, reg, T, M)) / (2 * h))
    fd_l = (_oracle_loss(fs, ft, w, lam + h, reg, T, M) - _oracle_loss(fs, ft, w, lam - h, reg, T, M)) / (2 * h)
    fd_w = torch.stack(fd_w)
    # and torch autograd through the FP64 restatement
    w64 = w.clone().requires_grad_(True)
    l64 = torch.tensor(lam, dtype=torch.float64, requires_grad=True)
    _oracle_loss(fs, ft, w64, l64, reg, T, M).backward()
    np.testing.assert_allclose(fd_w.numpy(), w64.grad.numpy(), rtol=1e-5)
    np.testing.assert_allclose(float(fd_l), float(l64.grad), rtol=1e-5)
    np.testing.assert_allclose(wg.grad.double().cpu().numpy(), fd_w.numpy(), rtol=2e-3, atol=1e-3 * float(fd_w.abs().max()))
    assert abs(float(lg.grad) - float(fd_l)) <= 2e-3 * abs(float(fd_l)) + 1e-6


def test_differentiation_form_full_size(device):
    # ER(4096, 0.004) with 3% noise; cost = lam * D(w) over degree,
    # clustering, mean neighbour degree and ego-net edges; 100 unrolled
    # iterations at eps = 0.01 with checkpoints every 10; loss = minus the
    # plan mass on the ground-truth mapping.  Checked against torch autograd
    # through the FP64 restatement on the device.
    from cugal_amd.adjacency import Adjacency
    from cugal_amd.synthetic import erdos_renyi

    inst = erdos_renyi(4096, 0.004, 0.03, seed=0)
    src, tgt = inst.source, inst.target
    n = max(max(src.nodes()), max(tgt.nodes())) + 1
    src.add_nodes_from(range(n))
    tgt.add_nodes_from(range(n))
    fs = structural_features(Adjacency.from_graph(src, device))
    ft = structural_features(Adjacency.from_graph(tgt, device))
    w = 0.5 + _rand(4, 11, dtype=torch.float32)
    lam, reg, T, k = 0.05, 0.01, 100, 10
    G = torch.zeros((n, n), dtype=torch.float64, device=device)
    G[torch.arange(n, device=device), torch.as_tensor(inst.source_mapping, device=device)] = -1.0
    wg = w.to(device).requires_grad_(True)
    lg = torch.tensor(lam, dtype=torch.float64, device=device, requires_grad=True)
    start, mid, end = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    start.record()
    plan = feature_sinkhorn_plan(fs, ft, wg, lg, reg, T, k)
    mid.record()
    (plan.double() * G).sum().backward()
    end.record()
    torch.cuda.synchronize()
    fwd_ms, bwd_ms = start.elapsed_time(mid), mid.elapsed_time(end)
    w64 = w.double().to(device).requires_grad_(True)
    l64 = torch.tensor(lam, dtype=torch.float64, device=device, requires_grad=True)
    C64 = oracle.feature_cost(fs.double(), ft.double(), w64, l64)
    C64.retain_grad()
    ref = oracle.unrolled_plan(C64, reg, T)
    (ref * G).sum().backward()
    print(f"differentiation form n={n}: forward {fwd_ms:.1f} ms, backward {bwd_ms:.1f} ms "
          f"(ratio {bwd_ms / fwd_ms:.2f}); dL/dw {wg.grad.tolist()} vs {w64.grad.tolist()}; "
          f"dL/dlam {float(lg.grad):.6g} vs {float(l64.grad):.6g}")
    rel_w = float((wg.grad.double() - w64.grad).abs().max() / w64.grad.abs().max())
    rel_l = abs(float(lg.grad) - float(l64.grad)) / abs(float(l64.grad))
    # measured 4.5e-7 and 1.5e-7: the contracted gradients average the
    # FP32 rounding of the n^2 adjoint entries
    assert rel_w < 1e-4 and rel_l < 1e-4, (rel_w, rel_l)
    ref = None
    # dL/dC itself, entry by entry, through the cost-level Function
    Cg = feature_cost(fs, ft, w.to(device), lam).requires_grad_(True)
    plan = sinkhorn_plan(Cg, reg, T, checkpoint_every=k)
    (plan.double() * G).sum().backward()
    got, want = Cg.grad.double(), C64.grad
    scale = float(want.abs().max())
    err = float((got - want).abs().max())
    print(f"dL/dC n={n}: max |error| {err:.3e} = {err / scale:.3e} of max |dL/dC|")
    assert err <= 1e-4 * scale
