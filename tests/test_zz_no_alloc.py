"""cg_update_quasi_permutation_log_static, the update entry point that takes
no workspace (its partial sums live in a fixed device array of the
library), against cg_update_quasi_permutation_log on the same inputs
through the C ABI: two calls in a row, the second on a smaller matrix,
each bitwise equal (the same kernels in the same order)."""

import pytest
import torch

from cugal_amd import _native as nat

pytestmark = pytest.mark.gpu


def _inputs(n, seed, device, ld=None):
    g = torch.Generator(device=device).manual_seed(seed)
    C = torch.rand((n, n), generator=g, device=device) * 3
    u = -torch.rand(n, generator=g, device=device)
    v = -torch.rand(n, generator=g, device=device)
    ld = ld or n
    P = torch.empty((n, ld), device=device)[:, :n]
    P.copy_(torch.rand((n, n), generator=g, device=device) / n)
    return C, u, v, P


def _with_workspace(P, C, u, v, alpha, reg, gap):
    n = P.shape[0]
    ws = torch.empty(int(nat.query("cg_update_workspace", n)), dtype=torch.uint8, device=P.device)
    nat.call("cg_update_quasi_permutation_log", nat.ptr(P), P.stride(0), nat.ptr(C), n, nat.ptr(u), nat.ptr(v),
             float(alpha), float(reg), nat.ptr(gap), nat.ptr(ws), ws.numel(), nat.stream_of(P.device))


def _static(P, C, u, v, alpha, reg, gap):
    n = P.shape[0]
    nat.call("cg_update_quasi_permutation_log_static", nat.ptr(P), P.stride(0), nat.ptr(C), n, nat.ptr(u),
             nat.ptr(v), float(alpha), float(reg), nat.ptr(gap), nat.stream_of(P.device))


def test_static_update_equals_the_workspace_form(device):
    # n = 4096: 2048 blocks of two rows, the array's full length of partial
    # sums; the second, smaller call (1000 blocks of one row) leaves the
    # first call's partials beyond its own, which it must not read.
    for n, ld, seed in ((4096, 4096 + 36, 1), (1000, 1000, 2)):
        C, u, v, P = _inputs(n, seed, device, ld)
        Q = P.clone()
        g1 = torch.empty(1, dtype=torch.float64, device=device)
        g2 = torch.empty(1, dtype=torch.float64, device=device)
        _with_workspace(P, C, u, v, 0.4, 0.7, g1)
        _static(Q, C, u, v, 0.4, 0.7, g2)
        torch.cuda.synchronize()
        assert torch.equal(P, Q), n
        assert torch.equal(g1, g2), (n, float(g1), float(g2))


def test_static_update_rejects_sizes_beyond_its_maximum(device):
    C, u, v, P = _inputs(8, 3, device)
    gap = torch.empty(1, dtype=torch.float64, device=device)
    with pytest.raises(nat.NativeError, match="outside"):
        nat.call("cg_update_quasi_permutation_log_static", nat.ptr(P), 8, nat.ptr(C), (1 << 30) + 1, nat.ptr(u),
                 nat.ptr(v), 0.5, 1.0, nat.ptr(gap), nat.stream_of(device))
