"""The precision form of the gradient (cg_gradient_mfma: 0/1 BF16 adjacency
times BF16 hi + lo P on the matrix cores, FP32 accumulation) against the
FP64 oracle (cugal/pred.py:38-74), with the error bound of the split stated:
each of the two products rounds its right/left operand to 16 significant
bits, so |G - G64| <= ~2 * 2^-16 * (2 A P B)_ij + FP32 accumulation."""

import time

import networkx as nx
import numpy as np
import pytest
import torch

from cugal_amd.adjacency import Adjacency
from cugal_amd.config import Config
from cugal_amd.feature_extraction import Features, node_features
from cugal_amd.pred import _gradient_into
from oracle import fugal as oracle

pytestmark = pytest.mark.gpu

SPLIT_REL = 4e-5  # 2 * 2^-16 = 3.05e-5, rounded up for the FP32 sums


def _case(device, n, p, seed):
    ga = nx.erdos_renyi_graph(n, p, seed=seed)
    gb = nx.erdos_renyi_graph(n, p, seed=seed + 100)
    A = Adjacency.from_graph(ga, device)
    B = Adjacency.from_graph(gb, device)
    gen = torch.Generator().manual_seed(seed)
    # Entries over several decades, so the lo half matters.
    P = torch.rand((n, n), generator=gen) ** 3
    P /= P.sum(1, keepdim=True)
    Ad = torch.tensor(nx.to_numpy_array(ga, nodelist=range(n)))
    Bd = torch.tensor(nx.to_numpy_array(gb, nodelist=range(n)))
    return A, B, P, Ad, Bd


def test_packed_adjacency_is_exact(device):
    n = 300  # padded to 512: word (k // 32) * 512 + i holds bit k % 32 of row i
    g = nx.erdos_renyi_graph(n, 0.05, seed=1)
    A = Adjacency.from_graph(g, device)
    words = A.packed_bits().cpu().numpy().view(np.uint32).reshape(512 // 32, 512)
    k = np.arange(512)
    bits = (words[k // 32, :] >> (k % 32)[:, None].astype(np.uint32)) & 1  # [k, i]
    want = np.zeros((512, 512), dtype=np.uint32)
    want[:n, :n] = nx.to_numpy_array(g, nodelist=range(n)).astype(np.uint32)
    assert np.array_equal(bits.T, want)


@pytest.mark.parametrize("sparse", [True, False])
@pytest.mark.parametrize("n,p", [(64, 0.2), (256, 0.05), (300, 0.05), (1000, 0.02), (2048, 0.3)])
def test_mfma_gradient_matches_oracle(device, sparse, n, p):
    A, B, P, Ad, Bd = _case(device, n, p, seed=n)
    cfg = Config(device="cuda", dtype=torch.float32, mu=0.5)
    feats = Features(node_features(A, cfg.mu), node_features(B, cfg.mu))
    lam = 3
    factor = 1.0 if sparse else 2.0
    out = torch.empty((n, n), device=device)
    _gradient_into(out, A, B, P.to(device), feats, lam, factor, precision="bf16x2")
    fs, ft = feats.source.double().cpu(), feats.target.double().cpu()
    D = torch.sqrt(((fs[:, None, :] - ft[None, :, :]) ** 2).sum(-1))
    Pd = P.double()
    ref = oracle.sparse_gradient(Ad, Bd, Pd, D, lam) if sparse else oracle.dense_gradient(Ad, Bd, Pd, D, lam)
    apb = 2.0 * (Ad @ Pd @ Bd)
    err = (out.double().cpu() - ref).abs()
    bound = SPLIT_REL * apb + 2e-6 * (D + factor * lam * (1 + 2 * Pd)) + 1e-7
    worst = float((err / (apb + 1e-30)).max())
    print(f"\nmfma gradient n={n} p={p}: max |err| {float(err.max()):.3e}, "
          f"max |err| / (2 A P B) {worst:.2e}")
    assert bool((err <= bound).all())


def test_mfma_and_csr_paths_agree(device):
    n = 1536
    A, B, P, _, _ = _case(device, n, 0.05, seed=4)
    Pd = P.to(device)
    out32 = torch.empty((n, n), device=device)
    out16 = torch.empty((n, n), device=device)
    _gradient_into(out32, A, B, Pd, None, 5, 1.0, precision="fp32")
    _gradient_into(out16, A, B, Pd, None, 5, 1.0, precision="bf16x2")
    apb = (out32 - 5.0 * (1 - 2 * Pd)).abs()  # = 2 A P B up to FP32 rounding
    assert bool(((out16 - out32).abs() <= SPLIT_REL * apb + 1e-6).all())


def test_mfma_closed_form_uniform_P(device):
    # P = J/n: A P B = d_A d_B^T / n; 1/n in BF16 hi + lo is exact for n a
    # power of two, so the result is exact up to FP32 accumulation.
    n = 1024
    ga = nx.watts_strogatz_graph(n, 6, 0.3, seed=2)
    gb = nx.barabasi_albert_graph(n, 3, seed=3)
    A, B = Adjacency.from_graph(ga, device), Adjacency.from_graph(gb, device)
    P = torch.full((n, n), 1.0 / n, device=device)
    out = torch.empty((n, n), device=device)
    _gradient_into(out, A, B, P, None, 0, 1.0, precision="bf16x2")
    da = torch.tensor([d for _, d in sorted(ga.degree())], dtype=torch.float64)
    db = torch.tensor([d for _, d in sorted(gb.degree())], dtype=torch.float64)
    want = -2.0 * torch.outer(da, db) / n
    np.testing.assert_allclose(out.double().cpu().numpy(), want.numpy(), rtol=1e-6, atol=1e-6)


def test_mfma_gradient_rate(device):
    # Timing only (the matrix-core rate of the two GEMMs at n = 8192).
    n = 8192
    ga = nx.erdos_renyi_graph(n, 0.01, seed=5)
    A = Adjacency.from_graph(ga, device)
    P = torch.rand((n, n), device=device)
    out = torch.empty((n, n), device=device)
    _gradient_into(out, A, A, P, None, 1, 1.0, precision="bf16x2")
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        _gradient_into(out, A, A, P, None, 1, 1.0, precision="bf16x2")
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    print(f"\nmfma gradient n={n}: {dt * 1e3:.2f} ms, {4 * n**3 / dt / 1e12:.0f} TFLOP/s algorithmic "
          f"({8 * n**3 / dt / 1e12:.0f} TFLOP/s issued with the hi + lo split)")
    assert torch.isfinite(out).all()


@pytest.mark.parametrize("p,want", [(0.002, "fp32"), (0.08, "bf16x2")])
def test_auto_precision_runs_the_chosen_kernel(device, p, want):
    # "auto" gives bit for bit the output of the kernel it selects.
    from cugal_amd.pred import gradient_precision_for

    n = 2048
    A, B, P, _, _ = _case(device, n, p, seed=9)
    assert gradient_precision_for(A, B, "auto") == want
    Pd = P.to(device)
    cfg = Config(device="cuda", dtype=torch.float32, mu=0.5)
    feats = Features(node_features(A, cfg.mu), node_features(B, cfg.mu))
    got = torch.empty((n, n), device=device)
    ref = torch.empty((n, n), device=device)
    _gradient_into(got, A, B, Pd, feats, 2, 1.0, precision="auto")
    _gradient_into(ref, A, B, Pd, feats, 2, 1.0, precision=want)
    assert torch.equal(got, ref)
