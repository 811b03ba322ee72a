"""Sinkhorn-Knopp projection of the Frank-Wolfe gradient, the surface of
cugal/sinkhorn.py:13-156.

``SinkhornState`` keeps the dual potentials u, v on the device; they
warm-start every solve of an alignment (cugal/pred.py:95).  The LOG method
runs entirely in ``cg_sinkhorn_log_solve`` and STANDARD in
``cg_sinkhorn_standard_solve``: the iterations, the column-marginal checks
every ``sinkhorn_eval_freq`` iterations, the threshold stop and Lehmann
momentum are all decided on the device, and the host synchronises once per
solve.  K (-C / reg, or exp(-C / reg)) is never materialised; ``solve``
returns it as a :class:`LogKernel` / :class:`StandardKernel` view of the cost.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat
from .config import Config, SinkhornMethod
from .profile import SinkhornProfile, TimeStamp

M_EPS = 1e-16


def can_use_cuda(config: Config) -> bool:
    """True when the HIP Sinkhorn applies (cugal/sinkhorn.py:16-17): FP32
    for both methods, FP64 for STANDARD."""
    if "cuda" not in config.device:
        return False
    if config.dtype == torch.float32:
        return True
    return config.dtype == torch.float64 and config.sinkhorn_method == SinkhornMethod.STANDARD


@dataclass
class LogKernel:
    """K = -cost / regularization, kept implicit."""

    cost: torch.Tensor
    regularization: float

    def materialize(self) -> torch.Tensor:
        return -self.cost / self.regularization


@dataclass
class StandardKernel:
    """K = exp(-cost / regularization), kept implicit."""

    cost: torch.Tensor
    regularization: float

    def materialize(self) -> torch.Tensor:
        return torch.exp(self.cost / -self.regularization)


def scale_kernel_matrix(K, u: torch.Tensor, v: torch.Tensor) -> torch.Tensor:
    """diag(u) K diag(v), in place on a tensor K (cugal/sinkhorn.py:20-23);
    a StandardKernel is materialised first."""
    if isinstance(K, StandardKernel):
        K = K.materialize()
    K *= v.reshape(1, -1)
    K *= u.reshape(-1, 1)
    return K


def scale_kernel_matrix_log(K, log_u: torch.Tensor, log_v: torch.Tensor) -> torch.Tensor:
    """exp(K + u_i + v_j), in place on a tensor K (cugal/sinkhorn.py:26-30);
    a LogKernel is materialised first."""
    if isinstance(K, LogKernel):
        K = K.materialize()
    K += log_u[:, None]
    K += log_v[None, :]
    torch.exp(K, out=K)
    return K


def marginal_error(K, u, v) -> float:
    """The STANDARD form's column-marginal error (cugal/sinkhorn.py:33-40).
    The device solve computes it inside each checking iteration and returns
    it in SinkhornProfile.errors; there is no standalone kernel for it."""
    raise NotImplementedError(
        "marginal_error is computed inside the device STANDARD solve (SinkhornProfile.errors); "
        "no standalone form is provided"
    )


def marginal_error_log(K, log_u: torch.Tensor, log_v: torch.Tensor) -> float:
    """sum_r |sum_c exp(K_rc + v_c + u_r) - 1| (the device kernel of
    cuda_kernels/sinkhorn_log.cu:68-98, FP64 accumulation)."""
    if isinstance(K, LogKernel):
        K = K.materialize()
    K = K.contiguous()
    n = K.shape[0]
    nat.require_device_tensor(K, "K", torch.float32)
    ws = nat.workspace(K.device, nat.query("cg_sinkhorn_log_marginal_workspace", n), "marginal")
    out = (np.zeros(1, dtype=np.float64))
    log_u, log_v = log_u.contiguous(), log_v.contiguous()
    nat.require_device_tensor(log_u, "log_u", torch.float32, n)
    nat.require_device_tensor(log_v, "log_v", torch.float32, n)
    nat.call(
        "cg_sinkhorn_log_marginal", nat.ptr(K), n, nat.ptr(log_u), nat.ptr(log_v),
        out.ctypes.data, nat.ptr(ws), ws.numel(), nat.stream_of(K.device),
    )
    return float(out[0])


def lehmann_momentum(config: Config, errors: list[float]) -> float:
    """Over-relaxation factor of cugal/sinkhorn.py:78-88 (host restatement
    of the rule the device applies)."""
    if config.sinkhorn_momentum_start is None or len(errors) * config.sinkhorn_eval_freq < config.sinkhorn_momentum_start:
        return 1.0
    index = config.sinkhorn_momentum_start // config.sinkhorn_eval_freq
    if len(errors) < index - 1:
        return 1.0
    ratio = min(errors[index - 1] / errors[index - 2], 0.99)
    if math.isnan(ratio):
        ratio = 0.99
    power = 1.0 / config.sinkhorn_eval_freq
    return 2.0 / (1.0 + math.sqrt(1 - ratio**power))


@dataclass
class SinkhornState:
    u: torch.Tensor
    v: torch.Tensor

    def __init__(self, n: int, config: Config):
        if "cuda" not in config.device:
            raise NotImplementedError("Sinkhorn runs on a ROCm device (config.device 'cuda')")
        match config.sinkhorn_method:
            case SinkhornMethod.LOG:
                vector = torch.zeros(n, device=config.device, dtype=config.dtype)
            case SinkhornMethod.STANDARD:
                vector = torch.full((n,), 1 / n, device=config.device, dtype=config.dtype)
            case _:
                raise NotImplementedError(
                    "SinkhornMethod.MIX has no solver in the reference (cugal/sinkhorn.py:146-156)"
                )
        self.u = vector.clone()
        self.v = vector

    def solve_log(self, C: torch.Tensor, config: Config, profile: SinkhornProfile | None = None):
        """Log-domain solve of cugal/sinkhorn.py:90-120 on the device."""
        profile = SinkhornProfile() if profile is None else profile
        if "cuda" not in config.device or config.dtype != torch.float32:
            raise NotImplementedError("the HIP log-Sinkhorn needs config.device 'cuda' and dtype torch.float32")
        C = C.contiguous()
        n = C.shape[0]
        nat.require_device_tensor(C, "cost", torch.float32, n * n)
        nat.require_device_tensor(self.u, "u", torch.float32, n)
        nat.require_device_tensor(self.v, "v", torch.float32, n)
        start_time = TimeStamp(config.device)
        errors, iterations, ms = solve_log_native(
            C, config.sinkhorn_regularization, self.u, self.v, config.sinkhorn_iterations,
            config.sinkhorn_threshold, config.sinkhorn_eval_freq, config.sinkhorn_momentum_start,
            fused=config.sinkhorn_fused,
        )
        profile.iteration_count = iterations
        profile.time = TimeStamp(config.device).elapsed_seconds(start_time)
        profile.errors = errors
        return LogKernel(C, config.sinkhorn_regularization), self.u, self.v

    def solve_standard(self, C: torch.Tensor, config: Config, profile: SinkhornProfile | None = None):
        """Scaling-form solve of cugal/sinkhorn.py:122-144 on the device, in
        FP32 or (config.dtype torch.float64, the reference's default) FP64."""
        profile = SinkhornProfile() if profile is None else profile
        if not can_use_cuda(config):
            raise NotImplementedError("the HIP Sinkhorn needs config.device 'cuda' and dtype torch.float32 "
                                      "(or torch.float64 for STANDARD)")
        C = C.contiguous()
        n = C.shape[0]
        dtype = config.dtype
        nat.require_device_tensor(C, "cost", dtype, n * n)
        nat.require_device_tensor(self.u, "u", dtype, n)
        nat.require_device_tensor(self.v, "v", dtype, n)
        start_time = TimeStamp(config.device)
        solve = solve_standard_native_f64 if dtype == torch.float64 else solve_standard_native
        errors, iterations = solve(
            C, config.sinkhorn_regularization, self.u, self.v, config.sinkhorn_iterations,
            config.sinkhorn_threshold, config.sinkhorn_eval_freq, config.sinkhorn_momentum_start,
            fused=config.sinkhorn_fused,
        )
        profile.iteration_count = iterations
        profile.time = TimeStamp(config.device).elapsed_seconds(start_time)
        profile.errors = errors
        return StandardKernel(C, config.sinkhorn_regularization), self.u, self.v

    def solve(self, C: torch.Tensor, config: Config, profile: SinkhornProfile | None = None):
        match config.sinkhorn_method:
            case SinkhornMethod.LOG:
                return self.solve_log(C, config, profile)
            case SinkhornMethod.STANDARD:
                return self.solve_standard(C, config, profile)
            case _:
                raise NotImplementedError("SinkhornMethod.MIX has no solver in the reference")


class SweepTiming:
    """Accumulates the device time of the C-reading Sinkhorn sweep kernels
    when enabled (used for roofline reporting)."""

    enabled = False
    total_ms = 0.0
    launches = 0

    @classmethod
    def reset(cls, enabled: bool) -> None:
        cls.enabled = enabled
        cls.total_ms = 0.0
        cls.launches = 0


def solve_log_native(
    C: torch.Tensor, regularization: float, u: torch.Tensor, v: torch.Tensor, max_iterations: int,
    threshold: float, eval_freq: int, momentum_start, fused: bool = True,
) -> tuple[list[float], int, float]:
    """Run cg_sinkhorn_log_solve in place on (u, v).  Returns (errors,
    iteration count, sweep milliseconds or 0)."""
    n = C.shape[0]
    dev = C.device
    ws = nat.workspace(dev, nat.query("cg_sinkhorn_workspace", n, int(fused), max_iterations, eval_freq), "sinkhorn")
    cap = max_iterations // max(1, eval_freq) + 4
    errors = np.zeros(cap, dtype=np.float64)
    err_count = np.zeros(1, dtype=np.int32)
    iters = np.zeros(1, dtype=np.int32)
    ms = np.zeros(1, dtype=np.float32)
    launches = np.zeros(1, dtype=np.int32)
    timing = SweepTiming.enabled
    nat.call(
        "cg_sinkhorn_log_solve", nat.ptr(C), n, float(regularization), nat.ptr(u), nat.ptr(v),
        int(max_iterations), float(threshold), int(eval_freq),
        -1 if momentum_start is None else int(momentum_start), int(bool(fused)),
        errors.ctypes.data, cap, err_count.ctypes.data, iters.ctypes.data,
        ms.ctypes.data if timing else None, launches.ctypes.data if timing else None,
        nat.ptr(ws), ws.numel(), nat.stream_of(dev),
    )
    if timing:
        SweepTiming.total_ms += float(ms[0])
        SweepTiming.launches += int(launches[0])
    count = int(min(err_count[0], cap))
    return [float(e) for e in errors[:count]], int(iters[0]), float(ms[0])


def solve_standard_native(
    C: torch.Tensor, regularization: float, u: torch.Tensor, v: torch.Tensor, max_iterations: int,
    threshold: float, eval_freq: int, momentum_start, fused: bool = True,
) -> tuple[list[float], int]:
    """Run cg_sinkhorn_standard_solve in place on (u, v).  Returns (errors,
    iteration count)."""
    n = C.shape[0]
    dev = C.device
    ws = nat.workspace(dev, nat.query("cg_sinkhorn_workspace", n, int(fused), max_iterations, eval_freq), "sinkhorn")
    cap = max_iterations // max(1, eval_freq) + 4
    errors = np.zeros(cap, dtype=np.float64)
    err_count = np.zeros(1, dtype=np.int32)
    iters = np.zeros(1, dtype=np.int32)
    nat.call(
        "cg_sinkhorn_standard_solve", nat.ptr(C), n, float(regularization), nat.ptr(u), nat.ptr(v),
        int(max_iterations), float(threshold), int(eval_freq),
        -1 if momentum_start is None else int(momentum_start), int(bool(fused)),
        errors.ctypes.data, cap, err_count.ctypes.data, iters.ctypes.data,
        nat.ptr(ws), ws.numel(), nat.stream_of(dev),
    )
    count = int(min(err_count[0], cap))
    return [float(e) for e in errors[:count]], int(iters[0])


def solve_standard_native_f64(
    C: torch.Tensor, regularization: float, u: torch.Tensor, v: torch.Tensor, max_iterations: int,
    threshold: float, eval_freq: int, momentum_start, fused: bool = False,
) -> tuple[list[float], int]:
    """Run cg_sinkhorn_standard_solve_f64 in place on FP64 (u, v); `fused`
    is accepted for symmetry (the FP64 solve always takes two sweeps).
    Returns (errors, iteration count)."""
    n = C.shape[0]
    dev = C.device
    ws = nat.workspace(dev, nat.query("cg_sinkhorn_standard_f64_workspace", n, max_iterations, eval_freq), "sinkhorn")
    cap = max_iterations // max(1, eval_freq) + 4
    errors = np.zeros(cap, dtype=np.float64)
    err_count = np.zeros(1, dtype=np.int32)
    iters = np.zeros(1, dtype=np.int32)
    nat.call(
        "cg_sinkhorn_standard_solve_f64", nat.ptr(C), n, float(regularization), nat.ptr(u), nat.ptr(v),
        int(max_iterations), float(threshold), int(eval_freq),
        -1 if momentum_start is None else int(momentum_start),
        errors.ctypes.data, cap, err_count.ctypes.data, iters.ctypes.data,
        nat.ptr(ws), ws.numel(), nat.stream_of(dev),
    )
    count = int(min(err_count[0], cap))
    return [float(e) for e in errors[:count]], int(iters[0])
