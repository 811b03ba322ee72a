"""ctypes binding of ``libcugal_hip.so``, the C ABI declared in
``include/cugal_hip.h``.

The library takes raw device pointers, sizes and a ``hipStream_t``; this module
turns torch tensors into those arguments, passes the current torch stream,
provides grow-only device workspaces from torch's caching allocator, and
raises :class:`NativeError` with the library's message on any non-zero status.

There is no fallback: if the library is missing or fails to load, every call
raises.
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcugal_hip.so")

_lib: Optional[ctypes.CDLL] = None

c_p = ctypes.c_void_p
c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_f = ctypes.c_float
c_d = ctypes.c_double
c_sz = ctypes.c_size_t

# name -> (restype, argtypes)
_SIGNATURES = {
    "cg_last_error": (ctypes.c_char_p, []),
    "cg_version": (c_int, []),
    "cg_csr_build_workspace": (c_sz, [c_i64, c_i64, c_int]),
    "cg_csr_build": (c_int, [c_p, c_i64, c_i64, c_int, c_p, c_p, c_p, c_sz, c_p]),
    "cg_adjacency_matmul": (c_int, [c_p, c_p, c_i64, c_p, c_p, c_int, c_p]),
    "cg_graph_features": (c_int, [c_p, c_p, c_i64, c_p, c_p, c_p, c_p]),
    "cg_average_neighbor_features": (c_int, [c_p, c_p, c_i64, c_p, c_p, c_p]),
    "cg_node_features_workspace": (c_sz, [c_i64]),
    "cg_node_features": (c_int, [c_p, c_p, c_i64, c_f, c_p, c_p, c_p, c_sz, c_p]),
    "cg_egonet_edges": (c_int, [c_p, c_p, c_i64, c_p, c_p]),
    "cg_edge_common_neighbours": (c_int, [c_p, c_p, c_i64, c_p, c_p]),
    "cg_fugal_features_workspace": (c_sz, [c_i64]),
    "cg_fugal_features": (c_int, [c_p, c_p, c_i64, c_f, c_p, c_p, c_sz, c_p]),
    "cg_add_distance": (c_int, [c_p, c_p, c_i64, c_p, c_p]),
    "cg_regularize": (c_int, [c_p, c_p, c_i64, c_int, c_p]),
    "cg_gradient_workspace": (c_sz, [c_i64]),
    "cg_gradient_padded_ld": (c_i64, [c_i64]),
    "cg_gradient": (c_int, [c_p, c_p, c_p, c_p, c_i64, c_p, c_i64, c_p, c_p, c_int, c_int, c_f, c_p, c_p, c_sz, c_p]),
    "cg_gradient_directed_workspace": (c_sz, [c_i64]),
    "cg_gradient_directed": (
        c_int, [c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_i64, c_p, c_i64, c_p, c_p, c_int, c_int, c_f, c_p, c_p, c_sz, c_p],
    ),
    "cg_permute_copy": (c_int, [c_p, c_i64, c_i64, c_p, c_p, c_p, c_i64, c_p]),
    "cg_adjacency_bits_words": (c_sz, [c_i64]),
    "cg_adjacency_pack_bits": (c_int, [c_p, c_p, c_i64, c_p, c_p]),
    "cg_gradient_mfma_workspace": (c_sz, [c_i64]),
    "cg_gradient_mfma": (c_int, [c_p, c_p, c_i64, c_p, c_i64, c_p, c_p, c_int, c_int, c_f, c_p, c_p, c_sz, c_p]),
    "cg_sinkhorn_log_step": (c_int, [c_p, c_i64, c_p, c_p, c_f, c_p]),
    "cg_sinkhorn_log_marginal_workspace": (c_sz, [c_i64]),
    "cg_sinkhorn_log_marginal": (c_int, [c_p, c_i64, c_p, c_p, c_p, c_p, c_sz, c_p]),
    "cg_sinkhorn_workspace": (c_sz, [c_i64, c_int, c_int, c_int]),
    "cg_sinkhorn_fused_supported": (c_int, [c_i64]),
    "cg_sinkhorn_log_solve": (
        c_int,
        [c_p, c_i64, c_f, c_p, c_p, c_int, c_d, c_int, c_int, c_int, c_p, c_int, c_p, c_p, c_p, c_p, c_p, c_sz, c_p],
    ),
    "cg_sinkhorn_standard_solve": (
        c_int, [c_p, c_i64, c_f, c_p, c_p, c_int, c_d, c_int, c_int, c_int, c_p, c_int, c_p, c_p, c_p, c_sz, c_p],
    ),
    "cg_sinkhorn_standard_f64_workspace": (c_sz, [c_i64, c_int, c_int]),
    "cg_sinkhorn_standard_solve_f64": (
        c_int, [c_p, c_i64, c_d, c_p, c_p, c_int, c_d, c_int, c_int, c_p, c_int, c_p, c_p, c_p, c_sz, c_p],
    ),
    "cg_update_quasi_permutation_f64": (
        c_int, [c_int, c_p, c_i64, c_p, c_i64, c_p, c_p, c_d, c_d, c_p, c_p, c_sz, c_p],
    ),
    "cg_feature_cost": (c_int, [c_p, c_p, c_p, c_f, c_i64, c_p, c_p]),
    "cg_sinkhorn_plan": (c_int, [c_p, c_i64, c_f, c_p, c_p, c_p, c_p, c_sz, c_p]),
    "cg_sinkhorn_checkpoint_floats": (c_sz, [c_i64, c_int, c_int]),
    "cg_sinkhorn_unrolled_workspace": (c_sz, [c_i64, c_int]),
    "cg_sinkhorn_unrolled_forward": (
        c_int, [c_p, c_i64, c_f, c_int, c_int, c_p, c_p, c_p, c_p, c_p, c_p, c_sz, c_p],
    ),
    "cg_sinkhorn_unrolled_backward": (
        c_int, [c_p, c_i64, c_f, c_int, c_int, c_p, c_p, c_p, c_p, c_p, c_p, c_f, c_p, c_p, c_sz, c_p],
    ),
    "cg_update_workspace": (c_sz, [c_i64]),
    "cg_update_quasi_permutation_log": (c_int, [c_p, c_i64, c_p, c_i64, c_p, c_p, c_f, c_f, c_p, c_p, c_sz, c_p]),
    "cg_update_quasi_permutation_standard": (c_int, [c_p, c_i64, c_p, c_i64, c_p, c_p, c_f, c_f, c_p, c_p, c_sz, c_p]),
    "cg_update_quasi_permutation_log_static": (c_int, [c_p, c_i64, c_p, c_i64, c_p, c_p, c_f, c_f, c_p, c_p]),
    "cg_lap_workspace": (c_sz, [c_i64]),
    "cg_lap_greedy": (c_int, [c_p, c_i64, c_int, c_p, c_p, c_sz, c_p]),
    "cg_lap_greedy_random": (c_int, [c_p, c_i64, c_int, c_p, c_p, c_p, c_p, c_sz, c_p]),
    "cg_lap_hungarian_workspace": (c_sz, [c_i64]),
    "cg_lap_hungarian": (c_int, [c_p, c_i64, c_int, c_p, c_p, c_p, c_sz, c_p]),
    "cg_alignment_counts_workspace": (c_sz, [c_i64, c_i64]),
    "cg_alignment_counts": (c_int, [c_p, c_p, c_i64, c_p, c_p, c_i64, c_p, c_p, c_i64, c_p, c_p, c_sz, c_p]),
}


class NativeError(RuntimeError):
    """A libcugal_hip entry point returned a non-zero status."""


def exported_symbols() -> list[str]:
    """Names the binding expects the library to export."""
    return list(_SIGNATURES)


def lib() -> ctypes.CDLL:
    """Load the library (once).  torch is imported first so that its HIP
    runtime is the one the library binds to (both use SONAME
    libamdhip64.so.7)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(
                f"{LIB_PATH} is not built: run `make -C cugal_amd/csrc` or __graft_entry__.build()"
            )
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def call(name: str, *args) -> None:
    """Call entry point `name`; raise NativeError on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().cg_last_error().decode(errors="replace")
        raise NativeError(f"{name} failed ({rc}): {msg}")


def query(name: str, *args):
    """Call a non-status entry point (workspace sizes, predicates)."""
    return getattr(lib(), name)(*args)


def ptr(t: Optional[torch.Tensor]) -> Optional[int]:
    """Device address of a tensor.  torch reports 0 for a tensor with no
    elements; a zero-element view of an allocated buffer (the column array of
    a graph without edges) passes the buffer's address instead, so the entry
    points' null checks accept it (nothing is read through it)."""
    if t is None:
        return None
    p = t.data_ptr()
    if p == 0 and t.numel() == 0:
        storage = t.untyped_storage()
        if storage.nbytes() > 0:
            return storage.data_ptr() + t.storage_offset() * t.element_size()
    return p


def row_stride(t: torch.Tensor, name: str) -> int:
    """Row stride (elements) of a 2-D tensor whose rows are contiguous, as
    the entry points that take a row stride expect; raises otherwise."""
    if t.dim() != 2 or t.stride(1) != 1 or t.stride(0) < t.shape[1]:
        raise ValueError(f"{name} must be a 2-D tensor with contiguous rows, got strides {tuple(t.stride())}")
    return int(t.stride(0))


def require_device_tensor(t: torch.Tensor, name: str, dtype: torch.dtype, numel: Optional[int] = None) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must live on a ROCm device (device 'cuda'), got {t.device}")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} must have {numel} elements, got {t.numel()}")


def stream_of(device: torch.device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


_workspaces: dict = {}
_workspaces_lock = threading.Lock()


def workspace(device: torch.device, nbytes: int, tag: str = "default") -> torch.Tensor:
    """Grow-only device scratch buffer (uint8) from torch's allocator, one
    per (device, calling thread, current stream, tag): calls from different
    host threads or on different streams never share one, so they may
    overlap; calls of one thread on one stream run in stream order."""
    dev = torch.device(device)
    key = (str(dev), threading.get_ident(), torch.cuda.current_stream(dev).cuda_stream, tag)
    with _workspaces_lock:
        buf = _workspaces.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)
            _workspaces[key] = buf
    return buf


def release_workspaces() -> None:
    _workspaces.clear()
