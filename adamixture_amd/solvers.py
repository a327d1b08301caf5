"""Batched SQP and BVLS solvers (drop-in for the reference's sqp_kernel and
bvls_kernel torch ops, adamixture/src/utils_c/cuda/{sqp,bvls}_kernel.cu)."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib

_QUEUE: dict[torch.device, torch.Tensor] = {}


def _queue(dev: torch.device) -> torch.Tensor:
    q = _QUEUE.get(dev)
    if q is None:
        q = _QUEUE[dev] = torch.zeros(1, dtype=torch.int64, device=dev)
    return q


def _f64c(t: torch.Tensor, dev) -> torch.Tensor:
    return t.to(device=dev, dtype=torch.float64).contiguous()


def v_kk(K: int, device) -> torch.Tensor:
    """br_qn_gpu.py:248-250: right singular vectors of ones(1, K) (CPU LAPACK),
    transposed."""
    _, _, vt = torch.linalg.svd(torch.ones((1, K), dtype=torch.float64), full_matrices=True)
    return vt.t().to(device=device, dtype=torch.float64).contiguous()


def _sqp(simplex: int, H, r, X, vkk, n: int, K: int) -> torch.Tensor:
    dev = X.device
    H, r, X = _f64c(H, dev), _f64c(r, dev), _f64c(X, dev)
    vk = _f64c(vkk, dev) if vkk is not None else None
    out = torch.empty((n, K), dtype=torch.float64, device=dev)
    q = _queue(dev)
    _lib.call("adx_sqp_solve", simplex, H.data_ptr(), r.data_ptr(), X.data_ptr(), _lib.ptr(vk), out.data_ptr(),
              n, K, q.data_ptr(), 8, _lib.stream_handle(dev))
    return out


def sqp_solve_q(XtX_q: torch.Tensor, Xtz_q: torch.Tensor, Q: torch.Tensor, v_kk: torch.Tensor, N: int,
                K: int) -> torch.Tensor:
    """sqp_kernel.cu:320-345 sqp_solve_q_cuda (Xtz_q is the negated gradient)."""
    if not 1 <= K <= 64:
        raise RuntimeError("K must be between 1 and 64")
    return _sqp(1, XtX_q, Xtz_q, Q, v_kk, N, K)


def sqp_solve_p(XtX_p: torch.Tensor, Xtz_p: torch.Tensor, P: torch.Tensor, M: int, K: int) -> torch.Tensor:
    """sqp_kernel.cu:347-370 sqp_solve_p_cuda (Xtz_p is the negated gradient)."""
    if not 1 <= K <= 64:
        raise RuntimeError("K must be between 1 and 64")
    return _sqp(0, XtX_p, Xtz_p, P, None, M, K)


def batch_bvls_bpp(A: torch.Tensor, B: torch.Tensor, lower: float, upper: float, *,
                   start: str = "inverse", Ainv: torch.Tensor | None = None) -> torch.Tensor:
    """bvls_kernel.cu:188-218 batch_bvls_bpp_cuda.  ``start="inverse"`` begins
    from X0 = B inv(A)^T like the CUDA op; ``start="solve"`` begins from the
    Gauss-Jordan solution like the Cython batch_bvls_bpp.  ``Ainv``: inv(A)
    on the device when the caller already has it (saves a host round trip)."""
    dev = B.device
    M, K = B.shape
    if K > 64:
        raise RuntimeError(f"K={K} exceeds MAX_K=64")
    if tuple(A.shape) != (K, K):
        raise RuntimeError("A must be K x K")
    A = _f64c(A, dev)
    B = _f64c(B, dev)
    if start != "inverse":
        Ainv = None
    elif Ainv is None:
        Ainv = _inv_host(A)
    else:
        Ainv = _f64c(Ainv, dev)
    X = torch.empty((M, K), dtype=torch.float64, device=dev)
    if Ainv is not None:
        nb = _lib.load().adx_bvls_workspace_bytes(K)
        ws = _bvls_workspace(dev, nb)
        _lib.call("adx_bvls_batch", A.data_ptr(), Ainv.data_ptr(), B.data_ptr(), X.data_ptr(), M, K, lower, upper,
                  ws.data_ptr(), nb, _lib.stream_handle(dev))
        return X
    q = _queue(dev)
    _lib.call("adx_bvls_batch", A.data_ptr(), None, B.data_ptr(), X.data_ptr(), M, K, lower, upper,
              q.data_ptr(), 8, _lib.stream_handle(dev))
    return X


def _inv_host(A: torch.Tensor) -> torch.Tensor:
    """inv(A) of the K x K matrix by LU with partial pivoting on the host (the
    matrix is tiny; a device LU would load the solver library on first use).
    A singular matrix raises RuntimeError, as torch.linalg.inv does."""
    try:
        inv = np.linalg.inv(A.cpu().numpy())
    except np.linalg.LinAlgError as exc:
        raise RuntimeError(f"batch_bvls_bpp: A is singular ({exc})") from exc
    return torch.from_numpy(inv).to(A.device).contiguous()


_BVLS_WS: dict = {}


def _bvls_workspace(dev, nbytes: int) -> torch.Tensor:
    """Per-device workspace of the shared-A BVLS path (the subset-inverse
    table); the first 8 bytes double as the work queue of the wavefront path."""
    w = _BVLS_WS.get(dev)
    if w is None or w.numel() < nbytes:
        w = _BVLS_WS[dev] = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=dev)
    return w
