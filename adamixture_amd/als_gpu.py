"""ALS warm start on the GPU (drop-in for adamixture/model/als_gpu.py).

All arithmetic is FP64.  The skinny products (Gram matrices and products
with the N x L / M x L SVD factors) run in libadx (adx_gemm_tn_f64,
adx_gemm_nn_f64), the K x K pseudo-inverses on the host, the per-row
bounded least squares in the batched BVLS kernel (one wavefront per row,
adx_bvls_batch).

The initial P is drawn with ``numpy.random.default_rng(seed).random((M, K))``
as the CPU reference does (als.py:33-40), which makes the GPU result
reproduce the CPU reference's numbers; the reference GPU path draws from
torch's device generator (als_gpu.py:69-86).
"""
from __future__ import annotations

import logging
import time

import numpy as np
import torch

from . import _lib
from .solvers import batch_bvls_bpp

log = logging.getLogger(__name__)


def mapP(P: torch.Tensor) -> None:
    """als_gpu.py:10-21."""
    P.clamp_(min=1e-5, max=1.0 - 1e-5)


def mapQ(Q: torch.Tensor) -> None:
    """als_gpu.py:23-36."""
    Q.clamp_(min=1e-5, max=1.0 - 1e-5)
    Q.div_(Q.sum(dim=1, keepdim=True))


def _host_inverses(a: np.ndarray, dev, with_inv: bool):
    """pinv (torch.linalg.pinv's cut-off) and, if asked, inv of a host K x K
    matrix, both moved to the device: one host copy of the Gram matrix serves
    the pseudo-inverse of the ALS update and the BVLS start X0 = B inv(A)^T."""
    pinv = torch.from_numpy(np.linalg.pinv(a, rcond=max(a.shape) * np.finfo(np.float64).eps)).to(dev)
    if not with_inv:
        return pinv, None
    try:
        inv = np.linalg.inv(a)
    except np.linalg.LinAlgError as exc:
        raise RuntimeError(f"batch_bvls_bpp: A is singular ({exc})") from exc
    return pinv, torch.from_numpy(inv).to(dev).contiguous()


def _pinv_host(A: torch.Tensor) -> torch.Tensor:
    """torch.linalg.pinv's cut-off (rtol = max(m, n) eps) on the host: the
    K x K matrix is tiny, and an LAPACK SVD on the CPU is ~10x faster than a
    device SVD launch sequence."""
    a = A.cpu().numpy()
    return torch.from_numpy(np.linalg.pinv(a, rcond=max(a.shape) * np.finfo(np.float64).eps)).to(A.device)


_TN_WORK: dict = {}


def _tn_work(dev: torch.device, a: int, b: int) -> torch.Tensor:
    need = _lib.ADX_GEMM_TN_BLOCKS * a * b
    w = _TN_WORK.get(dev)
    if w is None or w.numel() < need:
        w = _TN_WORK[dev] = torch.empty(need, dtype=torch.float64, device=dev)
    return w


def _tn(X: torch.Tensor, Y: torch.Tensor, alpha: float = 1.0) -> torch.Tensor:
    """alpha X^T Y for tall X (n x a), Y (n x b) (adx_gemm_tn_f64, with a
    workspace: fixed summation order, so ALS runs are reproducible)."""
    n, a = X.shape
    b = Y.shape[1]
    bc = min(64, 1024 // a)          # the kernel takes a, b <= 64 and a b <= 1024
    if b > bc:                       # column blocks of Y (strided views, no copies)
        return torch.cat([_tn(X, Y[:, c:c + bc], alpha) for c in range(0, b, bc)], dim=1)
    C = torch.empty((a, b), dtype=torch.float64, device=X.device)
    _lib.call("adx_gemm_tn_f64", X.data_ptr(), X.stride(0), Y.data_ptr(), Y.stride(0), n, a, b, alpha,
              C.data_ptr(), b, _tn_work(X.device, a, b).data_ptr(), _lib.stream_handle(X.device))
    return C


def _nn(X: torch.Tensor, W: torch.Tensor, alpha: float = 1.0, u: torch.Tensor | None = None,
        v: torch.Tensor | None = None) -> torch.Tensor:
    """alpha X W + u v^T for tall X (n x a) and small W (a x b) (adx_gemm_nn_f64)."""
    n, a = X.shape
    b = W.shape[1]
    W = W.contiguous()
    v = None if v is None else v.contiguous()
    C = torch.empty((n, b), dtype=torch.float64, device=X.device)
    for c in range(0, b, 16):        # the kernel takes b <= 16: column blocks of W, v and C
        bc = min(16, b - c)
        _lib.call("adx_gemm_nn_f64", X.data_ptr(), X.stride(0), n, a, bc, W[:, c:].data_ptr(), b, alpha,
                  _lib.ptr(u), _lib.ptr(None if v is None else v[c:]), 0.0, C[:, c:].data_ptr(), b,
                  _lib.stream_handle(X.device))
    return C


def ALS_gpu(U: torch.Tensor, S: torch.Tensor, V: torch.Tensor, f: torch.Tensor, seed: int, M: int, K: int,
            max_iter: int, tol: float, device, *, bvls_start: str = "inverse",
            P_init: torch.Tensor | None = None,
            Q_init: torch.Tensor | None = None) -> tuple[torch.Tensor, torch.Tensor]:
    """als_gpu.py:45-127.  ``bvls_start="solve"`` uses the Cython BVLS start
    (bvls.pyx) instead of the CUDA op's X0 = B inv(A)^T.  ``P_init`` and
    ``Q_init`` replace the seeded random P and its Q (als_gpu.py:85-96):
    passing the (P, Q) returned by an earlier call continues that call's
    iterations exactly.

    The same updates as the reference, evaluated in a cheaper order: with
    Z = U S, the P target (0.5 Z (V^T I_q) + f 1^T I_q) A_q is formed as
    Z (0.5 (V^T I_q) A_q) + f (1^T I_q A_q) by one pass over Z, and likewise
    for Q; the skinny products run in libadx (adx_gemm_tn_f64 / _nn_f64)."""
    t0 = time.time()
    dev = torch.device(device)
    dt = torch.float64
    Z = (U.to(dev, dt) * S.to(dev, dt)).contiguous()
    V = V.to(dev, dt).contiguous()
    f_w = f.to(dev, dt).contiguous()
    N = V.shape[0]
    ones_n = torch.ones(N, dtype=dt, device=dev)
    f_col = f_w.reshape(-1, 1)
    ones_col = ones_n.reshape(-1, 1)

    def q_side(P):
        """(A_p, 0.5 V (Z^T I_p) + 1 (f^T I_p)) without the final product by A_p."""
        A_p = _tn(P, P)
        I_p = _nn(P, _pinv_host(A_p))
        return A_p, _tn(Z, I_p), _tn(f_col, I_p)[0]

    if P_init is not None and Q_init is not None:
        P = P_init.to(dev, dt).contiguous().clone()
        Q = Q_init.to(dev, dt).contiguous().clone()
    else:
        P = torch.from_numpy(np.random.default_rng(seed).random(size=(M, K), dtype=np.float64)).to(dev)
        mapP(P)
        _, ZtI, fI = q_side(P)
        Q = _nn(V, ZtI, 0.5, ones_n, fI)
        mapQ(Q)
    Q0 = Q.clone()
    inv_start = bvls_start == "inverse"
    # Each half-step needs its K x K Gram matrix on the host (pinv, and inv
    # for the BVLS start).  The convergence value of an iteration travels
    # with the next iteration's A_q = Q^T Q, computed before the test (it
    # depends only on Q), so an iteration costs two host round trips.
    A_q = _tn(Q, Q)
    aq = A_q.cpu().numpy()
    for i in range(max_iter):
        pinv_q, inv_q = _host_inverses(aq, dev, inv_start)
        I_q = _nn(Q, pinv_q)
        VtI = _tn(V, I_q)
        s = _tn(ones_col, I_q)[0]
        P = batch_bvls_bpp(A_q, _nn(Z, 0.5 * (VtI @ A_q), 1.0, f_w, s @ A_q), 0.0, 1.0, start=bvls_start,
                           Ainv=inv_q)
        mapP(P)
        A_p = _tn(P, P)
        pinv_p, inv_p = _host_inverses(A_p.cpu().numpy(), dev, inv_start)
        I_p = _nn(P, pinv_p)
        ZtI, fI = _tn(Z, I_p), _tn(f_col, I_p)[0]
        Q = batch_bvls_bpp(A_p, _nn(V, 0.5 * (ZtI @ A_p), 1.0, ones_n, fI @ A_p), 0.0, 1.0, start=bvls_start,
                           Ainv=inv_p)
        mapQ(Q)
        rmse = torch.sqrt(torch.mean((Q - Q0) ** 2))
        A_q = _tn(Q, Q)
        both = torch.cat([rmse.reshape(1), A_q.reshape(-1)]).cpu().numpy()
        aq = both[1:].reshape(K, K)
        if both[0] < tol:
            log.info(f"        Convergence reached in iteration {i + 1}.")
            break
        Q0.copy_(Q)
    log.info(f"        Total ALS time={time.time() - t0:.3f}s\n")
    return P, Q
