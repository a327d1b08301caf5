"""ALS warm start on the GPU (drop-in for adamixture/model/als_gpu.py).

Small dense algebra (K x K Gram matrices, pseudo-inverses, skinny products)
runs in torch-ROCm in FP64; the per-row bounded least squares run in the
libadx batched BVLS kernel (one wavefront per row, adx_bvls_batch).

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

from .solvers import batch_bvls_bpp

log = logging.getLogger(__name__)


def mapP(P: torch.Tensor) -> None:
    """als_gpu.py:10-21."""
    P.clamp_(min=1e-5, max=1.0 - 1e-5)


def mapQ(Q: torch.Tensor) -> None:
    """als_gpu.py:23-36."""
    Q.clamp_(min=1e-5, max=1.0 - 1e-5)
    Q.div_(Q.sum(dim=1, keepdim=True))


def ALS_gpu(U: torch.Tensor, S: torch.Tensor, V: torch.Tensor, f: torch.Tensor, seed: int, M: int, K: int,
            max_iter: int, tol: float, device, *, bvls_start: str = "inverse") -> tuple[torch.Tensor, torch.Tensor]:
    """als_gpu.py:45-127.  ``bvls_start="solve"`` uses the Cython BVLS start
    (bvls.pyx) instead of the CUDA op's X0 = B inv(A)^T."""
    t0 = time.time()
    dev = torch.device(device)
    dt = torch.float64
    Z = (U.to(dev, dt) * S.to(dev, dt))
    V = V.to(dev, dt)
    f_w = f.to(dev, dt)
    P = torch.from_numpy(np.random.default_rng(seed).random(size=(M, K), dtype=np.float64)).to(dev)
    mapP(P)

    A_p = P.T @ P
    I_p = P @ torch.linalg.pinv(A_p)
    Q = 0.5 * (V @ (Z.T @ I_p)) + (f_w.reshape(1, -1) @ I_p)
    mapQ(Q)
    Q0 = Q.clone()
    for i in range(max_iter):
        A_q = Q.T @ Q
        I_q = Q @ torch.linalg.pinv(A_q)
        P_free = 0.5 * (Z @ (V.T @ I_q)) + torch.outer(f_w, I_q.sum(dim=0))
        P = batch_bvls_bpp(A_q, P_free @ A_q, 0.0, 1.0, start=bvls_start)
        mapP(P)
        A_p = P.T @ P
        I_p = P @ torch.linalg.pinv(A_p)
        Q_free = 0.5 * (V @ (Z.T @ I_p)) + (f_w.reshape(1, -1) @ I_p)
        Q = batch_bvls_bpp(A_p, Q_free @ A_p, 0.0, 1.0, start=bvls_start)
        mapQ(Q)
        rmse = torch.sqrt(torch.mean((Q - Q0) ** 2))
        if rmse < tol:
            log.info(f"        Convergence reached in iteration {i + 1}.")
            break
        Q0.copy_(Q)
    log.info(f"        Total ALS time={time.time() - t0:.3f}s\n")
    return P, Q
