"""Projection mode on the GPU: P fixed, only Q optimised (drop-in for
adamixture/src/projection.py:178-321 optimize_projection_gpu and :505-659
optimize_projection_original_gpu).

Adam-EM projection: each iteration is one fused EM pass at (P, Q) (the A / B
statistics of the P side come out of the same kernel and are ignored), the Q
half of the EM finalisation (adx_em_finalize with M = 0) and the Q half of
the Adam step (adx_adam_step with M = 0: clamp to [1e-5, 1 - 1e-5] and row
renormalisation, projection.py:275-290).  As in the joint Adam-EM driver, the
log-likelihood of a check is fused into the next pass, which runs at the same
(P, Q), so the trajectory, the lr schedule and the returned best Q are the
reference's.  The reference's third priming pass (:300) changes no state and
is not run.

SQP projection: per iteration two Q sweeps (adx_grad_hess on the Q side and
the batched wavefront SQP solver), ZAL extrapolation over the flattened N K
vector, clamp + row renormalisation, acceptance against the previous
iteration's FP64 log-likelihood, and the reference's stopping rules
(:576-656).
"""
from __future__ import annotations

import logging
import time

import torch

from .als_gpu import _nn, _tn
from .br_qn_gpu import _loglik, grad_hess
from .em_adam_gpu import EMAdamOptimizer, em_pass
from .genotypes import as_packed
from .solvers import sqp_solve_q, v_kk

log = logging.getLogger(__name__)


def _as_param(X: torch.Tensor, dev: torch.device) -> torch.Tensor:
    """X itself when it is already a contiguous FP64 tensor on ``dev`` (so Q is
    updated in place, as in the reference), else a converted copy."""
    return X.to(device=dev, dtype=torch.float64).contiguous()


def optimize_projection_gpu(G, P, Q, lr, beta1, beta2, reg_adam, max_iter, check, M, lr_decay, min_lr,
                            patience, tol_adam, device, chunk_size, threads_per_block, precision=None,
                            on_check=None) -> torch.Tensor:
    """projection.py:178-321.  ``precision`` selects the EM-pass form
    ("f64" default, "f32", "bf16"; Q, the Adam state and the log-likelihood
    stay FP64); ``on_check(it, logl)`` is called at every evaluation
    (it = 0 is the post-priming one)."""
    del chunk_size, threads_per_block
    dev = torch.device(device)
    P = _as_param(P, dev)
    Q = _as_param(Q, dev)
    G = as_packed(G, M, dev)
    opt = EMAdamOptimizer(P.shape, Q.shape, lr, beta1, beta2, reg_adam, dev, precision)

    ts = time.time()
    opt.accumulate(G, P, Q)
    opt.step_q(Q, opt.finalize_q(G, Q))
    log.info(f"    Priming done. ({time.time() - ts:.1f}s)\n")

    Q_best = Q.clone()
    L_best = None
    wait_lr = 0
    pending = 0            # iteration whose logL the next pass evaluates (0 = post-priming)

    def decide(it_done: int, L_cur: float) -> bool:
        """projection.py:303-319; True -> stop."""
        nonlocal L_best, wait_lr
        if on_check is not None:
            on_check(it_done, L_cur)
        if L_best is None:
            L_best = L_cur
            return False
        log.info(f"    Iteration {it_done}, Log-likelihood: {L_cur:.1f}")
        if L_cur > L_best + tol_adam:
            L_best = L_cur
            Q_best.copy_(Q)
            wait_lr = 0
        else:
            wait_lr += 1
            if wait_lr >= patience:
                old = opt.lr
                opt.lr = max(opt.lr * lr_decay, min_lr)
                log.info(f"    Plateau ({wait_lr} checks). Reducing lr: {old:.3e} → {opt.lr:.3e}")
                if opt.lr <= min_lr:
                    log.info("    Convergence reached.")
                    return True
                wait_lr = 0
        return False

    stopped = False
    for it in range(max_iter):
        ll = opt.accumulate(G, P, Q, loglik=pending is not None)
        if pending is not None:
            if decide(pending, ll.item()):
                stopped = True
                break
            pending = None
        opt.step_q(Q, opt.finalize_q(G, Q))
        if (it + 1) % check == 0:
            pending = it + 1
    if not stopped and pending is not None:
        decide(pending, em_pass(G, P, Q, loglik=True, stats=False).item())
    log.info(f"\n    Final log-likelihood (projection): {L_best:.1f}")
    return Q_best


def optimize_projection_original_gpu(G, P, Q, max_iter, K, M, N, tol, Q_hist, patience, device, chunk_size,
                                     threads_per_block, on_iter=None) -> torch.Tensor:
    """projection.py:505-659.  ``on_iter(it, logl)`` is called every iteration."""
    del chunk_size, threads_per_block
    dev = torch.device(device)
    P = _as_param(P, dev)
    Q = _as_param(Q, dev)
    G = as_packed(G, M, dev)
    vkk = v_kk(K, dev)
    dim = N * K
    U = torch.zeros((dim, Q_hist), dtype=torch.float64, device=dev)
    V = torch.zeros((dim, Q_hist), dtype=torch.float64, device=dev)
    ll_prev = -float("inf")
    ll_best = -float("inf")
    wait = 0
    Q_best = torch.empty_like(Q)

    def sweep(Qc: torch.Tensor) -> torch.Tensor:
        H, z = grad_hess(G, Qc, P, "q")
        return sqp_solve_q(H, z.neg_(), Qc, vkk, N, K)

    for it in range(1, max_iter + 1):
        t0 = time.time()
        Q1 = sweep(Q)
        Q2 = sweep(Q1)
        x, x1, x2 = Q.reshape(-1), Q1.reshape(-1), Q2.reshape(-1)
        col = (it - 1) % Q_hist
        U[:, col] = x1 - x
        V[:, col] = x2 - x1
        nc = min(it, Q_hist)
        Us, Vs = U[:, :nc], V[:, :nc]
        # skinny products (dim x nc, nc <= Q_hist) in libadx: the library
        # GEMMs take ~0.1 s per call at these shapes
        lhs = _tn(Us, Us - Vs)
        rhs = _tn(Us, (x - x1).unsqueeze(1))
        try:
            alpha = torch.linalg.solve(lhs, rhs)
        except RuntimeError:
            alpha = torch.linalg.lstsq(lhs, rhs).solution
        Q_qn = (x1 - _nn(Vs, alpha).squeeze(1)).view(N, K).clone()
        torch.clamp_(Q_qn, 1e-5, 1.0 - 1e-5)
        Q_qn.div_(Q_qn.sum(dim=1, keepdim=True))
        ll_qn = _loglik(G, P, Q_qn)
        if ll_qn > ll_prev:
            Q.copy_(Q_qn)
            ll_new = ll_qn
        else:
            Q.copy_(Q2)
            ll_new = _loglik(G, P, Q2)
        best_before = ll_best
        if ll_new > ll_best:
            ll_best = ll_new
            Q_best.copy_(Q)
        if best_before != -float("inf") and ll_new <= best_before + tol:
            wait += 1
        else:
            wait = 0
        if on_iter is not None:
            on_iter(it, ll_new)
        log.info(f"    Iteration {it}, Log-likelihood: {ll_new:.1f}, Time: {time.time() - t0:.3f}s")
        if abs(ll_new - ll_prev) < tol:
            log.info(f"    Converged at iteration {it}.")
            break
        if wait >= patience:
            log.info(f"    Converged at iteration {it} after {wait} iterations without improvement.")
            break
        ll_prev = ll_new
    log.info(f"\n    Final log-likelihood (projection): {ll_best:.1f}")
    return Q_best
