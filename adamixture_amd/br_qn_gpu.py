"""SQP block relaxation with ZAL quasi-Newton acceleration on the GPU
(drop-in for adamixture/model/br_qn_gpu.py).

Per iteration: two SQP sweeps (Q then P, twice), each a libadx gradient /
Hessian pass (adx_grad_hess) followed by the batched wavefront SQP solver
(adx_sqp_solve); the ZAL extrapolation over the flattened (M + N) K vector
and the q x q solve stay in torch, as in the reference; the candidate is
accepted by its FP64 log-likelihood (one fused libadx pass).
"""
from __future__ import annotations

import logging
import time

import torch

from . import _lib
from .als_gpu import _nn, _tn
from .em_adam_gpu import em_pass
from .genotypes import PackedGenotypes, as_packed, is_streamed
from .solvers import sqp_solve_p, sqp_solve_q, v_kk

log = logging.getLogger(__name__)


def grad_hess(G: PackedGenotypes, Q: torch.Tensor, P: torch.Tensor, side: str):
    """(XtX, Xtz) of compute_grad_hess_{Q,P}_gpu (br_qn_gpu.py:145-214),
    positive gradient convention."""
    M, K = P.shape
    N = Q.shape[0]
    rows = N if side == "q" else M
    if is_streamed(G):   # host-resident genotypes, band by band
        if side == "q":   # reduction over SNPs: band sums added
            XtX = torch.zeros((N, K, K), dtype=torch.float64, device=G.device)
            Xtz = torch.zeros((N, K), dtype=torch.float64, device=G.device)
            for m0, Mb, band in G.bands():
                h, z = grad_hess(band, Q, P[m0:m0 + Mb], side)
                XtX += h
                Xtz += z
            return XtX, Xtz
        XtX = torch.empty((M, K, K), dtype=torch.float64, device=G.device)
        Xtz = torch.empty((M, K), dtype=torch.float64, device=G.device)
        for m0, Mb, band in G.bands():   # per SNP: the band's rows
            XtX[m0:m0 + Mb], Xtz[m0:m0 + Mb] = grad_hess(band, Q, P[m0:m0 + Mb], side)
        return XtX, Xtz
    XtX = torch.empty((rows, K, K), dtype=torch.float64, device=G.device)
    Xtz = torch.empty((rows, K), dtype=torch.float64, device=G.device)
    _lib.call("adx_grad_hess", 0 if side == "q" else 1, G.data.data_ptr(), G.ldg, G.M, G.N, K,
              Q.contiguous().data_ptr(), P.contiguous().data_ptr(), XtX.data_ptr(), Xtz.data_ptr(),
              _lib.stream_handle(G.device))
    return XtX, Xtz


def compute_grad_hess_Q_gpu(G, Q, P, XtX_q, Xtz_q, M, chunk_size=0, unpacker=None, dtype=torch.float64):
    """br_qn_gpu.py:145-179 (fills the given buffers)."""
    H, z = grad_hess(as_packed(G, M, Q.device), Q, P, "q")
    XtX_q.copy_(H)
    Xtz_q.copy_(z)


def compute_grad_hess_P_gpu(G, Q, P, XtX_p, Xtz_p, M, chunk_size=0, unpacker=None, dtype=torch.float64):
    """br_qn_gpu.py:182-214 (fills the given buffers)."""
    H, z = grad_hess(as_packed(G, M, Q.device), Q, P, "p")
    XtX_p.copy_(H)
    Xtz_p.copy_(z)


def _mapPQ_gpu(P: torch.Tensor, Q: torch.Tensor) -> None:
    """br_qn_gpu.py:47-61."""
    torch.clamp_(P, 1e-5, 1.0 - 1e-5)
    torch.clamp_(Q, 1e-5, 1.0 - 1e-5)
    Q.div_(Q.sum(dim=1, keepdim=True))


def _loglik(G: PackedGenotypes, P, Q) -> float:
    return em_pass(G, P, Q, loglik=True, stats=False).item()


def _map_cpu(P: torch.Tensor, Q: torch.Tensor) -> None:
    """br_qn.py:159-160 (the CPU optimizer's quasi-Newton map): P clipped to
    the box, each Q row projected onto the simplex with pseudocount 1e-5."""
    torch.clamp_(P, 1e-5, 1.0 - 1e-5)
    _lib.call("adx_project_simplex", Q.data_ptr(), Q.shape[0], Q.shape[1], 1e-5, _lib.stream_handle(Q.device))


def _optimize_original_gpu_once(G, P, Q, max_iter, K, M, N, tol, Q_hist, patience, device, chunk_size,
                                threads_per_block, on_iter=None, snap=None, qn_map="gpu"):
    """br_qn_gpu.py:217-382.  ``snap(Q)``, if given, is applied to every Q
    the iteration produces -- after each SQP Q solve and after the
    quasi-Newton map -- as the supervised mode does (supervised.py:689, 697,
    741).  ``qn_map="cpu"`` maps the quasi-Newton point as the CPU optimizer
    does (br_qn.py:159-160), which the cross-validation polish runs."""
    snap = snap or (lambda X: None)
    map_pq = _mapPQ_gpu if qn_map == "gpu" else _map_cpu
    dev = torch.device(device)
    G = as_packed(G, M, dev)
    vkk = v_kk(K, dev)
    dim = M * K + N * K
    U = torch.zeros((dim, Q_hist), dtype=torch.float64, device=dev)
    V = torch.zeros((dim, Q_hist), dtype=torch.float64, device=dev)
    ll_prev = -float("inf")
    ll_best = -float("inf")
    wait = 0
    P_best = torch.empty_like(P)
    Q_best = torch.empty_like(Q)

    def sweep(Qc, Pc):
        Hq, zq = grad_hess(G, Qc, Pc, "q")
        Qn = sqp_solve_q(Hq, zq.neg_(), Qc, vkk, N, K)
        snap(Qn)
        Hp, zp = grad_hess(G, Qn, Pc, "p")
        Pn = sqp_solve_p(Hp, zp.neg_(), Pc, M, K)
        return Pn, Qn

    for it in range(1, max_iter + 1):
        t0 = time.time()
        P1, Q1 = sweep(Q, P)
        P2, Q2 = sweep(Q1, P1)
        x = torch.cat([P.reshape(-1), Q.reshape(-1)])
        x1 = torch.cat([P1.reshape(-1), Q1.reshape(-1)])
        x2 = torch.cat([P2.reshape(-1), Q2.reshape(-1)])
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
        x_qn = x1 - _nn(Vs, alpha).squeeze(1)
        P_qn = x_qn[:M * K].view(M, K).clone()
        Q_qn = x_qn[M * K:].view(N, K).clone()
        map_pq(P_qn, Q_qn)
        snap(Q_qn)
        ll_qn = _loglik(G, P_qn, Q_qn)
        if ll_qn > ll_prev:
            P.copy_(P_qn)
            Q.copy_(Q_qn)
            ll_new = ll_qn
        else:
            P.copy_(P2)
            Q.copy_(Q2)
            ll_new = _loglik(G, P2, Q2)
        best_before = ll_best
        if ll_new > ll_best:
            ll_best = ll_new
            P_best.copy_(P)
            Q_best.copy_(Q)
        if best_before != -float("inf") and ll_new <= best_before + tol:
            wait += 1
        else:
            wait = 0
        if on_iter is not None:
            on_iter(it, ll_new)
        log.info(f"    Iteration {it}, Log-likelihood: {ll_new:.1f}, Time: {time.time() - t0:.3f}s")
        if abs(ll_new - ll_prev) < tol:
            break
        if wait >= patience:
            break
        ll_prev = ll_new
    log.info(f"\n    Final log-likelihood: {ll_best:.1f}")
    return P_best, Q_best


def optimize_original_gpu(G, P, Q, max_iter, K, M, N, tol, Q_hist, patience, device, chunk_size,
                          threads_per_block, on_iter=None, snap=None):
    """br_qn_gpu.py:385-411 with the reference's OOM retry."""
    from .utils import MIN_GPU_CHUNK_SIZE, is_gpu_oom, reduce_gpu_chunk_or_raise
    cur = chunk_size
    P_attempt = torch.empty_like(P)
    Q_attempt = torch.empty_like(Q)
    while True:
        P_attempt.copy_(P)
        Q_attempt.copy_(Q)
        try:
            return _optimize_original_gpu_once(G, P_attempt, Q_attempt, max_iter, K, M, N, tol, Q_hist, patience,
                                               device, cur, threads_per_block, on_iter, snap)
        except RuntimeError as exc:
            if not is_gpu_oom(exc) or cur <= MIN_GPU_CHUNK_SIZE:
                raise
            cur = reduce_gpu_chunk_or_raise(cur, exc, "running SQP + ZAL QN")
