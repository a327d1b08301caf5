"""Adam-EM on the GPU through libadx (drop-in for adamixture/model/em_adam_gpu.py).

Signatures follow the reference module; the chunk/threads arguments are
accepted for compatibility but the work is one fused kernel launch per EM
pass (decode + H = P Q^T + G/H terms + A/B/T reductions), so no genotype
chunk is ever materialised.

Precision forms (``precision``):
  * ``"f64"`` (default): the reference's arithmetic (utils.py:235-250 returns
    float64 on CUDA); H and the A/B/T contractions on FP64 MFMA.
  * ``"f32"``: H and the contractions on FP32 MFMA from FP32 copies of P and
    Q; P, Q, the Adam state and the log-likelihood stay FP64.
  * ``"bf16"``: H (three-term hi/lo split) on BF16 MFMA, the EM weights in
    FP16 and the A/B/T contractions on FP16 MFMA, FP32 accumulation; P, Q,
    Adam state and log-likelihood FP64.
"""
from __future__ import annotations

import logging
import os
import time

import torch

from . import _lib
from .genotypes import PackedGenotypes, as_packed, is_streamed

log = logging.getLogger(__name__)

PRECISIONS = ("f64", "f32", "bf16")


def default_precision() -> str:
    p = os.environ.get("ADX_PRECISION", "f64")
    if p not in PRECISIONS:
        raise ValueError(f"ADX_PRECISION must be one of {PRECISIONS}")
    return p


def _check_params(P: torch.Tensor, Q: torch.Tensor) -> None:
    if P.dtype != torch.float64 or Q.dtype != torch.float64:
        raise TypeError("P and Q must be float64 tensors")
    if not (P.is_cuda and Q.is_cuda and P.is_contiguous() and Q.is_contiguous()):
        raise ValueError("P and Q must be contiguous CUDA tensors")
    if P.shape[1] != Q.shape[1]:
        raise ValueError("P and Q must have the same number of columns K")


class _Workspace:
    """Per-device scratch of the F32/F64 pass: log-likelihood partial sums
    and the T / A / B partial slices.  The pass needs the single-band size
    (adx_em_workspace_bytes) to run in one band; given less it runs in SNP
    bands whose slices fit.  The buffer is capped at ``cap(G)``: the larger
    of a tenth of the packed genotype matrix and ``min_cap`` bytes."""

    def __init__(self):
        self.buf: torch.Tensor | None = None
        self.ll = None
        self.min_cap = 256 << 20
        self.frac = 0.1

    def cap(self, G) -> int:
        return max(int(self.frac * G.host.numel() if is_streamed(G) else self.frac * G.data.numel()), self.min_cap)

    def shrink(self) -> bool:
        """Halve the cap after an out-of-memory error (False at the floor)."""
        if self.min_cap <= (32 << 20) and self.frac <= 1e-3:
            return False
        self.min_cap = max(self.min_cap // 2, 32 << 20)
        self.frac = max(self.frac / 2, 1e-3)
        self.buf = None
        return True

    def get(self, device, nbytes: int) -> tuple[torch.Tensor, torch.Tensor]:
        if self.buf is None or self.buf.numel() < nbytes or self.buf.device != device:
            self.buf = None                      # release before growing (the T slices can be GBs)
            self.buf = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=device)
        if self.ll is None or self.ll.device != device:
            self.ll = torch.zeros(1, dtype=torch.float64, device=device)
        return self.buf, self.ll


_WS = _Workspace()


def em_pass(G: PackedGenotypes, P, Q, A=None, B=None, T=None, *, dtype: int = _lib.ADX_F64,
            loglik: bool = False, stats: bool = True):
    """One fused pass.  Returns the log-likelihood tensor (1 element, FP64,
    on device) when ``loglik`` is set, else None."""
    K = P.shape[1]
    flags = (_lib.ADX_EM_STATS if stats else 0) | (_lib.ADX_EM_LOGLIK if loglik else 0)
    # K = 0 sizes only the log-likelihood partials (no T slices without stats)
    ws_bytes = min(_lib.load().adx_em_workspace_bytes(dtype, G.M, G.N, K if stats else 0), _WS.cap(G))
    ws, ll = _WS.get(G.device, ws_bytes)
    head = (dtype,)
    tail = (G.M, G.N, K, P.data_ptr(), P.stride(0), Q.data_ptr(), Q.stride(0), _lib.ptr(A), _lib.ptr(B),
            _lib.ptr(T), flags, _lib.ptr(ll) if loglik else None, ws.data_ptr(), ws_bytes)
    if is_streamed(G):
        _em_pass_streamed(G, head + (G.host.data_ptr(), G.ldg) + tail, dtype, K, flags, ws_bytes)
    else:
        _lib.call("adx_em_pass", *head, G.data.data_ptr(), G.ldg, *tail, _lib.stream_handle(G.device))
    return ll if loglik else None


def _em_pass_streamed(G, args, dtype: int, K: int, flags: int, ws_bytes: int) -> None:
    """adx_em_pass_host on a host-resident matrix: the staging holds the
    plan's band twice when G.staging_cap allows (then the result is
    bit-identical to the resident pass with the same workspace), else the
    bands shrink to what it holds."""
    import torch as _t
    band = _lib.load().adx_em_band_snps(dtype, G.M, G.N, K, flags, ws_bytes, 0)
    half = -(-((band + 3) // 4) * G.ldg // 256) * 256 if band > 0 else 0
    want = min(2 * half, max(G.staging_cap, 2 * 256 * ((G.ldg + 255) // 256)))
    buf = G.staging(want)
    if G.copy_stream is None:
        G.copy_stream = _t.cuda.Stream(device=G.device)
    _lib.call("adx_em_pass_host", *args, buf.data_ptr(), buf.numel(), _lib.stream_handle(G.device),
              G.copy_stream.cuda_stream)


class EMAdamOptimizer:
    """em_adam_gpu.py:121-224: Adam state, EM accumulators and targets."""

    def __init__(self, P_shape, Q_shape, lr: float, beta1: float, beta2: float, reg_adam: float,
                 device: torch.device, precision: str | None = None) -> None:
        self.device = torch.device(device)
        self.lr, self.beta1, self.beta2, self.reg_adam = lr, beta1, beta2, reg_adam
        self.precision = precision or default_precision()
        if self.precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        self.dtype = torch.float64
        self.t = 0.0
        M, K = P_shape
        N = Q_shape[0]
        z = lambda *s, dt=torch.float64: torch.zeros(*s, dtype=dt, device=self.device)  # noqa: E731
        self.m_P, self.v_P, self.m_Q, self.v_Q = z(M, K), z(M, K), z(N, K), z(N, K)
        low = self.precision in ("f32", "bf16")
        acc_dt = torch.float32 if low else torch.float64
        self.acc_dtype = _lib.ADX_F32 if low else _lib.ADX_F64
        self.A_accum, self.B_accum = z(M, K, dt=acc_dt), z(M, K, dt=acc_dt)
        self.T_accum = z(N, K, dt=acc_dt)
        self.P_EM, self.Q_EM = z(M, K), z(N, K)
        if self.precision == "f32":
            self.P_lo = torch.zeros(M, K, dtype=torch.float32, device=self.device)
            self.Q_lo = torch.zeros(N, K, dtype=torch.float32, device=self.device)
        if self.precision == "bf16":
            nbytes = _lib.load().adx_em_bf16_operand_bytes(M, N, K, None)
            self.operands = torch.zeros(nbytes, dtype=torch.uint8, device=self.device)

    def accumulate(self, G: PackedGenotypes, P: torch.Tensor, Q: torch.Tensor, loglik: bool = False):
        """EM statistics at (P, Q) into A/B/T; optionally the FP64 logL at the
        same point (fused into the same pass in the f64 form)."""
        if self.precision == "f64":
            if loglik and not is_streamed(G):
                # Two passes over the resident G: the fused statistics +
                # log-likelihood kernel needs more than 256 registers per lane
                # (one wave per SIMD) and takes longer than both (at 100k x
                # 500k, K=10: 233 ms against 150 + 34 ms).  Host-streamed G
                # keeps the fused pass (one copy of each band).
                em_pass(G, P, Q, self.A_accum, self.B_accum, self.T_accum, dtype=_lib.ADX_F64)
                return em_pass(G, P, Q, dtype=_lib.ADX_F64, loglik=True, stats=False)
            return em_pass(G, P, Q, self.A_accum, self.B_accum, self.T_accum, dtype=_lib.ADX_F64, loglik=loglik)
        if self.precision == "bf16":
            if is_streamed(G):
                raise NotImplementedError("the mixed-precision (bf16) pass needs the genotypes resident on the GPU; "
                                          "host-streamed genotypes run the f64 or f32 form")
            st = _lib.stream_handle(self.device)
            M, K = P.shape
            N = Q.shape[0]
            need = _lib.load().adx_em_bf16_operand_bytes(M, N, K, None)
            if self.operands.numel() < need:      # the tile shape changed since construction
                self.operands = torch.empty(need, dtype=torch.uint8, device=self.device)
            nb = self.operands.numel()
            _lib.call("adx_em_prep_bf16", P.data_ptr(), Q.data_ptr(), M, N, K, self.operands.data_ptr(), nb, st)
            _lib.call("adx_em_pass_bf16", G.data.data_ptr(), G.ldg, G.M, G.N, K, self.operands.data_ptr(), nb,
                      self.A_accum.data_ptr(), self.B_accum.data_ptr(), self.T_accum.data_ptr(), st)
            return em_pass(G, P, Q, dtype=_lib.ADX_F64, loglik=True, stats=False) if loglik else None
        self.P_lo.copy_(P)
        self.Q_lo.copy_(Q)
        em_pass(G, self.P_lo, self.Q_lo, self.A_accum, self.B_accum, self.T_accum, dtype=_lib.ADX_F32)
        return em_pass(G, P, Q, dtype=_lib.ADX_F64, loglik=True, stats=False) if loglik else None

    def finalize(self, G: PackedGenotypes, P: torch.Tensor, Q: torch.Tensor):
        _lib.call("adx_em_finalize", self.acc_dtype, P.data_ptr(), Q.data_ptr(), self.A_accum.data_ptr(),
                  self.B_accum.data_ptr(), self.T_accum.data_ptr(), G.qbat().data_ptr(), P.shape[0],
                  Q.shape[0], P.shape[1], self.P_EM.data_ptr(), self.Q_EM.data_ptr(),
                  _lib.stream_handle(self.device))
        return self.P_EM, self.Q_EM

    def finalize_q(self, G: PackedGenotypes, Q: torch.Tensor) -> torch.Tensor:
        """Q half of the EM finalisation only (projection.py:271-272 discards P_EM)."""
        _lib.call("adx_em_finalize", self.acc_dtype, None, Q.data_ptr(), None, None, self.T_accum.data_ptr(),
                  G.qbat().data_ptr(), 0, Q.shape[0], Q.shape[1], None, self.Q_EM.data_ptr(),
                  _lib.stream_handle(self.device))
        return self.Q_EM

    def run_em_step(self, G, P: torch.Tensor, Q: torch.Tensor, M: int, chunk_size: int = 0,
                    unpacker=None) -> tuple[torch.Tensor, torch.Tensor]:
        """em_adam_gpu.py:166-199: one full EM sweep -> (P_EM, Q_EM)."""
        del chunk_size, unpacker
        _check_params(P, Q)
        G = as_packed(G, M, self.device)
        self.accumulate(G, P, Q)
        return self.finalize(G, P, Q)

    def step(self, P: torch.Tensor, Q: torch.Tensor, P_target: torch.Tensor, Q_target: torch.Tensor) -> None:
        """em_adam_gpu.py:201-224: Adam toward the targets, clamp, renormalise."""
        _check_params(P, Q)
        self.t += 1.0
        _lib.call("adx_adam_step", P.data_ptr(), Q.data_ptr(), P_target.contiguous().data_ptr(),
                  Q_target.contiguous().data_ptr(), self.m_P.data_ptr(), self.v_P.data_ptr(),
                  self.m_Q.data_ptr(), self.v_Q.data_ptr(), P.shape[0], Q.shape[0], P.shape[1], self.t,
                  self.lr, self.beta1, self.beta2, self.reg_adam, None, None, 0,
                  _lib.stream_handle(self.device))

    def step_q(self, Q: torch.Tensor, Q_target: torch.Tensor) -> None:
        """projection.py:275-290: Adam on Q only, clamp, renormalise."""
        if Q.dtype != torch.float64 or not Q.is_contiguous():
            raise TypeError("Q must be a contiguous float64 tensor")
        self.t += 1.0
        _lib.call("adx_adam_step", None, Q.data_ptr(), None, Q_target.contiguous().data_ptr(), None, None,
                  self.m_Q.data_ptr(), self.v_Q.data_ptr(), 0, Q.shape[0], Q.shape[1], self.t, self.lr,
                  self.beta1, self.beta2, self.reg_adam, None, None, 0, _lib.stream_handle(self.device))


def loglikelihood_gpu(G, P: torch.Tensor, Q: torch.Tensor, M: int, N: int, batch_size: int = 0,
                      device: torch.device | None = None, threads_per_block: int = 0) -> float:
    """em_adam_gpu.py:250-279: FP64 log-likelihood, one fused pass."""
    del N, batch_size, threads_per_block
    P64 = P.to(dtype=torch.float64).contiguous()
    Q64 = Q.to(device=P64.device, dtype=torch.float64).contiguous()
    G = as_packed(G, M, device or P64.device)
    return em_pass(G, P64, Q64, loglik=True, stats=False).item()


def _optimize_parameters_gpu_once(G, P, Q, lr, beta1, beta2, reg_adam, max_iter, check, K, M, N, lr_decay,
                                  min_lr, patience, tol_adam, device, chunk_size, threads_per_block,
                                  precision=None, on_check=None, after_step=None):
    """em_adam_gpu.py:281-356 with the log-likelihood of each check fused into
    the next EM pass (which runs at the same (P, Q)): the decision taken at
    check iteration ``it`` happens before the Adam step of iteration it+1,
    so the trajectory, the lr schedule and the returned best (P, Q) are the
    reference's; at most one EM pass is computed and discarded at the end.
    ``on_check(it, logl)`` is called at every evaluation; ``after_step(P, Q)``
    after every Adam step (the supervised mode snaps labelled Q rows there)."""
    del K, N, chunk_size, threads_per_block
    device = torch.device(device)
    G = as_packed(G, M, device)
    opt = EMAdamOptimizer(P.shape, Q.shape, lr, beta1, beta2, reg_adam, device, precision)
    ts = time.time()
    opt.accumulate(G, P, Q)
    opt.finalize(G, P, Q)
    opt.step(P, Q, opt.P_EM, opt.Q_EM)
    if after_step is not None:
        after_step(P, Q)
    log.info(f"    Performed priming iteration... ({time.time() - ts:.1f}s)\n")

    P_best = P.clone()
    Q_best = Q.clone()
    L_best = None
    wait_lr = 0
    pending = 0          # iteration index whose logL the next pass evaluates (0 = priming)
    ts = time.time()

    def decide(it_done: int, L_cur: float) -> bool:
        """Checkpoint logic of em_adam_gpu.py:334-353; True -> stop."""
        nonlocal L_best, wait_lr
        if on_check is not None:
            on_check(it_done, L_cur)
        if L_best is None:            # the post-priming evaluation (:325-327)
            L_best = L_cur
            return False
        log.info(f"    Iteration {it_done}, Log-likelihood: {L_cur:.1f}")
        if L_cur > L_best + tol_adam:
            L_best = L_cur
            P_best.copy_(P)
            Q_best.copy_(Q)
            wait_lr = 0
        else:
            wait_lr += 1
            if wait_lr >= patience:
                old = opt.lr
                opt.lr = max(opt.lr * lr_decay, min_lr)
                log.info(f"    Plateau reached ({wait_lr} checks without beating best). "
                         f"Reducing lr: {old:.3e} → {opt.lr:.3e}")
                if opt.lr <= min_lr:
                    log.info("        Convergence reached.")
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
        opt.finalize(G, P, Q)
        opt.step(P, Q, opt.P_EM, opt.Q_EM)
        if after_step is not None:
            after_step(P, Q)
        if (it + 1) % check == 0:
            pending = it + 1
    if not stopped and pending is not None:
        ll = em_pass(G, P, Q, loglik=True, stats=False)
        decide(pending, ll.item())
    log.info(f"\n    Final log-likelihood: {L_best:.1f}")
    return P_best, Q_best


def optimize_parameters_gpu(G, P, Q, lr, beta1, beta2, reg_adam, max_iter, check, K, M, N, lr_decay, min_lr,
                            patience, tol_adam, device, chunk_size, threads_per_block, precision=None,
                            on_check=None, after_step=None):
    """em_adam_gpu.py:359-385: Adam-EM with the reference's OOM retry (the
    retry restarts from the unmodified P, Q; the fused pass has no chunk
    temporaries, so the halved chunk size changes nothing but is kept)."""
    from .utils import MIN_GPU_CHUNK_SIZE, is_gpu_oom, reduce_gpu_chunk_or_raise
    cur = chunk_size
    P_attempt = torch.empty_like(P)
    Q_attempt = torch.empty_like(Q)
    while True:
        P_attempt.copy_(P)
        Q_attempt.copy_(Q)
        try:
            return _optimize_parameters_gpu_once(G, P_attempt, Q_attempt, lr, beta1, beta2, reg_adam, max_iter,
                                                 check, K, M, N, lr_decay, min_lr, patience, tol_adam, device,
                                                 cur, threads_per_block, precision, on_check, after_step)
        except RuntimeError as exc:
            if not is_gpu_oom(exc):
                raise
            # the chunk size is the reference's knob (no chunk temporaries
            # here); what can shrink is the pass's partial-sum workspace
            if _WS.shrink():
                log.warning("    GPU out of memory while running Adam-EM: retrying with the EM pass's "
                            f"workspace capped at {_WS.min_cap >> 20} MiB.")
                continue
            if cur <= MIN_GPU_CHUNK_SIZE:
                raise
            cur = reduce_gpu_chunk_or_raise(cur, exc, "running Adam-EM")
