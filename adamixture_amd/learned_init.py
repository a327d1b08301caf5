"""Learned Q warm start (project extension; the reference initialises by
randomized SVD + ALS only, adamixture/src/adamixture.py:185-250).

A 3-layer MLP (64 -> 256 -> 256 -> 16, ReLU) maps each sample's
randomized-SVD features to a Q row: softmax(z + log max(q_als, 1e-5)) over
the first K outputs, i.e. a learned correction of the ALS row on the
simplex; the network is trained to map it to the Q the plain method
converges to.  P follows from Q by the ALS P-update (bounded least squares,
adamixture/model/als_gpu.py:99-106) on the exact centred genotypes (one
decode-fused pass).  The warm start is kept only if its FP64
log-likelihood beats the ALS start's; otherwise the plain SVD + ALS start
is returned (the fallback).

Training runs in FP32 on the libadx kernels (adx_mlp_gemm_f32, bias/ReLU,
softmax cross-entropy, Adam); inference is one fused BF16 MFMA kernel
(adx_mlp_infer_bf16) on weights packed once by adx_mlp_pack_bf16.

Features of sample i (row of X, FP32, 64 columns):
  [0, 16)   ALS Q row (K entries, zero-padded): the prior of the softmax
  [16, 32)  unconstrained least-squares Q row on the rank-2K SVD with the
            ALS P (the ALS Q-update before the bounds, als_gpu.py:108-113)
  [32, 48)  log(max(ALS Q, 1e-5)) / log(1e-5)  (in [0, 1]; 0 beyond K)
  48        fraction of the sample's rank-2K score energy in components K..2K-1
  49        log(|s_i| / median_i |s_i|), s_i = S * V[i] (rank 2K)
  50, 51    K / 16, 1 / K
  [52, 64)  singular-value profile S[k] / S[0], k = 1..12 (0 beyond 2K-1)
The three K-indexed blocks follow the ALS cluster order; training permutes
them together.
"""
from __future__ import annotations

import logging
import math
import time
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .als_gpu import _pinv_host
from .genotypes import PackedGenotypes

log = logging.getLogger(__name__)

D_IN, WIDTH, D_OUT = 64, 256, 16
NPARAM = D_IN * WIDTH + WIDTH + WIDTH * WIDTH + WIDTH + WIDTH * D_OUT + D_OUT
PACKED_BYTES = 2 * (WIDTH * D_IN + WIDTH * WIDTH + D_OUT * WIDTH) + 4 * (2 * WIDTH + D_OUT)
K_MAX = D_OUT
# ALS iterations behind the learned start's features: None = ALS to its own
# tolerance, as the plain start (up to 1000 iterations,
# adamixture/entry.py:68-69); an integer truncates it (the network was also
# trained in that mode, which pays off only while ALS dominates the start).
ALS_ITERS = None
# The fallback gate compares the two starts' FP64 logL over the leading
# SNPs only: ceil(M / GATE_FRACTION) of them, at least GATE_MIN_SNPS (all of
# G when that is M or more), in whole byte rows.  At 100,000 x 500,000 the
# two full passes cost 68 ms of a 0.3 s start; the band, 8.5 ms.
GATE_FRACTION = 8
GATE_MIN_SNPS = 20_000
# (name, rows, cols) of the flat parameter vector, in order.
BLOCKS = [("W1", D_IN, WIDTH), ("b1", 1, WIDTH), ("W2", WIDTH, WIDTH), ("b2", 1, WIDTH),
          ("W3", WIDTH, D_OUT), ("b3", 1, D_OUT)]
WEIGHTS_PATH = Path(__file__).resolve().parent / "learned_init_weights.npz"


def _offsets():
    out, o = {}, 0
    for name, r, c in BLOCKS:
        out[name] = (o, r, c)
        o += r * c
    return out


OFFSETS = _offsets()


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights and zero biases; the last layer is scaled by 0.1 so
    the untrained network stays close to the ALS prior."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, r, c in BLOCKS:
        if name.startswith("b"):
            parts.append(np.zeros(r * c))
        else:
            scale = np.sqrt(2.0 / r) * (0.1 if name == "W3" else 1.0)
            parts.append(rng.standard_normal(r * c) * scale)
    return np.concatenate(parts).astype(np.float32)


# ---------------------------------------------------------------------------
# features and the P that goes with a Q warm start
# ---------------------------------------------------------------------------

def _tall_tn(X: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
    """X^T Y for tall FP64 X, Y: on the GPU the skinny-product kernel
    (adx_gemm_tn_f64, split reduction in a fixed order; a library GEMM runs
    a 500,000-long reduction on a handful of workgroups)."""
    if X.is_cuda:
        from .als_gpu import _tn
        return _tn(X.contiguous(), Y.contiguous())
    return X.T @ Y


def _tall_nn(X: torch.Tensor, W: torch.Tensor) -> torch.Tensor:
    """X W for tall FP64 X and a small W (adx_gemm_nn_f64 on the GPU)."""
    if X.is_cuda:
        from .als_gpu import _nn
        return _nn(X.contiguous(), W.contiguous())
    return X @ W


def features(U2: torch.Tensor, S2: torch.Tensor, V2: torch.Tensor, f: torch.Tensor, P_als: torch.Tensor,
             Q_als: torch.Tensor, K: int) -> torch.Tensor:
    """Per-sample features (N x 64, FP32) from the rank-2K SVD (U2 M x 2K,
    S2, V2 N x 2K), the allele frequencies f and the ALS start (P_als, Q_als)."""
    if not 2 <= K <= K_MAX:
        raise ValueError(f"learned warm start supports 2 <= K <= {K_MAX} (got {K})")
    dev = Q_als.device
    dt = torch.float64
    N = Q_als.shape[0]
    U2, S2, V2 = U2.to(dev, dt), S2.to(dev, dt), V2.to(dev, dt)
    f = f.to(dev, dt)
    P = P_als.to(dev, dt)
    X = torch.zeros((N, D_IN), dtype=torch.float32, device=dev)
    Qa = Q_als.to(dev, dt)
    X[:, :K] = Qa.to(torch.float32)
    I_p = _tall_nn(P, _pinv_host(_tall_tn(P, P)))
    Z2 = (U2 * S2).contiguous()
    ZtI = _tall_tn(Z2, I_p)
    fI = _tall_tn(f.reshape(-1, 1).contiguous(), I_p)[0]
    X[:, 16:16 + K] = (0.5 * _tall_nn(V2.contiguous(), ZtI) + fI).to(torch.float32)
    X[:, 32:32 + K] = (torch.log(Qa.clamp_min(1e-5)) / math.log(1e-5)).to(torch.float32)
    s = V2 * S2
    e = (s * s).sum(dim=1)
    tail = (s[:, K:] * s[:, K:]).sum(dim=1) if s.shape[1] > K else torch.zeros_like(e)
    X[:, 48] = (tail / e.clamp_min(1e-30)).to(torch.float32)
    nrm = e.sqrt()
    X[:, 49] = torch.log(nrm.clamp_min(1e-30) / nrm.median().clamp_min(1e-30)).to(torch.float32)
    X[:, 50] = K / 16.0
    X[:, 51] = 1.0 / K
    prof = (S2[1:13] / S2[0]).to(torch.float32)
    X[:, 52:52 + prof.numel()] = prof
    return X


def p_from_q(U2: torch.Tensor, S2: torch.Tensor, V2: torch.Tensor, f: torch.Tensor, Q: torch.Tensor) -> torch.Tensor:
    """ALS P-update for a given Q (als_gpu.py:99-106): bounded least squares
    of the rank-2K centred genotypes on Q, then the [1e-5, 1-1e-5] map."""
    from .als_gpu import mapP
    from .solvers import batch_bvls_bpp
    dev = Q.device
    dt = torch.float64
    Z2 = U2.to(dev, dt) * S2.to(dev, dt)
    V2 = V2.to(dev, dt)
    f = f.to(dev, dt)
    A_q = Q.T @ Q
    I_q = Q @ _pinv_host(A_q)
    P_free = 0.5 * (Z2 @ (V2.T @ I_q)) + torch.outer(f, I_q.sum(dim=0))
    P = batch_bvls_bpp(A_q, P_free @ A_q, 0.0, 1.0)
    mapP(P)
    return P


def p_from_q_data(G: PackedGenotypes, f: torch.Tensor, Q: torch.Tensor) -> torch.Tensor:
    """The same P-update with the exact centred genotypes in place of the
    rank-2K SVD: one decode-fused pass X (Q pinv(Q^T Q)) (adx_svd_xw)."""
    from .als_gpu import mapP
    from .solvers import batch_bvls_bpp
    from .svd_gpu import xw
    dt = torch.float64
    Q = Q.to(dt).contiguous()
    A_q = _tall_tn(Q, Q)
    I_q = _tall_nn(Q, _pinv_host(A_q))
    Y = xw(G, f.to(torch.float32).contiguous(), I_q.to(torch.float32).contiguous()).to(dt)
    P_free = 0.5 * Y + torch.outer(f.to(dt), I_q.sum(dim=0))
    P = batch_bvls_bpp(A_q, _tall_nn(P_free, A_q), 0.0, 1.0)
    mapP(P)
    return P


# ---------------------------------------------------------------------------
# the network on the GPU
# ---------------------------------------------------------------------------

class WarmStartMLP:
    """FP32 master parameters, Adam state and the packed BF16 copy."""

    def __init__(self, device, params: np.ndarray | None = None, seed: int = 0):
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("WarmStartMLP runs on the GPU kernels of libadx only")
        _lib.load()
        w = init_params(seed) if params is None else np.asarray(params, np.float32)
        if w.shape != (NPARAM,):
            raise ValueError(f"expected {NPARAM} parameters, got {w.shape}")
        self.w = torch.from_numpy(w.copy()).to(self.device)
        self.g = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.t = 0
        self._packed = None
        self._bufs = {}

    # views into the flat vectors
    def _view(self, flat: torch.Tensor, name: str) -> torch.Tensor:
        o, r, c = OFFSETS[name]
        return flat[o:o + r * c]

    def _buf(self, name: str, shape, dtype=torch.float32) -> torch.Tensor:
        b = self._bufs.get(name)
        if b is None or b.shape != torch.Size(shape):
            b = torch.empty(shape, dtype=dtype, device=self.device)
            self._bufs[name] = b
        return b

    def _gemm(self, ta, tb, m, n, k, A, lda, B, ldb, C, ldc, beta=0.0, splits=1):
        _lib.call("adx_mlp_gemm_f32", ta, tb, m, n, k, 1.0, A.data_ptr(), lda, B.data_ptr(), ldb, beta,
                  C.data_ptr(), ldc, splits, _lib.stream_handle(self.device))

    def forward(self, X: torch.Tensor):
        """FP32 forward: returns (H1, H2, Z3) (Z3 = last-layer logits)."""
        B = X.shape[0]
        st = _lib.stream_handle(self.device)
        H1 = self._buf("H1", (B, WIDTH))
        H2 = self._buf("H2", (B, WIDTH))
        Z3 = self._buf("Z3", (B, D_OUT))
        self._gemm(0, 0, B, WIDTH, D_IN, X, D_IN, self._view(self.w, "W1"), WIDTH, H1, WIDTH)
        _lib.call("adx_mlp_bias_act", H1.data_ptr(), self._view(self.w, "b1").data_ptr(), B, WIDTH, WIDTH, 1, st)
        self._gemm(0, 0, B, WIDTH, WIDTH, H1, WIDTH, self._view(self.w, "W2"), WIDTH, H2, WIDTH)
        _lib.call("adx_mlp_bias_act", H2.data_ptr(), self._view(self.w, "b2").data_ptr(), B, WIDTH, WIDTH, 1, st)
        self._gemm(0, 0, B, D_OUT, WIDTH, H2, WIDTH, self._view(self.w, "W3"), D_OUT, Z3, D_OUT)
        _lib.call("adx_mlp_bias_act", Z3.data_ptr(), self._view(self.w, "b3").data_ptr(), B, D_OUT, D_OUT, 0, st)
        return H1, H2, Z3

    def probs(self, X: torch.Tensor, kvec: torch.Tensor) -> torch.Tensor:
        """FP32 softmax output (B x 16, zero beyond each row's K)."""
        X = X.contiguous()
        _, _, Z3 = self.forward(X)
        P = torch.empty((X.shape[0], D_OUT), dtype=torch.float32, device=self.device)
        _lib.call("adx_mlp_softmax_xent", Z3.data_ptr(), D_OUT, X.data_ptr(), D_IN, None, 0,
                  kvec.data_ptr(), X.shape[0], P.data_ptr(), None, None, _lib.stream_handle(self.device))
        return P

    def loss_and_grad(self, X: torch.Tensor, T: torch.Tensor, kvec: torch.Tensor) -> torch.Tensor:
        """Mean soft cross-entropy of the batch (device scalar); the gradient
        is left in ``self.g``."""
        X = X.contiguous()
        T = T.contiguous()
        B = X.shape[0]
        st = _lib.stream_handle(self.device)
        H1, H2, Z3 = self.forward(X)
        dZ3 = self._buf("dZ3", (B, D_OUT))
        loss = self._buf("loss", (B,))
        _lib.call("adx_mlp_softmax_xent", Z3.data_ptr(), D_OUT, X.data_ptr(), D_IN, T.data_ptr(), D_OUT,
                  kvec.data_ptr(), B, None, dZ3.data_ptr(), loss.data_ptr(), st)
        self.g.zero_()
        splits = max(1, min(64, B // 256))
        gv = lambda name: self._view(self.g, name)  # noqa: E731
        self._gemm(1, 0, WIDTH, D_OUT, B, H2, WIDTH, dZ3, D_OUT, gv("W3"), D_OUT, beta=1.0, splits=splits)
        _lib.call("adx_mlp_colsum", dZ3.data_ptr(), B, D_OUT, gv("b3").data_ptr(), st)
        dH2 = self._buf("dH2", (B, WIDTH))
        self._gemm(0, 1, B, WIDTH, D_OUT, dZ3, D_OUT, self._view(self.w, "W3"), D_OUT, dH2, WIDTH)
        _lib.call("adx_mlp_relu_bwd", dH2.data_ptr(), H2.data_ptr(), B * WIDTH, st)
        self._gemm(1, 0, WIDTH, WIDTH, B, H1, WIDTH, dH2, WIDTH, gv("W2"), WIDTH, beta=1.0, splits=splits)
        _lib.call("adx_mlp_colsum", dH2.data_ptr(), B, WIDTH, gv("b2").data_ptr(), st)
        dH1 = self._buf("dH1", (B, WIDTH))
        self._gemm(0, 1, B, WIDTH, WIDTH, dH2, WIDTH, self._view(self.w, "W2"), WIDTH, dH1, WIDTH)
        _lib.call("adx_mlp_relu_bwd", dH1.data_ptr(), H1.data_ptr(), B * WIDTH, st)
        self._gemm(1, 0, D_IN, WIDTH, B, X, D_IN, dH1, WIDTH, gv("W1"), WIDTH, beta=1.0, splits=splits)
        _lib.call("adx_mlp_colsum", dH1.data_ptr(), B, WIDTH, gv("b1").data_ptr(), st)
        return loss.mean()

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.t += 1
        _lib.call("adx_mlp_adam", self.w.data_ptr(), self.g.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                  NPARAM, lr, beta1, beta2, eps, self.t, _lib.stream_handle(self.device))
        self._packed = None

    def packed(self) -> torch.Tensor:
        if self._packed is None:
            self._packed = torch.empty(PACKED_BYTES, dtype=torch.uint8, device=self.device)
            _lib.call("adx_mlp_pack_bf16", self.w.data_ptr(), self._packed.data_ptr(),
                      _lib.stream_handle(self.device))
        return self._packed

    def infer(self, X: torch.Tensor, K: int) -> torch.Tensor:
        """BF16 inference: the FP64 Q warm start (N x K)."""
        X = X.to(torch.float32).contiguous()
        if X.shape[1] != D_IN:
            raise ValueError(f"features must have {D_IN} columns")
        Q = torch.empty((X.shape[0], K), dtype=torch.float64, device=self.device)
        _lib.call("adx_mlp_infer_bf16", X.data_ptr(), D_IN, X.shape[0], K, self.packed().data_ptr(), Q.data_ptr(),
                  K, _lib.stream_handle(self.device))
        return Q

    def params_numpy(self) -> np.ndarray:
        return self.w.cpu().numpy()

    def save(self, path, meta: dict | None = None) -> None:
        import json
        np.savez(path, params=self.params_numpy(), meta=np.array(json.dumps(meta or {})))

    @classmethod
    def load(cls, device, path=WEIGHTS_PATH) -> "WarmStartMLP":
        with np.load(path, allow_pickle=False) as z:
            return cls(device, params=z["params"])


def load_meta(path=WEIGHTS_PATH) -> dict:
    import json
    with np.load(path, allow_pickle=False) as z:
        return json.loads(str(z["meta"]))


# ---------------------------------------------------------------------------
# the warm start with its fallback
# ---------------------------------------------------------------------------

def gate_band(G):
    """The genotypes the fallback gate reads: the leading
    max(ceil(M / GATE_FRACTION), GATE_MIN_SNPS) SNPs in whole byte rows (a
    view of G), or G itself when that covers M or G is host-streamed."""
    from .genotypes import is_streamed
    m = max(-(-G.M // GATE_FRACTION), GATE_MIN_SNPS)
    rows = (m + 3) // 4
    if is_streamed(G) or 4 * rows >= G.M:
        return G
    return PackedGenotypes(G.data[:rows], 4 * rows, G.N)


def gate_loglik(G, P: torch.Tensor, Q: torch.Tensor) -> float:
    """FP64 logL of (P, Q) over the gate's SNPs (gate_band)."""
    from .em_adam_gpu import em_pass
    Gb = gate_band(G)
    return em_pass(Gb, P[: Gb.M].contiguous(), Q.contiguous(), loglik=True, stats=False).item()


def learned_warm_start(G: PackedGenotypes, U2, S2, V2, f, K: int, model: WarmStartMLP, seed: int, *,
                       als_iters: int | None = ALS_ITERS, max_als: int = 1000, tol_als: float = 1e-4):
    """Warm start from the rank-2K SVD (U2, S2, V2): ALS on its first K
    components (to tolerance, or ``als_iters`` iterations), the network's Q,
    and P from Q by one decode-fused pass over G.  Returns
    (P, Q, gate_logL, used_learned), gate_logL being the returned start's
    FP64 logL over the gate's SNPs (gate_band).  If the learned start's
    gate logL does not exceed that of the ALS start, the plain SVD + ALS
    start is returned (ALS continued to its tolerance first when it was
    truncated)."""
    from .als_gpu import ALS_gpu
    t0 = time.time()
    dev = G.device
    Uk, Sk, Vk = U2[:, :K].contiguous(), S2[:K].contiguous(), V2[:, :K].contiguous()
    n_als = max_als if als_iters is None else als_iters
    P_s, Q_s = ALS_gpu(Uk, Sk, Vk, f, seed, G.M, K, n_als, tol_als, dev)
    X = features(U2, S2, V2, f, P_s, Q_s, K)
    Q = model.infer(X, K)
    P = p_from_q_data(G, f, Q)
    ll = gate_loglik(G, P, Q)
    ll_s = gate_loglik(G, P_s, Q_s)
    used = ll > ll_s
    log.info(f"    Learned warm start gate logL {ll:.1f} vs SVD+ALS {ll_s:.1f} -> "
             f"{'learned' if used else 'fallback to SVD+ALS'} ({time.time() - t0:.3f}s)")
    if used:
        return P, Q, ll, True
    if als_iters is None:
        return P_s, Q_s, ll_s, False
    P_a, Q_a = ALS_gpu(Uk, Sk, Vk, f, seed, G.M, K, max(max_als - als_iters, 0), tol_als, dev, P_init=P_s,
                       Q_init=Q_s)
    return P_a, Q_a, gate_loglik(G, P_a, Q_a), False
