"""Supervised mode on the GPU (drop-in for adamixture/src/supervised.py:
init_q_supervised :20-46, init_p_supervised_packed :84-107, _snap_q_gpu
:308-339, optimize_supervised_gpu :342-462, optimize_supervised_original_gpu
:602-781; label parsing of adamixture/supervised_entry.py:147-188).

Samples with a known population (y > 0) keep a near-one-hot Q row: after every
Q update their rows are snapped back (adx_snap_labels, one small kernel on the
launch stream).  The Adam-EM and SQP + ZAL drivers are the joint ones with
that snap as a hook, so every other step -- the fused EM pass, the fused
log-likelihood of a check, the lr schedule, the quasi-Newton acceptance -- is
shared with the unsupervised path.  The P start is the labelled samples'
allele frequencies, computed from the packed matrix on the GPU
(adx_label_freq, bit-exact with the reference's host computation).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

from . import _lib
from .br_qn_gpu import optimize_original_gpu
from .em_adam_gpu import optimize_parameters_gpu
from .genotypes import PackedGenotypes, as_packed

SNAP_EPS = 1e-5


def parse_labels_file(path) -> list[str]:
    """supervised_entry.py:147-159: one label per non-empty line."""
    return [line.strip() for line in Path(path).open() if line.strip()]


def labels_to_y(raw_labels: list[str]) -> tuple[np.ndarray, dict[str, int]]:
    """supervised_entry.py:162-188: '-' -> 0 (unlabelled), other names ->
    1, 2, ... in order of first appearance."""
    pop_to_int: dict[str, int] = {}
    y = np.zeros(len(raw_labels), dtype=np.int32)
    for i, lbl in enumerate(raw_labels):
        if lbl != "-":
            y[i] = pop_to_int.setdefault(lbl, len(pop_to_int) + 1)
    return y, pop_to_int


def init_q_supervised(Q: np.ndarray, y: np.ndarray, K: int, eps: float = SNAP_EPS) -> None:
    """supervised.py:20-46 (host, N x K): labelled rows near-one-hot,
    unlabelled rows renormalised, in place."""
    y = np.asarray(y)
    lab = np.flatnonzero(y > 0)
    Q[lab] = eps / max(K - 1, 1)
    Q[lab, y[lab] - 1] = 1.0 - eps
    un = y == 0
    Q[un] /= Q[un].sum(axis=1, keepdims=True)


def _check_labels(y: np.ndarray, N: int, K: int) -> np.ndarray:
    y = np.asarray(y)
    if y.shape != (N,):
        raise ValueError(f"labels must have one entry per sample ({N}), got shape {y.shape}")
    if y.size and (y.min() < 0 or y.max() > K):
        raise ValueError(f"labels must lie in 0..{K}")
    return y.astype(np.int32, copy=False)


class Labels:
    """Device copy of the assignment vector y (int32, 0 = unlabelled) and the
    snap of supervised.py:308-339 as one kernel."""

    def __init__(self, y: np.ndarray, K: int, device, eps: float = SNAP_EPS):
        self.K, self.eps = K, eps
        self.device = torch.device(device)
        self.y = np.asarray(y).astype(np.int32)
        self.y_dev = torch.from_numpy(self.y).to(self.device)

    def snap(self, Q: torch.Tensor) -> None:
        if Q.dtype != torch.float64 or not Q.is_contiguous() or Q.shape != (self.y.size, self.K):
            raise TypeError("Q must be a contiguous float64 N x K tensor")
        _lib.call("adx_snap_labels", Q.data_ptr(), Q.shape[0], self.K, self.y_dev.data_ptr(), self.eps,
                  _lib.stream_handle(Q.device))


def label_freq_gpu(G: PackedGenotypes, y: np.ndarray, K: int, eps: float = SNAP_EPS) -> torch.Tensor:
    """P (M x K, FP64, on G's device) = init_p_supervised_packed of the
    reference, computed by adx_label_freq."""
    y = _check_labels(y, G.N, K)
    lab = np.flatnonzero(y > 0)
    order = lab[np.argsort(y[lab], kind="stable")].astype(np.int32)   # grouped by population, ascending
    offsets = np.zeros(K + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(np.bincount(y[lab] - 1, minlength=K)[:K])
    dev = G.device
    order_d = torch.from_numpy(order if order.size else np.zeros(1, np.int32)).to(dev)
    off_d = torch.from_numpy(offsets).to(dev)
    P = torch.empty((G.M, K), dtype=torch.float64, device=dev)
    _lib.call("adx_label_freq", G.data.data_ptr(), G.ldg, G.M, G.N, K, order_d.data_ptr(), off_d.data_ptr(), eps,
              P.data_ptr(), _lib.stream_handle(dev))
    return P


def init_p_supervised_packed(G, y: np.ndarray, K: int, M: int, eps: float = SNAP_EPS,
                             device=None) -> np.ndarray:
    """supervised.py:84-107: same result (bit for bit), computed on the GPU
    from the packed matrix; returned on the host like the reference's."""
    Gp = as_packed(G, M, device)
    return label_freq_gpu(Gp, y, K, eps).cpu().numpy()


def optimize_supervised_gpu(G, P, Q, y, lr, beta1, beta2, reg_adam, max_iter, check, M, lr_decay, min_lr,
                            patience, tol_adam, device, chunk_size, threads_per_block, precision=None,
                            on_check=None):
    """supervised.py:342-462: Adam-EM on P and Q with the labelled Q rows
    snapped after every Adam step (:404-406), priming included."""
    dev = torch.device(device)
    N, K = Q.shape
    labels = Labels(_check_labels(y, N, K), K, dev)
    P = P.to(device=dev, dtype=torch.float64).contiguous()
    Q = Q.to(device=dev, dtype=torch.float64).contiguous()
    return optimize_parameters_gpu(G, P, Q, lr, beta1, beta2, reg_adam, max_iter, check, K, M, N, lr_decay, min_lr,
                                   patience, tol_adam, dev, chunk_size, threads_per_block, precision=precision,
                                   on_check=on_check, after_step=lambda _P, Qc: labels.snap(Qc))


def optimize_supervised_original_gpu(G, P, Q, y, max_iter, K, M, N, tol, Q_hist, patience, device, chunk_size,
                                     threads_per_block, on_iter=None):
    """supervised.py:602-781: SQP + ZAL quasi-Newton on P and Q with the
    labelled Q rows snapped after each Q solve and after the quasi-Newton map."""
    dev = torch.device(device)
    labels = Labels(_check_labels(y, N, K), K, dev)
    P = P.to(device=dev, dtype=torch.float64).contiguous()
    Q = Q.to(device=dev, dtype=torch.float64).contiguous()
    return optimize_original_gpu(G, P, Q, max_iter, K, M, N, tol, Q_hist, patience, dev, chunk_size,
                                 threads_per_block, on_iter=on_iter, snap=labels.snap)
