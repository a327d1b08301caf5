"""Cross-validation on the GPU (drop-in for adamixture/src/cv.py:171-267
run_cross_validation; the reference runs it on the host over the unpacked
matrix, re-read from disk).

Per fold, on the device-resident packed matrix:
  1. the fold's held-out entries are masked in a copy of G (adx_cv_mask_list
     for the reference's shuffled-index folds, adx_cv_mask_hash for its
     hashed folds -- the same fold membership, entry for entry);
  2. P and Q are polished from the global fit by 5 SQP + ZAL iterations with
     early stopping disabled and the CPU optimizer's quasi-Newton map
     (br_qn.py:50-70, polish_sqp_qn), on the GPU kernels of the BR-QN path;
  3. the held-out entries of the original G are scored by their deviance at
     the polished (P, Q) (adx_cv_deviance_*; fixed summation order).
CV index = sum of the fold sums / number of observed genotypes.

Fold membership is the reference's: the shuffled path draws the same
permutation from numpy's default_rng(seed) on the host (it needs the flat
indices of the observed entries, built from the packed matrix), and the
hashed path evaluates splitmix64(m N + i + seed) mod n_folds on the device.
The reference switches to the hashed path when M N exceeds the int32 range;
``streaming`` overrides that choice.
"""
from __future__ import annotations

import ctypes
import logging

import numpy as np
import torch

from . import _lib
from .br_qn_gpu import _optimize_original_gpu_once
from .genotypes import PackedGenotypes, as_packed

log = logging.getLogger(__name__)

POLISH_ITERS = 5          # cv.py:146-168
POLISH_Q_HIST = 3


def _packed_input(G, M: int, N: int, device) -> PackedGenotypes:
    """The reference hands over the unpacked M x N matrix; packed matrices
    (ceil(M/4) x N, host or device) and PackedGenotypes are accepted too."""
    if isinstance(G, PackedGenotypes):
        return G
    if tuple(G.shape) == (M, N) and M != (M + 3) // 4:
        arr = G.detach().cpu().numpy() if isinstance(G, torch.Tensor) else np.asarray(G)
        return as_packed(pack_host(arr), M, device)
    return as_packed(G, M, device)


def pack_host(G: np.ndarray) -> np.ndarray:
    """Unpacked M x N codes (0, 1, 2, 3 = missing) -> ceil(M/4) x N bytes,
    SNP 4j + f in bits 2f of byte row j (fields past M are 0)."""
    M, N = G.shape
    pad = (-M) % 4
    Gx = np.concatenate([G, np.zeros((pad, N), np.uint8)]) if pad else G
    Gx = Gx.reshape(-1, 4, N).astype(np.uint8)
    return (Gx[:, 0] | (Gx[:, 1] << 2) | (Gx[:, 2] << 4) | (Gx[:, 3] << 6)).astype(np.uint8)


def observed_flat_host(Gp: PackedGenotypes) -> np.ndarray:
    """Flat indices m N + i of the observed genotypes in row-major order
    (np.flatnonzero(G != 3) of the reference, cv.py:214), as int64."""
    M, N = Gp.M, Gp.N
    host = Gp.data[:, :N].cpu().numpy()
    out = []
    rows = 1 << 14
    for j0 in range(0, host.shape[0], rows):
        blk = host[j0:j0 + rows]
        codes = np.stack([(blk >> (2 * f)) & 3 for f in range(4)], axis=1).reshape(-1, N)
        m_hi = min(M - 4 * j0, codes.shape[0])
        out.append(np.flatnonzero(codes[:m_hi] != 3) + 4 * j0 * N)
    return np.concatenate(out) if out else np.zeros(0, np.int64)


def _observed_count(Gp: PackedGenotypes) -> int:
    out = torch.zeros(2, dtype=torch.int64, device=Gp.device)
    _lib.call("adx_packed_allele_sum", Gp.data.data_ptr(), Gp.ldg, Gp.M, Gp.N, out.data_ptr(),
              _lib.stream_handle(Gp.device))
    return int(out[1].item()) // 2


def polish_gpu(Gm: PackedGenotypes, P: torch.Tensor, Q: torch.Tensor):
    """br_qn.py:50-70 polish_sqp_qn on the GPU: exactly POLISH_ITERS
    iterations (tol 0, patience n_iters + 1) from copies of (P, Q)."""
    M, K = P.shape
    N = Q.shape[0]
    return _optimize_original_gpu_once(Gm, P.clone(), Q.clone(), POLISH_ITERS, K, M, N, 0.0, POLISH_Q_HIST,
                                       POLISH_ITERS + 1, Gm.device, 0, 0, qn_map="cpu")


class FoldScorer:
    """Per-fold work of run_cross_validation on one device: mask, polish,
    deviance.  ``fold_sum(fold)`` returns the fold's deviance sum (0.0 for an
    empty fold)."""

    def __init__(self, Gp: PackedGenotypes, P: torch.Tensor, Q: torch.Tensor, n_folds: int, seed: int,
                 streaming: bool):
        self.Gp, self.P, self.Q = Gp, P, Q
        self.n_folds, self.seed, self.streaming = n_folds, seed, streaming
        dev = Gp.device
        self.st = _lib.stream_handle(dev)
        self.work = torch.empty(_lib.ADX_CV_BLOCKS, dtype=torch.float64, device=dev)
        self.out = torch.empty(1, dtype=torch.float64, device=dev)
        self.Gm = PackedGenotypes(torch.empty_like(Gp.data), Gp.M, Gp.N)
        if not streaming:
            flat = observed_flat_host(Gp).astype(np.int32)
            np.random.default_rng(seed).shuffle(flat)           # cv.py:27-42
            self.n_flat = flat.size
            self.flat_d = torch.from_numpy(flat).to(dev)

    def fold_sum(self, fold: int) -> float:
        Gp, Gm, M, N = self.Gp, self.Gm, self.Gp.M, self.Gp.N
        seed = ctypes.c_uint64(self.seed).value
        if self.streaming:
            _lib.call("adx_cv_mask_hash", Gp.data.data_ptr(), Gm.data.data_ptr(), Gp.ldg, M, N, seed, self.n_folds,
                      fold, self.st)
        else:
            n = len(range(fold, self.n_flat, self.n_folds))
            if n == 0:
                return 0.0
            Gm.data.copy_(Gp.data)
            _lib.call("adx_cv_mask_list", Gm.data.data_ptr(), Gm.ldg, M, N, self.flat_d.data_ptr(), 0, fold,
                      self.n_folds, n, self.st)
        Gm._qbat = None                                    # contents changed
        P_cv, Q_cv = polish_gpu(Gm, self.P, self.Q)
        K = self.P.shape[1]
        if self.streaming:
            _lib.call("adx_cv_deviance_hash", Gp.data.data_ptr(), Gp.ldg, M, N, seed, self.n_folds, fold,
                      P_cv.data_ptr(), Q_cv.data_ptr(), K, self.work.data_ptr(), self.out.data_ptr(), self.st)
        else:
            _lib.call("adx_cv_deviance_list", Gp.data.data_ptr(), Gp.ldg, M, N, self.flat_d.data_ptr(), 0, fold,
                      self.n_folds, n, P_cv.data_ptr(), Q_cv.data_ptr(), K, self.work.data_ptr(),
                      self.out.data_ptr(), self.st)
        return self.out.item()


def run_cross_validation(args, G, N: int, M: int, K: int, P_global, Q_global, *, device=None,
                         streaming: bool | None = None) -> float:
    """cv.py:171-267.  ``args`` needs ``cv`` (number of folds) and ``seed``."""
    n_folds = int(args.cv)
    dev = torch.device(device) if device is not None else (
        P_global.device if isinstance(P_global, torch.Tensor) and P_global.is_cuda else torch.device("cuda"))
    Gp = _packed_input(G, M, N, dev)
    if (Gp.M, Gp.N) != (M, N):
        raise ValueError(f"CV requires genotypes of shape {(M, N)}, got {(Gp.M, Gp.N)}.")
    P = torch.as_tensor(P_global).to(device=dev, dtype=torch.float64).contiguous()
    Q = torch.as_tensor(Q_global).to(device=dev, dtype=torch.float64).contiguous()
    if streaming is None:
        streaming = M * N > np.iinfo(np.int32).max
    n_observed = _observed_count(Gp)
    if n_observed == 0:
        raise ValueError("No non-missing genotypes available for cross-validation.")
    scorer = FoldScorer(Gp, P, Q, n_folds, int(args.seed), streaming)
    cv_sum = sum(scorer.fold_sum(fold) for fold in range(n_folds))
    return cv_sum / float(n_observed)
