"""Run driver for the GPU path (drop-in for adamixture/src/adamixture.py:22-602,
the `--device gpu` branches): one-time setup (frequencies + randomized SVD at
K_max), per-K initialisation (ALS on the sliced SVD, or random + EM) and
training (Adam-EM or SQP + ZAL quasi-Newton).

Host arrays follow the reference: P, Q are FP64 tensors on the device and
the returned (P, Q) are device tensors, written by ``utils.write_outputs``
as ADMIXTURE .P / .Q text.
"""
from __future__ import annotations

import logging

import numpy as np
import torch

from . import utils
from .als_gpu import ALS_gpu
from .em_adam_gpu import EMAdamOptimizer, optimize_parameters_gpu
from .genotypes import as_packed
from .svd_gpu import SVD_gpu

log = logging.getLogger(__name__)


def _require_gpu(device) -> torch.device:
    dev = torch.device(device)
    if dev.type != "cuda":
        raise ValueError("adamixture_amd implements the --device gpu path only (got device "
                         f"'{device}'); use the reference's CPU path for --device cpu")
    utils.load_extensions(dev)
    return dev


def setup(G, N: int, M: int, K_max: int, seed: int, power: int, tol_svd: float, chunk_size: int, device,
          original: bool = False, init_original: str = "em", q_hist: int = 3, compute_svd: bool = True):
    """adamixture.py:22-100 -> (device_obj, threads_per_block, f, U, S, V, G)."""
    dev = _require_gpu(device)
    tpb = utils.get_tuning_params(dev)
    algorithm = "brqn" if original else "adamem"
    if original and init_original == "em":
        G = utils.manage_gpu_memory(G, dev, M, N, K_max, chunk_size, algorithm, q_hist, include_initialization=False)
        return dev, tpb, None, None, None, None, G
    G = utils.manage_gpu_memory(G, dev, M, N, K_max, chunk_size, algorithm, q_hist)
    f = utils.calculate_frequencies_gpu(G, M, chunk_size, dev, tpb)
    U = S = V = None
    if compute_svd:
        U, S, V = SVD_gpu(G, N, M, f, K_max, seed, power, tol_svd, chunk_size, dev, tpb)
    return dev, tpb, f, U, S, V, G


def initialize_em_gpu(G, seed: int, M: int, N: int, K: int, device_obj, chunk_size: int,
                      threads_per_block: int, n_steps: int = 5, precision: str | None = None):
    """adamixture.py:140-182: numpy-seeded random P, Q then plain EM steps."""
    rng = np.random.default_rng(seed)
    P_np = np.clip(rng.random(size=(M, K), dtype=np.float64), 1e-5, 1.0 - 1e-5)
    Q_np = np.clip(rng.random(size=(N, K), dtype=np.float64), 1e-5, 1.0 - 1e-5)
    Q_np /= Q_np.sum(axis=1, keepdims=True)
    P = torch.from_numpy(P_np).to(device_obj)
    Q = torch.from_numpy(Q_np).to(device_obj)
    G = as_packed(G, M, device_obj)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.0, 0.0, 0.0, 0.0, device_obj, precision)
    for _ in range(n_steps):
        opt.run_em_step(G, P, Q, M)
        P.copy_(opt.P_EM)
        Q.copy_(opt.Q_EM)
    return P, Q


def initialize_k(G, N: int, M: int, K: int, f, seed: int, power: int, tol_svd: float, max_als: int,
                 tol_als: float, chunk_size: int, device_obj, threads_per_block: int, original: bool = False,
                 init_original: str = "em", em_init_steps: int = 5):
    """adamixture.py:185-250 -> (P, Q, initial logL); runs its own rank-K
    randomized SVD (not the K_max one of ``setup``)."""
    dev = _require_gpu(device_obj)
    logl_calc = utils.get_logl_calculator(dev)
    if original and init_original == "em":
        P, Q = initialize_em_gpu(G, seed, M, N, K, dev, chunk_size, threads_per_block, em_init_steps)
    else:
        U, S, V = SVD_gpu(G, N, M, f, K, seed, power, tol_svd, chunk_size, dev, threads_per_block)
        P, Q = ALS_gpu(U, S, V, f, seed, M, K, max_als, tol_als, dev)
    logl = logl_calc(G, P, Q, M, N, chunk_size, threads_per_block)
    log.info(f"    Initial log-likelihood for K={K}: {logl:.1f}.")
    return P, Q, logl


# SVD settings of a rank-2K SVD the learned start runs when the K_max SVD of
# setup has fewer than 2K components: the reference's defaults
# (adamixture/entry.py:65-66).
_WARM_SVD_POWER, _WARM_SVD_TOL = 5, 0.1
_WARM_MODELS: dict = {}


def _learned_start(G, N: int, M: int, K: int, U_max, S_max, V_max, f, seed: int, max_als: int, tol_als: float,
                   dev, chunk_size: int, threads_per_block: int):
    """(P, Q) of the learned warm start (learned_init.learned_warm_start,
    with its fallback to SVD + ALS), or None when the shipped network does
    not cover K or its weights are absent.  Features need a rank-2K SVD: the
    leading 2K components of setup's K_max SVD when it has them, else a
    rank-2K SVD of its own."""
    from .learned_init import K_MAX, WEIGHTS_PATH, WarmStartMLP, learned_warm_start
    if K > K_MAX or not WEIGHTS_PATH.exists():
        log.warning(f"    No learned warm start for K={K} (the network covers K <= {K_MAX}); using SVD + ALS.")
        return None
    model = _WARM_MODELS.get(dev)
    if model is None:
        model = _WARM_MODELS[dev] = WarmStartMLP.load(dev)
    if U_max is not None and U_max.shape[1] >= 2 * K:
        U2, S2, V2 = U_max[:, : 2 * K].contiguous(), S_max[: 2 * K].contiguous(), V_max[:, : 2 * K].contiguous()
    else:
        U2, S2, V2 = SVD_gpu(G, N, M, f, 2 * K, seed, _WARM_SVD_POWER, _WARM_SVD_TOL, chunk_size, dev,
                             threads_per_block)
    P, Q, _, used = learned_warm_start(as_packed(G, M, dev), U2, S2, V2, f, K, model, seed, max_als=max_als,
                                       tol_als=tol_als)
    log.info(f"    Warm start for K={K}: {'learned' if used else 'SVD + ALS (fallback)'}.")
    return P.contiguous(), Q.contiguous()


def train_k(G, N: int, M: int, K: int, U_max, S_max, V_max, f, seed: int, lr: float, beta1: float,
            beta2: float, reg_adam: float, max_iter: int, check: int, max_als: int, tol_als: float,
            lr_decay: float, min_lr: float, chunk_size: int, patience: int, tol_adam: float, device_obj,
            threads_per_block: int, original: bool = False, rtol: float = 1e-7, Q_hist: int = 3,
            init_original: str = "em", em_init_steps: int = 5, P_init=None, Q_init=None,
            precision: str | None = None, warm_start: str | None = None):
    """adamixture.py:253-352 (GPU branches; same parameters, order and
    defaults) -> (P, Q) device tensors.  The rank-K initialisation slices the
    K_max SVD of ``setup`` (U_max[:, :K], S_max[:K], V_max[:, :K]) as the
    reference does.  Project extensions: ``precision`` selects the Adam-EM
    pass form; ``warm_start="learned"`` starts from the learned warm start
    (with its SVD + ALS fallback) instead of the SVD + ALS start."""
    dev = _require_gpu(device_obj)
    if warm_start not in (None, "svd_als", "learned"):
        raise ValueError("warm_start must be None, 'svd_als' or 'learned'")
    if P_init is not None and Q_init is not None:
        P = torch.as_tensor(P_init, dtype=torch.float64, device=dev).contiguous()
        Q = torch.as_tensor(Q_init, dtype=torch.float64, device=dev).contiguous()
    else:
        start = None
        if warm_start == "learned" and not (original and init_original == "em"):
            start = _learned_start(G, N, M, K, U_max, S_max, V_max, f, seed, max_als, tol_als, dev, chunk_size,
                                   threads_per_block)
        if start is not None:
            P, Q = start
        elif original and init_original == "em":
            P, Q = initialize_em_gpu(G, seed, M, N, K, dev, chunk_size, threads_per_block, em_init_steps)
        else:
            U, S, V = U_max[:, :K].contiguous(), S_max[:K].contiguous(), V_max[:, :K].contiguous()
            P, Q = ALS_gpu(U, S, V, f, seed, M, K, max_als, tol_als, dev)
        logl = utils.get_logl_calculator(dev)(G, P, Q, M, N, chunk_size, threads_per_block)
        log.info(f"    Initial log-likelihood for K={K}: {logl:.1f}.")
    if original:
        from .br_qn_gpu import optimize_original_gpu
        return optimize_original_gpu(G, P, Q, max_iter, K, M, N, rtol, Q_hist, patience, dev, chunk_size,
                                     threads_per_block)
    return optimize_parameters_gpu(G, P, Q, lr, beta1, beta2, reg_adam, max_iter, check, K, M, N, lr_decay,
                                   min_lr, patience, tol_adam, dev, chunk_size, threads_per_block, precision)
