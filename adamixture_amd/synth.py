"""Synthetic admixture workloads (the benchmark configurations).

Parameters are drawn on the host with ``numpy.random.default_rng(seed)`` in
the order P, then Q; genotypes G ~ Binomial(2, Q P^T) and the missing mask
come from the counter-based device generator (adx_synth_packed) keyed by the
same seed, so any block can be regenerated on the CPU.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from .genotypes import PackedGenotypes, synthesize


@dataclass(frozen=True)
class Workload:
    name: str
    N: int
    M: int
    K: int
    p_model: str          # "uniform" or "bn" (Balding-Nichols drift)
    alpha: float          # Dirichlet concentration of Q rows
    miss: float
    fst: float = 0.1
    seed: int = 42


WORKLOADS = {
    "cpu_ref": Workload("cpu_ref", 1_000, 20_000, 3, "uniform", 0.5, 0.01),
    "gpu_f32": Workload("gpu_f32", 10_000, 100_000, 8, "bn", 0.3, 0.005),
    "biobank": Workload("biobank", 100_000, 500_000, 10, "bn", 0.2, 0.01),
    "multik": Workload("multik", 50_000, 200_000, 12, "bn", 0.3, 0.005),
    "learned_eval": Workload("learned_eval", 200_000, 600_000, 12, "bn", 0.1, 0.01),
}


def draw_params(w: Workload, K: int | None = None) -> tuple[np.ndarray, np.ndarray]:
    """(P, Q) as float64: P first, then Q, from default_rng(seed)."""
    K = w.K if K is None else K
    rng = np.random.default_rng(w.seed)
    if w.p_model == "uniform":
        P = rng.uniform(0.05, 0.95, size=(w.M, K))
    elif w.p_model == "bn":
        p = rng.uniform(0.05, 0.5, size=(w.M, 1))
        c = (1.0 - w.fst) / w.fst
        P = rng.beta(p * c, (1.0 - p) * c, size=(w.M, K))
        P = np.clip(P, 1e-5, 1.0 - 1e-5)
    else:
        raise ValueError(f"unknown P model {w.p_model}")
    Q = rng.dirichlet(np.full(K, w.alpha), size=w.N)
    return np.ascontiguousarray(P, dtype=np.float64), np.ascontiguousarray(Q, dtype=np.float64)


def make_genotypes(w: Workload, device: torch.device | str = "cuda", K: int | None = None,
                   params: tuple[np.ndarray, np.ndarray] | None = None) -> tuple[PackedGenotypes, np.ndarray, np.ndarray]:
    P, Q = draw_params(w, K) if params is None else params
    dev = torch.device(device)
    G = synthesize(torch.from_numpy(P).to(dev), torch.from_numpy(Q).to(dev), w.seed, w.miss)
    return G, P, Q
