"""Projection mode (P fixed, Q optimised) on the GPU against the oracle:
projection.py:178-321 (Adam-EM on Q) and :505-659 (SQP + ZAL on Q).

The FP64 Adam-EM projection is compared with the oracle's trajectory within
10x its chaos floor, measured as for the joint Adam-EM (the oracle re-run
with FP64-rounding-sized noise on every EM target, two seeds); the SQP
projection, a Newton-type iteration, is compared directly.  At a size the
oracle does not finish quickly the size-independent property is used: logL
is concave in Q for fixed P, so the projected Q maximises logL(P, .) -- it
beats the generating Q, the two algorithms agree, and the KKT conditions of
the box-clamped simplex hold (gradients from the GPU grad/Hessian pass)."""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd.br_qn_gpu import grad_hess
from adamixture_amd.em_adam_gpu import loglikelihood_gpu
from adamixture_amd.projection import optimize_projection_gpu, optimize_projection_original_gpu

pytestmark = pytest.mark.gpu

HP = dict(lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, check=5, lr_decay=0.5, min_lr=1e-4, patience=3,
          tol_adam=0.1)


def _random_q(N, K, seed):
    rng = np.random.default_rng(seed)
    Q = rng.random((N, K))
    return Q / Q.sum(axis=1, keepdims=True)


def _adam(G, P, Q0, M, dev, precision=None):
    hist = []
    Qt = torch.from_numpy(Q0.copy()).to(dev)
    Q = optimize_projection_gpu(G, torch.from_numpy(P).to(dev), Qt, HP["lr"], HP["beta1"], HP["beta2"],
                                HP["reg_adam"], 10000, HP["check"], M, HP["lr_decay"], HP["min_lr"],
                                HP["patience"], HP["tol_adam"], dev, 8192, 256, precision=precision,
                                on_check=lambda it, L: hist.append((it, L)))
    return Q.cpu().numpy(), hist, Qt


def _sqp(G, P, Q0, M, dev):
    hist = []
    N, K = Q0.shape
    Q = optimize_projection_original_gpu(G, torch.from_numpy(P).to(dev), torch.from_numpy(Q0.copy()).to(dev),
                                         10000, K, M, N, 0.1, 3, 3, dev, 8192, 256,
                                         on_iter=lambda it, L: hist.append((it, L)))
    return Q.cpu().numpy(), hist


def _floor(Gp, P, Q0, M, Qo, ohist):
    out = {"Q": 0.0, "ll": 0.0}
    L0 = np.array([L for _, L in ohist])
    for seed in (0, 1):
        h = []
        Qp = O.projection_gpu(Gp, P, Q0, M, history=h, noise=(seed, 1e-16))
        L1 = np.array([L for _, L in h])
        n = min(len(L0), len(L1))
        out["Q"] = max(out["Q"], np.abs(Qp - Qo).max())
        out["ll"] = max(out["ll"], float(np.max(np.abs(L1[:n] - L0[:n]) / np.abs(L0[:n]))))
    return out


def test_demo_adam_projection_f64_matches_oracle(cuda, fixtures):
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    P = fixtures["gpu_adamem_adamem_P"]
    Q0 = _random_q(*fixtures["gpu_adamem_adamem_Q"].shape, seed=3)
    print("case: demo Adam-EM projection f64, seed 3", flush=True)
    Q, hist, Qt = _adam(torch.from_numpy(Gp), P, Q0, M, cuda)
    ohist = []
    Qo = O.projection_gpu(Gp, P, Q0, M, history=ohist)
    floor = _floor(Gp, P, Q0, M, Qo, ohist)
    print(f"{len(hist)} checks; oracle chaos floor dQ {floor['Q']:.3e} logL rel {floor['ll']:.3e}; "
          f"GPU dQ {np.abs(Q - Qo).max():.3e}")
    assert [it for it, _ in hist] == [it for it, _ in ohist]
    np.testing.assert_allclose([L for _, L in hist], [L for _, L in ohist], rtol=max(1e-13, 10 * floor["ll"]))
    assert np.abs(Q - Qo).max() <= 10 * floor["Q"] + 1e-12
    # the caller's FP64 device Q is the one optimised in place (projection.py:190 "Updated in-place")
    assert not np.array_equal(Qt.cpu().numpy(), Q0)


@pytest.mark.parametrize("precision", ["f32", "bf16"])
def test_demo_adam_projection_low_precision(cuda, fixtures, precision):
    """FP32 / BF16 EM passes (FP64 Q, Adam and logL): the same maximiser to
    within the stopping tolerance."""
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    P = fixtures["gpu_adamem_adamem_P"]
    Q0 = _random_q(*fixtures["gpu_adamem_adamem_Q"].shape, seed=3)
    print(f"case: demo Adam-EM projection {precision}, seed 3", flush=True)
    Q, hist, _ = _adam(torch.from_numpy(Gp), P, Q0, M, cuda, precision)
    ohist = []
    Qo = O.projection_gpu(Gp, P, Q0, M, history=ohist)
    L, Lo = hist[-1][1], ohist[-1][1]
    print(f"{precision}: {len(hist)} checks (oracle {len(ohist)}); final logL rel diff {abs(L - Lo) / abs(Lo):.2e}; "
          f"dQ {np.abs(Q - Qo).max():.2e}")
    assert abs(L - Lo) / abs(Lo) < 1e-6
    assert np.abs(Q - Qo).max() < 5e-3


def test_demo_sqp_projection_matches_oracle(cuda, fixtures):
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    P = fixtures["gpu_brqn_brqn_P"]
    Q0 = _random_q(*fixtures["gpu_brqn_brqn_Q"].shape, seed=4)
    print("case: demo SQP projection, seed 4", flush=True)
    Q, hist = _sqp(torch.from_numpy(Gp), P, Q0, M, cuda)
    ohist = []
    Qo = O.projection_brqn_gpu(Gp, P, Q0, M, history=ohist)
    print(f"{len(hist)} iterations; dQ {np.abs(Q - Qo).max():.2e}")
    assert [it for it, _ in hist] == [it for it, _ in ohist]
    np.testing.assert_allclose([L for _, L in hist], [L for _, L in ohist], rtol=1e-10)
    assert np.abs(Q - Qo).max() < 1e-7


def test_projection_maximises_logl_at_scale(cuda):
    """N=4000, M=40000, K=6 synthetic genotypes, P = the generating P: both
    projections beat the generating Q, agree, and satisfy KKT."""
    from adamixture_amd.synth import Workload, make_genotypes
    w = Workload("proj", 4000, 40000, 6, "bn", 0.3, 0.01, fst=0.1, seed=91)
    G, P0, Qtrue = make_genotypes(w, cuda)
    P = np.clip(P0, 1e-5, 1 - 1e-5)
    Q0 = _random_q(w.N, w.K, seed=5)
    print("case: synthetic projection N=4000 M=40000 K=6", flush=True)
    Qa, ha, _ = _adam(G, P, Q0, w.M, cuda)
    Qb, hb = _sqp(G, P, Q0, w.M, cuda)
    Pd = torch.from_numpy(P).to(cuda)
    L_true = loglikelihood_gpu(G, Pd, torch.from_numpy(Qtrue).to(cuda), w.M, w.N)
    print(f"Adam {len(ha)} checks L={ha[-1][1]:.3f}; SQP {len(hb)} it L={hb[-1][1]:.3f}; L(Q_true)={L_true:.3f}; "
          f"|Qa-Qb| {np.abs(Qa - Qb).max():.2e}")
    assert hb[-1][1] >= L_true and ha[-1][1] >= L_true
    assert abs(ha[-1][1] - hb[-1][1]) < 1.0
    assert np.abs(Qa - Qb).max() < 5e-3
    _, z = grad_hess(G, torch.from_numpy(Qb).to(cuda), Pd, "q")
    z = z.cpu().numpy()
    inter = Qb > 2e-5
    lam = np.array([z[i][inter[i]].mean() for i in range(w.N)])
    scale = np.abs(z).mean()
    assert np.max(np.abs(np.where(inter, z - lam[:, None], 0.0))) / scale < 1e-4
    assert np.max(np.where(~inter, z - lam[:, None], -np.inf)) / scale < 1e-4
