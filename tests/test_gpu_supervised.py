"""Supervised mode on the GPU against the oracle (supervised.py:84-107,
:308-339, :342-462, :602-781).  The reference ships no labels file, so the
demo samples get synthetic labels: the argmax of the reference's own
converged Q where it exceeds 0.9 (93 of 105 samples), else unlabelled."""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import supervised as S
from adamixture_amd.genotypes import as_packed

pytestmark = pytest.mark.gpu

HP = dict(lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, check=5, lr_decay=0.5, min_lr=1e-4, patience=3,
          tol_adam=0.1)


def _labels(fixtures):
    Qf = fixtures["gpu_adamem_adamem_Q"]
    return np.where(Qf.max(axis=1) > 0.9, Qf.argmax(axis=1) + 1, 0).astype(np.int32)


def _start(fixtures, y, K):
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    P0 = O.init_p_supervised(Gp, y, K, M)
    Q0 = np.random.default_rng(1).random((len(y), K))
    O.init_q_supervised(Q0, y, K)
    return Gp, M, P0, Q0


@pytest.mark.parametrize("K", [7, 8])
def test_label_freq_bit_exact_on_demo(cuda, fixtures, K):
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    y = _labels(fixtures)                              # K = 8: population 8 has no labelled sample
    print(f"case: label_freq demo K={K}", flush=True)
    got = S.init_p_supervised_packed(torch.from_numpy(Gp), y, K, M, device=cuda)
    np.testing.assert_array_equal(got, O.init_p_supervised(Gp, y, K, M))


@pytest.mark.parametrize("N,M,K,frac", [(3000, 20001, 5, 0.6), (130, 7, 3, 1.0), (257, 1030, 4, 0.0)])
def test_label_freq_bit_exact_synthetic(cuda, N, M, K, frac):
    """Odd M (partial last byte row), N not a multiple of 64, all / no
    samples labelled, uneven population sizes."""
    rng = np.random.default_rng(N + M)
    G = rng.integers(0, 4, size=(M, N), dtype=np.uint8)
    Gp = O.pack(G)
    y = np.where(rng.random(N) < frac, rng.integers(1, K + 1, size=N), 0).astype(np.int32)
    print(f"case: label_freq synthetic N={N} M={M} K={K} labelled={int((y > 0).sum())}", flush=True)
    got = S.label_freq_gpu(as_packed(Gp, M, cuda), y, K).cpu().numpy()
    np.testing.assert_array_equal(got, O.init_p_supervised(Gp, y, K, M))


def test_snap_matches_oracle(cuda):
    rng = np.random.default_rng(3)
    N, K = 1000, 6
    y = np.where(rng.random(N) < 0.5, rng.integers(1, K + 1, size=N), 0).astype(np.int32)
    Q = rng.random((N, K))
    Qd = torch.from_numpy(Q.copy()).to(cuda)
    S.Labels(y, K, cuda).snap(Qd)
    O.snap_q(Q, y, K)
    np.testing.assert_array_equal(Qd.cpu().numpy(), Q)


def _floor(Gp, P0, Q0, M, y, Po, Qo, ohist):
    out = {"P": 0.0, "Q": 0.0, "ll": 0.0}
    L0 = np.array([L for _, L in ohist])
    for seed in (0, 1):
        h = []
        Pp, Qp = O.adamem_gpu(Gp, P0, Q0, M, history=h, noise=(seed, 1e-16), labels=y)
        L1 = np.array([L for _, L in h])
        n = min(len(L0), len(L1))
        out["P"] = max(out["P"], np.abs(Pp - Po).max())
        out["Q"] = max(out["Q"], np.abs(Qp - Qo).max())
        out["ll"] = max(out["ll"], float(np.max(np.abs(L1[:n] - L0[:n]) / np.abs(L0[:n]))))
    return out


def test_demo_supervised_adamem_f64_matches_oracle(cuda, fixtures):
    K = 7
    y = _labels(fixtures)
    Gp, M, P0, Q0 = _start(fixtures, y, K)
    print("case: demo supervised Adam-EM f64", flush=True)
    hist = []
    P, Q = S.optimize_supervised_gpu(torch.from_numpy(Gp), torch.from_numpy(P0).to(cuda),
                                     torch.from_numpy(Q0).to(cuda), y, HP["lr"], HP["beta1"], HP["beta2"],
                                     HP["reg_adam"], 10000, HP["check"], M, HP["lr_decay"], HP["min_lr"],
                                     HP["patience"], HP["tol_adam"], cuda, 8192, 256,
                                     on_check=lambda it, L: hist.append((it, L)))
    P, Q = P.cpu().numpy(), Q.cpu().numpy()
    ohist = []
    Po, Qo = O.adamem_gpu(Gp, P0, Q0, M, history=ohist, labels=y)
    floor = _floor(Gp, P0, Q0, M, y, Po, Qo, ohist)
    print(f"{len(hist)} checks; floor dP {floor['P']:.2e} dQ {floor['Q']:.2e} ll {floor['ll']:.2e}; "
          f"GPU dP {np.abs(P - Po).max():.2e} dQ {np.abs(Q - Qo).max():.2e}")
    assert [it for it, _ in hist] == [it for it, _ in ohist]
    np.testing.assert_allclose([L for _, L in hist], [L for _, L in ohist], rtol=max(1e-11, 10 * floor["ll"]))
    assert np.abs(P - Po).max() <= 10 * floor["P"] + 1e-12
    assert np.abs(Q - Qo).max() <= 10 * floor["Q"] + 1e-12
    lab = y > 0
    np.testing.assert_array_equal(Q[lab], Q0[lab])     # labelled rows stay snapped


def test_demo_supervised_sqp_matches_oracle(cuda, fixtures):
    K = 7
    y = _labels(fixtures)
    Gp, M, P0, Q0 = _start(fixtures, y, K)
    N = len(y)
    print("case: demo supervised SQP + ZAL", flush=True)
    hist = []
    P, Q = S.optimize_supervised_original_gpu(torch.from_numpy(Gp), torch.from_numpy(P0).to(cuda),
                                              torch.from_numpy(Q0).to(cuda), y, 10000, K, M, N, 0.1, 3, 3, cuda,
                                              8192, 256, on_iter=lambda it, L: hist.append(L))
    P, Q = P.cpu().numpy(), Q.cpu().numpy()
    ohist = []
    Po, Qo = O.brqn(Gp, P0, Q0, M, history=ohist, gpu_semantics=True, labels=y)
    print(f"{len(hist)} iterations (oracle {len(ohist)}); dP {np.abs(P - Po).max():.2e} dQ {np.abs(Q - Qo).max():.2e}")
    assert len(hist) == len(ohist)
    np.testing.assert_allclose(hist, [L for _, L in ohist], rtol=1e-10)
    np.testing.assert_allclose(P, Po, atol=1e-6)
    np.testing.assert_allclose(Q, Qo, atol=1e-6)
    lab = y > 0
    np.testing.assert_array_equal(Q[lab], Q0[lab])

