"""The drop-in driver end to end: setup -> train_k (and initialize_k) as the
reference's src/main.py:59-82 calls them (adamixture/src/adamixture.py:22-352),
on the reference's demo data (N=105, M=8451, K=7, seed 42, the settings of
tests/config.py), for Adam-EM and BR-QN, ALS and `--init em` starts, and a
K sweep over one K_max SVD.

Checks, per path:
  * the driver's result equals the module calls it stands for (ALS_gpu on
    the sliced SVD, then optimize_parameters_gpu -- bit for bit -- or
    optimize_original_gpu -- to 1e-6, its gradient passes add split partials
    with atomics), so the pinned kernel and optimizer tests apply;
  * the stages against the oracle or the reference's fixtures: frequencies
    bit-exact, SVD and ALS against the CPU fixtures (the GPU path draws its
    test matrix and ALS start with numpy as the CPU path does), the
    `--init em` start (numpy-seeded random P, Q and 5 EM steps,
    adamixture.py:140-182) against the oracle to 1e-10, and BR-QN from it
    against the oracle's trajectory;
  * the converged model against the reference's outputs at the reference's
    own tolerance (tests/config.py:24-26: 2e-2 after cluster alignment, logL
    within 50).
"""
import numpy as np
import pytest
import torch
from scipy.optimize import linear_sum_assignment

import oracle as O
from adamixture_amd import reader
from adamixture_amd.adamixture import initialize_em_gpu, initialize_k, setup, train_k
from adamixture_amd.als_gpu import ALS_gpu
from adamixture_amd.br_qn_gpu import optimize_original_gpu
from adamixture_amd.em_adam_gpu import optimize_parameters_gpu
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

SEED, POWER, TOL_SVD, CHUNK, MAX_ALS, TOL_ALS = 42, 5, 0.1, 8192, 1000, 1e-4
ADAM = (0.005, 0.80, 0.88, 1e-8)          # lr, beta1, beta2, reg_adam
MAX_ITER, CHECK, LR_DECAY, MIN_LR, PATIENCE, TOL = 10000, 5, 0.5, 1e-4, 3, 0.1


@pytest.fixture(scope="module")
def demo():
    Gt, N, M = reader.read_bed_packed(GOLDEN / "demo_data.bed")
    return Gt, N, M


def _train(G, N, M, K, U, S, V, f, dev, tpb, **kw):
    lr, b1, b2, eps = ADAM
    return train_k(G, N, M, K, U, S, V, f, SEED, lr, b1, b2, eps, kw.pop("max_iter", MAX_ITER), CHECK, MAX_ALS,
                   TOL_ALS, LR_DECAY, MIN_LR, CHUNK, PATIENCE, TOL, dev, tpb, **kw)


def _aligned(Qe, P, Q):
    """Columns of (P, Q) matched to Qe's clusters (Hungarian on |Q - Qe|)."""
    cost = np.abs(Qe[:, :, None] - Q[:, None, :]).sum(axis=0)
    _, perm = linear_sum_assignment(cost)
    return P[:, perm], Q[:, perm]


def _matches_fixture(fixtures, P, Q, Pe, Qe):
    P, Q = _aligned(Qe, P.cpu().numpy(), Q.cpu().numpy())
    assert np.isfinite(P).all() and P.min() >= 0 and P.max() <= 1
    np.testing.assert_allclose(Q.sum(1), 1.0, atol=1e-9)
    np.testing.assert_allclose(P, Pe, rtol=2e-2, atol=2e-2)
    np.testing.assert_allclose(Q, Qe, rtol=2e-2, atol=2e-2)
    G = fixtures["reader_G"]
    assert abs(O.loglik_cpu(G, P, Q) - O.loglik_cpu(G, Pe, Qe)) <= 50.0


def _same_brqn(P, Q, P2, Q2):
    """BR-QN's gradient / Hessian passes add their split partials with FP64
    atomics (order varies run to run), so two identical runs agree to the
    tolerance of the oracle-trajectory test (P, Q within 1e-6), not bit for
    bit."""
    np.testing.assert_allclose(P.cpu().numpy(), P2.cpu().numpy(), atol=1e-6)
    np.testing.assert_allclose(Q.cpu().numpy(), Q2.cpu().numpy(), atol=1e-6)


def _canon(U, V):
    U, V = U.copy(), V.copy()
    for c in range(U.shape[1]):
        p = np.argmax(np.abs(U[:, c]))
        if U[p, c] < 0:
            U[:, c] *= -1
            V[:, c] *= -1
    return U, V


def test_driver_adamem_als(cuda, demo, fixtures):
    Gt, N, M = demo
    K = 7
    dev, tpb, f, U, S, V, G = setup(Gt, N, M, K, SEED, POWER, TOL_SVD, CHUNK, "cuda:0")
    assert np.array_equal(f.cpu().numpy(), O.freq_gpu(fixtures["reader_G_packed"], M))
    Uc, Vc = _canon(U.cpu().numpy(), V.cpu().numpy())
    np.testing.assert_allclose(S.cpu().numpy(), fixtures["cpu_brqn_svd_S"], rtol=1e-4)
    np.testing.assert_allclose(Uc, fixtures["cpu_brqn_svd_U"], atol=1e-4)
    np.testing.assert_allclose(Vc, fixtures["cpu_brqn_svd_V"], atol=1e-4)
    P, Q = _train(G, N, M, K, U, S, V, f, dev, tpb)
    # the module calls the driver stands for, on the same inputs
    P_als, Q_als = ALS_gpu(U.contiguous(), S.contiguous(), V.contiguous(), f, SEED, M, K, MAX_ALS, TOL_ALS, dev)
    np.testing.assert_allclose(P_als.cpu().numpy(), fixtures["cpu_brqn_als_P"], atol=1e-6)
    lr, b1, b2, eps = ADAM
    P2, Q2 = optimize_parameters_gpu(G, P_als, Q_als, lr, b1, b2, eps, MAX_ITER, CHECK, K, M, N, LR_DECAY, MIN_LR,
                                     PATIENCE, TOL, dev, CHUNK, tpb)
    assert torch.equal(P, P2) and torch.equal(Q, Q2)
    # the CPU reference ran from this same (numpy-seeded) ALS start
    _matches_fixture(fixtures, P, Q, fixtures["cpu_adamem_adamem_P"], fixtures["cpu_adamem_adamem_Q"])


def test_driver_brqn_als(cuda, demo, fixtures):
    Gt, N, M = demo
    K = 7
    dev, tpb, f, U, S, V, G = setup(Gt, N, M, K, SEED, POWER, TOL_SVD, CHUNK, "cuda:0", original=True,
                                    init_original="als")
    assert U is not None and U.shape == (M, K)
    P, Q = _train(G, N, M, K, U, S, V, f, dev, tpb, original=True, rtol=TOL, Q_hist=3, init_original="als")
    P_als, Q_als = ALS_gpu(U.contiguous(), S.contiguous(), V.contiguous(), f, SEED, M, K, MAX_ALS, TOL_ALS, dev)
    P2, Q2 = optimize_original_gpu(G, P_als, Q_als, MAX_ITER, K, M, N, TOL, 3, PATIENCE, dev, CHUNK, tpb)
    _same_brqn(P, Q, P2, Q2)
    # the GPU optimizer's own semantics (quasi-Newton point mapped by clamp and
    # row renormalisation, br_qn_gpu.py:47-61) restated on the oracle, from
    # the same ALS start: same iterations, logL to 1e-10
    hist = []
    optimize_original_gpu(G, P_als, Q_als, MAX_ITER, K, M, N, TOL, 3, PATIENCE, dev, CHUNK, tpb,
                          on_iter=lambda it, L: hist.append(L))
    ohist = []
    Po, Qo = O.brqn(fixtures["reader_G_packed"], P_als.cpu().numpy(), Q_als.cpu().numpy(), M, history=ohist,
                    gpu_semantics=True)
    assert len(hist) == len(ohist)
    np.testing.assert_allclose(hist, [L for _, L in ohist], rtol=1e-10)
    np.testing.assert_allclose(P.cpu().numpy(), Po, atol=1e-6)
    # the CPU reference ran from this ALS start with its own quasi-Newton map
    # (box clip + simplex projection, br_qn.py:159-160): same likelihood to the
    # reference's tolerance, though 0.5% of P entries (near the bounds) settle
    # up to 0.07 apart
    Gu = fixtures["reader_G"]
    Pn, Qn = _aligned(fixtures["cpu_brqn_brqn_Q"], P.cpu().numpy(), Q.cpu().numpy())
    assert abs(O.loglik_cpu(Gu, Pn, Qn) - O.loglik_cpu(Gu, fixtures["cpu_brqn_brqn_P"], fixtures["cpu_brqn_brqn_Q"])) <= 50


def test_driver_brqn_em_init(cuda, demo, fixtures):
    """`--init em` (the reference's default for BR-QN): setup skips the SVD,
    train_k starts from numpy-seeded random P, Q refined by 5 plain EM steps."""
    Gt, N, M = demo
    K = 7
    Gp = fixtures["reader_G_packed"]
    dev, tpb, f, U, S, V, G = setup(Gt, N, M, K, SEED, POWER, TOL_SVD, CHUNK, "cuda:0", original=True,
                                    init_original="em")
    assert f is None and U is None and S is None and V is None
    P0, Q0 = initialize_em_gpu(G, SEED, M, N, K, dev, CHUNK, tpb, 5)
    # the start restated on the oracle: adamixture.py:155-179
    rng = np.random.default_rng(SEED)
    Po = np.clip(rng.random(size=(M, K)), 1e-5, 1 - 1e-5)
    Qo = np.clip(rng.random(size=(N, K)), 1e-5, 1 - 1e-5)
    Qo /= Qo.sum(axis=1, keepdims=True)
    for _ in range(5):
        Po, Qo = O.em_final_gpu(Po, Qo, *O.em_stats_gpu(Gp, Po, Qo, M))
    np.testing.assert_allclose(P0.cpu().numpy(), Po, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(Q0.cpu().numpy(), Qo, rtol=1e-10, atol=1e-13)
    P, Q = _train(G, N, M, K, U, S, V, f, dev, tpb, original=True, rtol=TOL, Q_hist=3, init_original="em")
    P2, Q2 = optimize_original_gpu(G, P0, Q0, MAX_ITER, K, M, N, TOL, 3, PATIENCE, dev, CHUNK, tpb)
    _same_brqn(P, Q, P2, Q2)
    hist = []
    P3, Q3 = optimize_original_gpu(G, P0, Q0, MAX_ITER, K, M, N, TOL, 3, PATIENCE, dev, CHUNK, tpb,
                                   on_iter=lambda it, L: hist.append(L))
    ohist = []
    O.brqn(Gp, P0.cpu().numpy(), Q0.cpu().numpy(), M, history=ohist, gpu_semantics=True)
    print(f"em-init brqn: iterations gpu {len(hist)} oracle {len(ohist)}, logL {hist[-1]:.6f} vs {ohist[-1][1]:.6f}")
    assert len(hist) == len(ohist)
    np.testing.assert_allclose(hist, [L for _, L in ohist], rtol=1e-10)


def test_driver_k_sweep_slices_the_k_max_svd(cuda, demo, fixtures):
    """One SVD at K_max = 7, then K = 4 and K = 7 from its leading columns
    (adamixture.py:295-298), each equal to ALS on the slice + the optimizer;
    initialize_k instead runs its own rank-K SVD."""
    Gt, N, M = demo
    dev, tpb, f, U, S, V, G = setup(Gt, N, M, 7, SEED, POWER, TOL_SVD, CHUNK, "cuda:0")
    lr, b1, b2, eps = ADAM
    for K in (4, 7):
        P, Q = _train(G, N, M, K, U, S, V, f, dev, tpb, max_iter=30)
        assert P.shape == (M, K) and Q.shape == (N, K)
        P_als, Q_als = ALS_gpu(U[:, :K].contiguous(), S[:K].contiguous(), V[:, :K].contiguous(), f, SEED, M, K,
                               MAX_ALS, TOL_ALS, dev)
        P2, Q2 = optimize_parameters_gpu(G, P_als, Q_als, lr, b1, b2, eps, 30, CHECK, K, M, N, LR_DECAY, MIN_LR,
                                         PATIENCE, TOL, dev, CHUNK, tpb)
        assert torch.equal(P, P2) and torch.equal(Q, Q2)
    P4, Q4, ll4 = initialize_k(G, N, M, 4, f, SEED, POWER, TOL_SVD, MAX_ALS, TOL_ALS, CHUNK, dev, tpb)
    assert P4.shape == (M, 4) and np.isfinite(ll4)
    assert ll4 == pytest.approx(O.loglik_gpu(fixtures["reader_G_packed"], P4.cpu().numpy(), Q4.cpu().numpy(), M),
                                rel=1e-12)


def test_driver_learned_warm_start(cuda, demo, fixtures):
    """train_k(..., warm_start="learned") (project extension): K = 3 takes the
    leading 2K = 6 components of the K_max = 7 SVD, K = 7 runs its own rank-14
    SVD; each equals learned_warm_start + the optimizer on the same inputs,
    and converges to a model within the reference's tolerance of the plain
    start's (logL within 50, adamixture tests/config.py:24-26)."""
    from adamixture_amd.genotypes import as_packed
    from adamixture_amd.learned_init import WarmStartMLP, learned_warm_start
    from adamixture_amd.svd_gpu import SVD_gpu
    Gt, N, M = demo
    dev, tpb, f, U, S, V, G = setup(Gt, N, M, 7, SEED, POWER, TOL_SVD, CHUNK, "cuda:0")
    lr, b1, b2, eps = ADAM
    model = WarmStartMLP.load(dev)
    Gp = fixtures["reader_G_packed"]
    for K in (3, 7):
        print(f"case: learned warm start K={K}", flush=True)
        P, Q = _train(G, N, M, K, U, S, V, f, dev, tpb, warm_start="learned")
        if 2 * K <= U.shape[1]:
            U2, S2, V2 = U[:, : 2 * K].contiguous(), S[: 2 * K].contiguous(), V[:, : 2 * K].contiguous()
        else:
            U2, S2, V2 = SVD_gpu(G, N, M, f, 2 * K, SEED, POWER, TOL_SVD, CHUNK, dev, tpb)
        P0, Q0, _, used = learned_warm_start(as_packed(G, M, dev), U2, S2, V2, f, K, model, SEED)
        P2, Q2 = optimize_parameters_gpu(G, P0.contiguous(), Q0.contiguous(), lr, b1, b2, eps, MAX_ITER, CHECK, K, M,
                                         N, LR_DECAY, MIN_LR, PATIENCE, TOL, dev, CHUNK, tpb)
        assert torch.equal(P, P2) and torch.equal(Q, Q2)
        Pn, Qn = P.cpu().numpy(), Q.cpu().numpy()
        assert np.isfinite(Pn).all() and Pn.min() >= 1e-5 - 1e-15 and Pn.max() <= 1 - 1e-5 + 1e-15
        np.testing.assert_allclose(Qn.sum(1), 1.0, atol=1e-9)
        Pp, Qp = _train(G, N, M, K, U, S, V, f, dev, tpb)
        ll_w = O.loglik_gpu(Gp, Pn, Qn, M)
        ll_p = O.loglik_gpu(Gp, Pp.cpu().numpy(), Qp.cpu().numpy(), M)
        print(f"K={K}: learned start used {used}; converged logL {ll_w:.2f} vs plain {ll_p:.2f}", flush=True)
        assert abs(ll_w - ll_p) <= 50.0
