"""Pins of the CPU oracle (no GPU).

Each test checks an oracle function against known-answer data the reference's
own tests hold (tests/golden, copied from adamixture/demo) or against a
property the mathematics fixes, chosen so that a plausible slip (dropped
term, wrong sign or index, transposed operand) fails.  The planted-mistake
tests show that the brute-force checks do catch such slips.
"""
import numpy as np
import pytest
from scipy.optimize import lsq_linear

import oracle as O
from tests.conftest import GOLDEN


def _demo_bed():
    N = sum(1 for _ in open(GOLDEN / "demo_data.fam"))
    raw = np.fromfile(GOLDEN / "demo_data.bed", dtype=np.uint8)
    nb = (N + 3) // 4
    M = (raw.size - 3) // nb
    return raw[3:].reshape(M, nb), M, N


# ---------------------------------------------------------------- layout ----

def test_reader_fixtures_exact(fixtures):
    bed, M, N = _demo_bed()
    G = O.bed_to_unpacked(bed, M, N)
    Gp = O.bed_to_packed(bed, M, N)
    assert O.mean_packed(Gp, M) < 0.5          # demo data is not flipped
    assert np.array_equal(G, fixtures["reader_G"])
    assert np.array_equal(Gp, fixtures["reader_G_packed"])
    assert np.array_equal(O.unpack(Gp, M), G)
    assert np.array_equal(O.pack(G), Gp)


def test_package_reader_matches_fixture(fixtures, tmp_path):
    from adamixture_amd import reader
    Gp, N, M = reader.read_bed_packed(GOLDEN / "demo_data.bed")
    assert (N, M) == (105, 8451)
    assert np.array_equal(Gp.numpy(), fixtures["reader_G_packed"])
    # write/read round trip of a flipped, ragged matrix (M % 4 = 3, N % 4 = 1)
    rng = np.random.default_rng(1)
    G = rng.choice(4, size=(23, 13), p=[0.1, 0.2, 0.6, 0.1]).astype(np.uint8)
    reader.write_bed(tmp_path / "x", G)
    Gp2, N2, M2 = reader.read_bed_packed(tmp_path / "x.bed")
    expect = np.where(G == 3, 3, 2 - G).astype(np.uint8)        # mean >= 0.5 -> flipped
    assert (N2, M2) == (13, 23)
    assert np.array_equal(O.unpack(Gp2.numpy(), M2), expect)
    assert np.all((Gp2.numpy()[-1] >> 6) == 0)                   # padding field zeroed


def test_unpack_bruteforce():
    rng = np.random.default_rng(0)
    M, N = 37, 29
    Gp = rng.integers(0, 256, size=((M + 3) // 4, N), dtype=np.uint8)
    want = np.array([[(Gp[m // 4, i] >> (2 * (m % 4))) & 3 for i in range(N)] for m in range(M)], np.uint8)
    assert np.array_equal(O.unpack(Gp, M), want)
    assert np.array_equal(O.unpack(Gp, M, 5, 11), want[5:16])
    f = rng.random(M).astype(np.float32)
    c = O.unpack_center(Gp, f, M, 3, 40)
    ref = np.where(want[3:] == 3, 0.0, want[3:].astype(np.float32) - 2.0 * f[3:, None]).astype(np.float32)
    assert np.array_equal(c, ref)


def test_flip_and_mean_closed_form():
    rng = np.random.default_rng(2)
    M, N = 22, 17
    G = rng.integers(0, 4, size=(M, N)).astype(np.uint8)
    Gp = O.pack(G)
    obs = G != 3
    assert O.mean_packed(Gp, M) == pytest.approx(G[obs].sum() / (2 * obs.sum()), rel=0, abs=1e-15)
    Gp_pad = Gp.copy()
    Gp_pad[-1] |= 0b11110000          # garbage in the padding fields (M % 4 = 2)
    O.flip_packed(Gp_pad, M)
    assert np.array_equal(O.unpack(Gp_pad, M), np.where(G == 3, 3, 2 - G))
    assert np.all(Gp_pad[-1] >> 4 == 0)


def test_freq_and_qbat_closed_form():
    rng = np.random.default_rng(3)
    M, N = 41, 33
    G = rng.choice(4, size=(M, N), p=[0.3, 0.3, 0.3, 0.1]).astype(np.uint8)
    G[7] = 3                                   # an all-missing SNP
    Gp = O.pack(G)
    obs = G != 3
    s = np.where(obs, G, 0).sum(1)
    c = obs.sum(1)
    want = np.where(c > 0, s / np.maximum(2 * c, 1), 0).astype(np.float32)
    assert np.array_equal(O.freq_gpu(Gp, M), want)
    f_cpu = O.freq_cpu(G)
    assert np.isnan(f_cpu[7]) and np.allclose(np.delete(f_cpu, 7), np.delete(want, 7))
    assert np.array_equal(O.qbat(Gp, M), 2.0 * obs.sum(0))


def test_synth_moments():
    rng = np.random.default_rng(4)
    M, N, K = 64, 2000, 3
    P = rng.uniform(0.05, 0.95, size=(M, K))
    Q = rng.dirichlet(np.ones(K), size=N)
    Gp = O.synth_packed(P, Q, seed=42, miss_rate=0.02)
    G = O.unpack(Gp, M)
    miss = G == 3
    assert abs(miss.mean() - 0.02) < 0.004
    H = P @ Q.T
    z = (np.where(miss, 0, G) - 2 * H * ~miss).sum() / np.sqrt((2 * H * (1 - H) * ~miss).sum())
    assert abs(z) < 4.0
    # a block regenerated on its own is bit-identical
    assert np.array_equal(O.synth_packed(P, Q, 42, 0.02, row0=5, rows=3), Gp[5:8])


# ------------------------------------------------------------ Adam-EM -------

def _tiny(seed=5, M=9, N=7, K=3, miss=0.15):
    rng = np.random.default_rng(seed)
    P = rng.uniform(0.02, 0.98, size=(M, K))
    Q = rng.dirichlet(np.ones(K), size=N)
    G = rng.choice(4, size=(M, N), p=[(1 - miss) / 3] * 3 + [miss]).astype(np.uint8)
    return G, P, Q


def _brute_stats(G, P, Q, drop_colsum=False):
    M, K = P.shape
    N = Q.shape[0]
    A = np.zeros((M, K)); B = np.zeros((M, K)); T = np.zeros((N, K)); qb = np.zeros(N)
    for m in range(M):
        for i in range(N):
            g = int(G[m, i])
            if g == 3:
                continue
            h = min(max(sum(P[m, k] * Q[i, k] for k in range(K)), 1e-5), 1 - 1e-5)
            a, b = g / h, (2 - g) / (1 - h)
            for k in range(K):
                A[m, k] += a * Q[i, k]
                B[m, k] += b * Q[i, k]
                T[i, k] += (a - b) * P[m, k] + (0.0 if drop_colsum else b)
            qb[i] += 2
    return A, B, T, qb


def test_em_stats_bruteforce():
    G, P, Q = _tiny()
    A, B, T, qb = O.em_stats_gpu(O.pack(G), P, Q, G.shape[0])
    Ab, Bb, Tb, qbb = _brute_stats(G, P, Q)
    for x, y in ((A, Ab), (B, Bb), (T, Tb), (qb, qbb)):
        np.testing.assert_allclose(x, y, rtol=1e-13, atol=0)


def test_planted_mistake_caught():
    G, P, Q = _tiny()
    _, _, T, _ = O.em_stats_gpu(O.pack(G), P, Q, G.shape[0])
    _, _, Tb, _ = _brute_stats(G, P, Q, drop_colsum=True)
    assert not np.allclose(T, Tb, rtol=1e-6)


def test_em_final_closed_form():
    G, P, Q = _tiny(K=4)
    A, B, T, qb = O.em_stats_gpu(O.pack(G), P, Q, G.shape[0])
    PE, QE = O.em_final_gpu(P, Q, A, B, T, qb)
    np.testing.assert_allclose(PE, A * P / ((A - B) * P + B), rtol=1e-14)
    QEb = Q * T / qb[:, None]
    np.testing.assert_allclose(QE, QEb / QEb.sum(1, keepdims=True), rtol=1e-14)
    # EM identity: sum_k Q_ik T_ik = q_bat_i when no clamp is active
    np.testing.assert_allclose((Q * T).sum(1), qb, rtol=1e-12)


def test_loglik_bruteforce_and_binomial():
    from scipy.stats import binom
    G, P, Q = _tiny(M=11, N=8)
    Gp = O.pack(G)
    H = P @ Q.T
    obs = G != 3
    want = np.sum(np.where(obs, G * np.log(H) + (2 - G) * np.log1p(-H), 0.0))
    assert O.loglik_gpu(Gp, P, Q, G.shape[0]) == pytest.approx(want, rel=1e-14)
    assert O.loglik_cpu(G, P, Q) == pytest.approx(want, rel=1e-14)
    # equals the binomial log-pmf minus the log binomial coefficients
    lp = np.sum(np.where(obs, binom.logpmf(np.where(obs, G, 0), 2, H) - np.log(np.where(G == 1, 2.0, 1.0)), 0))
    assert want == pytest.approx(lp, rel=1e-12)


def test_plain_em_ascends():
    rng = np.random.default_rng(6)
    M, N, K = 200, 60, 3
    P = rng.uniform(0.05, 0.95, size=(M, K))
    Q = rng.dirichlet(np.ones(K), size=N)
    Gp = O.synth_packed(P, Q, 7, 0.05)
    P0 = np.clip(rng.random((M, K)), 1e-5, 1 - 1e-5)
    Q0 = rng.dirichlet(np.ones(K), size=N)
    last = O.loglik_gpu(Gp, P0, Q0, M)
    for _ in range(15):
        A, B, T, qb = O.em_stats_gpu(Gp, P0, Q0, M)
        P0, Q0 = O.em_final_gpu(P0, Q0, A, B, T, qb)
        cur = O.loglik_gpu(Gp, P0, Q0, M)
        assert cur >= last - 1e-9 * abs(last)
        last = cur


def test_adamem_gpu_semantics_matches_reference_fixture(fixtures):
    """outputs/gpu/adamem: the CUDA FP64 run from its own ALS init."""
    G, Gp = fixtures["reader_G"], fixtures["reader_G_packed"]
    M = G.shape[0]
    P, Q = O.adamem_gpu(Gp, fixtures["gpu_adamem_als_P"], fixtures["gpu_adamem_als_Q"], M)
    Pe, Qe = fixtures["gpu_adamem_adamem_P"], fixtures["gpu_adamem_adamem_Q"]
    np.testing.assert_allclose(P, Pe, rtol=2e-2, atol=2e-2)
    np.testing.assert_allclose(Q, Qe, rtol=2e-2, atol=2e-2)
    assert abs(O.loglik_cpu(G, P, Q) - O.loglik_cpu(G, Pe, Qe)) <= 50.0
    assert np.abs(P - Pe).max() < 5e-3        # much tighter than the reference's 2e-2


def test_adamem_cpu_semantics_matches_reference_fixture(fixtures):
    G = fixtures["reader_G"]
    P, Q = O.adamem_cpu(G, fixtures["cpu_brqn_als_P"], fixtures["cpu_brqn_als_Q"])
    Pe, Qe = fixtures["cpu_adamem_adamem_P"], fixtures["cpu_adamem_adamem_Q"]
    np.testing.assert_allclose(P, Pe, rtol=2e-2, atol=2e-2)
    np.testing.assert_allclose(Q, Qe, rtol=2e-2, atol=2e-2)
    assert abs(O.loglik_cpu(G, P, Q) - O.loglik_cpu(G, Pe, Qe)) <= 50.0


# ------------------------------------------------------------ Adam step -----

ADAM_SETTINGS = [(0.005, 0.80, 0.88, 1e-8),      # the reference's defaults (entry.py:35-47)
                 (0.05, 0.70, 0.95, 1e-3)]


def _oracle_adam_gpu(P0, Q0, P_t, Q_t, lr, b1, b2, eps):
    M, K = P0.shape
    opt = O.AdamEMGPU(M, Q0.shape[0], K, lr, b1, b2, eps)
    P, Q = P0.copy(), Q0.copy()
    for tp, tq in zip(P_t, Q_t):
        opt.step(P, Q, tp, tq)
    return P, Q, opt.mP, opt.vP, opt.mQ, opt.vQ


def _oracle_adam_cpu(P0, Q0, P_t, Q_t, lr, b1, b2, eps):
    from oracle import oracle as OO
    M, K = P0.shape
    N = Q0.shape[0]
    P, Q = P0.copy(), Q0.copy()
    mP, vP, mQ, vQ = np.zeros_like(P), np.zeros_like(P), np.zeros_like(Q), np.zeros_like(Q)
    for t, (tp, tq) in enumerate(zip(P_t, Q_t), start=1):
        tp, tq = np.ascontiguousarray(tp), np.ascontiguousarray(tq)
        OO.lib().oc_adamP_cpu(OO._p(P), OO._p(tp), OO._p(mP), OO._p(vP), lr, b1, b2, eps, t, M, K)
        OO.lib().oc_adamQ_cpu(OO._p(Q), OO._p(tq), OO._p(mQ), OO._p(vQ), lr, b1, b2, eps, t, N, K)
    return P, Q, mP, vP, mQ, vQ


@pytest.mark.parametrize("semantics", ["gpu", "cpu"])
@pytest.mark.parametrize("setting", ADAM_SETTINGS)
def test_adam_em_step_is_kingma_ba_textbook(semantics, setting):
    """Three Adam-EM steps of the oracle (em_adam_gpu.py:11-42, :201-224 and
    its Cython twin em.pyx:92-209) against Kingma & Ba written out by hand
    (tests/textbook.py: closed-form moment sums, t incremented before use,
    bias corrections, clamp of P and Q, then the Q row renormalisation), on
    a case whose steps cross both clamp bounds.  P, Q and all four moment
    buffers must agree to rounding."""
    from tests import textbook as TB
    lr, b1, b2, eps = setting
    P0, Q0, P_t, Q_t = TB.adam_em_case()
    run = _oracle_adam_gpu if semantics == "gpu" else _oracle_adam_cpu
    got = run(P0, Q0, P_t, Q_t, lr, b1, b2, eps)
    ref = TB.adam_em_steps(P0, Q0, P_t, Q_t, lr, b1, b2, eps)
    for name, g, r in zip(("P", "Q", "mP", "vP", "mQ", "vQ"), got, ref):
        np.testing.assert_allclose(g, np.array(r), rtol=1e-12, atol=1e-15, err_msg=name)
    # the case exercises both bounds and the renormalisation
    assert got[0].max() == TB.HI and got[0].min() == TB.LO
    assert abs(got[1][0].sum() - 1.0) < 1e-15 and got[1][0, 1] < 1e-5


@pytest.mark.parametrize("variant", ["no_bias_correction", "swapped_betas", "renorm_before_clamp",
                                     "eps_inside_sqrt", "t_from_zero_at_step_two"])
def test_adam_em_textbook_check_catches_planted_slips(variant):
    """Each planted slip in the hand-written Adam step moves P or Q far
    outside the 1e-12 tolerance of the check above, so an oracle (or a
    kernel) carrying that slip would fail it."""
    from tests import textbook as TB
    lr, b1, b2, eps = ADAM_SETTINGS[1] if variant == "eps_inside_sqrt" else ADAM_SETTINGS[0]
    P0, Q0, P_t, Q_t = TB.adam_em_case()
    got = _oracle_adam_gpu(P0, Q0, P_t, Q_t, lr, b1, b2, eps)
    bad = TB.adam_em_steps(P0, Q0, P_t, Q_t, lr, b1, b2, eps, variant=variant)
    dev = max(np.abs(got[0] - np.array(bad[0])).max(), np.abs(got[1] - np.array(bad[1])).max())
    assert dev > 1e-9, f"{variant}: deviation {dev:.2e} would pass the 1e-12 check"


# ------------------------------------------------------------ init ----------

def _centred_rows(G):
    f = O.freq_cpu(G)

    def rows(s, n):
        X = G[s:s + n].astype(np.float32) - 2.0 * f[s:s + n, None]
        X[G[s:s + n] == 3] = 0.0
        return X
    return f, rows


def _canon(U, V):
    U, V = U.copy(), V.copy()
    for c in range(U.shape[1]):
        p = np.argmax(np.abs(U[:, c]))
        if U[p, c] < 0:
            U[:, c] *= -1
            V[:, c] *= -1
    return U, V


def test_rsvd_matches_fixture_and_dense_svd(fixtures):
    G = fixtures["reader_G"]
    M, N = G.shape
    f, rows = _centred_rows(G)
    U, S, V = O.rsvd(rows, M, N, 7, 42, 5, 0.1, 8192)
    U, V = _canon(U, V)
    np.testing.assert_allclose(S, fixtures["cpu_brqn_svd_S"], rtol=1e-5)
    np.testing.assert_allclose(U, fixtures["cpu_brqn_svd_U"], atol=1e-5)
    np.testing.assert_allclose(V, fixtures["cpu_brqn_svd_V"], atol=1e-5)
    np.testing.assert_allclose(U.T @ U, np.eye(7), atol=1e-4)
    # with the power iterations run to convergence the result is the dense SVD
    Sd = np.linalg.svd(rows(0, M).astype(np.float64), compute_uv=False)[:7]
    _, S2, _ = O.rsvd(rows, M, N, 7, 42, 40, 0.0, 8192)
    np.testing.assert_allclose(S2, Sd, rtol=1e-4)


def test_als_matches_fixture(fixtures):
    G = fixtures["reader_G"]
    M, N = G.shape
    f = O.freq_cpu(G)
    P, Q = O.als(fixtures["cpu_brqn_svd_U"], fixtures["cpu_brqn_svd_S"], fixtures["cpu_brqn_svd_V"], f, 42,
                 M, N, 7, 1000, 1e-4)
    np.testing.assert_allclose(P, fixtures["cpu_brqn_als_P"], atol=1e-6)
    np.testing.assert_allclose(Q, fixtures["cpu_brqn_als_Q"], atol=1e-6)


@pytest.mark.parametrize("variant", ["cpu", "gpu"])
def test_bvls_kkt_and_lsq_linear(variant):
    rng = np.random.default_rng(8)
    K = 6
    X = rng.normal(size=(30, K))
    A = X.T @ X + 0.1 * np.eye(K)
    B = rng.normal(scale=2.0, size=(40, K)) @ A
    sol = (O.bvls_cpu if variant == "cpu" else O.bvls_gpu)(A, B, 0.0, 1.0)
    L = np.linalg.cholesky(A)
    for r in range(B.shape[0]):
        ref = lsq_linear(L.T, np.linalg.solve(L, B[r]), bounds=(0, 1), method="bvls", tol=1e-12).x
        np.testing.assert_allclose(sol[r], ref, atol=1e-7)
        grad = A @ sol[r] - B[r]
        free = (sol[r] > 1e-9) & (sol[r] < 1 - 1e-9)
        assert np.all(np.abs(grad[free]) < 1e-7)
        assert np.all(grad[sol[r] <= 1e-9] >= -1e-7) and np.all(grad[sol[r] >= 1 - 1e-9] <= 1e-7)


def test_project_simplex_closed_form():
    rng = np.random.default_rng(9)
    for _ in range(50):
        K = rng.integers(2, 12)
        b = rng.normal(size=K)
        got = O.project_simplex(b, 1e-5)
        # Euclidean projection onto {x >= pc, sum x = 1} by bisection on tau
        lo, hi = b.min() - 1.0, b.max()
        for _ in range(200):
            tau = 0.5 * (lo + hi)
            s = np.maximum(b - 1e-5 - tau, 0).sum() + K * 1e-5
            lo, hi = (tau, hi) if s > 1 else (lo, tau)
        np.testing.assert_allclose(got, np.maximum(b - 1e-5 - tau, 0) + 1e-5, atol=1e-10)


def test_sqp_box_qp_is_kkt_optimal():
    rng = np.random.default_rng(10)
    K = 5
    for _ in range(30):
        X = rng.normal(size=(20, K))
        H = X.T @ X + 0.5 * np.eye(K)
        par = rng.uniform(0.05, 0.95, size=K)
        r = rng.normal(scale=3.0, size=K)
        tab = np.zeros((K + 1, K + 1))
        tab[:K, :K] = H
        tab[:K, K] = r
        tab[K, :K] = r
        delta, it = O.qp_sweep(tab, par, K, 0)
        assert it > 0
        x = par + delta
        # maximise r.d - d'Hd/2 s.t. 0 <= par + d <= 1: gradient r - H d
        grad = r - H @ delta
        free = (x > 1e-5 + 1e-7) & (x < 1 - 1e-5 - 1e-7)
        assert np.all(np.abs(grad[free]) < 1e-6)
        assert np.all(x >= -1e-9) and np.all(x <= 1 + 1e-9)


def test_grad_hess_finite_differences():
    rng = np.random.default_rng(11)
    M, N, K = 15, 9, 3
    P = rng.uniform(0.1, 0.9, size=(M, K))
    Q = rng.dirichlet(np.ones(K), size=N)
    Gp = O.synth_packed(P, Q, 3, 0.1)
    Hq, zq = O.grad_hess_q(Gp, Q, P, M)
    Hp, zp = O.grad_hess_p(Gp, Q, P, M)
    eps = 1e-6
    i, k = 4, 1
    Qp, Qm = Q.copy(), Q.copy()
    Qp[i, k] += eps
    Qm[i, k] -= eps
    d = (O.loglik_gpu(Gp, P, Qp, M) - O.loglik_gpu(Gp, P, Qm, M)) / (2 * eps)
    # z_q = sum_j a p_j + b (1 - p_j) = dlogL/dQ_ik + sum_j b_j
    Gu = O.unpack(Gp, M)[:, i].astype(float)
    obs = Gu != 3
    h = P @ Q[i]
    b = np.where(obs, (2 - Gu) / (1 - h), 0.0)
    assert zq[i, k] - b.sum() == pytest.approx(d, rel=1e-5)
    j = 6
    Pp, Pm = P.copy(), P.copy()
    Pp[j, k] += eps
    Pm[j, k] -= eps
    d = (O.loglik_gpu(Gp, Pp, Q, M) - O.loglik_gpu(Gp, Pm, Q, M)) / (2 * eps)
    assert zp[j, k] == pytest.approx(d, rel=1e-5)
    # Hessian: minus the second derivative of logL along P_jk
    e2 = 1e-4
    Pp[j, k] = P[j, k] + e2
    Pm[j, k] = P[j, k] - e2
    d2 = (O.loglik_gpu(Gp, Pp, Q, M) - 2 * O.loglik_gpu(Gp, P, Q, M) + O.loglik_gpu(Gp, Pm, Q, M)) / e2**2
    assert Hp[j, k, k] == pytest.approx(-d2, rel=1e-5)


def test_brqn_matches_reference_fixture(fixtures):
    Gp = fixtures["reader_G_packed"]
    M = fixtures["reader_G"].shape[0]
    P, Q = O.brqn(Gp, fixtures["cpu_brqn_als_P"], fixtures["cpu_brqn_als_Q"], M)
    np.testing.assert_allclose(P, fixtures["cpu_brqn_brqn_P"], atol=1e-8)
    np.testing.assert_allclose(Q, fixtures["cpu_brqn_brqn_Q"], atol=1e-8)


# ------------------------------------------------------------ projection ----

def _kkt_residual(Gp, P, Q, M):
    """Relative KKT residuals of max_Q logL(P, Q) on the box-clamped simplex:
    with z the (positive) Q gradient, interior entries share one multiplier
    per sample and entries at the lower bound may not exceed it."""
    _, z = O.grad_hess_q(Gp, Q, P, M)
    inter = Q > 2e-5
    lam = np.array([z[i][inter[i]].mean() for i in range(Q.shape[0])])
    scale = np.abs(z).mean()
    r_int = np.max(np.abs(np.where(inter, z - lam[:, None], 0.0))) / scale
    r_bnd = np.max(np.where(~inter, z - lam[:, None], -np.inf)) / scale
    return r_int, r_bnd


def test_projection_priming_is_the_q_half_of_adamem(fixtures):
    """With P fixed the first Adam step on Q sees the same EM target as the
    joint Adam-EM priming step (em_adam_gpu.py:311-322 vs projection.py:
    293-300), so both must give the same Q bit for bit."""
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    P0, Q0 = fixtures["gpu_adamem_als_P"], fixtures["gpu_adamem_als_Q"]
    _, Qj = O.adamem_gpu(Gp, P0, Q0, M, max_iter=0)
    Qp = O.projection_gpu(Gp, P0, Q0, M, max_iter=0)
    np.testing.assert_array_equal(Qp, Qj)


def test_projection_reaches_the_concave_maximiser(fixtures):
    """logL is concave in Q for fixed P, so from a random start both
    projection algorithms must reach the same maximiser: KKT conditions hold,
    the two agree, and at the fixture P of the reference's own converged GPU
    run (outputs/gpu/brqn) they reproduce its fixture Q."""
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    P, Qs = fixtures["gpu_brqn_brqn_P"], fixtures["gpu_brqn_brqn_Q"]
    rng = np.random.default_rng(42)
    Q0 = rng.random(Qs.shape)
    Q0 /= Q0.sum(axis=1, keepdims=True)
    ha, hb = [], []
    Qa = O.projection_gpu(Gp, P, Q0, M, history=ha)
    Qb = O.projection_brqn_gpu(Gp, P, Q0, M, history=hb)
    r_int, r_bnd = _kkt_residual(Gp, P, Qb, M)
    assert r_int < 1e-5 and r_bnd < 1e-5
    r_int, r_bnd = _kkt_residual(Gp, P, Qa, M)
    assert r_int < 5e-4 and r_bnd < 5e-4                 # Adam stops at tol 0.1 in logL
    assert np.abs(Qa - Qb).max() < 2e-3
    assert np.abs(Qb - Qs).max() < 1e-3
    Ls = O.loglik_gpu(Gp, P, Qs, M)
    assert hb[-1][1] >= Ls - 1e-3 and ha[-1][1] >= Ls - 0.01
    assert ha[-1][1] > ha[0][1] and len(ha) > 3 and len(hb) > 2
    # the KKT check separates: a maximiser paired with a column-swapped P
    # (a transposed-operand slip) or the random start fails it by far
    assert _kkt_residual(Gp, P[:, ::-1].copy(), Qb, M)[0] > 1e-3
    assert _kkt_residual(Gp, P, Q0, M)[0] > 1e-3


# ------------------------------------------------------------ supervised ----

def _demo_labels(fixtures, frac=0.9):
    """Synthetic labels for the demo samples (the reference ships none): the
    fixture Q's argmax where it exceeds ``frac``, else unlabelled."""
    Qf = fixtures["gpu_adamem_adamem_Q"]
    return np.where(Qf.max(axis=1) > frac, Qf.argmax(axis=1) + 1, 0).astype(np.int32)


def test_init_p_supervised_packed_equals_unpacked_and_bruteforce(fixtures):
    G, Gp = fixtures["reader_G"], fixtures["reader_G_packed"]
    M = G.shape[0]
    K = 8                                             # population 8 has no labelled sample -> 0.5
    y = _demo_labels(fixtures)
    P = O.init_p_supervised(Gp, y, K, M)
    np.testing.assert_array_equal(P, O.init_p_supervised_unpacked(G, y, K, M))
    assert np.all(P[:, 7] == 0.5)
    for m in range(0, M, 397):
        for k in range(7):
            obs = [int(G[m, i]) for i in range(G.shape[1]) if y[i] == k + 1 and G[m, i] != 3]
            want = sum(obs) / (2.0 * len(obs)) if obs else 0.5
            assert P[m, k] == min(max(want, 1e-5), 1 - 1e-5)


def test_snap_init_q_and_labels_closed_form():
    y, ids = O.labels_to_y(["a", "-", "b", "a", "-", "c"])
    assert y.tolist() == [1, 0, 2, 1, 0, 3] and ids == {"a": 1, "b": 2, "c": 3}
    Q = np.random.default_rng(0).random((6, 3))
    O.init_q_supervised(Q, y, 3)
    np.testing.assert_array_equal(Q[0], [1 - 1e-5, 5e-6, 5e-6])
    np.testing.assert_array_equal(Q[5], [5e-6, 5e-6, 1 - 1e-5])
    np.testing.assert_allclose(Q.sum(axis=1), 1.0, rtol=1e-15)


def test_supervised_reduces_to_unsupervised_without_labels(fixtures):
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    P0, Q0 = fixtures["gpu_adamem_als_P"], fixtures["gpu_adamem_als_Q"]
    none = np.zeros(Q0.shape[0], dtype=np.int32)
    for a, b in zip(O.adamem_gpu(Gp, P0, Q0, M, labels=none, max_iter=40), O.adamem_gpu(Gp, P0, Q0, M, max_iter=40)):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(O.brqn(Gp, P0, Q0, M, labels=none, gpu_semantics=True, max_iter=4),
                    O.brqn(Gp, P0, Q0, M, gpu_semantics=True, max_iter=4)):
        np.testing.assert_array_equal(a, b)


def test_supervised_all_labelled_maximises_over_p(fixtures):
    """Every sample labelled: Q never leaves its snapped rows, and P must be
    the maximiser of logL(., Q) (concave in P for fixed Q) -- the KKT
    conditions of the box hold for the SQP result, the Adam result is close
    to it, and both start from the labelled frequencies."""
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    N, K = fixtures["gpu_adamem_adamem_Q"].shape
    y = (fixtures["gpu_adamem_adamem_Q"].argmax(axis=1) + 1).astype(np.int32)
    P0 = O.init_p_supervised(Gp, y, K, M)
    Q0 = np.random.default_rng(1).random((N, K))
    O.init_q_supervised(Q0, y, K)
    Pa, Qa = O.adamem_gpu(Gp, P0, Q0, M, labels=y)
    Pb, Qb = O.brqn(Gp, P0, Q0, M, labels=y, gpu_semantics=True)
    np.testing.assert_array_equal(Qa, Q0)
    np.testing.assert_array_equal(Qb, Q0)
    _, z = O.grad_hess_p(Gp, Qb, Pb, M)
    sc = np.abs(z).mean()
    inter = (Pb > 2e-5) & (Pb < 1 - 2e-5)
    assert np.abs(np.where(inter, z, 0.0)).max() / sc < 1e-6
    assert np.where(Pb <= 2e-5, z, -np.inf).max() <= 0 and np.where(Pb >= 1 - 2e-5, -z, -np.inf).max() <= 0
    assert np.abs(Pa - Pb).max() < 2e-3
    # the check separates: the starting P is off the maximiser
    _, z0 = O.grad_hess_p(Gp, Qb, P0, M)
    assert np.abs(np.where(inter, z0, 0.0)).max() / sc > 1e-4


def test_supervised_partial_labels_keep_snapped_rows(fixtures):
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    N, K = fixtures["gpu_adamem_adamem_Q"].shape
    y = _demo_labels(fixtures)
    P0 = O.init_p_supervised(Gp, y, K, M)
    Q0 = np.random.default_rng(1).random((N, K))
    O.init_q_supervised(Q0, y, K)
    h = []
    P, Q = O.adamem_gpu(Gp, P0, Q0, M, labels=y, history=h)
    lab = y > 0
    np.testing.assert_array_equal(Q[lab], Q0[lab])
    assert not np.allclose(Q[~lab], Q0[~lab])
    assert h[-1][1] > h[0][1]


# -------------------------------------------------------------- VCF / BCF ----

def _bcf_stream(samples, records, gt_idx=None):
    """A decompressed BCF 2.2 byte stream.  records: (contig index, GT values
    as an int array of shape N x n, numpy dtype int8 / int16)."""
    import struct
    fmt = '##FORMAT=<ID=GT,Number=1,Type=String,Description="Genotype"' + \
        (f",IDX={gt_idx}>" if gt_idx is not None else ">")
    text = "\n".join(["##fileformat=VCFv4.3", "##contig=<ID=1>", "##contig=<ID=2>", fmt,
                      "\t".join(["#CHROM", "POS", "ID", "REF", "ALT", "QUAL", "FILTER", "INFO", "FORMAT",
                                 *samples])]) + "\n\0"
    out = b"BCF\x02\x02" + struct.pack("<I", len(text)) + text.encode()
    key = gt_idx if gt_idx is not None else 0
    for contig, gt in records:
        n_s, n = gt.shape
        shared = struct.pack("<iiifII", contig, 100, 1, 0.0, 0, (1 << 24) | n_s)
        t = {1: 1, 2: 2}[gt.dtype.itemsize]
        indiv = bytes([0x11, key, (n << 4) | t]) + gt.astype(gt.dtype.newbyteorder("<")).tobytes()
        out += struct.pack("<II", len(shared), len(indiv)) + shared + indiv
    return out


def test_vcf_bcf_oracles_match_reader_fixture(fixtures):
    """The reference's tests/test_reader.py reads every format of the demo
    data to one fixture (demo_data.G.expected)."""
    import gzip
    G = fixtures["reader_G"]
    v, N, M = O.vcf_to_unpacked(gzip.decompress((GOLDEN / "demo_data.vcf.gz").read_bytes()))
    assert (N, M) == (105, 8451) and np.array_equal(v, G)
    for name in ("demo_data.bcf", "demo_data.bcf.gz"):
        data = (GOLDEN / name).read_bytes()
        while data[:2] == b"\x1f\x8b":
            data = gzip.decompress(data)
        b, N, M = O.bcf_to_unpacked(data)
        assert np.array_equal(b, G)


def test_pgen_oracle_matches_reader_fixture(fixtures):
    """The demo .pgen holds records of types 0 (1673), 1 (272), 2 (43), 3 (1)
    and 4 (6462); decoded they must give the reference's reader fixture."""
    data = (GOLDEN / "demo_data.pgen").read_bytes()
    G, N, M = O.pgen_to_unpacked(data)
    assert (N, M) == (105, 8451)
    np.testing.assert_array_equal(G, fixtures["reader_G"])
    _, _, starts, lens, types = O.pgen_header(data)
    assert starts[-1] + lens[-1] == len(data)
    assert sorted(set((types & 7).tolist())) == [0, 1, 2, 3, 4]


def test_vcf_gt_known_answers():
    """Codes written out from the reference parser's rules
    (snp_reader.pyx:305-375): '.', '-' or no digits -> missing; haploid a ->
    2a (missing if a > 1); diploid a1 + a2, missing if > 2; fields a short
    line lacks -> missing."""
    head = b"##fileformat=VCFv4.3\n#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\ta\tb\tc\td\te\n"
    body = (b"1\t1\t.\tA\tC\t.\t.\t.\tGT\t0/0\t0|1\t1/1\t./.\t0/.\n"
            b"1\t2\t.\tA\tC\t.\t.\t.\tGT:DP\t1/2:3\t0/2:1\t.|1\t-\t\n"
            b"2\t3\t.\tA\tC\t.\t.\t.\tGT\t0\t1\t2\t10/0\t1:9\n"
            b"2\t4\t.\tA\tC\t.\t.\t.\tGT\t1/0\t0/0\n"
            b"X\t5\t.\tA\tC\t.\t.\t.\tGT\t0/1\t1/1\t0/0\t0/0\t0/0")
    want = np.array([[0, 1, 2, 3, 3], [3, 2, 3, 3, 3], [0, 2, 3, 3, 2], [1, 0, 3, 3, 3], [1, 2, 0, 0, 0]],
                    dtype=np.uint8)
    G, N, M = O.vcf_to_unpacked(head + body)
    assert (N, M) == (5, 5)
    np.testing.assert_array_equal(G, want)
    G, _, M = O.vcf_to_unpacked(head + body, keep=lambda t: t in (b"1", b"2"))
    np.testing.assert_array_equal(G, want[:4])


def test_bcf_known_answers():
    """BCF GT values are (allele + 1) << 1 | phased; 0 = missing allele,
    -127 (int8) / -32767 (int16) = vector end; the sum of the two alleles
    (or the single allele of a haploid vector), 3 where one is missing
    (snp_reader.pyx:1025-1079)."""
    enc = lambda a, ph=0: ((a + 1) << 1) | ph                                  # noqa: E731
    r1 = np.array([[enc(0), enc(0)], [enc(0), enc(1, 1)], [enc(1), enc(1)], [0, enc(0)], [enc(1), -127]], np.int8)
    r2 = np.array([[enc(1)], [enc(0)], [0], [enc(2)], [enc(0)]], np.int8)
    r3 = np.array([[enc(0), enc(2)], [enc(1), enc(1)], [enc(0), 0], [enc(1), enc(0)], [-32767, -32767]], np.int16)
    data = _bcf_stream(list("abcde"), [(0, r1), (1, r2), (0, r3)])
    G, N, M = O.bcf_to_unpacked(data)
    np.testing.assert_array_equal(G, [[0, 1, 2, 3, 3], [1, 0, 3, 2, 0], [2, 2, 3, 1, 3]])
    G, _, M = O.bcf_to_unpacked(_bcf_stream(list("abcde"), [(0, r1), (1, r2)], gt_idx=5))
    np.testing.assert_array_equal(G, [[0, 1, 2, 3, 3], [1, 0, 3, 2, 0]])


# ------------------------------------------------------- cross-validation ----

def test_splitmix64_published_outputs():
    """cv.py's fold hash is splitmix64: its first three outputs from state 0
    are the generator's published test values."""
    C1 = 0x9E3779B97F4A7C15
    got = [int(O.splitmix64(x)) for x in (0, C1, (2 * C1) % 2**64)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_cv_folds_partition_observed_entries(fixtures):
    G = fixtures["reader_G"]
    obs = np.flatnonzero(G != 3)
    fold = O.cv_hash_folds(obs, 42, 5)
    counts = np.bincount(fold, minlength=5)
    assert counts.sum() == obs.size and fold.min() >= 0 and fold.max() < 5
    assert np.all(np.abs(counts - obs.size / 5) < 5 * np.sqrt(obs.size / 5))
    assert not np.array_equal(fold, O.cv_hash_folds(obs, 43, 5))          # the seed changes the folds
    sh = O.cv_shuffled_entries(G, 42)
    parts = [sh[f::5] for f in range(5)]
    assert np.array_equal(np.sort(np.concatenate(parts)), obs) and sh.dtype == np.int32


def test_cv_deviance_is_heldout_loglik_identity(fixtures):
    """sum of g log(g/mu) + (2-g) log((2-g)/(2-mu)) with mu = 2h equals
    -logL(held-out entries) - 2 ln 2 x #heterozygous held-out entries
    (logL pinned above); a brute-force loop checks single entries."""
    G = fixtures["reader_G"]
    M, N = G.shape
    P, Q = fixtures["cpu_brqn_brqn_P"], fixtures["cpu_brqn_brqn_Q"]
    obs = np.flatnonzero(G != 3)
    held = obs[O.cv_hash_folds(obs, 7, 5) == 2]
    codes = G.ravel()[held]
    dev = O.cv_deviance_sum(codes, held, P, Q, N)
    Gh = np.full_like(G, 3)
    Gh.ravel()[held] = codes
    ll = O.loglik_gpu(O.pack(Gh), P, Q, M)
    assert abs(dev - (-ll - 2 * np.log(2) * np.count_nonzero(codes == 1))) < 1e-9 * abs(dev)
    import math
    for e in held[:50]:
        m, i = divmod(int(e), N)
        g = float(G[m, i])
        mu = min(max(2 * sum(Q[i, k] * P[m, k] for k in range(Q.shape[1])), 1e-10), 2 - 1e-10)
        want = (g * math.log(g / mu) if g > 0 else 0.0) + ((2 - g) * math.log((2 - g) / (2 - mu)) if g < 2 else 0.0)
        assert abs(O.cv_deviance_sum([G[m, i]], [e], P, Q, N) - want) < 1e-12 * max(1.0, abs(want))


def test_cv_polish_runs_exactly_five_iterations(fixtures):
    """br_qn.py:50-70 with cv.py's n_iters = 5: tol 0 and patience 6 disable
    every early stop (as the reference's own test_cv.py asserts of the call)."""
    Gp, M = fixtures["reader_G_packed"], fixtures["reader_G"].shape[0]
    h = []
    O.brqn(Gp, fixtures["cpu_brqn_brqn_P"], fixtures["cpu_brqn_brqn_Q"], M, max_iter=5, tol=0.0, Q_hist=3,
           patience=6, history=h)
    assert [it for it, _ in h] == [1, 2, 3, 4, 5]


# ------------------------------------------------------- learned warm start ----

def _mlp_case(seed=3, rows=7):
    from oracle import mlp_oracle as MO
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(MO.NPARAM) * 0.08
    X = rng.random((rows, MO.D0))
    X[:, :MO.D3] = rng.dirichlet(np.ones(MO.D3), size=rows)
    kvec = rng.integers(2, MO.D3 + 1, size=rows)
    T = np.zeros((rows, MO.D3))
    for r, k in enumerate(kvec):
        T[r, :k] = rng.dirichlet(np.ones(k))
    return MO, w, X, kvec, T


def test_mlp_oracle_gradient_finite_differences():
    """Every parameter block, including ReLU units on both sides of zero:
    central differences of the loss match backward() (kink-free points)."""
    MO, w, X, kvec, T = _mlp_case()
    g = MO.backward(w, X, kvec, T)
    rng = np.random.default_rng(0)
    off = 0
    checked = 0
    for name, shp in MO.SHAPES:
        n = int(np.prod(shp))
        for idx in rng.choice(n, size=12, replace=False):
            j = off + idx
            e = 1e-6
            wp, wm = w.copy(), w.copy()
            wp[j] += e
            wm[j] -= e
            fd = (MO.loss(wp, X, kvec, T) - MO.loss(wm, X, kvec, T)) / (2 * e)
            assert abs(fd - g[j]) <= 1e-7 + 1e-5 * abs(g[j]), (name, idx, fd, g[j])
            checked += 1
        off += n
    assert checked == 72
    _, (_, h1, h2, _, _) = MO.forward(w, X, kvec)
    assert (h1 == 0).any() and (h1 > 0).any() and (h2 == 0).any() and (h2 > 0).any()


def test_mlp_oracle_softmax_closed_form():
    """With zero weights the output is the renormalised prior and the
    gradient w.r.t. b3 is mean(p - t) over the rows (masked columns zero)."""
    MO, w, X, kvec, T = _mlp_case(seed=5)
    w = np.zeros_like(w)
    p, _ = MO.forward(w, X, kvec)
    for r, k in enumerate(kvec):
        prior = np.maximum(X[r, :k], 1e-5)
        np.testing.assert_allclose(p[r, :k], prior / prior.sum(), rtol=1e-12)
        assert np.all(p[r, k:] == 0)
    g = MO.unflatten(MO.backward(w, X, kvec, T))
    mask = np.arange(MO.D3)[None, :] < kvec[:, None]
    np.testing.assert_allclose(g["b3"], np.where(mask, p - T, 0).mean(0), atol=1e-15)
    assert np.all(g["W1"] == 0) and np.all(g["W2"] == 0)


def test_mlp_oracle_adam_textbook():
    """Two Adam steps against Kingma & Ba, Algorithm 1, written out by hand;
    the first bias-corrected step moves each weight by lr sign(g)."""
    from oracle import mlp_oracle as MO
    g1, g2 = np.array([0.3, -2.0, 1e-3]), np.array([-0.1, 0.5, 0.0])
    w0 = np.array([1.0, 2.0, 3.0])
    w1, m, v = MO.adam(w0, g1, np.zeros(3), np.zeros(3), 1, 0.01)
    np.testing.assert_allclose(w1, w0 - 0.01 * g1 / (np.abs(g1) + 1e-8), rtol=1e-12)
    w2, m, v = MO.adam(w1, g2, m, v, 2, 0.01)
    m_ref = 0.9 * 0.1 * g1 + 0.1 * g2
    v_ref = 0.999 * 0.001 * g1 ** 2 + 0.001 * g2 ** 2
    np.testing.assert_allclose(w2, w1 - 0.01 * (m_ref / (1 - 0.81)) / (np.sqrt(v_ref / (1 - 0.998001)) + 1e-8),
                               rtol=1e-12)


def test_mlp_oracle_bf16_rounding():
    """Round-to-nearest-even on the 16-bit boundary, ties to even."""
    from oracle import mlp_oracle as MO
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)
    x = np.array([one, one + ulp / 2, one + ulp * 1.5, one + ulp * 0.75, -3.0], np.float32)
    np.testing.assert_array_equal(MO._bf16(x), [1.0, 1.0, 1.0 + 2 * float(ulp), 1.0 + float(ulp), -3.0])


def test_learned_features_host_logic():
    """Feature layout of the warm start on a tiny exact case (torch on the
    CPU): ALS prior block, the unconstrained least-squares block reproduces
    Q when the low-rank model is exact, energy fraction, K encodings."""
    import torch
    from adamixture_amd.learned_init import D_IN, features
    rng = np.random.default_rng(1)
    M, N, K = 60, 40, 3
    P = rng.uniform(0.1, 0.9, (M, K))
    Q = rng.dirichlet(np.ones(K), size=N)
    X = P @ Q.T                                   # expected genotype / 2
    f = X.mean(axis=1)
    C = 2 * X - 2 * f[:, None]                    # centred expected genotypes
    U, S, Vt = np.linalg.svd(C, full_matrices=False)
    r = 2 * K
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))  # noqa: E731
    F = features(t(U[:, :r]), t(S[:r]), t(Vt[:r].T), t(f), t(P), t(Q), K).numpy()
    assert F.shape == (N, D_IN) and F.dtype == np.float32
    np.testing.assert_allclose(F[:, :K], Q, rtol=1e-6)
    assert np.all(F[:, K:16] == 0)
    np.testing.assert_allclose(F[:, 16:16 + K], Q, atol=1e-5)   # exact model: LS returns Q
    np.testing.assert_allclose(F[:, 32:32 + K], np.log(np.maximum(Q, 1e-5)) / np.log(1e-5), rtol=1e-5)
    assert np.all(F[:, 32 + K:48] == 0)
    assert np.all(F[:, 48] < 1e-6)                # rank-(K-1) centred data: no tail energy
    np.testing.assert_allclose(F[:, 50], K / 16)
    np.testing.assert_allclose(F[:, 51], 1 / K)
    np.testing.assert_allclose(F[0, 52:52 + r - 1], (S[1:r] / S[0]).astype(np.float32), rtol=1e-5)


def test_map_pq_gpu_pins_the_gpu_quasi_newton_map():
    """br_qn_gpu.py:47-61 (the one step where the GPU-semantics BR-QN driver
    differs from the pinned CPU driver): clip both to [1e-5, 1 - 1e-5], then
    divide Q rows by their sums.  Feasible points are fixed points; the
    bounds are the reference's; clipping comes before the division."""
    P = np.array([[1e-5, 0.5, 1 - 1e-5], [-1.0, 0.0, 2.0]])
    Q = np.array([[0.5, 0.25, 0.25], [3.0, 0.5, -1.0], [2.0, 2.0, 2.0]])
    Pm, Qm = O.map_pq_gpu(P, Q)
    np.testing.assert_array_equal(Pm, [[1e-5, 0.5, 1 - 1e-5], [1e-5, 1e-5, 1 - 1e-5]])
    np.testing.assert_array_equal(Qm[0], Q[0])                        # on the simplex already
    want = np.array([1 - 1e-5, 0.5, 1e-5]) / (1 - 1e-5 + 0.5 + 1e-5)  # clip, then divide
    np.testing.assert_allclose(Qm[1], want, rtol=1e-15)
    np.testing.assert_allclose(Qm[2], [1 / 3] * 3, rtol=1e-15)
    np.testing.assert_allclose(Qm.sum(axis=1), 1.0, rtol=1e-15)
    assert np.all(P == np.array([[1e-5, 0.5, 1 - 1e-5], [-1.0, 0.0, 2.0]]))  # inputs untouched


def test_fast_oracle_build_leaves_the_fp_mode_alone():
    """Loading the fast-math CPU-baseline build must not turn on
    flush-to-zero / denormals-are-zero for the process (a -ffast-math link
    of a shared library would): bench.py loads it before the
    time-to-target race, whose host-side eigensolves and pseudo-inverses
    would then differ."""
    import subprocess
    import sys
    code = ("import numpy as np, oracle; oracle.lib_fast(); "
            "print(float(np.float32(1e-40) * np.float32(1)), float(np.float64(1e-310) * 1.0))")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True,
                         cwd=str(__import__("pathlib").Path(__file__).resolve().parents[1])).stdout.split()
    assert float(out[-2]) > 0.0 and float(out[-1]) > 0.0
