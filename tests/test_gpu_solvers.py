"""Batched SQP / BVLS kernels against the oracle (FP64, tolerance 1e-9: the
solvers take the same pivot/active-set decisions, only the order of the
FP64 sums in the sweeps and eliminations differs)."""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import solvers

pytestmark = pytest.mark.gpu


def _spd_batch(rng, n, K, cond=1e4):
    """SPD K x K matrices with condition number <= cond (uneven spectra)."""
    Hs = np.empty((n, K, K))
    for i in range(n):
        U, _ = np.linalg.qr(rng.normal(size=(K, K)))
        ev = np.exp(rng.uniform(0, np.log(cond), size=K))
        Hs[i] = (U * ev) @ U.T
    return Hs


@pytest.mark.parametrize("K", [2, 3, 7, 12, 16])
def test_sqp_q_matches_oracle(cuda, K):
    rng = np.random.default_rng(K)
    n = 3000
    H = _spd_batch(rng, n, K) * rng.uniform(1, 1e3, size=(n, 1, 1))
    z = rng.normal(scale=50.0, size=(n, K))               # gradients: uneven active sets
    X = rng.dirichlet(np.full(K, 0.3), size=n)
    vkk = O.v_kk(K)
    want = O.sqp_q(H, z, X, vkk)
    dev = cuda
    got = solvers.sqp_solve_q(torch.from_numpy(H).to(dev), torch.from_numpy(-z).to(dev), torch.from_numpy(X).to(dev),
                              solvers.v_kk(K, dev), n, K).cpu().numpy()
    np.testing.assert_allclose(got, want, atol=1e-9)
    np.testing.assert_allclose(got.sum(1), 1.0, atol=1e-12)


@pytest.mark.parametrize("K", [1, 4, 10, 16])
def test_sqp_p_matches_oracle(cuda, K):
    rng = np.random.default_rng(100 + K)
    n = 4000
    H = _spd_batch(rng, n, K) * 10.0
    z = rng.normal(scale=30.0, size=(n, K))
    X = rng.uniform(1e-5, 1 - 1e-5, size=(n, K))
    want = O.sqp_p(H, z, X)
    got = solvers.sqp_solve_p(torch.from_numpy(H).to(cuda), torch.from_numpy(-z).to(cuda),
                              torch.from_numpy(X).to(cuda), n, K).cpu().numpy()
    np.testing.assert_allclose(got, want, atol=1e-9)


def test_sqp_from_real_gradients(cuda):
    """Hessians/gradients of an admixture log-likelihood (oracle grad_hess)."""
    rng = np.random.default_rng(7)
    M, N, K = 400, 120, 5
    P = rng.uniform(0.05, 0.95, size=(M, K))
    Q = rng.dirichlet(np.full(K, 0.5), size=N)
    Gp = O.synth_packed(P, Q, 3, 0.02)
    P0 = np.clip(rng.random((M, K)), 1e-5, 1 - 1e-5)
    Q0 = rng.dirichlet(np.ones(K), size=N)
    Hq, zq = O.grad_hess_q(Gp, Q0, P0, M)
    want = O.sqp_q(Hq, zq, Q0, O.v_kk(K))
    got = solvers.sqp_solve_q(torch.from_numpy(Hq).to(cuda), torch.from_numpy(-zq).to(cuda),
                              torch.from_numpy(Q0).to(cuda), solvers.v_kk(K, cuda), N, K).cpu().numpy()
    np.testing.assert_allclose(got, want, atol=1e-9)
    Hp, zp = O.grad_hess_p(Gp, Q0, P0, M)
    want = O.sqp_p(Hp, zp, P0)
    got = solvers.sqp_solve_p(torch.from_numpy(Hp).to(cuda), torch.from_numpy(-zp).to(cuda),
                              torch.from_numpy(P0).to(cuda), M, K).cpu().numpy()
    np.testing.assert_allclose(got, want, atol=1e-9)


@pytest.mark.parametrize("K", [1, 3, 8, 12, 16, 33])
@pytest.mark.parametrize("start", ["inverse", "solve"])
def test_bvls_matches_oracle(cuda, K, start):
    rng = np.random.default_rng(K)
    X = rng.normal(size=(4 * K + 5, K))
    A = X.T @ X + 0.05 * np.eye(K)
    B = rng.normal(scale=1.5, size=(5000, K)) @ A
    want = (O.bvls_gpu if start == "inverse" else O.bvls_cpu)(A, B, 0.0, 1.0)
    got = solvers.batch_bvls_bpp(torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda), 0.0, 1.0,
                                 start=start).cpu().numpy()
    np.testing.assert_allclose(got, want, atol=1e-9)
    assert got.min() >= -1e-12 and got.max() <= 1 + 1e-12


@pytest.mark.parametrize("K", [4, 10, 12])
def test_bvls_thread_path_matches_wavefront_path(cuda, K):
    """The shared-A path (subset-inverse table, one thread per problem) and
    the wavefront path (Gauss-Jordan per free set) on ALS-like problems where
    many variables end on a bound: the same active sets, solutions equal to
    rounding."""
    from adamixture_amd import _lib
    rng = np.random.default_rng(100 + K)
    Qm = rng.dirichlet(np.full(K, 0.3), size=2000)
    A = Qm.T @ Qm
    B = rng.normal(0.5, 0.8, size=(20000, K)) @ A
    Ad, Bd = torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda)
    fast = solvers.batch_bvls_bpp(Ad, Bd, 0.0, 1.0).cpu().numpy()
    Ainv = torch.linalg.inv(Ad).contiguous()
    slow = torch.empty_like(Bd)
    q = torch.zeros(1, dtype=torch.int64, device=cuda)
    _lib.call("adx_bvls_batch", Ad.data_ptr(), Ainv.data_ptr(), Bd.data_ptr(), slow.data_ptr(), B.shape[0], K, 0.0,
              1.0, q.data_ptr(), 8, _lib.stream_handle(cuda))
    slow = slow.cpu().numpy()
    at_bound = np.mean((fast <= 1e-12) | (fast >= 1 - 1e-12))
    print(f"K={K}: {at_bound:.2f} of the entries on a bound; max |fast - slow| {np.abs(fast - slow).max():.2e}")
    assert at_bound > 0.2
    np.testing.assert_array_equal(fast <= 1e-12, slow <= 1e-12)
    np.testing.assert_allclose(fast, slow, rtol=0, atol=1e-10)
