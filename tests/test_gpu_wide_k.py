"""K = 17 .. 64, the reference GPU path's limit (adamixture/entry.py:299-303,
MAX_K = 64): every form of the EM pass, the log-likelihood, the BR-QN
gradient / Hessian, and the initializer's column-split products, against the
CPU oracle through the C ABI, on ragged M and N.

Tolerances are those of the K <= 16 tests: FP64 1e-12 relative (summation
order only), FP32 2e-4, mixed precision 1e-3 of the largest entry;
gradient / Hessian 1e-11.
"""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import _lib
from adamixture_amd import genotypes as GT
from adamixture_amd.br_qn_gpu import grad_hess
from adamixture_amd.em_adam_gpu import EMAdamOptimizer, em_pass

pytestmark = pytest.mark.gpu

WIDE = [  # (M, N, K, missing rate)
    (1037, 523, 17, 0.02),
    (777, 301, 32, 0.05),
    (517, 259, 64, 0.01),
]


def _case(M, N, K, miss, seed, dev):
    rng = np.random.default_rng(seed)
    P = np.clip(rng.beta(0.6, 1.2, size=(M, K)), 1e-5, 1 - 1e-5)
    Q = rng.dirichlet(np.full(K, 0.4), size=N)
    G = GT.synthesize(torch.from_numpy(P).to(dev), torch.from_numpy(Q).to(dev), seed + 100, miss)
    return G, G.to_host_packed(), P, Q


def _close(got, want, rtol):
    g = got.double().cpu().numpy()
    np.testing.assert_allclose(g, want, rtol=rtol, atol=rtol * np.abs(want).max())
    return np.abs(g - want).max() / np.abs(want).max()


@pytest.mark.parametrize("M,N,K,miss", WIDE)
def test_em_pass_f64_wide_k(cuda, M, N, K, miss):
    G, Gp, P, Q = _case(M, N, K, miss, 1, cuda)
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    z = lambda *s: torch.zeros(*s, dtype=torch.float64, device=cuda)  # noqa: E731
    A, B, T = z(M, K), z(M, K), z(N, K)
    ll = em_pass(G, Pd, Qd, A, B, T, dtype=_lib.ADX_F64, loglik=True)
    Ao, Bo, To, _ = O.em_stats_gpu(Gp, P, Q, M)
    for got, want in ((A, Ao), (B, Bo), (T, To)):
        _close(got, want, 1e-12)
    llo = O.loglik_gpu(Gp, P, Q, M)
    assert ll.item() == pytest.approx(llo, rel=1e-12)
    assert em_pass(G, Pd, Qd, dtype=_lib.ADX_F64, loglik=True, stats=False).item() == pytest.approx(llo, rel=1e-12)


@pytest.mark.parametrize("M,N,K,miss", WIDE)
def test_em_pass_f32_wide_k(cuda, M, N, K, miss):
    G, Gp, P, Q = _case(M, N, K, miss, 2, cuda)
    P32 = torch.from_numpy(P).to(cuda, torch.float32)
    Q32 = torch.from_numpy(Q).to(cuda, torch.float32)
    z = lambda *s: torch.zeros(*s, dtype=torch.float32, device=cuda)  # noqa: E731
    A, B, T = z(M, K), z(M, K), z(N, K)
    em_pass(G, P32, Q32, A, B, T, dtype=_lib.ADX_F32)
    Ao, Bo, To, _ = O.em_stats_gpu(Gp, P32.double().cpu().numpy(), Q32.double().cpu().numpy(), M)
    for got, want in ((A, Ao), (B, Bo), (T, To)):
        _close(got, want, 2e-4)


@pytest.mark.parametrize("M,N,K,miss", WIDE)
def test_em_pass_bf16_wide_k(cuda, M, N, K, miss):
    G, Gp, P, Q = _case(M, N, K, miss, 3, cuda)
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision="bf16")
    opt.accumulate(G, Pd, Qd)
    Ao, Bo, To, _ = O.em_stats_gpu(Gp, P, Q, M)
    errs = [_close(got, want, 1e-3) for got, want in ((opt.A_accum, Ao), (opt.B_accum, Bo), (opt.T_accum, To))]
    print(f"bf16 K={K}: max rel err A {errs[0]:.2e} B {errs[1]:.2e} T {errs[2]:.2e}")


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_em_step_and_adam_wide_k_match_oracle(cuda, precision):
    """Two EM + Adam steps at K = 20 through the optimizer (finalize, Adam,
    clamp, renormalise), deterministic on repeat."""
    M, N, K = 901, 333, 20
    G, Gp, P, Q = _case(M, N, K, 0.01, 4, cuda)
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision=precision)
    ref = O.AdamEMGPU(M, N, K, 0.005, 0.8, 0.88, 1e-8)
    Po, Qo = P.copy(), Q.copy()
    tol = 1e-11 if precision == "f64" else 1e-4
    for _ in range(2):
        PE1, QE1 = (x.clone() for x in opt.run_em_step(G, Pd, Qd, M))
        PE, QE = opt.run_em_step(G, Pd, Qd, M)
        assert torch.equal(PE, PE1) and torch.equal(QE, QE1)
        PEo, QEo = ref.run_em_step(Gp, Po, Qo, M)
        np.testing.assert_allclose(PE.cpu().numpy(), PEo, rtol=tol, atol=tol * 1e-3)
        np.testing.assert_allclose(QE.cpu().numpy(), QEo, rtol=tol, atol=tol * 1e-3)
        opt.step(Pd, Qd, PE, QE)
        ref.step(Po, Qo, PE.cpu().numpy(), QE.cpu().numpy())
        np.testing.assert_allclose(Pd.cpu().numpy(), Po, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("K", [17, 32, 64])
def test_grad_hess_wide_k_matches_oracle(cuda, K):
    rng = np.random.default_rng(K)
    M, N = 389, 213
    P = rng.uniform(0.02, 0.98, size=(M, K))
    Q = rng.dirichlet(np.full(K, 0.5), size=N)
    Gp = O.synth_packed(P, Q, 5, 0.05)
    G = GT.as_packed(torch.from_numpy(Gp), M, cuda)
    P0 = np.clip(rng.random((M, K)), 1e-5, 1 - 1e-5)
    Q0 = rng.dirichlet(np.ones(K), size=N)
    Pd, Qd = torch.from_numpy(P0).to(cuda), torch.from_numpy(Q0).to(cuda)
    for side, fn in (("q", O.grad_hess_q), ("p", O.grad_hess_p)):
        H, z = grad_hess(G, Qd, Pd, side)
        Ho, zo = fn(Gp, Q0, P0, M)
        np.testing.assert_allclose(H.cpu().numpy(), Ho, rtol=1e-11, atol=1e-11 * np.abs(Ho).max())
        np.testing.assert_allclose(z.cpu().numpy(), zo, rtol=1e-11, atol=1e-11 * np.abs(zo).max())
        Hh = H.cpu().numpy()
        assert np.array_equal(Hh, np.swapaxes(Hh, 1, 2))


def test_svd_products_wider_than_one_launch(cuda):
    """L = max(K + 10, 20) = 74 at K = 64: the decode-fused products split
    into column blocks of at most 64 and equal the dense products."""
    from adamixture_amd.svd_gpu import xty, xw
    rng = np.random.default_rng(5)
    M, N, L = 611, 203, 74
    P = rng.uniform(0.05, 0.95, size=(M, 3))
    Q = rng.dirichlet(np.ones(3), size=N)
    Gp = O.synth_packed(P, Q, 1, 0.03)
    G = GT.as_packed(torch.from_numpy(Gp), M, cuda)
    f = O.freq_gpu(Gp, M)
    X = O.unpack_center(Gp, f, M, 0, M).astype(np.float64)
    W = rng.normal(size=(N, L)).astype(np.float32)
    Y = rng.normal(size=(M, L)).astype(np.float32)
    fd = torch.from_numpy(f).to(cuda)
    np.testing.assert_allclose(xw(G, fd, torch.from_numpy(W).to(cuda)).cpu().numpy(), X @ W, rtol=1e-4, atol=1e-3)
    np.testing.assert_allclose(xty(G, fd, torch.from_numpy(Y).to(cuda)).cpu().numpy(), X.T @ Y, rtol=1e-4,
                               atol=1e-3)


def test_als_products_wider_than_one_launch(cuda):
    """The ALS skinny products at K = 64 (X^T Y in column blocks of a b <=
    1024, X W + u v^T in blocks of 16 columns) equal the dense products."""
    from adamixture_amd.als_gpu import _nn, _tn
    rng = np.random.default_rng(6)
    n, K = 5003, 64
    X = rng.normal(size=(n, K))
    Y = rng.normal(size=(n, K))
    W = rng.normal(size=(K, K))
    u, v = rng.normal(size=n), rng.normal(size=K)
    t = lambda a: torch.from_numpy(a).to(cuda)  # noqa: E731
    np.testing.assert_allclose(_tn(t(X), t(Y), 0.5).cpu().numpy(), 0.5 * X.T @ Y, rtol=1e-11, atol=1e-10)
    np.testing.assert_allclose(_nn(t(X), t(W), 2.0, t(u), t(v)).cpu().numpy(), 2.0 * X @ W + np.outer(u, v),
                               rtol=1e-11, atol=1e-10)
