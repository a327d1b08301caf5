"""Randomized SVD, ALS, gradient/Hessian and BR-QN on the GPU against the
oracle and the reference's demo fixtures (adamixture/demo/outputs)."""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import reader
from adamixture_amd.als_gpu import ALS_gpu
from adamixture_amd.br_qn_gpu import grad_hess, optimize_original_gpu
from adamixture_amd.genotypes import as_packed
from adamixture_amd.svd_gpu import SVD_gpu, xty, xw
from adamixture_amd.utils import calculate_frequencies_gpu
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _canon(U, V):
    U, V = U.copy(), V.copy()
    for c in range(U.shape[1]):
        p = np.argmax(np.abs(U[:, c]))
        if U[p, c] < 0:
            U[:, c] *= -1
            V[:, c] *= -1
    return U, V


@pytest.fixture(scope="module")
def demo(cuda):
    Gt, N, M = reader.read_bed_packed(GOLDEN / "demo_data.bed")
    return as_packed(Gt, M, cuda), N, M


def test_svd_gemms_match_dense(cuda):
    rng = np.random.default_rng(0)
    M, N, L = 1037, 411, 20
    P = rng.uniform(0.05, 0.95, size=(M, 3))
    Q = rng.dirichlet(np.ones(3), size=N)
    Gp = O.synth_packed(P, Q, 1, 0.03)
    G = as_packed(torch.from_numpy(Gp), M, cuda)
    f = O.freq_gpu(Gp, M)
    X = O.unpack_center(Gp, f, M, 0, M).astype(np.float64)
    W = rng.normal(size=(N, L)).astype(np.float32)
    Y = rng.normal(size=(M, L)).astype(np.float32)
    fd = torch.from_numpy(f).to(cuda)
    got_y = xw(G, fd, torch.from_numpy(W).to(cuda)).cpu().numpy()
    got_a = xty(G, fd, torch.from_numpy(Y).to(cuda)).cpu().numpy()
    np.testing.assert_allclose(got_y, X @ W, rtol=1e-4, atol=1e-3)
    np.testing.assert_allclose(got_a, X.T @ Y, rtol=1e-4, atol=1e-3)


def _col_err(got, want):
    """Largest error of a column relative to that column's largest entry."""
    return float((np.abs(got - want).max(0) / np.maximum(np.abs(want).max(0), 1e-300)).max())


@pytest.mark.parametrize("M,N,L,scale", [(1037, 411, 20, 1.0), (20003, 3001, 1, 1.0), (20003, 3001, 33, 1.0),
                                         (4001, 2049, 64, 1.0), (4001, 2049, 20, 1e-6), (4001, 2049, 20, 3e5),
                                         (5, 17, 20, 1.0)])
def test_svd_products_h2_match_fp32_accuracy(cuda, M, N, L, scale):
    """The FP16 two-term forms against the FP64 dense products: within 2e-6 of
    each column's largest entry, and not more than twice the FP32 forms'
    error (plus 2e-7) -- ragged M and N, L = 1 .. 64, one and several
    reduction splits, operands scaled far from 1 (the power-of-two scale)."""
    rng = np.random.default_rng(M + L)
    P = rng.uniform(0.05, 0.95, size=(M, 3))
    Q = rng.dirichlet(np.ones(3), size=N)
    Gp = O.synth_packed(P, Q, 3, 0.03)
    G = as_packed(torch.from_numpy(Gp), M, cuda)
    f = O.freq_gpu(Gp, M)
    X = O.unpack_center(Gp, f, M, 0, M).astype(np.float64)
    W = (rng.normal(size=(N, L)) * scale).astype(np.float32)
    Y = (rng.normal(size=(M, L)) * scale).astype(np.float32)
    fd = torch.from_numpy(f).to(cuda)
    Wd, Yd = torch.from_numpy(W).to(cuda), torch.from_numpy(Y).to(cuda)
    want_y, want_a = X @ W.astype(np.float64), X.T @ Y.astype(np.float64)
    print(f"case: h2 products M={M} N={N} L={L} scale={scale}", flush=True)
    for prod, fn, Z, want in (("xw", xw, Wd, want_y), ("xty", xty, Yd, want_a)):
        e_h2 = _col_err(fn(G, fd, Z, "h2").cpu().numpy(), want)
        e_32 = _col_err(fn(G, fd, Z, "f32").cpu().numpy(), want)
        print(f"{prod}: h2 {e_h2:.2e}  f32 {e_32:.2e}", flush=True)
        assert e_h2 <= 2e-6 and e_h2 <= 2 * e_32 + 2e-7, (prod, e_h2, e_32)


def test_svd_products_h2_zero_operand(cuda):
    """An all-zero operand (scale 1) gives exact zeros."""
    rng = np.random.default_rng(9)
    M, N, L = 999, 300, 20
    Gp = O.synth_packed(rng.uniform(0.05, 0.95, size=(M, 2)), rng.dirichlet(np.ones(2), size=N), 1, 0.02)
    G = as_packed(torch.from_numpy(Gp), M, cuda)
    fd = torch.from_numpy(O.freq_gpu(Gp, M)).to(cuda)
    assert not xw(G, fd, torch.zeros(N, L, device=cuda), "h2").any()
    assert not xty(G, fd, torch.zeros(M, L, device=cuda), "h2").any()


def test_init_products_are_deterministic(cuda):
    """The split reductions of the SVD GEMMs (several splits at this shape)
    and the ALS X^T Y products are summed in a fixed order: repeated calls
    agree bit for bit."""
    from adamixture_amd import _lib
    from adamixture_amd.als_gpu import _tn
    rng = np.random.default_rng(1)
    M, N, L = 20000, 3000, 20
    P = rng.uniform(0.05, 0.95, size=(M, 4))
    Q = rng.dirichlet(np.ones(4), size=N)
    Gp = O.synth_packed(P, Q, 2, 0.02)
    G = as_packed(torch.from_numpy(Gp), M, cuda)
    assert _lib.load().adx_svd_work_floats(0, M, N, L) > 0 and _lib.load().adx_svd_work_floats(1, M, N, L) > 0
    fd = torch.from_numpy(O.freq_gpu(Gp, M)).to(cuda)
    W = torch.from_numpy(rng.normal(size=(N, L)).astype(np.float32)).to(cuda)
    Y = torch.from_numpy(rng.normal(size=(M, L)).astype(np.float32)).to(cuda)
    for p in ("h2", "f32"):
        assert torch.equal(xw(G, fd, W, p), xw(G, fd, W, p)) and torch.equal(xty(G, fd, Y, p), xty(G, fd, Y, p))
    Xd = torch.from_numpy(rng.normal(size=(300000, 10))).to(cuda)
    Yd = torch.from_numpy(rng.normal(size=(300000, 12))).to(cuda)
    C1, C2 = _tn(Xd, Yd), _tn(Xd, Yd)
    assert torch.equal(C1, C2)
    np.testing.assert_allclose(C1.cpu().numpy(), Xd.cpu().numpy().T @ Yd.cpu().numpy(), rtol=1e-10, atol=1e-9)


def test_svd_matches_cpu_reference_fixture(demo, fixtures):
    G, N, M = demo
    f = calculate_frequencies_gpu(G, M, 8192, G.device)
    np.testing.assert_array_equal(f.cpu().numpy(), O.freq_gpu(fixtures["reader_G_packed"], M))
    U, S, V = SVD_gpu(G, N, M, f, 7, 42, 5, 0.1, 8192, G.device)
    U, V = _canon(U.cpu().numpy(), V.cpu().numpy())
    np.testing.assert_allclose(S.cpu().numpy(), fixtures["cpu_brqn_svd_S"], rtol=1e-4)
    np.testing.assert_allclose(U, fixtures["cpu_brqn_svd_U"], atol=1e-4)
    np.testing.assert_allclose(V, fixtures["cpu_brqn_svd_V"], atol=1e-4)


def test_svd_device_test_matrix(demo):
    """omega="device" (a seeded torch device generator, as the reference GPU
    path draws it): reruns are identical, and the leading singular values
    agree with the host-drawn SVD's (a different random start of the same
    power iteration)."""
    G, N, M = demo
    f = calculate_frequencies_gpu(G, M, 8192, G.device)
    U1, S1, V1 = SVD_gpu(G, N, M, f, 7, 42, 5, 0.1, 8192, G.device, omega="device")
    U2, S2, V2 = SVD_gpu(G, N, M, f, 7, 42, 5, 0.1, 8192, G.device, omega="device")
    assert torch.equal(U1, U2) and torch.equal(S1, S2) and torch.equal(V1, V2)
    Uh, Sh, Vh = SVD_gpu(G, N, M, f, 7, 42, 5, 0.1, 8192, G.device)
    np.testing.assert_allclose(S1[:3].cpu().numpy(), Sh[:3].cpu().numpy(), rtol=1e-3)
    with pytest.raises(ValueError, match="omega"):
        SVD_gpu(G, N, M, f, 7, 42, 5, 0.1, 8192, G.device, omega="gpu")


@pytest.mark.parametrize("start", ["solve", "inverse"])
def test_als_matches_cpu_reference_fixture(demo, fixtures, start):
    G, N, M = demo
    f = calculate_frequencies_gpu(G, M, 8192, G.device)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(G.device)  # noqa: E731
    P, Q = ALS_gpu(t(fixtures["cpu_brqn_svd_U"]), t(fixtures["cpu_brqn_svd_S"]), t(fixtures["cpu_brqn_svd_V"]),
                   f, 42, M, 7, 1000, 1e-4, G.device, bvls_start=start)
    np.testing.assert_allclose(P.cpu().numpy(), fixtures["cpu_brqn_als_P"], atol=1e-6)
    np.testing.assert_allclose(Q.cpu().numpy(), fixtures["cpu_brqn_als_Q"], atol=1e-6)


@pytest.mark.parametrize("K", [3, 7, 9, 10, 11, 13, 16])
def test_grad_hess_matches_oracle(cuda, K):
    rng = np.random.default_rng(K)
    M, N = 523, 300
    P = rng.uniform(0.02, 0.98, size=(M, K))
    Q = rng.dirichlet(np.full(K, 0.5), size=N)
    Gp = O.synth_packed(P, Q, 5, 0.05)
    G = as_packed(torch.from_numpy(Gp), M, cuda)
    P0 = np.clip(rng.random((M, K)), 1e-5, 1 - 1e-5)
    Q0 = rng.dirichlet(np.ones(K), size=N)
    Pd, Qd = torch.from_numpy(P0).to(cuda), torch.from_numpy(Q0).to(cuda)
    for side, fn in (("q", O.grad_hess_q), ("p", O.grad_hess_p)):
        H, z = grad_hess(G, Qd, Pd, side)
        Ho, zo = fn(Gp, Q0, P0, M)
        np.testing.assert_allclose(H.cpu().numpy(), Ho, rtol=1e-11, atol=1e-11 * np.abs(Ho).max())
        np.testing.assert_allclose(z.cpu().numpy(), zo, rtol=1e-11, atol=1e-11 * np.abs(zo).max())


def test_brqn_demo_matches_oracle_and_reference_fixture(demo, fixtures):
    G, N, M = demo
    K = 7
    P0, Q0 = fixtures["gpu_brqn_als_P"], fixtures["gpu_brqn_als_Q"]
    hist = []
    P, Q = optimize_original_gpu(G, torch.from_numpy(P0).to(G.device), torch.from_numpy(Q0).to(G.device), 10000,
                                 K, M, N, 0.1, 3, 3, G.device, 8192, 256, on_iter=lambda it, L: hist.append(L))
    P, Q = P.cpu().numpy(), Q.cpu().numpy()
    ohist = []
    Po, Qo = O.brqn(fixtures["reader_G_packed"], P0, Q0, M, history=ohist, gpu_semantics=True)
    print(f"brqn iterations gpu {len(hist)} oracle {len(ohist)}; max|dP| {np.abs(P - Po).max():.2e} "
          f"max|dQ| {np.abs(Q - Qo).max():.2e}; logL {hist[-1]:.6f} vs {ohist[-1][1]:.6f}")
    assert len(hist) == len(ohist)
    np.testing.assert_allclose(hist, [L for _, L in ohist], rtol=1e-10)
    np.testing.assert_allclose(P, Po, atol=1e-6)
    np.testing.assert_allclose(Q, Qo, atol=1e-6)
    Pe, Qe = fixtures["gpu_brqn_brqn_P"], fixtures["gpu_brqn_brqn_Q"]
    np.testing.assert_allclose(P, Pe, rtol=2e-2, atol=2e-2)
    np.testing.assert_allclose(Q, Qe, rtol=2e-2, atol=2e-2)
    Gu = fixtures["reader_G"]
    assert abs(O.loglik_cpu(Gu, P, Q) - O.loglik_cpu(Gu, Pe, Qe)) <= 50.0
