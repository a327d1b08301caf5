"""Kernels of the learned warm start (adx_mlp_*) against the FP64 oracle
(oracle/mlp_oracle.py): GEMM forms, gradients, Adam, BF16 inference."""
import numpy as np
import pytest
import torch

from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _case(seed, rows):
    rng = np.random.default_rng(seed)
    w = (rng.standard_normal(MO.NPARAM) * 0.08).astype(np.float32)
    X = rng.random((rows, MO.D0)).astype(np.float32)
    kvec = rng.integers(2, MO.D3 + 1, size=rows).astype(np.int32)
    T = np.zeros((rows, MO.D3), np.float32)
    for r, k in enumerate(kvec):
        X[r, :MO.D3] = 0
        X[r, :k] = rng.dirichlet(np.ones(k))
        T[r, :k] = rng.dirichlet(np.ones(k))
    return w, X, kvec, T


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("splits", [1, 5])
def test_mlp_gemm_forms(cuda, ta, tb, splits):
    from adamixture_amd import _lib
    rng = np.random.default_rng(10 * ta + tb + splits)
    m, n, k = 70, 130, 333
    A = rng.standard_normal((k, m) if ta else (m, k)).astype(np.float32)
    B = rng.standard_normal((n, k) if tb else (k, n)).astype(np.float32)
    C0 = rng.standard_normal((m, n)).astype(np.float32)
    ref = (A.T if ta else A).astype(np.float64) @ (B.T if tb else B).astype(np.float64)
    beta = 1.0 if splits > 1 else 0.5
    ref = 0.75 * ref + beta * C0
    dA, dB, dC = (torch.from_numpy(x).to(cuda) for x in (A, B, C0))
    print(f"gemm ta={ta} tb={tb} splits={splits}", flush=True)
    _lib.call("adx_mlp_gemm_f32", ta, tb, m, n, k, 0.75, dA.data_ptr(), A.shape[1], dB.data_ptr(), B.shape[1],
              beta, dC.data_ptr(), n, splits, _lib.stream_handle(cuda))
    np.testing.assert_allclose(dC.cpu().numpy(), ref, rtol=0, atol=2e-4 * np.sqrt(k))


def test_mlp_gemm_rejects_split_without_beta_one(cuda):
    from adamixture_amd import _lib
    x = torch.zeros((64, 64), device=cuda)
    with pytest.raises(RuntimeError, match="beta = 1"):
        _lib.call("adx_mlp_gemm_f32", 0, 0, 64, 64, 64, 1.0, x.data_ptr(), 64, x.data_ptr(), 64, 0.0,
                  x.data_ptr(), 64, 4, _lib.stream_handle(cuda))


def test_mlp_gradients_match_oracle(cuda):
    """FP32 forward/backward kernels vs the FP64 oracle gradient (pinned by
    finite differences in test_oracle): per parameter block, max error
    <= 1e-4 of the block's max |gradient|."""
    from adamixture_amd.learned_init import WarmStartMLP
    w, X, kvec, T = _case(7, 1500)
    model = WarmStartMLP(cuda, params=w)
    loss = model.loss_and_grad(torch.from_numpy(X).to(cuda), torch.from_numpy(T).to(cuda),
                               torch.from_numpy(kvec).to(cuda))
    g_gpu = model.g.cpu().numpy().astype(np.float64)
    g_ref = MO.backward(w.astype(np.float64), X, kvec, T)
    l_ref = MO.loss(w.astype(np.float64), X, kvec, T)
    assert abs(float(loss) - l_ref) <= 1e-5 * abs(l_ref)
    off = 0
    for name, shp in MO.SHAPES:
        n = int(np.prod(shp))
        a, b = g_gpu[off:off + n], g_ref[off:off + n]
        err = np.max(np.abs(a - b))
        print(f"grad {name}: max err {err:.3e} max |g| {np.max(np.abs(b)):.3e}", flush=True)
        assert err <= 1e-4 * np.max(np.abs(b)) + 1e-9, name
        off += n
    probs = model.probs(torch.from_numpy(X).to(cuda), torch.from_numpy(kvec).to(cuda)).cpu().numpy()
    p_ref, _ = MO.forward(w.astype(np.float64), X, kvec)
    np.testing.assert_allclose(probs, p_ref, atol=2e-6)


def test_mlp_adam_matches_oracle(cuda):
    from adamixture_amd.learned_init import WarmStartMLP
    w, X, kvec, T = _case(8, 600)
    model = WarmStartMLP(cuda, params=w)
    wr, m, v = w.astype(np.float64), np.zeros(MO.NPARAM), np.zeros(MO.NPARAM)
    for t in (1, 2, 3):
        model.loss_and_grad(torch.from_numpy(X).to(cuda), torch.from_numpy(T).to(cuda),
                            torch.from_numpy(kvec).to(cuda))
        g = model.g.cpu().numpy().astype(np.float64)      # same gradient on both sides
        model.adam_step(1e-3)
        wr, m, v = MO.adam(wr, g, m, v, t, 1e-3)
        wr = wr.astype(np.float32).astype(np.float64)      # the GPU keeps FP32 weights
    np.testing.assert_allclose(model.w.cpu().numpy(), wr, rtol=0, atol=2e-6)


@pytest.mark.parametrize("K", [2, 7, 12, 16])
def test_mlp_infer_bf16_matches_oracle(cuda, K):
    """Fused BF16 MFMA inference vs the oracle's BF16 restatement (same
    roundings, FP64 sums) and vs the FP32 network."""
    from adamixture_amd.learned_init import WarmStartMLP
    w, X, _, _ = _case(20 + K, 1000 + 37)            # ragged last tile
    rng = np.random.default_rng(K)
    X[:, :MO.D3] = 0
    X[:, :K] = rng.dirichlet(np.full(K, 0.5), size=X.shape[0])
    model = WarmStartMLP(cuda, params=w)
    dX = torch.from_numpy(X).to(cuda)
    print(f"infer K={K}", flush=True)
    Q = model.infer(dX, K).cpu().numpy()
    assert Q.shape == (X.shape[0], K)
    np.testing.assert_allclose(Q.sum(1), 1.0, atol=1e-12)
    assert Q.min() >= 0.999e-5          # clamped at 1e-5 in FP32, then renormalised
    Q_ref = MO.infer_bf16(w, X, K)
    err = np.max(np.abs(Q - Q_ref))
    kv = np.full(X.shape[0], K, np.int32)
    P32 = model.probs(dX, torch.from_numpy(kv).to(cuda)).cpu().numpy()[:, :K]
    err32 = np.max(np.abs(Q - P32))
    print(f"infer K={K}: max |Q - oracle bf16| {err:.2e}, max |Q - fp32| {err32:.2e}", flush=True)
    assert err < 5e-3
    assert err32 < 3e-2


def test_mlp_training_reduces_loss(cuda):
    """Fifty Adam steps on a fixed batch with a learnable target (a teacher
    network's output) cut the cross-entropy gap to the teacher by half."""
    from adamixture_amd.learned_init import WarmStartMLP, init_params
    _, X, kvec, _ = _case(30, 2048)
    teacher = (init_params(99) * 3.0).astype(np.float32)
    T, _ = MO.forward(teacher.astype(np.float64), X, kvec)
    T = T.astype(np.float32)
    ent = -np.sum(np.where(T > 0, T * np.log(np.maximum(T, 1e-30)), 0)) / X.shape[0]
    model = WarmStartMLP(cuda, seed=1)
    dX, dT, dk = (torch.from_numpy(a).to(cuda) for a in (X, T, kvec))
    l0 = float(model.loss_and_grad(dX, dT, dk)) - ent
    for _ in range(50):
        model.loss_and_grad(dX, dT, dk)
        model.adam_step(3e-3)
    l1 = float(model.loss_and_grad(dX, dT, dk)) - ent
    print(f"training gap {l0:.4f} -> {l1:.4f}", flush=True)
    assert l1 < 0.5 * l0
