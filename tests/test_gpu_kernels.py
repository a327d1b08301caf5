"""Parity of every libadx kernel with the CPU oracle, through the C ABI.

Integer/byte work is compared bit-exactly; FP64 EM statistics to 1e-12
relative (only the summation order differs: MFMA tiles and atomics vs the
oracle's sequential sums); the FP32 form to 2e-4 relative (FP32 MFMA).
"""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import _lib
from adamixture_amd import genotypes as GT
from adamixture_amd.em_adam_gpu import EMAdamOptimizer, em_pass

pytestmark = pytest.mark.gpu

SHAPES = [  # (M, N, K, missing rate): ragged M and N, K from 1 to 16
    (37, 29, 3, 0.05),
    (1001, 777, 7, 0.01),
    (4101, 130, 16, 0.02),
    (12, 64, 1, 0.0),
    (3000, 1000, 10, 0.3),
]


def _params(M, N, K, seed=0):
    rng = np.random.default_rng(seed)
    P = np.clip(rng.beta(0.6, 1.2, size=(M, K)), 1e-5, 1 - 1e-5)
    Q = rng.dirichlet(np.full(K, 0.4), size=N)
    return P, Q


def _device(P, Q, seed, miss, dev):
    return GT.synthesize(torch.from_numpy(P).to(dev), torch.from_numpy(Q).to(dev), seed, miss)


# accumulator row held by lane group g, element j (adx_common.h)
ROW = {0: lambda g, j: 4 * g + j, 1: lambda g, j: g + 4 * j}


def test_mfma_lane_layout(cuda):
    rng = np.random.default_rng(1)
    for kind, dt, tol in ((0, np.float32, 1e-6), (1, np.float64, 1e-14)):
        A = rng.normal(size=(16, 4)).astype(dt)
        B = rng.normal(size=(4, 16)).astype(dt)
        want = A.astype(np.float64) @ B.astype(np.float64)
        a, b = torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda)
        raw = torch.zeros(256, dtype=a.dtype, device=cuda)
        _lib.call("adx_probe_mfma", kind + 2, a.data_ptr(), b.data_ptr(), raw.data_ptr(), _lib.stream_handle(cuda))
        raw = raw.cpu().numpy()
        # which (row, col) of A @ B each raw accumulator holds
        where = [np.unravel_index(np.argmin(np.abs(want - raw[l * 4 + j])), want.shape)
                 for l in range(64) for j in range(4)]
        print(f"kind {kind}: lane 16 holds rows {[where[16 * 4 + j][0] for j in range(4)]}")
        for l in range(64):
            for j in range(4):
                assert where[l * 4 + j] == (ROW[kind](l // 16, j), l % 16), (kind, l, j, where[l * 4 + j])
        d = torch.zeros((16, 16), dtype=a.dtype, device=cuda)
        _lib.call("adx_probe_mfma", kind, a.data_ptr(), b.data_ptr(), d.data_ptr(), _lib.stream_handle(cuda))
        np.testing.assert_allclose(d.cpu().numpy(), want, rtol=tol, atol=tol)


@pytest.mark.parametrize("kind,dt", [(4, torch.bfloat16), (5, torch.float16)])
def test_mfma_16x16x32_lane_layout(cuda, kind, dt):
    """16x16x32 BF16 and FP16 MFMA: operands A[l%16][8(l/16)+i],
    B[8(l/16)+i][l%16], accumulator element j of lane l = D[4(l/16)+j][l%16]."""
    rng = np.random.default_rng(2)
    A = rng.integers(-8, 9, size=(16, 32)).astype(np.float32)
    B = rng.integers(-8, 9, size=(32, 16)).astype(np.float32)
    A[np.arange(16), np.arange(16)] += 50.0 * (1 + np.arange(16))   # make rows of A @ B distinct
    # the probe rounds its FP32 inputs to the operand type (the diagonal above
    # 256 is not exact in BF16's 8 bits); integer products and sums are exact
    # in FP32
    A = torch.from_numpy(A).to(dt).to(torch.float32).numpy()
    want = A.astype(np.float64) @ B.astype(np.float64)
    raw = torch.zeros(256, dtype=torch.float32, device=cuda)
    a, b = torch.from_numpy(A).to(cuda), torch.from_numpy(B).to(cuda)
    _lib.call("adx_probe_mfma", kind, a.data_ptr(), b.data_ptr(), raw.data_ptr(), _lib.stream_handle(cuda))
    raw = raw.cpu().numpy().astype(np.float64)
    for l in range(64):
        for j in range(4):
            assert raw[l * 4 + j] == want[4 * (l // 16) + j, l % 16], (l, j)


@pytest.mark.parametrize("M,N,K,miss", SHAPES)
def test_synth_bit_exact(cuda, M, N, K, miss):
    P, Q = _params(M, N, K)
    G = _device(P, Q, 42, miss, cuda)
    host = G.data.cpu().numpy()
    assert np.array_equal(host[:, :N], O.synth_packed(P, Q, 42, miss))
    assert np.all(host[:, N:] == 0xFF)


@pytest.mark.parametrize("M,N,K,miss", SHAPES[:3])
def test_unpack_freq_qbat_bit_exact(cuda, M, N, K, miss):
    P, Q = _params(M, N, K, 1)
    G = _device(P, Q, 7, miss, cuda)
    Gp = G.to_host_packed()
    for s, c in ((0, M), (5, 11), (M - 3, 100)):
        assert np.array_equal(GT.unpack_chunk(G, s, c).cpu().numpy(), O.unpack(Gp, M, s, c))
    f = torch.rand(M, dtype=torch.float32, device=cuda)
    got = GT.unpack_center_chunk(G, f, 4, 50).cpu().numpy()
    assert np.array_equal(got, O.unpack_center(Gp, f.cpu().numpy(), M, 4, 50))
    assert np.array_equal(GT.allele_frequencies(G).cpu().numpy(), O.freq_gpu(Gp, M))
    assert np.array_equal(G.qbat().cpu().numpy(), O.qbat(Gp, M))
    # an unpadded host tensor is repacked into the padded layout bit-exactly
    G2 = GT.as_packed(torch.from_numpy(Gp), M, cuda)
    assert G2.ldg % 64 == 0 and np.array_equal(G2.data.cpu().numpy(), G.data.cpu().numpy())


def _zeros(dev, dt, *shape):
    return torch.zeros(*shape, dtype=dt, device=dev)


@pytest.mark.parametrize("M,N,K,miss", SHAPES)
def test_em_pass_f64_matches_oracle(cuda, M, N, K, miss):
    P, Q = _params(M, N, K, 2)
    G = _device(P, Q, 11, miss, cuda)
    Gp = G.to_host_packed()
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    A, B, T = _zeros(cuda, torch.float64, M, K), _zeros(cuda, torch.float64, M, K), _zeros(cuda, torch.float64, N, K)
    ll = em_pass(G, Pd, Qd, A, B, T, dtype=_lib.ADX_F64, loglik=True)
    Ao, Bo, To, qb = O.em_stats_gpu(Gp, P, Q, M)
    np.testing.assert_allclose(A.cpu().numpy(), Ao, rtol=1e-12, atol=1e-12 * np.abs(Ao).max())
    np.testing.assert_allclose(B.cpu().numpy(), Bo, rtol=1e-12, atol=1e-12 * np.abs(Bo).max())
    np.testing.assert_allclose(T.cpu().numpy(), To, rtol=1e-12, atol=1e-12 * np.abs(To).max())
    assert ll.item() == pytest.approx(O.loglik_gpu(Gp, P, Q, M), rel=1e-12)
    ll2 = em_pass(G, Pd, Qd, dtype=_lib.ADX_F64, loglik=True, stats=False)
    assert ll2.item() == pytest.approx(O.loglik_gpu(Gp, P, Q, M), rel=1e-12)


@pytest.mark.parametrize("M,N,K,miss", SHAPES)
def test_em_pass_f32_matches_oracle(cuda, M, N, K, miss):
    P, Q = _params(M, N, K, 3)
    G = _device(P, Q, 13, miss, cuda)
    Gp = G.to_host_packed()
    P32 = torch.from_numpy(P).to(cuda, torch.float32)
    Q32 = torch.from_numpy(Q).to(cuda, torch.float32)
    A, B, T = _zeros(cuda, torch.float32, M, K), _zeros(cuda, torch.float32, M, K), _zeros(cuda, torch.float32, N, K)
    em_pass(G, P32, Q32, A, B, T, dtype=_lib.ADX_F32)
    Ao, Bo, To, _ = O.em_stats_gpu(Gp, P32.double().cpu().numpy(), Q32.double().cpu().numpy(), M)
    for got, want in ((A, Ao), (B, Bo), (T, To)):
        np.testing.assert_allclose(got.double().cpu().numpy(), want, rtol=2e-4, atol=2e-5 * np.abs(want).max())


def test_em_pass_all_missing_rows_and_columns(cuda):
    M, N, K = 260, 200, 4
    P, Q = _params(M, N, K, 4)
    G = _device(P, Q, 5, 0.0, cuda)
    Gp = G.to_host_packed()
    Gu = O.unpack(Gp, M)
    Gu[:, 17] = 3        # a sample with no observed genotype
    Gu[101] = 3          # a SNP with no observed genotype
    Gp = O.pack(Gu)
    G = GT.as_packed(torch.from_numpy(Gp), M, cuda)
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision="f64")
    PE, QE = opt.run_em_step(G, Pd, Qd, M)
    PEo, QEo = O.em_final_gpu(P, Q, *O.em_stats_gpu(Gp, P, Q, M))
    np.testing.assert_allclose(PE.cpu().numpy(), PEo, rtol=1e-11, atol=1e-14)
    np.testing.assert_allclose(QE.cpu().numpy(), QEo, rtol=1e-11, atol=1e-14)
    assert np.isfinite(PE.cpu().numpy()).all() and np.isfinite(QE.cpu().numpy()).all()


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_run_em_step_and_adam_step_match_oracle(cuda, precision):
    M, N, K = 1500, 700, 6
    P, Q = _params(M, N, K, 5)
    G = _device(P, Q, 17, 0.01, cuda)
    Gp = G.to_host_packed()
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision=precision)
    ref = O.AdamEMGPU(M, N, K, 0.005, 0.8, 0.88, 1e-8)
    Po, Qo = P.copy(), Q.copy()
    tol = 1e-11 if precision == "f64" else 1e-4
    for _ in range(3):
        PE, QE = opt.run_em_step(G, Pd, Qd, M)
        PEo, QEo = ref.run_em_step(Gp, Po, Qo, M)
        np.testing.assert_allclose(PE.cpu().numpy(), PEo, rtol=tol, atol=tol * 1e-3)
        np.testing.assert_allclose(QE.cpu().numpy(), QEo, rtol=tol, atol=tol * 1e-3)
        opt.step(Pd, Qd, PE, QE)
        ref.step(Po, Qo, PE.cpu().numpy(), QE.cpu().numpy())   # same targets: isolates the Adam step
        np.testing.assert_allclose(Pd.cpu().numpy(), Po, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(Qd.cpu().numpy(), Qo, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("setting", [(0.005, 0.80, 0.88, 1e-8), (0.05, 0.70, 0.95, 1e-3)])
def test_adam_step_is_kingma_ba_textbook(cuda, setting):
    """adx_adam_step (em_adam_gpu.py:11-42, :201-224) over three steps
    against Kingma & Ba written out by hand (tests/textbook.py), on a case
    whose steps cross both clamp bounds; P, Q and the moments to rounding."""
    from tests import textbook as TB
    lr, b1, b2, eps = setting
    P0, Q0, P_t, Q_t = TB.adam_em_case()
    M, K = P0.shape
    N = Q0.shape[0]
    opt = EMAdamOptimizer(P0.shape, Q0.shape, lr, b1, b2, eps, cuda, precision="f64")
    Pd, Qd = torch.from_numpy(P0.copy()).to(cuda), torch.from_numpy(Q0.copy()).to(cuda)
    for tp, tq in zip(P_t, Q_t):
        opt.step(Pd, Qd, torch.from_numpy(tp).to(cuda), torch.from_numpy(tq).to(cuda))
    ref = TB.adam_em_steps(P0, Q0, P_t, Q_t, lr, b1, b2, eps)
    got = (Pd, Qd, opt.m_P, opt.v_P, opt.m_Q, opt.v_Q)
    for name, g, r in zip(("P", "Q", "mP", "vP", "mQ", "vQ"), got, ref):
        np.testing.assert_allclose(g.cpu().numpy(), np.array(r), rtol=1e-12, atol=1e-15, err_msg=name)


@pytest.mark.parametrize("M,N,K,miss", SHAPES)
def test_em_pass_bf16_error_vs_oracle(cuda, M, N, K, miss):
    """Mixed-precision form against the FP64 oracle at the same (P, Q).
    Stated bound: 1e-3 relative on A, B, T (H from the three-term BF16 split
    is good to ~5e-5; the rest is the rounding of a, b, Q and P to FP16,
    averaged over the contraction)."""
    P, Q = _params(M, N, K, 3)
    G = _device(P, Q, 13, miss, cuda)
    Gp = G.to_host_packed()
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision="bf16")
    opt.accumulate(G, Pd, Qd)
    Ao, Bo, To, _ = O.em_stats_gpu(Gp, P, Q, M)
    errs = []
    for got, want in ((opt.A_accum, Ao), (opt.B_accum, Bo), (opt.T_accum, To)):
        g64 = got.double().cpu().numpy()
        errs.append(np.abs(g64 - want).max() / np.abs(want).max())
        np.testing.assert_allclose(g64, want, rtol=1e-3, atol=1e-3 * np.abs(want).max())
    print(f"bf16 M={M} N={N} K={K}: max rel err A {errs[0]:.2e} B {errs[1]:.2e} T {errs[2]:.2e}")


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16"])
def test_em_pass_is_deterministic(cuda, precision):
    """T, A and B are reduced through partial slices summed in a fixed order,
    so two passes over the same input give bit-identical statistics."""
    from adamixture_amd.em_adam_gpu import EMAdamOptimizer
    from adamixture_amd.synth import Workload, make_genotypes
    w = Workload("det", 3000, 20000, 6, "bn", 0.3, 0.01, fst=0.1, seed=77)
    G, P0, Q0 = make_genotypes(w, cuda)
    P = torch.from_numpy(np.clip(P0 * 0.9 + 0.05, 1e-5, 1 - 1e-5)).to(cuda)
    Q = torch.from_numpy(Q0).to(cuda)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision)
    out = []
    for _ in range(2):
        PE, QE = opt.run_em_step(G, P, Q, G.M)
        out.append((PE.clone(), QE.clone()))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
