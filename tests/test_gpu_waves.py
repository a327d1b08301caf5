"""EM passes with 8-wave workgroups (adx_set_em_tiles waves = 8: a workgroup
owns 16 cw 8 SNPs and writes half as many partial-T slices as with 4 waves):
every form against the oracle at ragged shapes, K from 1 to 16, with the
bounds of the 4-wave tests (FP64 1e-12, FP32 2e-4, BF16 1e-3 relative), and
bit-identical run to run."""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import _lib, autotune
from adamixture_amd import genotypes as GT
from adamixture_amd.em_adam_gpu import EMAdamOptimizer, em_pass

pytestmark = pytest.mark.gpu

SHAPES = [  # (M, N, K, missing rate): M below, at and past one workgroup's SNPs
    (37, 29, 3, 0.05),
    (1001, 777, 7, 0.01),
    (4101, 130, 16, 0.02),
    (12, 64, 1, 0.0),
    (3000, 1000, 10, 0.3),
    (2049, 200, 12, 0.01),
]


@pytest.fixture
def tiles():
    saved = {p: autotune.get_tiles(p) for p in ("f32", "f64", "bf16")}
    yield autotune.set_tiles
    for p, t in saved.items():
        autotune.set_tiles(p, *t)


def _case(M, N, K, miss, seed, dev):
    rng = np.random.default_rng(seed)
    P = np.clip(rng.beta(0.6, 1.2, size=(M, K)), 1e-5, 1 - 1e-5)
    Q = rng.dirichlet(np.full(K, 0.4), size=N)
    G = GT.synthesize(torch.from_numpy(P).to(dev), torch.from_numpy(Q).to(dev), seed + 7, miss)
    return P, Q, G


def _zeros(dev, dt, *shape):
    return torch.zeros(shape, dtype=dt, device=dev)


@pytest.mark.parametrize("M,N,K,miss", SHAPES)
def test_f64_pass_with_8_waves_matches_oracle(cuda, tiles, M, N, K, miss):
    tiles("f64", 4, 8, 4096)
    print(f"case: f64 waves 8 M={M} N={N} K={K}", flush=True)
    P, Q, G = _case(M, N, K, miss, 2, cuda)
    Gp = G.to_host_packed()
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    Ao, Bo, To, _ = O.em_stats_gpu(Gp, P, Q, M)
    for loglik in (False, True):         # the fused log-likelihood pass runs 4-wave workgroups
        A, B, T = (_zeros(cuda, torch.float64, M, K), _zeros(cuda, torch.float64, M, K),
                   _zeros(cuda, torch.float64, N, K))
        ll = em_pass(G, Pd, Qd, A, B, T, dtype=_lib.ADX_F64, loglik=loglik)
        for got, want in ((A, Ao), (B, Bo), (T, To)):
            np.testing.assert_allclose(got.cpu().numpy(), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        if loglik:
            assert ll.item() == pytest.approx(O.loglik_gpu(Gp, P, Q, M), rel=1e-12)


@pytest.mark.parametrize("cw", [4, 8])
@pytest.mark.parametrize("M,N,K,miss", SHAPES)
def test_f32_pass_with_8_waves_matches_oracle(cuda, tiles, cw, M, N, K, miss):
    tiles("f32", cw, 8, 8192)
    print(f"case: f32 cw {cw} waves 8 M={M} N={N} K={K}", flush=True)
    P, Q, G = _case(M, N, K, miss, 3, cuda)
    Gp = G.to_host_packed()
    P32 = torch.from_numpy(P).to(cuda, torch.float32)
    Q32 = torch.from_numpy(Q).to(cuda, torch.float32)
    A, B, T = _zeros(cuda, torch.float32, M, K), _zeros(cuda, torch.float32, M, K), _zeros(cuda, torch.float32, N, K)
    em_pass(G, P32, Q32, A, B, T, dtype=_lib.ADX_F32)
    Ao, Bo, To, _ = O.em_stats_gpu(Gp, P32.double().cpu().numpy(), Q32.double().cpu().numpy(), M)
    for got, want in ((A, Ao), (B, Bo), (T, To)):
        np.testing.assert_allclose(got.double().cpu().numpy(), want, rtol=2e-4, atol=2e-5 * np.abs(want).max())


@pytest.mark.parametrize("cw", [4, 8])
@pytest.mark.parametrize("M,N,K,miss", SHAPES)
def test_bf16_pass_with_8_waves_matches_oracle(cuda, tiles, cw, M, N, K, miss):
    tiles("bf16", cw, 8, 8192)
    print(f"case: bf16 cw {cw} waves 8 M={M} N={N} K={K}", flush=True)
    P, Q, G = _case(M, N, K, miss, 3, cuda)
    Gp = G.to_host_packed()
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision="bf16")
    opt.accumulate(G, Pd, Qd)
    Ao, Bo, To, _ = O.em_stats_gpu(Gp, P, Q, M)
    for got, want in ((opt.A_accum, Ao), (opt.B_accum, Bo), (opt.T_accum, To)):
        np.testing.assert_allclose(got.double().cpu().numpy(), want, rtol=1e-3, atol=1e-3 * np.abs(want).max())


@pytest.mark.parametrize("precision,cw", [("f64", 4), ("f32", 8), ("bf16", 8), ("bf16", 4)])
def test_8_wave_pass_is_deterministic(cuda, tiles, precision, cw):
    from adamixture_amd.synth import Workload, make_genotypes
    tiles(precision, cw, 8, 4096)
    print(f"case: deterministic {precision} cw {cw} waves 8", flush=True)
    w = Workload("det", 3000, 20000, 6, "bn", 0.3, 0.01, fst=0.1, seed=78)
    G, P0, Q0 = make_genotypes(w, cuda)
    P = torch.from_numpy(np.clip(P0 * 0.9 + 0.05, 1e-5, 1 - 1e-5)).to(cuda)
    Q = torch.from_numpy(Q0).to(cuda)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision)
    out = []
    for _ in range(2):
        PE, QE = opt.run_em_step(G, P, Q, G.M)
        out.append((PE.clone(), QE.clone()))
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])
