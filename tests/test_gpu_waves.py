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


def _stats(G, P, Q, dtype, ws_bytes, mode):
    """Statistics-only pass through the C ABI: resident with a workspace of
    ws_bytes, host-streamed with the same workspace, or allocation-free."""
    dev = torch.device("cuda:0")
    M, K = P.shape
    N = Q.shape[0]
    tdt = torch.float64 if dtype == _lib.ADX_F64 else torch.float32
    Pd, Qd = torch.from_numpy(P).to(dev, tdt), torch.from_numpy(Q).to(dev, tdt)
    A, B, T = (torch.zeros(r, K, dtype=tdt, device=dev) for r in (M, M, N))
    args = (M, N, K, Pd.data_ptr(), K, Qd.data_ptr(), K, A.data_ptr(), B.data_ptr(), T.data_ptr(),
            _lib.ADX_EM_STATS, None)
    st = _lib.stream_handle(dev)
    if mode == "noalloc":
        _lib.call("adx_em_pass_noalloc", dtype, G.data.data_ptr(), G.ldg, *args, st)
    else:
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
        if mode == "streamed":
            band = _lib.load().adx_em_band_snps(dtype, M, N, K, _lib.ADX_EM_STATS, ws_bytes, 0)
            nst = 2 * (-(-((band + 3) // 4) * G.ldg // 256) * 256)
            stg = torch.empty(nst, dtype=torch.uint8, device=dev)
            cs = torch.cuda.Stream(device=dev)
            _lib.call("adx_em_pass_host", dtype, G.host.data_ptr(), G.ldg, *args, ws.data_ptr(), ws_bytes,
                      stg.data_ptr(), nst, st, cs.cuda_stream)
        else:
            _lib.call("adx_em_pass", dtype, G.data.data_ptr(), G.ldg, *args, ws.data_ptr(), ws_bytes, st)
    torch.cuda.synchronize(dev)
    return A, B, T


@pytest.mark.parametrize("dtype,cw", [(_lib.ADX_F64, 4), (_lib.ADX_F32, 8)])
@pytest.mark.parametrize("frac", [1, 4])
def test_8_wave_streamed_pass_equals_resident(cuda, tiles, dtype, cw, frac):
    """The SNP-band plan and the host-streamed pass at 8 waves: with the same
    workspace (one band, or a quarter of it: several bands) the streamed
    statistics are bit-identical to the resident ones, and match the oracle."""
    tiles("f64" if dtype == _lib.ADX_F64 else "f32", cw, 8, 4096)
    M, N, K = 12003, 1501, 9
    print(f"case: 8-wave streamed dtype={dtype} frac={frac}", flush=True)
    P, Q, Gd = _case(M, N, K, 0.02, 21, cuda)
    Gh = GT.HostPackedGenotypes.from_packed(Gd.to_host_packed(), M, cuda)
    ws = _lib.load().adx_em_workspace_bytes(dtype, M, N, K) // frac
    res = _stats(Gd, P, Q, dtype, ws, "resident")
    strm = _stats(Gh, P, Q, dtype, ws, "streamed")
    assert all(torch.equal(a, b) for a, b in zip(res, strm))
    tdt = np.float64 if dtype == _lib.ADX_F64 else np.float32
    Pc, Qc = P.astype(tdt).astype(np.float64), Q.astype(tdt).astype(np.float64)
    tol = 1e-12 if dtype == _lib.ADX_F64 else 2e-4
    for got, want in zip(res, O.em_stats_gpu(Gd.to_host_packed(), Pc, Qc, M)[:3]):
        np.testing.assert_allclose(got.double().cpu().numpy(), want, rtol=tol, atol=tol * np.abs(want).max())


def test_8_wave_noalloc_pass_equals_workspace_pass(cuda, tiles):
    """adx_em_pass_noalloc plans its bands for the 8-wave shape too: equal to
    adx_em_pass with a workspace of the fixed scratch's size, twice in a row
    (the second call on a smaller shape)."""
    tiles("f32", 8, 8, 4096)
    for M, N, K in ((20011, 2001, 7), (3001, 333, 5)):
        print(f"case: 8-wave noalloc M={M} N={N} K={K}", flush=True)
        P, Q, Gd = _case(M, N, K, 0.01, 22, cuda)
        ref = _stats(Gd, P, Q, _lib.ADX_F32, int(_lib.load().adx_em_workspace_bytes(_lib.ADX_F32, M, N, K)),
                     "resident")
        got = _stats(Gd, P, Q, _lib.ADX_F32, 0, "noalloc")
        for a, b in zip(got, ref):
            np.testing.assert_allclose(a.cpu().numpy(), b.cpu().numpy(), rtol=1e-5, atol=1e-5 * float(b.abs().max()))
