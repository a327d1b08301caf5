"""Host-streamed genotypes: when G does not fit the GPU the memory manager
keeps it in pinned host memory (the reference keeps it on the CPU then,
adamixture/src/utils.py:425-428, and copies each byte-row chunk per use,
:538-557), and the passes stream its SNP bands through a device staging
buffer.

With the same workspace (hence the same SNP-band plan) the streamed EM pass
gives A, B, T and the log-likelihood bit-identical to the resident pass; so
does a whole Adam-EM run.  The per-SNP ops (frequencies, unpack) and q_bat
(integer sums) are bit-exact band by band; the SVD products and the BR-QN
gradient / Hessian sum their SNP reductions band by band (equal to rounding).
"""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import _lib, utils
from adamixture_amd import genotypes as GT
from adamixture_amd.br_qn_gpu import grad_hess
from adamixture_amd.em_adam_gpu import EMAdamOptimizer, optimize_parameters_gpu
from adamixture_amd.svd_gpu import xty, xw

pytestmark = pytest.mark.gpu


def _case(M, N, K, seed, dev, miss=0.02):
    rng = np.random.default_rng(seed)
    P = np.clip(rng.beta(0.6, 1.2, size=(M, K)), 1e-5, 1 - 1e-5)
    Q = rng.dirichlet(np.full(K, 0.4), size=N)
    Gd = GT.synthesize(torch.from_numpy(P).to(dev), torch.from_numpy(Q).to(dev), seed + 5, miss)
    Gh = GT.HostPackedGenotypes.from_packed(Gd.to_host_packed(), M, dev)
    return Gd, Gh, P, Q


def _pass(G, P, Q, dtype, ws_bytes, staging_bytes=None):
    dev = G.device
    M, K = P.shape
    N = Q.shape[0]
    tdt = torch.float64 if dtype == _lib.ADX_F64 else torch.float32
    Pd, Qd = torch.from_numpy(P).to(dev, tdt), torch.from_numpy(Q).to(dev, tdt)
    A, B, T = (torch.zeros(r, K, dtype=tdt, device=dev) for r in (M, M, N))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    ll = torch.zeros(1, dtype=torch.float64, device=dev)
    loglik = dtype == _lib.ADX_F64
    flags = _lib.ADX_EM_STATS | (_lib.ADX_EM_LOGLIK if loglik else 0)
    args = (M, N, K, Pd.data_ptr(), K, Qd.data_ptr(), K, A.data_ptr(), B.data_ptr(), T.data_ptr(), flags,
            ll.data_ptr() if loglik else None, ws.data_ptr(), ws_bytes)
    st = _lib.stream_handle(dev)
    if GT.is_streamed(G):
        if staging_bytes is None:   # two halves holding the plan's band
            band = _lib.load().adx_em_band_snps(dtype, M, N, K, flags, ws_bytes, 0)
            staging_bytes = 2 * (-(-((band + 3) // 4) * G.ldg // 256) * 256)
        stg = torch.empty(staging_bytes, dtype=torch.uint8, device=dev)
        cs = torch.cuda.Stream(device=dev)
        _lib.call("adx_em_pass_host", dtype, G.host.data_ptr(), G.ldg, *args, stg.data_ptr(), staging_bytes, st,
                  cs.cuda_stream)
    else:
        _lib.call("adx_em_pass", dtype, G.data.data_ptr(), G.ldg, *args, st)
    torch.cuda.synchronize(dev)
    return (A, B, T), (ll.item() if loglik else None)


@pytest.mark.parametrize("dtype", [_lib.ADX_F64, _lib.ADX_F32])
@pytest.mark.parametrize("frac", [1, 5])
def test_streamed_pass_is_bit_identical_to_resident(cuda, dtype, frac):
    """frac: workspace = one-band size / frac (1: one band; 5: several)."""
    M, N, K = 12003, 1501, 9
    Gd, Gh, P, Q = _case(M, N, K, 1, cuda)
    ws = _lib.load().adx_em_workspace_bytes(dtype, M, N, K) // frac
    print(f"streamed dtype={dtype} frac={frac}", flush=True)
    (A, B, T), ll = _pass(Gd, P, Q, dtype, ws)
    (A2, B2, T2), ll2 = _pass(Gh, P, Q, dtype, ws)
    assert torch.equal(A, A2) and torch.equal(B, B2) and torch.equal(T, T2)
    assert ll == ll2


def test_streamed_pass_with_small_staging_matches_oracle(cuda):
    """A staging that holds less than the plan's band narrows the bands
    (results equal to the oracle to the summation-order bound)."""
    M, N, K = 9001, 777, 6
    Gd, Gh, P, Q = _case(M, N, K, 2, cuda)
    ws = _lib.load().adx_em_workspace_bytes(_lib.ADX_F64, M, N, K)
    (A, B, T), ll = _pass(Gh, P, Q, _lib.ADX_F64, ws, staging_bytes=2 * 300 * Gh.ldg)
    Gp = Gd.to_host_packed()
    for got, want in zip((A, B, T), O.em_stats_gpu(Gp, P, Q, M)[:3]):
        np.testing.assert_allclose(got.cpu().numpy(), want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert ll == pytest.approx(O.loglik_gpu(Gp, P, Q, M), rel=1e-12)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_streamed_adam_em_run_is_bit_identical(cuda, precision):
    M, N, K = 8000, 1200, 5
    Gd, Gh, P, Q = _case(M, N, K, 3, cuda)
    runs = []
    for G in (Gd, Gh):
        hist = []
        Pr, Qr = optimize_parameters_gpu(G, torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda), 0.005, 0.8,
                                         0.88, 1e-8, 25, 5, K, M, N, 0.5, 1e-4, 3, 0.1, cuda, 8192, 256,
                                         precision=precision, on_check=lambda it, L, h=hist: h.append((it, L)))
        runs.append((Pr, Qr, hist))
    assert torch.equal(runs[0][0], runs[1][0]) and torch.equal(runs[0][1], runs[1][1])
    assert runs[0][2] == runs[1][2]


def test_streamed_bf16_form_refuses(cuda):
    M, N, K = 2000, 300, 4
    Gd, Gh, P, Q = _case(M, N, K, 4, cuda)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision="bf16")
    with pytest.raises(NotImplementedError, match="resident"):
        opt.accumulate(Gh, torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda))


def test_streamed_genotype_ops_band_by_band(cuda):
    """Staging capped at 37 byte rows, so every op runs in several bands."""
    M, N, K = 3001, 517, 6
    Gd, Gh, P, Q = _case(M, N, K, 5, cuda)
    Gh.staging_cap = 37 * Gh.ldg
    assert torch.equal(GT.allele_frequencies(Gh), GT.allele_frequencies(Gd))
    assert torch.equal(Gh.qbat(), Gd.qbat())
    for s, c in ((0, 100), (5, 11), (2990, 50), (1001, 400)):
        assert torch.equal(GT.unpack_chunk(Gh, s, c), GT.unpack_chunk(Gd, s, c))
    f = GT.allele_frequencies(Gd)
    assert torch.equal(GT.unpack_center_chunk(Gh, f, 7, 300), GT.unpack_center_chunk(Gd, f, 7, 300))
    rng = np.random.default_rng(0)
    W = torch.from_numpy(rng.normal(size=(N, 20)).astype(np.float32)).to(cuda)
    Y = torch.from_numpy(rng.normal(size=(M, 20)).astype(np.float32)).to(cuda)
    np.testing.assert_allclose(xw(Gh, f, W).cpu().numpy(), xw(Gd, f, W).cpu().numpy(), rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(xty(Gh, f, Y).cpu().numpy(), xty(Gd, f, Y).cpu().numpy(), rtol=1e-5, atol=1e-3)
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    for side in ("q", "p"):
        H1, z1 = grad_hess(Gd, Qd, Pd, side)
        H2, z2 = grad_hess(Gh, Qd, Pd, side)
        np.testing.assert_allclose(H2.cpu().numpy(), H1.cpu().numpy(), rtol=1e-12, atol=1e-12 * H1.abs().max().item())
        np.testing.assert_allclose(z2.cpu().numpy(), z1.cpu().numpy(), rtol=1e-12, atol=1e-12 * z1.abs().max().item())


def test_driver_streams_when_the_matrix_does_not_fit(cuda, monkeypatch):
    """setup -> train_k with the free memory faked below G + buffers: G stays
    on the host, and the run converges to the resident run's likelihood."""
    from adamixture_amd import reader
    from adamixture_amd.adamixture import setup, train_k
    from tests.conftest import GOLDEN
    Gt, N, M = reader.read_bed_packed(GOLDEN / "demo_data.bed")
    K = 7
    args = (42, 0.005, 0.8, 0.88, 1e-8, 10000, 5, 1000, 1e-4, 0.5, 1e-4, 8192, 3, 0.1)
    dev, tpb, f, U, S, V, Gr = setup(Gt.clone(), N, M, K, 42, 5, 0.1, 8192, "cuda:0")
    Pr, Qr = train_k(Gr, N, M, K, U, S, V, f, *args, dev, tpb)
    buffers = utils.estimate_gpu_bytes(M, N, K, "adamem")
    monkeypatch.setattr(utils, "get_free_gpu_memory", lambda d: (buffers + 200 * 128) / 2**20 / 0.95)
    dev, tpb, f2, U2, S2, V2, Gs = setup(Gt.clone(), N, M, K, 42, 5, 0.1, 8192, "cuda:0")
    assert GT.is_streamed(Gs) and not hasattr(Gs, "data")
    assert torch.equal(f2, f)
    Ps, Qs = train_k(Gs, N, M, K, U2, S2, V2, f2, *args, dev, tpb)
    Gp = Gr.to_host_packed()
    lr, ls = O.loglik_gpu(Gp, Pr.cpu().numpy(), Qr.cpu().numpy(), M), O.loglik_gpu(Gp, Ps.cpu().numpy(),
                                                                                   Qs.cpu().numpy(), M)
    print(f"resident logL {lr:.6f}, streamed logL {ls:.6f}")
    assert ls == pytest.approx(lr, rel=1e-6)
