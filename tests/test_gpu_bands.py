"""The F32 / F64 EM pass in SNP bands: given less workspace than one band
over all SNPs needs, adx_em_pass runs band after band, each band's T / A / B
partial slices in one reused workspace region and its log-likelihood in one
slot per band.  Banded statistics equal the oracle's (the reference's
em_batch_math, em_adam_gpu.py:46-78, summed over chunks at :187-197) and the
single-band pass's to the summation-order bound, and are deterministic."""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import _lib
from adamixture_amd import genotypes as GT

pytestmark = pytest.mark.gpu


def _case(M, N, K, seed, dev, miss=0.02):
    rng = np.random.default_rng(seed)
    P = np.clip(rng.beta(0.6, 1.2, size=(M, K)), 1e-5, 1 - 1e-5)
    Q = rng.dirichlet(np.full(K, 0.4), size=N)
    G = GT.synthesize(torch.from_numpy(P).to(dev), torch.from_numpy(Q).to(dev), seed + 7, miss)
    return G, P, Q


def _pass(G, P, Q, dtype, ws_bytes, loglik):
    dev = G.device
    M, K = P.shape
    N = Q.shape[0]
    tdt = torch.float64 if dtype == _lib.ADX_F64 else torch.float32
    Pd = torch.from_numpy(P).to(dev, tdt)
    Qd = torch.from_numpy(Q).to(dev, tdt)
    A, B, T = (torch.zeros(r, K, dtype=tdt, device=dev) for r in (M, M, N))
    ws = torch.empty(max(ws_bytes, 8), dtype=torch.uint8, device=dev)
    ll = torch.zeros(1, dtype=torch.float64, device=dev)
    flags = _lib.ADX_EM_STATS | (_lib.ADX_EM_LOGLIK if loglik else 0)
    _lib.call("adx_em_pass", dtype, G.data.data_ptr(), G.ldg, M, N, K, Pd.data_ptr(), K, Qd.data_ptr(), K,
              A.data_ptr(), B.data_ptr(), T.data_ptr(), flags, ll.data_ptr() if loglik else None, ws.data_ptr(),
              ws_bytes, _lib.stream_handle(dev))
    return [x.double().cpu().numpy() for x in (A, B, T)], ll.item()


@pytest.mark.parametrize("K", [7, 20])
@pytest.mark.parametrize("frac", [4, 40, 100])
def test_banded_f64_pass_matches_oracle_and_single_band(cuda, K, frac):
    """frac: the workspace is the single-band size / frac (4: a few bands;
    40 and 100: one 256-SNP workgroup column per band, with fewer sample
    splits, down to one at 100)."""
    M, N = 6001, 1037
    G, P, Q = _case(M, N, K, K + frac, cuda)
    Gp = G.to_host_packed()
    full = _lib.load().adx_em_workspace_bytes(_lib.ADX_F64, M, N, K)
    one, ll1 = _pass(G, P, Q, _lib.ADX_F64, full, True)
    band, llb = _pass(G, P, Q, _lib.ADX_F64, full // frac, True)
    band2, llb2 = _pass(G, P, Q, _lib.ADX_F64, full // frac, True)
    Ao, Bo, To, _ = O.em_stats_gpu(Gp, P, Q, M)
    llo = O.loglik_gpu(Gp, P, Q, M)
    for got, ref, want in zip(band, one, (Ao, Bo, To)):
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    assert llb == pytest.approx(llo, rel=1e-12) and ll1 == pytest.approx(llo, rel=1e-12)
    for x, y in zip(band, band2):                 # deterministic for a given workspace
        assert np.array_equal(x, y)
    assert llb == llb2


def test_banded_f32_pass_matches_single_band(cuda):
    M, N, K = 9000, 700, 10
    G, P, Q = _case(M, N, K, 3, cuda)
    full = _lib.load().adx_em_workspace_bytes(_lib.ADX_F32, M, N, K)
    one, _ = _pass(G, P, Q, _lib.ADX_F32, full, False)
    band, _ = _pass(G, P, Q, _lib.ADX_F32, full // 8, False)
    for got, ref in zip(band, one):
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max())


def test_loglik_needs_a_workspace_that_fits_one_band(cuda):
    """Too small a workspace for the log-likelihood is an error, not a
    silent wrong value; the statistics alone fall back to atomics."""
    M, N, K = 2000, 300, 5
    G, P, Q = _case(M, N, K, 9, cuda)
    with pytest.raises(RuntimeError, match="ADX_EM_LOGLIK needs a workspace"):
        _pass(G, P, Q, _lib.ADX_F64, 64, True)
    Gp = G.to_host_packed()
    got, _ = _pass(G, P, Q, _lib.ADX_F64, 64, False)
    for x, want in zip(got, O.em_stats_gpu(Gp, P, Q, M)[:3]):
        np.testing.assert_allclose(x, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_optimizer_workspace_is_capped(cuda):
    """The Adam-EM optimizer's FP64 pass asks for at most max(G / 10, 256 MiB)
    of workspace; with the cap lowered it runs banded and its EM targets
    equal the uncapped run's to the summation-order bound."""
    from adamixture_amd import em_adam_gpu as E
    M, N, K = 12000, 900, 6
    G, P, Q = _case(M, N, K, 11, cuda)
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    full = _lib.load().adx_em_workspace_bytes(_lib.ADX_F64, M, N, K)
    assert E._WS.cap(G) >= 256 << 20
    opt = E.EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision="f64")
    PE1, QE1 = (x.clone() for x in opt.run_em_step(G, Pd, Qd, M))
    saved = (E._WS.min_cap, E._WS.frac)
    try:
        E._WS.min_cap, E._WS.frac = full // 6, 0.0
        E._WS.buf = None
        assert E._WS.cap(G) < full
        PE2, QE2 = opt.run_em_step(G, Pd, Qd, M)
        np.testing.assert_allclose(PE2.cpu().numpy(), PE1.cpu().numpy(), rtol=1e-12, atol=1e-15)
        np.testing.assert_allclose(QE2.cpu().numpy(), QE1.cpu().numpy(), rtol=1e-12, atol=1e-15)
        assert E._WS.buf.numel() <= full // 6
    finally:
        E._WS.min_cap, E._WS.frac = saved
        E._WS.buf = None
