"""adx_em_pass_noalloc (the allocation-free variant of adx_em_pass for callers
that cannot allocate device memory) against adx_em_pass with a caller
workspace, through the C ABI: two calls in a row, the second on a smaller
input, in both dtypes with the log-likelihood; a shape that needs SNP bands
inside the fixed scratch; and the refusal of a shape beyond it."""
import numpy as np
import pytest
import torch

from adamixture_amd import _lib
from adamixture_amd import genotypes as GT

pytestmark = pytest.mark.gpu


def _case(M, N, K, seed, dev):
    rng = np.random.default_rng(seed)
    P = np.clip(rng.beta(0.6, 1.2, size=(M, K)), 1e-5, 1 - 1e-5)
    Q = rng.dirichlet(np.full(K, 0.4), size=N)
    Pd, Qd = torch.from_numpy(P).to(dev), torch.from_numpy(Q).to(dev)
    return GT.synthesize(Pd, Qd, seed + 3, 0.01), Pd, Qd


def _run(noalloc, G, P, Q, dtype):
    dev = G.device
    M, K = P.shape
    N = Q.shape[0]
    tdt = torch.float64 if dtype == _lib.ADX_F64 else torch.float32
    P, Q = P.to(tdt).contiguous(), Q.to(tdt).contiguous()
    A, B, T = (torch.zeros(r, K, dtype=tdt, device=dev) for r in (M, M, N))
    ll = torch.zeros(1, dtype=torch.float64, device=dev)
    loglik = dtype == _lib.ADX_F64
    flags = _lib.ADX_EM_STATS | (_lib.ADX_EM_LOGLIK if loglik else 0)
    head = (dtype, G.data.data_ptr(), G.ldg, M, N, K, P.data_ptr(), K, Q.data_ptr(), K, A.data_ptr(), B.data_ptr(),
            T.data_ptr(), flags, ll.data_ptr() if loglik else None)
    st = _lib.stream_handle(dev)
    if noalloc:
        _lib.call("adx_em_pass_noalloc", *head, st)
    else:
        nb = _lib.load().adx_em_workspace_bytes(dtype, M, N, K)
        ws = torch.empty(nb, dtype=torch.uint8, device=dev)
        _lib.call("adx_em_pass", *head, ws.data_ptr(), nb, st)
    torch.cuda.synchronize(dev)
    return [x.double().cpu().numpy() for x in (A, B, T)], (ll.item() if loglik else None)


def _same(got, ref, rtol):
    (ga, gl), (ra, rl) = got, ref
    for x, y in zip(ga, ra):
        np.testing.assert_allclose(x, y, rtol=rtol, atol=rtol * np.abs(y).max())
    if rl is not None:
        assert gl == pytest.approx(rl, rel=1e-12)


@pytest.mark.parametrize("dtype,rtol", [(_lib.ADX_F64, 1e-12), (_lib.ADX_F32, 1e-5)])
def test_noalloc_pass_equals_workspace_pass_twice(cuda, dtype, rtol):
    """Two calls in a row on one GPU, the second with a smaller input: the
    fixed scratch keeps the first call's contents, which the second must not
    read."""
    for M, N, K, seed in ((5003, 1201, 12, 1), (2001, 333, 5, 2)):
        G, P, Q = _case(M, N, K, seed, cuda)
        print(f"noalloc dtype={dtype} M={M} N={N} K={K}", flush=True)
        _same(_run(True, G, P, Q, dtype), _run(False, G, P, Q, dtype), rtol)


def test_noalloc_pass_runs_in_bands_within_the_scratch(cuda):
    """A shape whose one-band partial sums exceed the fixed scratch runs in
    SNP bands inside it and still equals the one-band pass (FP64, to the
    summation-order bound), twice with the same result."""
    M, N, K = 4096, 200000, 10
    full = _lib.load().adx_em_workspace_bytes(_lib.ADX_F64, M, N, K)
    assert full > _lib.ADX_EM_NOALLOC_SCRATCH_BYTES
    G, P, Q = _case(M, N, K, 3, cuda)
    print(f"noalloc banded M={M} N={N} K={K}", flush=True)
    first = _run(True, G, P, Q, _lib.ADX_F64)
    _same(first, _run(False, G, P, Q, _lib.ADX_F64), 1e-12)
    again = _run(True, G, P, Q, _lib.ADX_F64)
    assert all(np.array_equal(x, y) for x, y in zip(first[0], again[0])) and first[1] == again[1]


def test_noalloc_refuses_a_shape_beyond_the_scratch(cuda):
    """One SNP band at N = 3,000,000, K = 64 needs N K x 8 = 1.5 GB of T
    partial sums: refused on the host before any device work."""
    dev = cuda
    buf = torch.zeros(1 << 20, dtype=torch.uint8, device=dev)
    M, N, K = 256, 3_000_000, 64
    ldg = (N + 63) // 64 * 64
    ll = torch.zeros(1, dtype=torch.float64, device=dev)
    rc = _lib.load().adx_em_pass_noalloc(_lib.ADX_F64, buf.data_ptr(), ldg, M, N, K, buf.data_ptr(), K,
                                         buf.data_ptr(), K, buf.data_ptr(), buf.data_ptr(), buf.data_ptr(),
                                         _lib.ADX_EM_STATS | _lib.ADX_EM_LOGLIK, ll.data_ptr(),
                                         _lib.stream_handle(dev))
    assert rc != 0 and b"fixed scratch" in _lib.load().adx_last_error()
