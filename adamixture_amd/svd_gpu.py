"""Randomized SVD initializer on the GPU (drop-in for adamixture/model/svd_gpu.py).

The streamed GEMMs with the centred genotype matrix run in libadx, decode
fused, in one of two arithmetics (``PRODUCTS``):
  * "h2" (default): adx_svd_xw_h2 / adx_svd_xty_h2, FP16 MFMA on two-term
    FP16 splits of the operands with FP32 accumulation -- the FP32 forms'
    accuracy at a quarter of the matrix-core time;
  * "f32": adx_svd_xw / adx_svd_xty, FP32 MFMA, the reference's SGEMM
    arithmetic (svd_gpu.py:104-107, :133-138).
The eigSVD steps keep the tall matrices on the device: the L x L Gram
matrix and U = C V S^-1 are FP64 device products and only the L x L
eigen-decomposition (L = max(k + 10, 20)) runs on the host in LAPACK.

The Gaussian test matrix is drawn on the host with
``numpy.random.default_rng(seed).standard_normal((M, L), float32)`` exactly
as svd.py:52-62 does, so the GPU result reproduces the CPU reference's
numbers (``omega="host"``, the default).  ``omega="device"`` draws it from a
seeded torch device generator as the reference GPU path does
(svd_gpu.py:68-71): deterministic on one GPU model, without the host draw
(~0.05 s at M=500k, L=20), but not reproducible on other devices.
"""
from __future__ import annotations

import logging
import time

import numpy as np
import torch

from . import _lib
from .genotypes import PackedGenotypes, as_packed, is_streamed

log = logging.getLogger(__name__)


# Product arithmetic of xw / xty: "h2" (FP16 two-term splits on FP16 MFMA)
# or "f32" (FP32 MFMA).
PRODUCTS = "h2"


def _gram(Cd: torch.Tensor) -> np.ndarray:
    """Cd^T Cd (L x L, FP64, fixed summation order) for a tall FP64 Cd."""
    from .als_gpu import _tn
    L = Cd.shape[1]
    rows = [_tn(Cd[:, r:r + 64], Cd) for r in range(0, L, 64)]   # the kernel takes a <= 64
    return torch.cat(rows, dim=0).cpu().numpy()


def _times_small(Cd: torch.Tensor, W: np.ndarray) -> torch.Tensor:
    """Cd W (FP64) for a tall Cd (n x L) and a small host W (L x b)."""
    from .als_gpu import _nn
    Wd = torch.from_numpy(np.ascontiguousarray(W)).to(Cd.device)
    out = None
    for r in range(0, Cd.shape[1], 64):                           # a <= 64 per product
        part = _nn(Cd[:, r:r + 64], Wd[r:r + 64])
        out = part if out is None else out + part
    return out


def eigSVD_dev(C: torch.Tensor):
    """eigSVD (svd.py:13-28, svd_gpu.py:11-39) with C (n x L, FP32) on the
    device: the Gram matrix C^T C and U = C V S^-1 are FP64 device products
    (adx_gemm_tn_f64 / _nn_f64), the L x L eigh runs on the host.  Returns U
    (device FP32) and host S (FP32), V (FP64), descending order."""
    Cd = C.double()
    D, V = np.linalg.eigh(_gram(Cd))
    V = np.ascontiguousarray(V[:, ::-1])
    S = np.sqrt(D[::-1])
    U = _times_small(Cd, V / S).float()
    return U, S.astype(np.float32), V


_WORK: dict = {}


def _work(G: PackedGenotypes, nbytes: int):
    """Per-device scratch of at least nbytes (None for 0); the caching
    allocator returns 256-byte aligned blocks."""
    if nbytes <= 0:
        return None
    w = _WORK.get(G.device)
    if w is None or w.numel() < nbytes:
        _WORK[G.device] = None
        w = _WORK[G.device] = torch.empty(nbytes, dtype=torch.uint8, device=G.device)
    return w


def _products(products):
    p = PRODUCTS if products is None else products
    if p not in ("h2", "f32"):
        raise ValueError(f"unknown SVD product arithmetic {p!r} (h2 or f32)")
    return p


# Columns per decode-fused GEMM launch (the kernels take L <= 64; the
# rank-L test matrix, L = max(k + 10, 20), is wider for k > 54).
MAX_L = 64


def xw(G: PackedGenotypes, f: torch.Tensor, W: torch.Tensor, products: str | None = None) -> torch.Tensor:
    """Y = X W (M x L) with X the centred genotypes, W (N x L) FP32."""
    p = _products(products)
    L = W.shape[1]
    if L > MAX_L:   # column blocks: each column of Y depends on one column of W only
        return torch.cat([xw(G, f, W[:, c:c + MAX_L].contiguous(), p) for c in range(0, L, MAX_L)], dim=1)
    if is_streamed(G):   # host-resident genotypes: rows of Y band by band
        Y = torch.empty((G.M, L), dtype=torch.float32, device=G.device)
        for m0, Mb, band in G.bands():
            Y[m0:m0 + Mb] = xw(band, f[m0:m0 + Mb], W, p)
        return Y
    W = W.contiguous()
    st = _lib.stream_handle(G.device)
    if p == "h2":
        nb = _lib.load().adx_svd_h2_work_bytes(0, G.M, G.N, L)
        Y = torch.empty((G.M, L), dtype=torch.float32, device=G.device)
        _lib.call("adx_svd_xw_h2", G.data.data_ptr(), G.ldg, G.M, G.N, f.data_ptr(), W.data_ptr(), L,
                  Y.data_ptr(), _lib.ptr(_work(G, nb)), nb, st)
        return Y
    Y = torch.zeros((G.M, L), dtype=torch.float32, device=G.device)
    _lib.call("adx_svd_xw", G.data.data_ptr(), G.ldg, G.M, G.N, f.data_ptr(), W.data_ptr(), L,
              Y.data_ptr(), _lib.ptr(_work(G, 4 * _lib.load().adx_svd_work_floats(0, G.M, G.N, L))), st)
    return Y


def xty(G: PackedGenotypes, f: torch.Tensor, Y: torch.Tensor, products: str | None = None) -> torch.Tensor:
    """A = X^T Y (N x L) with Y (M x L) FP32."""
    p = _products(products)
    L = Y.shape[1]
    if L > MAX_L:
        return torch.cat([xty(G, f, Y[:, c:c + MAX_L].contiguous(), p) for c in range(0, L, MAX_L)], dim=1)
    if is_streamed(G):   # host-resident genotypes: the SNP reduction summed band by band
        A = torch.zeros((G.N, L), dtype=torch.float32, device=G.device)
        for m0, Mb, band in G.bands():
            A += xty(band, f[m0:m0 + Mb], Y[m0:m0 + Mb], p)
        return A
    Y = Y.contiguous()
    st = _lib.stream_handle(G.device)
    if p == "h2":
        nb = _lib.load().adx_svd_h2_work_bytes(1, G.M, G.N, L)
        A = torch.empty((G.N, L), dtype=torch.float32, device=G.device)
        _lib.call("adx_svd_xty_h2", G.data.data_ptr(), G.ldg, G.M, G.N, f.data_ptr(), Y.data_ptr(), L,
                  A.data_ptr(), _lib.ptr(_work(G, nb)), nb, st)
        return A
    A = torch.zeros((G.N, L), dtype=torch.float32, device=G.device)
    _lib.call("adx_svd_xty", G.data.data_ptr(), G.ldg, G.M, G.N, f.data_ptr(), Y.data_ptr(), L,
              A.data_ptr(), _lib.ptr(_work(G, 4 * _lib.load().adx_svd_work_floats(1, G.M, G.N, L))), st)
    return A


def _test_matrix(M: int, L: int, seed: int, dev, omega: str) -> torch.Tensor:
    if omega == "host":
        rng = np.random.default_rng(seed)
        return torch.from_numpy(rng.standard_normal(size=(M, L), dtype=np.float32)).to(dev)
    if omega == "device":
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        return torch.randn((M, L), dtype=torch.float32, generator=gen, device=dev)
    raise ValueError(f"omega must be 'host' or 'device' (got {omega!r})")


def _SVD_gpu_once(G, N: int, M: int, f: torch.Tensor, k: int, seed: int, power: int, tol: float,
                  chunk_size: int, device, threads_per_block: int = 0, omega: str = "host"):
    """svd.py:30-140 / svd_gpu.py:40-154 (dashSVD with dynamic shift)."""
    t0 = time.time()
    dev = torch.device(device)
    G = as_packed(G, M, dev)
    f = f.to(device=dev, dtype=torch.float32).contiguous()
    L = max(k + 10, 20)
    omega = _test_matrix(M, L, seed, dev, omega)

    accum = xty(G, f, omega)
    del omega
    orth, _, _ = eigSVD_dev(accum)
    alpha = 0.0
    sv = np.zeros(L, dtype=np.float32)
    for i in range(power):
        proj = xw(G, f, orth)
        accum = xty(G, f, proj)
        del proj
        accum -= np.float32(alpha) * orth
        orth, s_vals, _ = eigSVD_dev(accum)
        if i > 0:
            s_cur = s_vals + alpha
            rel = np.abs(s_cur - sv[:len(s_cur)]) / np.maximum(s_cur, 1e-12)
            if np.max(rel[:k]) < tol:
                log.info(f"        Converged at iteration {i}.")
                break
            sv[:len(s_cur)] = s_cur
        else:
            sv[:len(s_vals)] = s_vals + alpha
        if s_vals[-1] > alpha:
            alpha = 0.5 * (s_vals[-1] + alpha)
    proj = xw(G, f, orth)
    u, s_all, v = eigSVD_dev(proj)
    U = u[:, :k].contiguous()
    S = torch.from_numpy(np.ascontiguousarray(s_all[:k])).to(dev)
    V = _times_small(orth.double(), v[:, :k]).float()
    log.info(f"\n    Total time for SVD={time.time() - t0:.3f}s")
    return U, S, V


def SVD_gpu(G, N: int, M: int, f: torch.Tensor, k: int, seed: int, power: int, tol: float, chunk_size: int,
            device, threads_per_block: int = 0, *, omega: str = "host"):
    """svd_gpu.py:157-178 with the OOM retry (no chunk temporaries here, so a
    retry can only come from the O(M L) buffers)."""
    from .utils import MIN_GPU_CHUNK_SIZE, is_gpu_oom, reduce_gpu_chunk_or_raise
    cur = chunk_size
    while True:
        try:
            return _SVD_gpu_once(G, N, M, f, k, seed, power, tol, cur, device, threads_per_block, omega)
        except RuntimeError as exc:
            if not is_gpu_oom(exc) or cur <= MIN_GPU_CHUNK_SIZE:
                raise
            cur = reduce_gpu_chunk_or_raise(cur, exc, "running SVD")
