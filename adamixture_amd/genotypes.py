"""Device-resident SNP-packed genotypes and the genotype-level GPU ops.

Layout (the reference's, adamixture/src/utils_c/cython/snp_reader.pyx:76-120):
``ceil(M/4)`` byte rows, byte ``[j, i]`` holds SNPs ``4j..4j+3`` of sample
``i`` at bits ``2k..2k+1``; codes 0/1/2 = allele count, 3 = missing.  On the
device every row is padded to ``ldg = roundup(N, 64)`` bytes (padding bytes
0xFF, i.e. missing) so the EM kernels can stream 16-byte aligned 64-sample
sub-tiles without bounds checks.  At N = 100,000, M = 500,000 the matrix is
12.5 GB, resident in HBM.
"""
from __future__ import annotations

import weakref

import numpy as np
import torch

from . import _lib


def padded_ld(N: int) -> int:
    return ((N + 63) // 64) * 64


class PackedGenotypes:
    """SNP-packed genotype matrix on one GPU with 64-padded rows."""

    def __init__(self, data: torch.Tensor, M: int, N: int):
        if data.dtype != torch.uint8 or data.dim() != 2 or not data.is_cuda:
            raise ValueError("PackedGenotypes needs a 2-D uint8 CUDA tensor")
        if data.shape[0] != (M + 3) // 4 or data.shape[1] < N or data.stride(1) != 1:
            raise ValueError("PackedGenotypes: data must be ceil(M/4) x ldg with ldg >= N")
        if data.stride(0) % 16 != 0 or data.data_ptr() % 16 != 0:
            raise ValueError("PackedGenotypes: rows must be 16-byte aligned")
        self.data = data
        self.M = int(M)
        self.N = int(N)
        self.ldg = int(data.stride(0))
        self._qbat: torch.Tensor | None = None

    # The reference's helpers test packedness through the shape
    # (adamixture/src/utils.py:483-494); expose the logical shape.
    @property
    def shape(self) -> tuple[int, int]:
        return (self.data.shape[0], self.N)

    @property
    def device(self) -> torch.device:
        return self.data.device

    def qbat(self) -> torch.Tensor:
        """2 x observed genotypes per sample (FP64), computed once."""
        if self._qbat is None:
            q = torch.empty(self.N, dtype=torch.float64, device=self.device)
            _lib.call("adx_qbat", self.data.data_ptr(), self.ldg, self.M, self.N, q.data_ptr(),
                      _lib.stream_handle(self.device))
            self._qbat = q
        return self._qbat

    def to_host_packed(self) -> np.ndarray:
        """Unpadded ceil(M/4) x N numpy copy."""
        return self.data[:, : self.N].cpu().numpy().copy()


class HostPackedGenotypes:
    """SNP-packed genotype matrix kept in pinned host memory (the padded
    layout) for a GPU it does not fit -- the reference keeps G on the host
    then and copies each byte-row chunk per use (adamixture/src/utils.py:
    425-428, :538-557).  The EM and log-likelihood passes stream its SNP
    bands through a device staging buffer (adx_em_pass_host); the other
    genotype-level ops run band by band on a staged copy (``bands``).

    It has no ``data`` attribute: code that is not stream-aware and would
    hand a kernel ``G.data`` raises instead of passing a host pointer."""

    def __init__(self, host: torch.Tensor, M: int, N: int, device, staging_cap: int = 1 << 30):
        if host.dtype != torch.uint8 or host.dim() != 2 or host.is_cuda:
            raise ValueError("HostPackedGenotypes needs a 2-D uint8 host tensor")
        if host.shape[0] != (M + 3) // 4 or host.stride(0) % 64 != 0 or host.stride(1) != 1:
            raise ValueError("HostPackedGenotypes: host must be ceil(M/4) x ldg, ldg a multiple of 64")
        self.host = host
        self.M = int(M)
        self.N = int(N)
        self.ldg = int(host.stride(0))
        self.device = torch.device(device)
        self.staging_cap = int(staging_cap)     # device bytes for staged rows
        self._staging: torch.Tensor | None = None
        self._qbat: torch.Tensor | None = None
        self.copy_stream = None

    @classmethod
    def from_packed(cls, G, M: int, device, staging_cap: int = 1 << 30) -> "HostPackedGenotypes":
        """Pinned padded copy of a ceil(M/4) x N packed host matrix."""
        if isinstance(G, torch.Tensor):
            G = G.cpu().numpy()
        G = np.ascontiguousarray(G, dtype=np.uint8)
        Mb, N = G.shape
        if Mb != (M + 3) // 4:
            raise ValueError(f"packed genotypes have {Mb} byte rows, expected ceil({M}/4)")
        ldg = padded_ld(N)
        host = torch.empty((Mb, ldg), dtype=torch.uint8, pin_memory=True)
        h = host.numpy()
        h[:, N:] = 0xFF
        h[:, :N] = G
        return cls(host, M, N, device, staging_cap)

    @property
    def shape(self) -> tuple[int, int]:
        return (self.host.shape[0], self.N)

    def staging(self, nbytes: int) -> torch.Tensor:
        """Device staging buffer of at least nbytes (grown on demand)."""
        if self._staging is None or self._staging.numel() < nbytes:
            self._staging = None
            self._staging = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
        return self._staging

    def band_rows(self) -> int:
        return max(1, self.staging_cap // self.ldg)

    def bands(self, rows: int | None = None):
        """Yield (m0, Mb, band) with band a PackedGenotypes view of byte rows
        m0 / 4 .. staged on the device (one staging buffer, reused in stream
        order: each copy is enqueued after the previous band's kernels)."""
        rows = rows or self.band_rows()
        Mb_total = self.host.shape[0]
        buf = self.staging(min(rows, Mb_total) * self.ldg).view(-1)
        for r0 in range(0, Mb_total, rows):
            r1 = min(r0 + rows, Mb_total)
            dst = buf[: (r1 - r0) * self.ldg].view(r1 - r0, self.ldg)
            dst.copy_(self.host[r0:r1], non_blocking=True)
            m0 = 4 * r0
            yield m0, min(4 * r1, self.M) - m0, PackedGenotypes(dst, min(4 * r1, self.M) - m0, self.N)

    def stage_rows(self, r0: int, r1: int) -> torch.Tensor:
        """Byte rows r0 .. r1 - 1 copied to the device (stream order)."""
        buf = self.staging((r1 - r0) * self.ldg).view(-1)[: (r1 - r0) * self.ldg].view(r1 - r0, self.ldg)
        buf.copy_(self.host[r0:r1], non_blocking=True)
        return buf

    def qbat(self) -> torch.Tensor:
        """2 x observed genotypes per sample (FP64, integer-valued: the band
        sums are exact)."""
        if self._qbat is None:
            q = torch.zeros(self.N, dtype=torch.float64, device=self.device)
            for _, _, band in self.bands():
                q += band.qbat()
            self._qbat = q
        return self._qbat

    def to_host_packed(self) -> np.ndarray:
        return self.host[:, : self.N].numpy().copy()


def is_streamed(G) -> bool:
    return isinstance(G, HostPackedGenotypes)


# Device copies keyed by the identity of the source tensor; an entry is
# dropped when its source dies.  (A tensor cannot key a WeakKeyDictionary: a
# second lookup with the same tensor compares keys with ==, which for
# tensors is elementwise and raises.)
_CACHE: dict[int, tuple[weakref.ref, dict]] = {}


def _per_source(G: torch.Tensor) -> dict:
    """The cache slot of source tensor ``G`` (same dict for the same object)."""
    k = id(G)
    ent = _CACHE.get(k)
    if ent is not None and ent[0]() is G:
        return ent[1]

    def drop(r, k=k):
        cur = _CACHE.get(k)
        if cur is not None and cur[0] is r:
            del _CACHE[k]

    per: dict = {}
    _CACHE[k] = (weakref.ref(G, drop), per)
    return per


def as_packed(G, M: int, device: torch.device | str | None = None) -> PackedGenotypes:
    """Return ``G`` as PackedGenotypes on ``device``.

    Accepts PackedGenotypes, or a ceil(M/4) x N uint8 numpy array / torch
    tensor (host or device) as produced by the reference reader.  Device
    copies of a given tensor are cached so repeated calls through the
    reference-style API do not repeat the upload.
    """
    if isinstance(G, (PackedGenotypes, HostPackedGenotypes)):
        return G
    if isinstance(G, np.ndarray):
        G = torch.from_numpy(np.ascontiguousarray(G))
    if not isinstance(G, torch.Tensor) or G.dtype != torch.uint8 or G.dim() != 2:
        raise TypeError("genotypes must be a 2-D uint8 packed matrix")
    Mb, N = G.shape
    if Mb != (M + 3) // 4:
        raise ValueError(f"packed genotypes have {Mb} byte rows, expected ceil({M}/4)")
    dev = torch.device(device) if device is not None else (G.device if G.is_cuda else torch.device("cuda"))
    per_src = _per_source(G)
    key = (G.data_ptr(), str(dev), M, G._version)
    hit = per_src.get(key)
    if hit is not None:
        return hit
    ldg = padded_ld(N)
    out = torch.empty((Mb, ldg), dtype=torch.uint8, device=dev)
    src = G if G.is_cuda and G.device == dev else G.to(dev)
    if src.stride(1) != 1:
        src = src.contiguous()
    _lib.call("adx_repack_padded", src.data_ptr(), src.stride(0), Mb, N, out.data_ptr(), ldg,
              _lib.stream_handle(dev))
    pg = PackedGenotypes(out, M, N)
    per_src.clear()
    per_src[key] = pg
    return pg


def _staged(G, m_start: int, rows: int):
    """(device pointer, byte_offset) of the rows holding SNPs m_start ..
    m_start + rows - 1: G itself when resident, else those byte rows staged
    from the host (pack2bit.cu:83-110's byte_offset form)."""
    if not is_streamed(G):
        return G.data.data_ptr(), 0
    r0, r1 = m_start // 4, (m_start + rows + 3) // 4
    return G.stage_rows(r0, r1).data_ptr(), r0


def unpack_chunk(G, m_start: int, chunk: int) -> torch.Tensor:
    """pack2bit.cu:83-110 semantics: codes of SNPs m_start .. m_start+C'-1."""
    rows = max(0, min(chunk, G.M - m_start))
    out = torch.empty((rows, G.N), dtype=torch.uint8, device=G.device)
    if rows:
        ptr, off = _staged(G, m_start, rows)
        _lib.call("adx_unpack_u8", ptr, G.ldg, G.M, G.N, m_start, rows, off, out.data_ptr(),
                  _lib.stream_handle(G.device))
    return out


def unpack_center_chunk(G, f: torch.Tensor, m_start: int, chunk: int) -> torch.Tensor:
    """pack2bit.cu:112-141 semantics: g - 2 f[m] as FP32, 0 where missing."""
    rows = max(0, min(chunk, G.M - m_start))
    out = torch.empty((rows, G.N), dtype=torch.float32, device=G.device)
    f = f.to(device=G.device, dtype=torch.float32).contiguous()
    if rows:
        ptr, off = _staged(G, m_start, rows)
        _lib.call("adx_unpack_center_f32", ptr, G.ldg, f.data_ptr(), G.M, G.N, m_start, rows, off,
                  out.data_ptr(), _lib.stream_handle(G.device))
    return out


def allele_frequencies(G) -> torch.Tensor:
    """utils.py:781-829 semantics, one pass over the packed matrix (FP32;
    per SNP, so a host-streamed matrix gives the same values band by band)."""
    f = torch.empty(G.M, dtype=torch.float32, device=G.device)
    if is_streamed(G):
        for m0, Mb, band in G.bands():
            _lib.call("adx_freq_f32", band.data.data_ptr(), band.ldg, Mb, band.N, f[m0:].data_ptr(),
                      _lib.stream_handle(G.device))
        return f
    _lib.call("adx_freq_f32", G.data.data_ptr(), G.ldg, G.M, G.N, f.data_ptr(), _lib.stream_handle(G.device))
    return f


def synthesize(P: torch.Tensor, Q: torch.Tensor, seed: int, miss_rate: float) -> PackedGenotypes:
    """Counter-based synthetic genotypes G ~ Binomial(2, Q P^T) on the device."""
    P = P.to(dtype=torch.float64).contiguous()
    Q = Q.to(device=P.device, dtype=torch.float64).contiguous()
    M, K = P.shape
    N = Q.shape[0]
    ldg = padded_ld(N)
    out = torch.empty(((M + 3) // 4, ldg), dtype=torch.uint8, device=P.device)
    _lib.call("adx_synth_packed", P.data_ptr(), Q.data_ptr(), M, N, K, seed & 0xFFFFFFFFFFFFFFFF, miss_rate,
              out.data_ptr(), ldg, _lib.stream_handle(P.device))
    return PackedGenotypes(out, M, N)
