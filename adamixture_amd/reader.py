"""PLINK .bed/.bim/.fam reader producing the SNP-packed layout.

Follows adamixture/src/snp_reader.py:353-443 (_read_bed, packed=True) and
:747-757 (flip rule) with the kernels of snp_reader.pyx:76-120 (BED code map
{00->2, 01->3 missing, 10->1, 11->0}) and :215-302 (mean / flip, padding
fields of the last byte zeroed), vectorised in numpy over SNP blocks.
"""
from __future__ import annotations

import logging
from math import ceil
from pathlib import Path

import numpy as np
import torch

log = logging.getLogger(__name__)

_BED_LUT = np.array([2, 3, 1, 0], dtype=np.uint8)
# per BED byte: the 4 sample codes it holds
_BYTE_CODES = np.stack([_BED_LUT[(np.arange(256) >> (2 * s)) & 3] for s in range(4)], axis=1).astype(np.uint8)
_FLIP = np.array([2, 1, 0, 3], dtype=np.uint8)


def _keep_chrom(chrom: str, chrom_mode: str, autosomes: int, specific) -> bool:
    """snp_reader.py _keep_chromosome: 'all' keeps every row; 'autosomes'
    keeps 1..autosomes (or the listed chromosomes)."""
    if chrom_mode == "all":
        return True
    c = chrom.lower().removeprefix("chr")
    if specific is not None:
        return c in {str(s).lower().removeprefix("chr") for s in specific}
    return c.isdigit() and 1 <= int(c) <= autosomes


def bed_to_packed(bed: np.ndarray, N: int, block: int = 65536) -> np.ndarray:
    """M x ceil(N/4) BED rows -> ceil(M/4) x N SNP-packed bytes."""
    M = bed.shape[0]
    Mb = (M + 3) // 4
    out = np.zeros((Mb, N), dtype=np.uint8)
    for s in range(0, M, block):
        e = min(M, s + block)
        codes = _BYTE_CODES[bed[s:e]].reshape(e - s, -1)[:, :N]          # (rows, N)
        pad = (-(e - s)) % 4
        if pad:
            codes = np.concatenate([codes, np.zeros((pad, N), dtype=np.uint8)])
        c4 = codes.reshape(-1, 4, N)
        out[s // 4: s // 4 + c4.shape[0]] = c4[:, 0] | (c4[:, 1] << 2) | (c4[:, 2] << 4) | (c4[:, 3] << 6)
    return out


def packed_mean(Gp: np.ndarray, M: int) -> float:
    """snp_reader.pyx:215-251 get_mean_packed."""
    s = c = 0
    for k in range(4):
        rows = Gp if k < (M % 4 or 4) else Gp[:-1]
        codes = (rows >> (2 * k)) & 3
        obs = codes != 3
        s += int(codes[obs].sum(dtype=np.int64))
        c += 2 * int(obs.sum())
    return 0.0 if c == 0 else s / c


def flip_packed(Gp: np.ndarray, M: int) -> None:
    """snp_reader.pyx:254-302 flip_packed (in place)."""
    tab = np.zeros(256, dtype=np.uint8)
    b = np.arange(256)
    for k in range(4):
        tab |= (_FLIP[(b >> (2 * k)) & 3] << (2 * k)).astype(np.uint8)
    last = M % 4 or 4
    mask = np.uint8(sum(3 << (2 * k) for k in range(last)))
    Gp[:] = tab[Gp]
    Gp[-1] &= mask


def read_bed_packed(path: str | Path, chrom_mode: str = "all", autosomes: int = 22,
                    specific_chrom=None) -> tuple[torch.Tensor, int, int]:
    """Read base.bed/.bim/.fam -> (packed uint8 tensor ceil(M/4) x N, N, M)."""
    path = Path(path)
    base = path.with_suffix("") if path.suffix == ".bed" else path
    with open(f"{base}.fam") as f:
        N = sum(1 for _ in f)
    with open(f"{base}.bim") as f:
        keep = np.array([_keep_chrom(line.split()[0], chrom_mode, autosomes, specific_chrom)
                         for line in f if line.strip()], dtype=bool)
    nb = ceil(N / 4)
    raw = np.fromfile(f"{base}.bed", dtype=np.uint8)
    if raw.size < 3 or raw[0] != 0x6C or raw[1] != 0x1B or raw[2] != 0x01:
        raise ValueError(f"{base}.bed is not a SNP-major PLINK bed file")
    if (raw.size - 3) % nb:
        raise ValueError("bim file doesn't match!")
    bed = raw[3:].reshape(-1, nb)
    if bed.shape[0] != keep.size:
        raise ValueError("bim file doesn't match!")
    bed = bed[keep]
    M = bed.shape[0]
    if M == 0 or N == 0:
        raise ValueError("genotype data contains 0 SNPs or 0 samples after filtering")
    Gp = bed_to_packed(bed, N)
    if packed_mean(Gp, M) >= 0.5:
        log.info("    Flipping genotype encoding (packed).")
        flip_packed(Gp, M)
    return torch.from_numpy(Gp), N, M


def read_bed_gpu(path: str | Path, device="cuda", chrom_mode: str = "all", autosomes: int = 22,
                 specific_chrom=None, block_bytes: int = 1 << 29):
    """read_bed_packed with the transpose, mean and flip on the GPU
    (adx_bed_to_packed / adx_packed_allele_sum / adx_flip_packed): the .bed
    is memory-mapped, the kept SNP rows are copied to the device in blocks of
    about ``block_bytes`` and transposed straight into the padded device
    layout.  Returns (PackedGenotypes, N, M)."""
    from . import _lib
    from .genotypes import PackedGenotypes, padded_ld
    path = Path(path)
    base = path.with_suffix("") if path.suffix == ".bed" else path
    dev = torch.device(device)
    with open(f"{base}.fam") as f:
        N = sum(1 for _ in f)
    with open(f"{base}.bim") as f:
        keep = np.array([_keep_chrom(line.split()[0], chrom_mode, autosomes, specific_chrom)
                         for line in f if line.strip()], dtype=bool)
    nb = ceil(N / 4)
    raw = np.memmap(f"{base}.bed", dtype=np.uint8, mode="c")       # copy-on-write: writable views, file untouched
    if raw.size < 3 or raw[0] != 0x6C or raw[1] != 0x1B or raw[2] != 0x01:
        raise ValueError(f"{base}.bed is not a SNP-major PLINK bed file")
    if (raw.size - 3) % nb or (raw.size - 3) // nb != keep.size:
        raise ValueError("bim file doesn't match!")
    bed = raw[3:].reshape(-1, nb)
    idx = np.flatnonzero(keep)
    M = idx.size
    if M == 0 or N == 0:
        raise ValueError("genotype data contains 0 SNPs or 0 samples after filtering")
    ldg = padded_ld(N)
    data = torch.empty(((M + 3) // 4, ldg), dtype=torch.uint8, device=dev)
    block_snps = max(4, (block_bytes // nb) // 4 * 4)
    st = _lib.stream_handle(dev)
    contiguous = M == keep.size
    for s in range(0, M, block_snps):
        e = min(M, s + block_snps)
        rows = np.asarray(bed[s:e]) if contiguous else bed[idx[s:e]]
        dbed = torch.from_numpy(np.ascontiguousarray(rows)).to(dev)
        _lib.call("adx_bed_to_packed", dbed.data_ptr(), nb, e - s, N, data.data_ptr(), ldg, s // 4, st)
    G = PackedGenotypes(data, M, N)
    flip_gpu_if_needed(G)
    return G, N, M


def flip_gpu_if_needed(G) -> bool:
    """snp_reader.py:747-757 on the device: flip 0 <-> 2 when the mean allele
    frequency of the observed genotypes is >= 0.5.  Returns whether it did."""
    from . import _lib
    st = _lib.stream_handle(G.device)
    sums = torch.empty(2, dtype=torch.int64, device=G.device)     # uint64 counters, < 2^63
    _lib.call("adx_packed_allele_sum", G.data.data_ptr(), G.ldg, G.M, G.N, sums.data_ptr(), st)
    s_all, c_all = (int(x) for x in sums.cpu().tolist())
    if c_all and s_all / c_all >= 0.5:
        log.info("    Flipping genotype encoding (packed).")
        _lib.call("adx_flip_packed", G.data.data_ptr(), G.ldg, G.M, G.N, st)
        G._qbat = None
        return True
    return False


def write_bed(base: str | Path, G: np.ndarray) -> None:
    """Write an unpacked M x N genotype matrix (codes 0/1/2, 3 missing) as
    PLINK .bed/.bim/.fam (code map inverse of the reader's)."""
    base = Path(base)
    M, N = G.shape
    enc = np.array([3, 2, 0, 1], dtype=np.uint8)[G]          # 0->11, 1->10, 2->00, 3->01
    pad = (-N) % 4
    if pad:
        enc = np.concatenate([enc, np.zeros((M, pad), dtype=np.uint8)], axis=1)
    e4 = enc.reshape(M, -1, 4)
    packed = e4[..., 0] | (e4[..., 1] << 2) | (e4[..., 2] << 4) | (e4[..., 3] << 6)
    with open(f"{base}.bed", "wb") as f:
        f.write(bytes([0x6C, 0x1B, 0x01]))
        f.write(packed.astype(np.uint8).tobytes())
    with open(f"{base}.bim", "w") as f:
        for j in range(M):
            f.write(f"1\trs{j}\t0\t{j + 1}\tA\tG\n")
    with open(f"{base}.fam", "w") as f:
        for i in range(N):
            f.write(f"fam{i} ind{i} 0 0 0 -9\n")
