"""VCF and BCF readers producing the device SNP-packed layout (drop-in for
the VCF / BCF branches of adamixture/src/snp_reader.py:445-682 with the
Cython parsers of adamixture/src/utils_c/cython/snp_reader.pyx:305-1108).

The host does what is inherently sequential -- decompression, finding line
or record boundaries, the chromosome filter, locating each BCF record's GT
vector -- and streams the file in blocks of whole records to the device,
where adx_vcf_gt / adx_bcf_gt decode every genotype straight into the packed
matrix (fields preset to missing).  The flip rule (mean allele frequency >=
0.5) runs on the device afterwards, as for .bed input.

Compression: ``.gz`` (gzip or BGZF) as the reference; ``.zst`` needs the
``zstandard`` module, which the reference also imports on demand -- where it
is missing a clear error is raised.
"""
from __future__ import annotations

import gzip
import io
import logging
import struct
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .genotypes import PackedGenotypes, padded_ld

log = logging.getLogger(__name__)

BLOCK_BYTES = 1 << 27           # host -> device text block (whole lines)


# ----------------------------------------------------------------- filter ----

def normalize_specific_chrom(specific) -> list | None:
    """snp_reader.py:190-221 _normalize_specific_chrom: comma/space separated
    items; digits (optionally 'chr'-prefixed) become ints."""
    if specific is None:
        return None
    if isinstance(specific, (int, str)):
        raw = [str(specific)]
    elif isinstance(specific, (list, tuple, set)):
        raw = [str(x) for x in specific]
    else:
        return None
    out: list = []
    for item in raw:
        for p in (q.strip() for q in item.replace(",", " ").split() if q.strip()):
            bare = p[3:] if p.lower().startswith("chr") and p[3:].isdigit() else p
            v = int(bare) if bare.isdigit() else p
            if v not in out:
                out.append(v)
    return out or None


def _allowed_sets(norm: list) -> tuple[set, set]:
    """snp_reader.pyx:579-597: the numbers and lower-case names accepted."""
    nums: set = set()
    strs: set = set()
    for item in norm:
        if isinstance(item, int):
            nums.add(item)
            strs.update({str(item), f"chr{item}"})
        else:
            s = str(item).strip().lower()
            strs.add(s)
            if s.startswith("chr"):
                strs.add(s[3:])
                if s[3:].isdigit():
                    nums.add(int(s[3:]))
            else:
                strs.add(f"chr{s}")
                if s.isdigit():
                    nums.add(int(s))
    return nums, strs


class ChromFilter:
    """snp_reader.pyx:498-552 _keep_chromosome_line on a CHROM token."""

    def __init__(self, chrom_mode: str, autosomes: int, specific=None):
        if chrom_mode not in ("all", "autosomes"):
            raise ValueError("chrom_mode must be 'all' or 'autosomes'")
        self.keep_all = chrom_mode == "all"
        self.autosomes = autosomes
        self.spec = normalize_specific_chrom(specific) if chrom_mode == "autosomes" else None
        if self.spec:
            self.nums, self.strs = _allowed_sets(self.spec)
        elif chrom_mode == "autosomes" and autosomes < 1:
            raise ValueError("autosomes must be at least 1")
        self._memo: dict = {}

    def __call__(self, token: bytes) -> bool:
        hit = self._memo.get(token)
        if hit is None:
            hit = self._memo[token] = self._keep(token)
        return hit

    def _keep(self, token: bytes) -> bool:
        if self.keep_all:
            return True
        body = token[3:] if len(token) >= 3 and token[:3].lower() == b"chr" else token
        pure = all(48 <= c <= 57 for c in body)
        num = int(body) if pure and body else 0
        if self.spec:
            if pure and num > 0 and num in self.nums:
                return True
            t = token.decode("utf-8", errors="ignore").strip().lower()
            return t in self.strs or (t.startswith("chr") and t[3:] in self.strs)
        return pure and 1 <= num <= self.autosomes

    def warn(self, skipped: int) -> None:
        if skipped <= 0:
            return
        if self.spec:
            log.warning(f"        Warning: Skipped {skipped} SNPs outside specific chromosomes {self.spec}.")
        elif not self.keep_all:
            log.warning(f"        Warning: Skipped {skipped} SNPs outside autosomes 1..{self.autosomes}.")


# -------------------------------------------------------------------- I/O ----

def open_binary(path) -> io.BufferedIOBase:
    """snp_reader.pyx:11-17 _open_vcf_file."""
    p = str(path)
    if p.endswith(".gz"):
        return gzip.open(p, "rb")
    if p.endswith(".zst"):
        try:
            import zstandard
        except ImportError as exc:
            raise RuntimeError(f"{p}: reading .zst input needs the 'zstandard' module, which is not installed") \
                from exc
        return io.BufferedReader(zstandard.open(p, "rb"))
    return open(p, "rb")


def _line_blocks(fh, block_bytes: int):
    """Blocks of whole lines; every block ends with '\\n' (one is appended to
    a last line that lacks it)."""
    rest = b""
    while True:
        chunk = fh.read(block_bytes)
        if not chunk:
            if rest:
                yield rest + b"\n"
            return
        buf = rest + chunk
        cut = buf.rfind(b"\n")
        if cut < 0:
            rest = buf
            continue
        yield buf[:cut + 1]
        rest = buf[cut + 1:]


def _lines(block: bytes):
    """(start, end) of every line of a block; end is the index of its '\\n'."""
    nl = np.flatnonzero(np.frombuffer(block, dtype=np.uint8) == 10)
    starts = np.empty_like(nl)
    starts[0] = 0
    starts[1:] = nl[:-1] + 1
    return starts, nl


def _device_matrix(M: int, N: int, dev) -> PackedGenotypes:
    ldg = padded_ld(N)
    data = torch.full(((M + 3) // 4, ldg), 0xFF, dtype=torch.uint8, device=dev)
    return PackedGenotypes(data, M, N)


def _finish(G: PackedGenotypes, flip: bool = True) -> PackedGenotypes:
    """Zero the padding fields of the last byte row and apply the flip rule
    (snp_reader.py:747-757) on the device."""
    st = _lib.stream_handle(G.device)
    _lib.call("adx_clear_tail_fields", G.data.data_ptr(), G.ldg, G.M, G.N, st)
    if flip and G.M > 0:
        from .reader import flip_gpu_if_needed
        flip_gpu_if_needed(G)
    return G


# -------------------------------------------------------------------- VCF ----

def read_vcf_gpu(path, device="cuda", chrom_mode: str = "all", autosomes: int = 22, specific_chrom=None,
                 block_bytes: int = BLOCK_BYTES, flip: bool = True) -> tuple[PackedGenotypes, int, int]:
    """snp_reader.pyx:710-827 read_vcf_file_packed (+ the flip of
    snp_reader.py:747-757): returns (G, N, M) with G on ``device``."""
    dev = torch.device(device)
    keep = ChromFilter(chrom_mode, autosomes, specific_chrom)
    N, M, skipped = 0, 0, 0
    with open_binary(path) as fh:                       # pass 1: samples, kept lines
        for block in _line_blocks(fh, block_bytes):
            starts, ends = _lines(block)
            for s, e in zip(starts.tolist(), ends.tolist()):
                if block[s:s + 1] == b"#":
                    if block[s:s + 6] in (b"#CHROM", b"#chrom"):
                        N = len(block[s:e].split(b"\t")) - 9
                    continue
                tab = block.find(b"\t", s, e)
                if keep(block[s:tab if tab >= 0 else e]):
                    M += 1
                else:
                    skipped += 1
    if N <= 0 or M <= 0:
        raise ValueError("Invalid or empty VCF file")
    keep.warn(skipped)
    G = _device_matrix(M, N, dev)
    st = _lib.stream_handle(dev)
    m0 = 0
    with open_binary(path) as fh:                       # pass 2: decode on the device
        for block in _line_blocks(fh, block_bytes):
            starts, ends = _lines(block)
            sel = []
            for k, (s, e) in enumerate(zip(starts.tolist(), ends.tolist())):
                if block[s:s + 1] == b"#":
                    continue
                tab = block.find(b"\t", s, e)
                if keep(block[s:tab if tab >= 0 else e]):
                    sel.append(k)
            if not sel:
                continue
            sel = np.asarray(sel)
            text = torch.from_numpy(np.frombuffer(block, dtype=np.uint8).copy()).to(dev)
            ls = torch.from_numpy(starts[sel].astype(np.int64)).to(dev)
            le = torch.from_numpy(ends[sel].astype(np.int64)).to(dev)
            _lib.call("adx_vcf_gt", text.data_ptr(), ls.data_ptr(), le.data_ptr(), len(sel), m0, N,
                      G.data.data_ptr(), G.ldg, st)
            m0 += len(sel)
            torch.cuda.current_stream(dev).synchronize()        # the block buffers are freed next
    if m0 != M:
        raise ValueError(f"VCF variant count mismatch: expected {M}, parsed {m0}")
    return _finish(G, flip), N, M


# -------------------------------------------------------------------- BCF ----

def bcf_bytes(path) -> bytes:
    """Decompressed BCF stream: the extension's compression first (as
    snp_reader.py _materialize_binary), then BGZF / gzip (as
    _read_bgzf_or_gzip)."""
    with open_binary(path) as fh:
        data = fh.read()
    if data[:2] == b"\x1f\x8b":
        data = gzip.decompress(data)
    return data


def _typed_int(d: bytes, pos: int) -> tuple[int, int]:
    """snp_reader.pyx:829-847 _bcf_parse_typed_int."""
    b = d[pos]
    t, n = b & 0x0F, (b >> 4) & 0x0F
    pos += 1
    if n == 15:
        _, pos = _typed_int(d, pos)
    if t == 1:
        return struct.unpack_from("<b", d, pos)[0], pos + 1
    if t == 2:
        return struct.unpack_from("<h", d, pos)[0], pos + 2
    if t == 3:
        return struct.unpack_from("<i", d, pos)[0], pos + 4
    return 0, pos


def _descriptor(d: bytes, pos: int) -> tuple[int, int, int]:
    """snp_reader.pyx:849-857 _bcf_parse_descriptor -> (count, bytes per value, values pos)."""
    b = d[pos]
    t, n = b & 0x0F, (b >> 4) & 0x0F
    pos += 1
    if n == 15:
        n, pos = _typed_int(d, pos)
    return n, (1 if t in (1, 7) else (2 if t == 2 else 4)), pos


def bcf_index(data: bytes, keep: ChromFilter):
    """Header and per-kept-record GT location (snp_reader.pyx:859-1023):
    returns (N, rec_starts, rec_ends, gt_off, gt_desc, skipped) with record
    spans and GT offsets absolute in ``data`` and gt_desc = n | size << 16
    (-1: no usable GT, all missing)."""
    if not data or len(data) < 9:
        raise ValueError("Invalid or empty BCF data")
    if data[:5] not in (b"BCF\x02\x02", b"BCF\x02\x01"):
        raise ValueError("Invalid BCF magic header")
    l_text = struct.unpack_from("<I", data, 5)[0]
    if 9 + l_text > len(data):
        raise ValueError("Malformed BCF: header text truncated")
    header = data[9:9 + l_text].decode("utf-8", errors="ignore")
    contigs: dict[int, str] = {}
    samples: list[str] = []
    gt_idx = -1
    for line in header.splitlines():
        if line.startswith("##contig=") and "ID=" in line:
            contigs[len(contigs)] = line.split("ID=")[1].split(",")[0].split(">")[0]
        elif line.startswith(("#CHROM", "#chrom")):
            parts = line.split("\t")
            samples = parts[9:] if len(parts) > 9 else []
    for line in header.splitlines():
        if line.startswith("##FORMAT=") and "ID=GT" in line:
            if "IDX=" in line:
                try:
                    gt_idx = int(line.split("IDX=")[1].split(",")[0].split(">")[0])
                except ValueError:
                    pass
            break
    N = len(samples)
    if N <= 0:
        raise ValueError("BCF header contains no samples")
    starts, ends, offs, descs = [], [], [], []
    skipped = 0
    off = 9 + l_text
    while off + 8 <= len(data):
        l_shared, l_indiv = struct.unpack_from("<II", data, off)
        base = off + 8
        if base + 4 > len(data):
            break
        chrom_id = struct.unpack_from("<i", data, base)[0]
        name = contigs.get(chrom_id, str(chrom_id)).encode("utf-8")
        nxt = off + 8 + l_shared + l_indiv
        if keep(name):
            n_fmt = struct.unpack_from("<I", data, base + 20)[0] >> 24
            pos = base + l_shared
            gt_off, nv, sz = -1, 2, 1
            for _ in range(n_fmt):
                key, pos = _typed_int(data, pos)
                n, size, vpos = _descriptor(data, pos)
                if key == gt_idx or gt_idx < 0:
                    gt_off, nv, sz = vpos, n, size
                    break
                pos = vpos + N * n * size
            ok = gt_off >= 0 and nv >= 1 and gt_off + N * nv * sz <= len(data)
            starts.append(off)
            ends.append(min(max(nxt, gt_off + N * nv * sz if ok else nxt), len(data)))
            offs.append(gt_off if ok else 0)
            descs.append((nv | (sz << 16)) if ok else -1)
        else:
            skipped += 1
        off = nxt
    if not starts:
        raise ValueError("No variants found in BCF file after chromosome filtering")
    return (N, np.asarray(starts, np.int64), np.asarray(ends, np.int64), np.asarray(offs, np.int64),
            np.asarray(descs, np.int32), skipped)


def read_bcf_gpu(path, device="cuda", chrom_mode: str = "all", autosomes: int = 22, specific_chrom=None,
                 block_bytes: int = BLOCK_BYTES, flip: bool = True) -> tuple[PackedGenotypes, int, int]:
    """snp_reader.pyx:1082-1108 read_bcf_file_packed (+ flip): returns
    (G, N, M) with G on ``device``."""
    dev = torch.device(device)
    keep = ChromFilter(chrom_mode, autosomes, specific_chrom)
    data = bcf_bytes(path)
    N, rs, re_, go, gd, skipped = bcf_index(data, keep)
    keep.warn(skipped)
    M = len(rs)
    G = _device_matrix(M, N, dev)
    st = _lib.stream_handle(dev)
    raw = np.frombuffer(data, dtype=np.uint8)
    r0 = 0
    while r0 < M:                                       # record blocks of about block_bytes
        r1 = r0 + 1
        while r1 < M and re_[r1] - rs[r0] <= block_bytes:
            r1 += 1
        lo, hi = int(rs[r0]), int(re_[r1 - 1])
        buf = torch.from_numpy(raw[lo:hi].copy()).to(dev)
        off = torch.from_numpy(np.where(gd[r0:r1] >= 0, go[r0:r1] - lo, 0)).to(dev)
        desc = torch.from_numpy(gd[r0:r1].copy()).to(dev)
        _lib.call("adx_bcf_gt", buf.data_ptr(), off.data_ptr(), desc.data_ptr(), r1 - r0, r0, N, G.data.data_ptr(),
                  G.ldg, st)
        torch.cuda.current_stream(dev).synchronize()
        r0 = r1
    return _finish(G, flip), N, M


# ------------------------------------------------------------------- PGEN ----

def pgen_layout(data: bytes):
    """Header of a variable-width (storage mode 0x10) PLINK 2 .pgen: returns
    (M, N, record starts, record lengths, record types).  Byte 11 holds the
    record-type width (4 or 8 bits) and the record-length width (1-4 bytes);
    each variant block of 2^16 variants stores its types, lengths and
    optional provisional-reference flags after the block offsets.  The
    reference reads .pgen through pgenlib (snp_reader.pyx:1110-1135); stored
    allele counts (multiallelic variants) are not supported here."""
    if len(data) < 12 or data[:2] != b"\x6c\x1b":
        raise ValueError("not a PLINK 2 .pgen file")
    if data[2] != 0x10:
        raise ValueError(f".pgen storage mode {data[2]:#x} is not supported (only variable-width, 0x10)")
    M = int.from_bytes(data[3:7], "little")
    N = int.from_bytes(data[7:11], "little")
    ctl = data[11]
    if ctl & 0x30:
        raise ValueError(".pgen with stored allele counts (multiallelic variants) is not supported")
    nblk = (M + 65535) // 65536
    pos = 12 + 8 * nblk
    offs = np.frombuffer(data, "<u8", nblk, 12).astype(np.int64)
    lb = (ctl & 3) + 1
    types, lens = [], []
    for b in range(nblk):
        cnt = min(65536, M - 65536 * b)
        if (ctl >> 2) & 1:
            t = np.frombuffer(data, np.uint8, cnt, pos)
            pos += cnt
        else:
            v = np.frombuffer(data, np.uint8, (cnt + 1) // 2, pos)
            pos += (cnt + 1) // 2
            t = np.stack([v & 15, v >> 4], axis=1).reshape(-1)[:cnt]
        raw = np.frombuffer(data, np.uint8, cnt * lb, pos).reshape(cnt, lb).astype(np.int64)
        pos += cnt * lb
        if (ctl >> 6) == 3:
            pos += (cnt + 7) // 8
        types.append(t)
        lens.append((raw << (8 * np.arange(lb))).sum(axis=1))
    types, lens = np.concatenate(types).astype(np.uint8), np.concatenate(lens)
    starts = np.empty(M, np.int64)
    for b in range(nblk):
        s, e = 65536 * b, min(M, 65536 * (b + 1))
        starts[s:e] = offs[b] + np.concatenate([[0], np.cumsum(lens[s:e])[:-1]])
        end = offs[b] + lens[s:e].sum()
        if end > len(data) or (b + 1 < nblk and end != offs[b + 1]):
            raise ValueError(".pgen record lengths do not match the block offsets")
    if M and starts[0] < pos:
        raise ValueError(".pgen records overlap the header")
    return M, N, starts, lens, types


def _pgen_variant_chroms(pgen_path: Path) -> list[bytes]:
    """CHROM column of the .pvar (or .bim) next to the .pgen, plain or .gz."""
    stem = str(pgen_path)
    for ext in (".pgen.gz", ".pgen.zst", ".pgen"):
        if stem.endswith(ext):
            stem = stem[:-len(ext)]
            break
    for ext in (".pvar", ".pvar.gz", ".pvar.zst", ".bim", ".bim.gz", ".bim.zst"):
        cand = Path(stem + ext)
        if cand.exists():
            with open_binary(cand) as fh:
                return [ln.split()[0] for ln in fh.read().splitlines() if ln.strip() and not ln.startswith(b"#")]
    raise FileNotFoundError(f"Variant file (.pvar or .bim) not found for {stem}")


def read_pgen_gpu(path, device="cuda", chrom_mode: str = "all", autosomes: int = 22, specific_chrom=None,
                  flip: bool = True) -> tuple[PackedGenotypes, int, int]:
    """PGEN hardcalls decoded on the GPU (adx_pgen_decode: non-LD records,
    then the LD-compressed ones against their decoded base), the variants
    the .pvar/.bim CHROM filter keeps gathered by adx_select_snps, then the
    flip rule (snp_reader.py:592-682 semantics)."""
    dev = torch.device(device)
    keep = ChromFilter(chrom_mode, autosomes, specific_chrom)
    with open_binary(path) as fh:
        data = fh.read()
    M_all, N, starts, lens, types = pgen_layout(data)
    chroms = _pgen_variant_chroms(Path(str(path)))
    if len(chroms) != M_all:
        raise ValueError(f"variant file lists {len(chroms)} variants, the .pgen {M_all}")
    kept = np.flatnonzero([keep(c) for c in chroms])
    keep.warn(M_all - kept.size)
    if kept.size == 0 or N == 0:
        raise ValueError("No variants found in dataset matching chromosome filters.")
    t = types & 7
    if np.any((t == 5) | (t == 6) | (t == 7)):
        raise ValueError(f".pgen record types {sorted(set(t[t >= 5].tolist()))} are not supported")
    need = np.where(t == 0, (N + 3) // 4, np.where(t == 1, 1 + (N + 7) // 8 + 1, 1))
    if np.any(lens < need):
        raise ValueError(".pgen record shorter than its genotype track")
    ld = (t == 2) | (t == 3)
    last = np.where(~ld, np.arange(M_all), -1)
    base = np.maximum.accumulate(last)
    if np.any(ld & (base < 0)):
        raise ValueError(".pgen LD-compressed record without a preceding base record")
    idb = 1 if N <= 1 << 8 else (2 if N <= 1 << 16 else (3 if N <= 1 << 24 else 4))
    Gall = _device_matrix(M_all, N, dev)
    st = _lib.stream_handle(dev)
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)           # noqa: E731
    buf = up(np.frombuffer(data, np.uint8).copy())
    d_st, d_len, d_t, d_base = up(starts), up(lens), up(types), up(base.astype(np.int64))
    for sel in (np.flatnonzero(~ld), np.flatnonzero(ld)):
        if sel.size:
            lst = up(sel.astype(np.int64))
            _lib.call("adx_pgen_decode", buf.data_ptr(), d_st.data_ptr(), d_len.data_ptr(), d_t.data_ptr(),
                      d_base.data_ptr(), lst.data_ptr(), sel.size, N, idb, Gall.data.data_ptr(), Gall.ldg, st)
    if kept.size == M_all:
        G = Gall
    else:
        G = _device_matrix(kept.size, N, dev)
        idx = up(kept.astype(np.int64))
        _lib.call("adx_select_snps", Gall.data.data_ptr(), Gall.ldg, idx.data_ptr(), kept.size, N, G.data.data_ptr(),
                  st)
    torch.cuda.current_stream(dev).synchronize()
    return _finish(G, flip), N, G.M


# ------------------------------------------------------------- dispatcher ----

def read_data(tr_file, packed: bool = True, chunk_size: int = 4096, chrom_mode: str = "autosomes",
              autosomes: int = 22, specific_chrom=None, verbose: bool = True, device="cuda"):
    """utils.py:161-205 / snp_reader.py:685-770 read_data for the GPU path:
    .bed, .vcf, .bcf and .pgen (plain or .gz) go straight to the device-resident
    packed matrix (returned as PackedGenotypes); ``packed=False`` returns the
    unpacked M x N uint8 matrix on the host.  A filter that leaves no SNP
    exits with status 1, as the reference does.  ``chunk_size`` is accepted
    for compatibility (blocks are sized in bytes here)."""
    import sys
    del chunk_size
    names = Path(str(tr_file)).name.split(".")[1:]
    exts = {f".{e}" for e in names}
    kw = dict(chrom_mode=chrom_mode, autosomes=autosomes, specific_chrom=specific_chrom)
    if chrom_mode not in ("all", "autosomes"):
        raise ValueError("chrom_mode must be 'all' or 'autosomes'")
    try:
        if ".bed" in exts:
            from .reader import read_bed_gpu
            path = Path(str(tr_file))
            if path.suffix in (".gz", ".zst"):
                raise RuntimeError(f"{path}: compressed .bed input is read by the host reader only")
            G, N, M = read_bed_gpu(path, device, **kw)
        elif ".vcf" in exts:
            G, N, M = read_vcf_gpu(tr_file, device, **kw)
        elif ".bcf" in exts:
            G, N, M = read_bcf_gpu(tr_file, device, **kw)
        elif ".pgen" in exts:
            G, N, M = read_pgen_gpu(tr_file, device, **kw)
        else:
            log.error("    Invalid format. Unrecognized file format. Make sure file ends with .bed, .pgen, "
                      ".vcf or .bcf .")
            sys.exit(1)
    except ValueError as err:
        log.error(f"    Error: {err}\n    Please check if your chromosome filter (--chrom_mode, --specific_chrom, "
                  "--autosomes) excluded all variants, or if the input dataset has no variants.")
        sys.exit(1)
    if verbose:
        log.info(f"    Data contains {N} samples and {M} SNPs.")
    if packed:
        return G, N, M
    out = torch.empty((M, N), dtype=torch.uint8, device=G.device)
    _lib.call("adx_unpack_u8", G.data.data_ptr(), G.ldg, M, N, 0, M, 0, out.data_ptr(), _lib.stream_handle(G.device))
    return out.cpu().numpy(), N, M
