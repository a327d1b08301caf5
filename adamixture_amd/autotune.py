"""Tile-shape autotuner for the EM passes (project extension; the reference
has fixed launch shapes, adamixture/src/utils_c/cuda/*.cu).

The EM kernels take three launch parameters, read by libadx at every launch
(adx_set_em_tiles): ``cw``, the number of 16-SNP column tiles a wavefront
owns (register footprint vs. reuse of the sample-side operands); ``waves``,
the wavefronts per workgroup (4 or 8: a workgroup owns 16 cw waves SNPs and
writes one partial-T slice for them, so 8 waves halve that traffic at the
same occupancy); and ``target_wgs``, the number of workgroups the sample
range is split to reach (grid fill vs. the A/B partial slices of the
sample-side reduction).  ``tune_em``
times each candidate on the caller's own genotypes with HIP events on the
launch stream and keeps the fastest; results are cached per (device,
precision, M, N, K) in memory and, if ``ADX_TUNE_CACHE`` names a file, on
disk as JSON.
"""
from __future__ import annotations

import ctypes
import json
import logging
import os
from pathlib import Path

import torch

from . import _lib
from .genotypes import PackedGenotypes

log = logging.getLogger(__name__)

DTYPE_TAG = {"f32": _lib.ADX_F32, "f64": _lib.ADX_F64, "bf16": _lib.ADX_BF16}
TARGETS = (1024, 2048, 4096, 8192)
CANDIDATES = {
    "f32": [(cw, wv, t) for cw in (4, 8) for wv in (4, 8) for t in TARGETS],
    "f64": [(4, wv, t) for wv in (4, 8) for t in TARGETS],
    "bf16": [(cw, wv, t) for cw in (4, 8) for wv in (4, 8) for t in TARGETS],
}
_MEM: dict[str, tuple[int, int, int]] = {}


def get_tiles(precision: str) -> tuple[int, int, int]:
    """The current (cw, waves, target_wgs) of ``precision``."""
    cw, wv, tg = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int64(0)
    _lib.call("adx_get_em_tiles", DTYPE_TAG[precision], ctypes.addressof(cw), ctypes.addressof(wv),
              ctypes.addressof(tg))
    return cw.value, wv.value, tg.value


def set_tiles(precision: str, cw: int, waves: int, target_wgs: int) -> None:
    _lib.call("adx_set_em_tiles", DTYPE_TAG[precision], int(cw), int(waves), int(target_wgs))


def tiles_label(tiles) -> str:
    """'cw/waves/target_wgs', the key of a tile shape in timing tables."""
    return "/".join(str(int(x)) for x in tiles)


def _key(device, precision: str, M: int, N: int, K: int) -> str:
    name = torch.cuda.get_device_name(device).replace(" ", "_")
    return f"{name}|{precision}|M={M}|N={N}|K={K}"


def _disk_cache() -> Path | None:
    p = os.environ.get("ADX_TUNE_CACHE")
    return Path(p) if p else None


def _load_disk() -> dict:
    p = _disk_cache()
    if p and p.exists():
        try:
            return json.loads(p.read_text())
        except (OSError, ValueError):
            return {}
    return {}


def _store_disk(key: str, value) -> None:
    p = _disk_cache()
    if not p:
        return
    data = _load_disk()
    data[key] = value
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(json.dumps(data, indent=1, sort_keys=True))


def _stats_pass(G: PackedGenotypes, P, Q, precision: str, bufs: dict) -> None:
    """One EM-statistics pass of ``precision`` into scratch accumulators."""
    from .em_adam_gpu import em_pass
    M, K = P.shape
    N = Q.shape[0]
    st = _lib.stream_handle(G.device)
    if precision == "bf16":
        nb = bufs["ops"].numel()
        _lib.call("adx_em_pass_bf16", G.data.data_ptr(), G.ldg, M, N, K, bufs["ops"].data_ptr(), nb,
                  bufs["A"].data_ptr(), bufs["B"].data_ptr(), bufs["T"].data_ptr(), st)
    elif precision == "f32":
        em_pass(G, bufs["P32"], bufs["Q32"], bufs["A"], bufs["B"], bufs["T"], dtype=_lib.ADX_F32)
    else:
        em_pass(G, P, Q, bufs["A"], bufs["B"], bufs["T"], dtype=_lib.ADX_F64)


def tune_em(G: PackedGenotypes, P: torch.Tensor, Q: torch.Tensor, precision: str, *, repeats: int = 2,
            candidates=None) -> dict:
    """Time every candidate tile shape on (G, P, Q); set and return the best.
    Returns {"best": (cw, waves, target_wgs), "ms": {"cw/waves/target": ms, ...}}."""
    M, K = P.shape
    N = Q.shape[0]
    dev = G.device
    acc_dt = torch.float64 if precision == "f64" else torch.float32
    bufs = {"A": torch.zeros((M, K), dtype=acc_dt, device=dev), "B": torch.zeros((M, K), dtype=acc_dt, device=dev),
            "T": torch.zeros((N, K), dtype=acc_dt, device=dev)}
    before = get_tiles(precision)
    cands = candidates or CANDIDATES[precision]
    if precision == "bf16":
        nb = 0                                      # the operand buffer holds per-tile partial sums
        for tiles in cands:
            set_tiles(precision, *tiles)
            nb = max(nb, _lib.load().adx_em_bf16_operand_bytes(M, N, K, None))
        set_tiles(precision, *before)
        bufs["ops"] = torch.zeros(nb, dtype=torch.uint8, device=dev)
        _lib.call("adx_em_prep_bf16", P.data_ptr(), Q.data_ptr(), M, N, K, bufs["ops"].data_ptr(), nb,
                  _lib.stream_handle(dev))
    elif precision == "f32":
        bufs["P32"] = P.to(torch.float32).contiguous()
        bufs["Q32"] = Q.to(torch.float32).contiguous()
    timings = {}
    for tiles in cands:
        set_tiles(precision, *tiles)
        _stats_pass(G, P, Q, precision, bufs)          # warm-up (code load, caches)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(repeats):
            _stats_pass(G, P, Q, precision, bufs)
        e1.record()
        e1.synchronize()
        timings[tiles_label(tiles)] = e0.elapsed_time(e1) / repeats
    best = min(timings, key=timings.get)
    cw, wv, tg = (int(x) for x in best.split("/"))
    set_tiles(precision, cw, wv, tg)
    log.info(f"    autotune {precision} M={M} N={N} K={K}: best cw={cw} waves={wv} target_wgs={tg} "
             f"({timings[best]:.3f} ms; default {tiles_label(before)}: "
             f"{timings.get(tiles_label(before), float('nan')):.3f} ms)")
    return {"best": (cw, wv, tg), "ms": timings, "default": before}


def ensure_tuned(G: PackedGenotypes, P: torch.Tensor, Q: torch.Tensor, precision: str) -> tuple[int, int, int]:
    """Apply the cached tile shape for this problem, tuning it on first use."""
    M, K = P.shape
    key = _key(G.device, precision, M, Q.shape[0], K)
    hit = _MEM.get(key) or (tuple(_load_disk()[key]) if key in _load_disk() else None)
    if hit is not None and len(hit) != 3:      # an entry from before the waves parameter
        hit = None
    if hit is None:
        hit = tune_em(G, P, Q, precision)["best"]
        _store_disk(key, list(hit))
    _MEM[key] = hit
    set_tiles(precision, *hit)
    return hit
