"""Device utilities with the signatures of adamixture/src/utils.py.

The unpack / log-likelihood / frequency seams call libadx; the OOM and
output helpers keep the reference's behaviour.
"""
from __future__ import annotations

import gc
import logging
import random
from collections.abc import Callable
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .genotypes import (HostPackedGenotypes, PackedGenotypes, allele_frequencies, as_packed, unpack_center_chunk,
                        unpack_chunk)

log = logging.getLogger(__name__)

MIN_GPU_CHUNK_SIZE = 2
MIN_CPU_CHUNK_SIZE = 2


def is_gpu_oom(exc: BaseException) -> bool:
    """utils.py:19-43: OOM recognised by exception type or message."""
    message = str(exc).lower()
    return isinstance(exc, torch.cuda.OutOfMemoryError) or "out of memory" in message


def reduce_gpu_chunk_or_raise(chunk_size: int, exc: BaseException, context: str) -> int:
    """utils.py:62-89: halve the chunk after an OOM and release cached memory."""
    if chunk_size <= MIN_GPU_CHUNK_SIZE:
        raise exc
    new = max(chunk_size // 2, MIN_GPU_CHUNK_SIZE)
    exc.__traceback__ = None
    gc.collect()
    if torch.cuda.is_available():
        torch.cuda.empty_cache()
    log.warning(f"    GPU out of memory while {context}: retrying with chunk_size={new}.")
    return new


def get_dtype(device: torch.device) -> torch.dtype:
    """utils.py:235-250: FP64 for the GPU path."""
    return torch.float64


def get_tuning_params(device: torch.device) -> int:
    """utils.py:207-233.  Kept for signature compatibility; the fused kernels
    pick their own launch shapes (256-thread workgroups, wave64)."""
    return 256


def load_extensions(device: torch.device) -> None:
    """utils.py:431-480 JIT-compiled three CUDA extensions; here the
    prebuilt libadx.so is loaded (raises if it is missing)."""
    if torch.device(device).type == "cuda":
        _lib.load()


def is_packed_genotype_tensor(G, M: int) -> bool:
    """utils.py:483-494."""
    return isinstance(G, PackedGenotypes) or (isinstance(G, torch.Tensor) and G.ndim == 2 and G.size(0) != M)


def get_unpacker(device: torch.device, threads_per_block: int = 0) -> Callable:
    """utils.py:497-559: (G, start, C, M) -> C x N uint8 codes on the device."""
    dev = torch.device(device)

    def unpack(G, start_idx: int, actual_chunk_size: int, M: int) -> torch.Tensor:
        return unpack_chunk(as_packed(G, M, dev), start_idx, actual_chunk_size)

    return unpack


def get_centering_unpacker(device: torch.device, threads_per_block: int = 0) -> Callable:
    """utils.py:713-779: (G, f, start, C, M) -> C x N centred FP32 chunk."""
    dev = torch.device(device)

    def unpack_center(G, f: torch.Tensor, start_idx: int, actual_chunk_size: int, M: int) -> torch.Tensor:
        return unpack_center_chunk(as_packed(G, M, dev), f, start_idx, actual_chunk_size)

    return unpack_center


def get_logl_calculator(device: torch.device) -> Callable:
    """utils.py:605-658: (G, P, Q, M, N, batch, tpb) -> FP64 log-likelihood."""
    from .em_adam_gpu import loglikelihood_gpu
    dev = torch.device(device)

    def logl(G, P, Q, M, N, batch_size=0, threads_per_block=0) -> float:
        return loglikelihood_gpu(G, P, Q, M, N, batch_size, dev, threads_per_block)

    return logl


def calculate_frequencies_gpu(G, M: int, chunk_size: int, device_obj: torch.device,
                              threads_per_block: int = 0) -> torch.Tensor:
    """utils.py:832-861: per-SNP allele frequency (FP32), one fused pass."""
    return allele_frequencies(as_packed(G, M, device_obj))


def get_free_gpu_memory(device: torch.device) -> float:
    """utils.py:290-308: MB available to new tensors."""
    device = torch.device(device)
    torch.cuda.synchronize(device)
    free, _ = torch.cuda.mem_get_info(device)
    return round((free + torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)) / 2**20, 2)


def estimate_gpu_bytes(M: int, N: int, K: int, algorithm: str, q_hist: int = 3,
                       include_initialization: bool = True) -> int:
    """Peak device bytes besides G for the fused kernels.  Unlike the
    reference estimate (utils.py:360-417) there are no C x N chunk
    temporaries: only O((M + N) K) buffers."""
    L = max(K + 10, 20)
    svd = M * L * 4 + 2 * N * L * 4 + M * L * 4
    als = 5 * (M + N) * K * 8
    if algorithm == "adamem":
        # + the FP64 pass's partial-sum workspace: the single-band T, A and B
        # slices, capped at max(G / 10, 256 MiB) (the pass then runs in SNP bands)
        g_bytes = ((M + 3) // 4) * ((N + 63) // 64 * 64)
        ws = min(_lib.load().adx_em_workspace_bytes(_lib.ADX_F64, M, N, K), max(g_bytes // 10, 256 << 20))
        opt = (8 * M * K + 7 * N * K + N) * 8 + 2 * (M + N) * K * 4 + ws
    elif algorithm == "brqn":
        dim = (M + N) * K
        opt = ((M + N) * K * K + (M + N) * K) * 8 + 3 * dim * q_hist * 8 + 9 * dim * 8
    else:
        raise ValueError(f"Unknown algorithm for GPU memory estimate: {algorithm}")
    return max(svd, als, opt) if include_initialization else opt


# Device bytes the host-streamed path stages genotype rows in (two halves of
# the EM pass's band, or one band of the other passes).
STREAM_STAGING_BYTES = 1 << 30


def manage_gpu_memory(G, device: torch.device, M: int, N: int, K: int, chunk_size: int, algorithm: str,
                      q_hist: int = 3, include_initialization: bool = True):
    """utils.py:311-429: place G on the device when it and the working
    buffers fit 95% of free memory (PackedGenotypes, padded layout), else
    keep it in pinned host memory and stream its byte rows through a device
    staging buffer (HostPackedGenotypes; the reference keeps G on the CPU
    then, :425-428).  If not even the buffers and the staging fit, a
    RuntimeError naming "out of memory" is raised."""
    dev = torch.device(device)
    if dev.type != "cuda":
        return G
    if isinstance(G, (PackedGenotypes, HostPackedGenotypes)):
        return G
    g_bytes = ((M + 3) // 4) * ((N + 63) // 64 * 64)
    buffers = estimate_gpu_bytes(M, N, K, algorithm, q_hist, include_initialization)
    avail = get_free_gpu_memory(dev) * 0.95 * 2**20
    if g_bytes + buffers <= avail:
        log.info("    Moving genotype matrix to GPU...")
        return as_packed(G, M, dev)
    staging = min(STREAM_STAGING_BYTES, max(avail - buffers, 0))
    if staging < 4 * ((N + 63) // 64 * 64):
        raise RuntimeError(f"out of memory: genotypes ({g_bytes / 2**30:.1f} GiB) do not fit and the buffers need "
                           f"{buffers / 2**30:.1f} GiB of {avail / 2**30:.1f} GiB free")
    log.info("    Genotype matrix too large for GPU, keeping it on the host (streamed in SNP bands)...")
    return HostPackedGenotypes.from_packed(G, M, dev, staging_cap=int(staging))


def write_outputs(Q: np.ndarray, run_name: str, K: int, out_path, P: np.ndarray | None = None) -> None:
    """utils.py:252-273: ADMIXTURE-compatible .Q / .P text files
    (numpy.savetxt defaults: '%.18e', space-delimited)."""
    out_path = Path(out_path)
    np.savetxt(out_path / f"{run_name}.{K}.Q", Q, delimiter=" ")
    if P is not None:
        np.savetxt(out_path / f"{run_name}.{K}.P", P, delimiter=" ")


def set_seed(seed: int) -> None:
    """utils.py:276-287."""
    np.random.seed(seed)
    random.seed(seed)
    torch.manual_seed(seed)
    if torch.cuda.is_available():
        torch.cuda.manual_seed_all(seed)
