"""ctypes binding of libadx.so (the gfx950 kernels, C ABI in include/adx.h).

The library is built in-tree by ``adamixture_amd._build.build_library`` and
loaded after ``import torch`` so that it resolves ``libamdhip64.so.7`` to the
HIP runtime torch already loaded (one runtime per process).  There is no
fallback: if the library is missing, every GPU entry point raises.
"""
from __future__ import annotations

import ctypes
import threading
from pathlib import Path

import torch  # noqa: F401  (must precede loading libadx: shared HIP runtime)

LIB_PATH = Path(__file__).resolve().parent / "libadx.so"

ADX_F32 = 0
ADX_F64 = 1
ADX_BF16 = 2
ADX_EM_STATS = 1
ADX_EM_LOGLIK = 2
ADX_CV_BLOCKS = 4096          # include/adx.h
ADX_GEMM_TN_BLOCKS = 512      # include/adx.h
ADX_EM_NOALLOC_SCRATCH_BYTES = 128 << 20   # include/adx.h

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_flt = ctypes.c_float
_u64 = ctypes.c_uint64
_p = ctypes.c_void_p

# name -> (restype, argtypes); mirrors include/adx.h
SIGNATURES = {
    "adx_last_error": (ctypes.c_char_p, []),
    "adx_unpack_u8": (_int, [_p, _i64, _i64, _i64, _i64, _i64, _i64, _p, _p]),
    "adx_unpack_center_f32": (_int, [_p, _i64, _p, _i64, _i64, _i64, _i64, _i64, _p, _p]),
    "adx_freq_f32": (_int, [_p, _i64, _i64, _i64, _p, _p]),
    "adx_qbat": (_int, [_p, _i64, _i64, _i64, _p, _p]),
    "adx_repack_padded": (_int, [_p, _i64, _i64, _i64, _p, _i64, _p]),
    "adx_synth_packed": (_int, [_p, _p, _i64, _i64, _int, _u64, _dbl, _p, _i64, _p]),
    "adx_em_workspace_bytes": (_i64, [_int, _i64, _i64, _int]),
    "adx_em_pass": (_int, [_int, _p, _i64, _i64, _i64, _int, _p, _i64, _p, _i64, _p, _p, _p, _int, _p,
                           _p, _i64, _p]),
    "adx_em_band_snps": (_i64, [_int, _i64, _i64, _int, _int, _i64, _i64]),
    "adx_em_pass_host": (_int, [_int, _p, _i64, _i64, _i64, _int, _p, _i64, _p, _i64, _p, _p, _p, _int, _p, _p,
                                _i64, _p, _i64, _p, _p]),
    "adx_em_pass_noalloc": (_int, [_int, _p, _i64, _i64, _i64, _int, _p, _i64, _p, _i64, _p, _p, _p, _int,
                                   _p, _p]),
    "adx_em_finalize": (_int, [_int, _p, _p, _p, _p, _p, _p, _i64, _i64, _int, _p, _p, _p]),
    "adx_adam_step": (_int, [_p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _int, _dbl, _dbl, _dbl, _dbl,
                             _dbl, _p, _p, _i64, _p]),
    "adx_em_bf16_operand_bytes": (_i64, [_i64, _i64, _int, _p]),
    "adx_em_prep_bf16": (_int, [_p, _p, _i64, _i64, _int, _p, _i64, _p]),
    "adx_em_pass_bf16": (_int, [_p, _i64, _i64, _i64, _int, _p, _i64, _p, _p, _p, _p]),
    "adx_svd_xw": (_int, [_p, _i64, _i64, _i64, _p, _p, _int, _p, _p, _p]),
    "adx_svd_xty": (_int, [_p, _i64, _i64, _i64, _p, _p, _int, _p, _p, _p]),
    "adx_svd_work_floats": (_i64, [_int, _i64, _i64, _int]),
    "adx_svd_xw_h2": (_int, [_p, _i64, _i64, _i64, _p, _p, _int, _p, _p, _i64, _p]),
    "adx_svd_xty_h2": (_int, [_p, _i64, _i64, _i64, _p, _p, _int, _p, _p, _i64, _p]),
    "adx_svd_h2_work_bytes": (_i64, [_int, _i64, _i64, _int]),
    "adx_grad_hess": (_int, [_int, _p, _i64, _i64, _i64, _int, _p, _p, _p, _p, _p]),
    "adx_sqp_solve": (_int, [_int, _p, _p, _p, _p, _p, _i64, _int, _p, _i64, _p]),
    "adx_bvls_workspace_bytes": (_i64, [_int]),
    "adx_bvls_batch": (_int, [_p, _p, _p, _p, _i64, _int, _dbl, _dbl, _p, _i64, _p]),
    "adx_probe_mfma": (_int, [_int, _p, _p, _p, _p]),
    "adx_adam_step_sched": (_int, [_p, _p, _p, _p, _p, _p, _p, _p, _i64, _i64, _int, _p, _dbl, _dbl, _dbl, _p, _p,
                                   _i64, _p]),
    "adx_bed_to_packed": (_int, [_p, _i64, _i64, _i64, _p, _i64, _i64, _p]),
    "adx_packed_allele_sum": (_int, [_p, _i64, _i64, _i64, _p, _p]),
    "adx_flip_packed": (_int, [_p, _i64, _i64, _i64, _p]),
    "adx_label_freq": (_int, [_p, _i64, _i64, _i64, _int, _p, _p, _dbl, _p, _p]),
    "adx_snap_labels": (_int, [_p, _i64, _int, _p, _dbl, _p]),
    "adx_cv_mask_list": (_int, [_p, _i64, _i64, _i64, _p, _int, _i64, _i64, _i64, _p]),
    "adx_cv_mask_hash": (_int, [_p, _p, _i64, _i64, _i64, _u64, _int, _int, _p]),
    "adx_cv_deviance_list": (_int, [_p, _i64, _i64, _i64, _p, _int, _i64, _i64, _i64, _p, _p, _int, _p, _p, _p]),
    "adx_cv_deviance_hash": (_int, [_p, _i64, _i64, _i64, _u64, _int, _int, _p, _p, _int, _p, _p, _p]),
    "adx_project_simplex": (_int, [_p, _i64, _int, _dbl, _p]),
    "adx_vcf_gt": (_int, [_p, _p, _p, _i64, _i64, _i64, _p, _i64, _p]),
    "adx_bcf_gt": (_int, [_p, _p, _p, _i64, _i64, _i64, _p, _i64, _p]),
    "adx_clear_tail_fields": (_int, [_p, _i64, _i64, _i64, _p]),
    "adx_pgen_decode": (_int, [_p, _p, _p, _p, _p, _p, _i64, _i64, _int, _p, _i64, _p]),
    "adx_select_snps": (_int, [_p, _i64, _p, _i64, _i64, _p, _p]),
    "adx_gemm_tn_f64": (_int, [_p, _i64, _p, _i64, _i64, _int, _int, _dbl, _p, _i64, _p, _p]),
    "adx_gemm_nn_f64": (_int, [_p, _i64, _i64, _int, _int, _p, _i64, _dbl, _p, _p, _dbl, _p, _i64, _p]),
    "adx_set_em_tiles": (_int, [_int, _int, _int, _i64]),
    "adx_get_em_tiles": (_int, [_int, _p, _p, _p]),
    "adx_mlp_gemm_f32": (_int, [_int, _int, _i64, _i64, _i64, _flt, _p, _i64, _p, _i64, _flt, _p, _i64, _int, _p]),
    "adx_mlp_bias_act": (_int, [_p, _p, _i64, _i64, _i64, _int, _p]),
    "adx_mlp_relu_bwd": (_int, [_p, _p, _i64, _p]),
    "adx_mlp_colsum": (_int, [_p, _i64, _i64, _p, _p]),
    "adx_mlp_softmax_xent": (_int, [_p, _i64, _p, _i64, _p, _i64, _p, _i64, _p, _p, _p, _p]),
    "adx_mlp_adam": (_int, [_p, _p, _p, _p, _i64, _flt, _flt, _flt, _flt, _int, _p]),
    "adx_mlp_pack_bf16": (_int, [_p, _p, _p]),
    "adx_mlp_infer_bf16": (_int, [_p, _i64, _i64, _int, _p, _p, _i64, _p]),
}

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


def load() -> ctypes.CDLL:
    """Load and bind libadx.so; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise RuntimeError(
                    f"libadx.so not found at {LIB_PATH}; build it with "
                    "`python -c 'import adamixture_amd._build as b; b.build_library()'`")
            lib = ctypes.CDLL(str(LIB_PATH))
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
        return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = load().adx_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({status}): {msg}")


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args), name)


def stream_handle(device: torch.device | None = None) -> int:
    """HIP stream of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()
