"""The C ABI of libadx.so: it loads, and exports every function include/adx.h
declares with the argument count the Python binding uses (no GPU needed)."""
import re

from tests.conftest import ROOT


def _declared():
    src = (ROOT / "include" / "adx.h").read_text()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*(?:const\s+)?[\w\s\*]+?\b(adx_\w+)\s*\(([^)]*)\)\s*;", src, flags=re.M):
        args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        out[m.group(1)] = len(args)
    return out


def test_header_declares_entry_points():
    decl = _declared()
    assert {"adx_em_pass", "adx_em_finalize", "adx_adam_step", "adx_unpack_u8", "adx_freq_f32"} <= set(decl)


def test_library_exports_every_declared_symbol():
    from adamixture_amd import _lib
    lib = _lib.load()
    decl = _declared()
    for name, nargs in decl.items():
        assert hasattr(lib, name), f"{name} declared in adx.h but not exported"
        assert name in _lib.SIGNATURES, f"{name} has no ctypes binding"
        assert len(_lib.SIGNATURES[name][1]) == nargs, f"{name}: binding has wrong arity"
    assert set(_lib.SIGNATURES) == set(decl)


def test_errors_are_reported_without_a_gpu():
    from adamixture_amd import _lib
    lib = _lib.load()
    rc = lib.adx_em_pass(0, None, 16, 10, 10, 99, None, 99, None, 99, None, None, None, 1, None, None, 0, None)
    assert rc != 0
    assert b"K must be" in lib.adx_last_error()


def test_eight_wave_tiles_halve_the_partial_t_slices():
    """Host-side sizing only: with 8-wave workgroups (16 cw 8 SNPs each) the
    per-workgroup T slices of the F32/F64 workspace and of the BF16 operand
    buffer are half those of 4-wave workgroups; passes that carry the
    log-likelihood keep 4 waves, and the workspace covers both."""
    from adamixture_amd import _lib
    lib = _lib.load()
    M, N, K = 100_000, 20_000, 10
    saved = {}
    for tag in (_lib.ADX_F32, _lib.ADX_F64, _lib.ADX_BF16):
        cw, wv, tg = _lib.ctypes.c_int(), _lib.ctypes.c_int(), _lib.ctypes.c_int64()
        assert lib.adx_get_em_tiles(tag, _lib.ctypes.byref(cw), _lib.ctypes.byref(wv), _lib.ctypes.byref(tg)) == 0
        saved[tag] = (cw.value, wv.value, tg.value)
    try:
        for tag, esz in ((_lib.ADX_F32, 4), (_lib.ADX_F64, 8)):
            sizes = {}
            for wv in (4, 8):
                assert lib.adx_set_em_tiles(tag, 4, wv, 64) == 0      # one sample split: slices are T only
                sizes[wv] = lib.adx_em_workspace_bytes(tag, M, N, K)
            t4 = -(-M // 256) * N * K * esz
            assert sizes[4] >= t4 and sizes[4] - t4 < 2 * 2 * M * K * esz + (1 << 20)
            assert sizes[8] == sizes[4]                      # sized for the 4-wave log-likelihood pass too
            # the statistics-only band plan at 8 waves fits half the T slices
            band = lib.adx_em_band_snps(tag, M, N, K, _lib.ADX_EM_STATS, t4 // 2 + 2 * 2 * M * K * esz + (1 << 20), 0)
            assert band == M
        b = {}
        for wv in (4, 8):
            assert lib.adx_set_em_tiles(_lib.ADX_BF16, 8, wv, 64) == 0
            b[wv] = lib.adx_em_bf16_operand_bytes(M, N, K, None)
        assert b[4] - b[8] == (-(-M // 512) - (-(-M // 1024))) * N * K * 4
        assert lib.adx_set_em_tiles(_lib.ADX_BF16, 8, 6, 64) != 0
    finally:
        for tag, t in saved.items():
            lib.adx_set_em_tiles(tag, *t)
