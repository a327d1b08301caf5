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
