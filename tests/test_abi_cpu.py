"""The C-ABI library loads and exports every symbol include/nqcd_mi355x.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "nqcd_mi355x.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nqcd_mi355x_\w+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    from nqcd_mi355x import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.EXPORTED_SYMBOLS)


def test_abi_version_and_error_without_device():
    from nqcd_mi355x import _lib
    L = _lib.lib()
    assert L.nqcd_mi355x_abi_version() == _lib.ABI_VERSION


def test_library_is_gfx950():
    """The shared object carries gfx950 code objects (built by nqcdynamics.jl_amd/build.py)."""
    from nqcd_mi355x import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_autotune_entry_point_rejects_null_handle():
    """nqcd_mi355x_autotune checks its arguments before any device work; its result struct matches
    the header's layout (four int32 then two doubles)."""
    from nqcd_mi355x import _lib
    L = _lib.lib()
    r = _lib.TuneResult()
    assert ctypes.sizeof(_lib.TuneResult) == 32
    assert L.nqcd_mi355x_autotune(None, 0.0, 1.0, None, ctypes.c_uint64(1), ctypes.byref(r)) == _lib.ERR_ARG
