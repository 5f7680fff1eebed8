"""CPU-side checks of the C ABI: the library builds, loads without a GPU and
exports every entry point declared in include/frovedis_hip.h.  No compute
calls are made here."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "frovedis_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fhip_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def libpath():
    from frovedis_amd import build as b
    if not os.path.exists(b.LIB):
        b.build()
    return b.LIB


def test_header_declares_entry_points():
    names = _declared()
    assert "fhip_dev_radix_sort" in names and "fhip_last_error" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol(libpath):
    L = ctypes.CDLL(libpath)
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"


def test_last_error_is_empty_string_initially(libpath):
    L = ctypes.CDLL(libpath)
    L.fhip_last_error.restype = ctypes.c_char_p
    assert L.fhip_last_error() == b""


def test_package_imports_and_fails_loudly_without_library(tmp_path, monkeypatch):
    import frovedis_amd
    monkeypatch.setattr(frovedis_amd, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(frovedis_amd, "_lib", None)
    with pytest.raises(frovedis_amd.FhipError):
        frovedis_amd.lib()


def _build_shim(out):
    import subprocess
    cmd = ["g++", "-std=c++11", "-O2", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "shim_test.cpp"),
           "-L", os.path.join(ROOT, "frovedis_amd"), "-lfrovedis_hip",
           "-Wl,-rpath," + os.path.join(ROOT, "frovedis_amd"), "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)


def test_cpp_dropin_headers_compile_as_cxx11(libpath, tmp_path):
    # the reference builds with -std=c++11 (Makefile.in.x86:3-4)
    _build_shim(tmp_path / "shim_test")
    assert (tmp_path / "shim_test").exists()


def test_prototypes_cover_every_declared_entry_point():
    # each declaration maps to ctypes types (sizes as size_t, pointers as
    # 64-bit), so no Python integer reaches the library as a 32-bit int
    from frovedis_amd import abi
    protos = abi.prototypes()
    assert sorted(protos) == _declared()
    res, args = protos["fhip_dev_radix_sort"]
    assert res is ctypes.c_int
    assert args == [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                    ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    assert protos["fhip_last_error"] == (ctypes.c_char_p, [])


def test_loaded_library_rejects_a_wrongly_typed_argument(libpath):
    import frovedis_amd
    L = frovedis_amd.lib()
    with pytest.raises(ctypes.ArgumentError):
        L.fhip_dev_radix_sort(3, 1.5, 0, None, 0, 0, 0, None)
    with pytest.raises(TypeError):  # too few arguments
        L.fhip_dev_radix_sort(3, None, 0)
