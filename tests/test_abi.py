"""The C-ABI library loads and exports every entry point include/fastdose_hip.h declares.

No compute calls: this runs without a GPU.
"""
import ctypes as C
import re
from pathlib import Path

import fastdose_amd as fd
from fastdose_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "fastdose_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fd_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert len(names) >= 25
    assert "fd_ctx_create" in names and "fd_resample_csr" in names


def test_library_exports_every_declared_symbol():
    lib = fd.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"


def test_struct_layouts_match_header():
    # POD sizes the header implies
    assert C.sizeof(fd.Volume) == 48
    assert C.sizeof(fd.Beam) == 4 * 18
    assert C.sizeof(fd.Spectrum) == 4 + 4 * 4 * 24
    assert C.sizeof(fd.Kernel) == 8 + 8 * 4 * 24
    assert C.sizeof(fd.Bundle) == 4 * 16
    assert C.sizeof(fd.CSR) == 48
    assert C.sizeof(fd._lib.ServeOpts) == 32
    assert C.sizeof(fd._lib.ServeStats) == 4 * 8 + 6 * 8 + 8 + 8
    assert C.sizeof(fd._lib.Tuning) == 16


def test_version_and_error_strings():
    lib = fd.load()
    assert b"gfx950" in lib.fd_version()
    assert lib.fd_calc_range(None, None) == _lib.FD_ERR_ARG
    assert b"null" in lib.fd_last_error()
