"""The C-ABI library loads on a host without a GPU and exports every function that
include/moao_hip.h declares; the Python loader binds each of them (CPU only, no compute)."""
import ctypes
import os
import re

from conftest import ROOT
from moao_amd import lib as L

HEADER = os.path.join(ROOT, "include", "moao_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(moao_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_parses():
    names = declared()
    assert "moao_hip_reconstructor" in names and "moao_hip_matcov_tile" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    so = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in declared() if not hasattr(so, n)]
    assert not missing, missing


def test_loader_binds_every_declared_symbol():
    L.lib()
    assert sorted(set(declared()) - set(L.exported_symbols())) == []


def test_version_and_error_strings():
    assert b"gfx950" in L.lib().moao_version()
    assert isinstance(L.lib().moao_last_error(), bytes)
