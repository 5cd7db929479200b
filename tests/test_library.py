"""The C-ABI library: built for gfx950, loads on a CPU-only host and exports
every function the public headers declare (no device calls here)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "ahf_amd", "lib", "libahf_amd.so")
HEADERS = ["ahf_gpu.h", "ahf_dropin.h"]


def declared(header):
    path = os.path.join(ROOT, "include", header)
    if not os.path.exists(path):
        return []
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:extern\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", txt, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while", "sizeof", "defined")))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from ahf_amd import build
        build.build_gpu()
    return ctypes.CDLL(LIB)


@pytest.mark.parametrize("header", HEADERS)
def test_exports_every_declared_symbol(lib, header):
    names = declared(header)
    if header == "ahf_gpu.h":
        assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{header}: not exported: {missing}"


def test_built_for_gfx950(lib):
    # the embedded offload bundle names its target
    blob = open(LIB, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_python_mirror_fails_loudly_without_library(tmp_path, monkeypatch):
    import ahf_amd
    monkeypatch.setattr(ahf_amd, "LIBDIR", str(tmp_path))
    monkeypatch.setattr(ahf_amd, "_lib", None)
    with pytest.raises(ahf_amd.NativeLibraryMissing):
        ahf_amd.load()
