"""The C-ABI libraries load without a GPU and export every function their headers declare."""
import ctypes
import os
import re

import pytest

from conftest import REPO

LIBS = {"pxgpu.h": "libpxgpu.so", "pxgen.h": "libpxgen.so"}


def declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(px\w+)\s*\(", text, flags=re.M)))


@pytest.mark.parametrize("header", sorted(LIBS))
def test_library_exports_declared_functions(header):
    path = os.path.join(REPO, "pixeltrack-standalone_amd", "lib", LIBS[header])
    assert os.path.exists(path), f"{path} not built"
    lib = ctypes.CDLL(path)
    names = declared(header)
    assert len(names) >= 10, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"{LIBS[header]} lacks {missing}"


def test_header_is_plain_c():
    """The boundary header compiles as C (no C++ or torch types leak into it)."""
    import subprocess
    for h in LIBS:
        r = subprocess.run(["gcc", "-std=c99", "-fsyntax-only", "-x", "c", os.path.join(REPO, "include", h)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_missing_library_fails_loudly(monkeypatch):
    """With the HIP library absent the product path raises instead of falling back."""
    from pxtrack import _lib
    monkeypatch.setattr(_lib, "LIB_DIR", "/nonexistent")
    with pytest.raises(OSError):
        _lib.load_library("libpxgpu_absent.so")
