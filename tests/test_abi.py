"""The C-ABI library loads (no GPU needed to load) and exports every entry
point that include/c3d.h declares; the ctypes binding covers them all."""
import ctypes
import os
import re
import subprocess

import pytest

from channel3d_parallel_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(c3d_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header_symbols():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libc3d.so not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared("c3d.h")
    assert names, "no entry points parsed"
    for n in names:
        assert hasattr(lib, n), f"{n} declared in c3d.h but not exported"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\bT (c3d_[a-z0-9_]+)", out))
    assert set(names) <= exported


def test_ctypes_binding_covers_header():
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(declared("c3d.h")) == bound


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()
