"""The C-ABI library loads and exports every symbol the headers under
include/ declare (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from kelp_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in sorted(os.listdir(inc)):
        if fn.endswith(".h"):
            text = re.sub(r"/\*.*?\*/", "", open(os.path.join(inc, fn)).read(), flags=re.S)
            names.update(re.findall(r"\b(kelp_\w+)\s*\(", text))
    return sorted(names)


def test_header_symbols_exported():
    lib = ctypes.CDLL(_native.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(_native.SIGNATURES)


def test_version_and_error_without_gpu():
    lib = _native.load()
    assert lib.kelp_version() >= 1
    assert isinstance(lib.kelp_last_error(), bytes)
    # a NULL context is rejected with a message, no device touched
    assert lib.kelp_synchronize(None) == -1
    assert b"NULL" in lib.kelp_last_error()
