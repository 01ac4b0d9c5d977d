"""The C-ABI library loads and exports every entry point include/*.h declares
(no GPU needed: nothing is called)."""
import ctypes as ct
import os
import re

from theses_code_amd import _native as N

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    inc = os.path.join(REPO, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            src = open(os.path.join(inc, f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(frm_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = ct.CDLL(N.LIB_PATH)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_every_declared_symbol():
    assert declared_symbols() <= set(N.SIGNATURES), declared_symbols() - set(N.SIGNATURES)


def test_strerror_without_gpu():
    lib = N.lib()
    assert lib.frm_strerror(0) == b"ok"
    assert lib.frm_strerror(-4).startswith(b"capacity")
