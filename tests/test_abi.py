"""libzkg loads and exports every entry point include/zkg.h declares (CPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "zkg.h")
LIB = os.path.join(ROOT, "ziros_amd", "libzkg.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(zkg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("zkg_ctx_create", "zkg_coset_lde_batch", "zkg_merkle_commit", "zkg_fri_commit_phase",
                 "zkg_commit_batch", "zkg_serve_submit", "zkg_last_error"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzkg.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="libzkg.so not built")
def test_binding_covers_header():
    from ziros_amd import _lib

    assert set(declared()) == set(_lib.SIGNATURES)
    lib = _lib.load()
    assert lib.zkg_abi_version() == 1
    assert isinstance(lib.zkg_last_error(), bytes)  # thread-local; earlier calls in this process may have set it
