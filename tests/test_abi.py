"""The C-ABI library loads and exports every symbol include/klampt_gpu.h
declares (no compute calls: no GPU needed)."""
import ctypes
import os
import re

from klampt_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "klampt_gpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kg_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(_lib.SIGNATURES) == _declared()


def test_errors_without_device_calls():
    L = _lib.lib()
    # NULL handles are rejected with an error code and a message
    assert L.kg_geom_collide(None, None, None) == -1
    assert b"NULL" in L.kg_last_error()
    assert L.kg_version() >= 1
