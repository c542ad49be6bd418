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


def test_edge_sample_index_matches_oracle_bisection_order():
    """The closed-form k-th bisection sample (kg_edge_sample_index, used by
    the device edge rounds) equals the oracle's queue-built level order."""
    import numpy as np
    import oracle
    L = _lib.lib()
    rng = np.random.default_rng(0)
    ns = list(range(0, 700)) + [1023, 1024, 1025, 4095, 4096, 4097] + list(rng.integers(2, 200000, 12))
    for n in ns:
        n = int(n)
        order = oracle.bisection_order(n)
        got = np.array([L.kg_edge_sample_index(n, k) for k in range(max(0, n - 1))], np.int32)
        assert np.array_equal(got, order), n
        assert L.kg_edge_sample_index(n, -1) == -1 and L.kg_edge_sample_index(n, max(0, n - 1)) == -1
