"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and rejects bad arguments with status codes."""
import ctypes as C

import piro_amd
from piro_amd import lib


def test_every_declared_symbol_is_exported():
    syms = piro_amd.declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert missing == []


def test_host_side_helpers_match_reference_formulas():
    # src/operatorlaplacian.cpp:33 and src/operatorgradient.cpp:25
    import numpy as np
    h = np.float32(1.0) / np.float32(30)
    assert lib.piro_lap_norm(1.0, 30) == np.float32(float(h) ** 2)
    assert lib.piro_grad_norm(1.0, 30) == np.float32(2.0) * h
    v = C.c_int64()
    assert lib.piro_lap_nnz(32, 32, 32, C.byref(v)) == 0 and v.value == 223232
    assert lib.piro_lap_nnz(256, 256, 256, C.byref(v)) == 0 and v.value == 117047296


def test_errors_are_status_codes_not_crashes():
    v = C.c_int64()
    assert lib.piro_lap_nnz(0, 4, 4, C.byref(v)) == 1
    assert b"dims" in lib.piro_last_error()
    assert lib.piro_grad_nnz(4, 4, 4, 3, C.byref(v)) == 1
    ts = C.c_int()
    # the shipped hashtable.ini fit turns negative near N ~ 18652 (finding 11)
    assert lib.piro_ht_table_size(140000, 20000, 3.3608, -0.5552, -0.0143, 0.7, C.byref(ts)) == 1
    assert lib.piro_ht_table_size(725, 125, 3.3608, -0.5552, -0.0143, 0.7, C.byref(ts)) == 0
    assert ts.value > 725
    # a null context is rejected
    assert lib.piro_fe_step_csr(None, 4, 4, 4, None, None, None, 0, 0.1, None, None, None) == 1


def test_variant_table():
    n = lib.piro_fe_variant_count()
    assert n >= 4
    names = [lib.piro_fe_variant_name(i).decode() for i in range(n)]
    assert len(set(names)) == n
