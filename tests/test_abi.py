"""The C-ABI library loads and exports every symbol include/mcstar_hip.h declares (no GPU calls)."""
import ctypes


def test_library_exports_declared_symbols(mcs):
    lib = mcs.lib()
    declared = mcs.declared_symbols()
    assert len(declared) >= 35
    missing = [s for s in declared if not hasattr(lib, s)]
    assert missing == []


def test_error_strings(mcs):
    lib = mcs.lib()
    assert lib.mcs_error_string(0) == b"MCS_OK"
    assert lib.mcs_error_string(5) == b"MCS_ERR_CAPACITY"
    assert lib.mcs_error_string(99) == b"<unknown>"


def test_argument_validation_needs_no_device(mcs):
    lib = mcs.lib()
    p = ctypes.c_void_p()
    assert lib.mcs_fft_plan_create(ctypes.byref(p), (ctypes.c_int * 3)(0, 1, 1), 32, 0) == mcs.ERR_INVALID
    assert lib.mcs_fft_plan_create(ctypes.byref(p), (ctypes.c_int * 3)(8, 1, 1), 16, 0) == mcs.ERR_INVALID
    assert lib.mcs_fft_plan_create(None, (ctypes.c_int * 3)(8, 1, 1), 32, 0) == mcs.ERR_INVALID
    assert lib.mcs_psi3_split(None, None, 0.0, 0, (ctypes.c_int * 3)(8, 1, 1), 32) == mcs.ERR_INVALID
    assert lib.mcs_bond_order_knn(None, 0, 8, 1.0, 40, None, None, None, None, 0) == mcs.ERR_INVALID
    assert b"bad argument" in lib.mcs_last_error()
