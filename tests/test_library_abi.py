"""CPU-side checks of the product library: it loads and exports every entry point
declared in include/*.h (no compute call -- there may be no GPU here)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "accrqa_amd", "libaccrqa_amd.so")

# Mangled names of the reference's C++ API (include/AccRQA_functions.hpp:77-362),
# which include/AccRQA_amd.hpp redeclares.
CXX_SYMBOLS = [
    "_Z9accrqa_RRPfS_mPiiS0_iS_i15Accrqa_Distance19Accrqa_CompPlatformP12Accrqa_Error",
    "_Z9accrqa_RRPdS_mPiiS0_iS_i15Accrqa_Distance19Accrqa_CompPlatformP12Accrqa_Error",
    "_Z10accrqa_DETPfS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
    "_Z10accrqa_DETPdS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
    "_Z10accrqa_LAMPfS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
    "_Z10accrqa_LAMPdS_mPiiS0_iS0_iS_i15Accrqa_Distancei19Accrqa_CompPlatformP12Accrqa_Error",
    "_Z34accrqa_DET_output_size_in_elementsiiii",
    "_Z34accrqa_LAM_output_size_in_elementsiiii",
    "_Z33accrqa_RR_output_size_in_elementsiii",
    "_Z18accrqa_print_errorP12Accrqa_Error",
    "_Z9accrqa_RPPcPfmiif15Accrqa_DistanceP12Accrqa_Error",
    "_Z9accrqa_RPPcPdmiid15Accrqa_DistanceP12Accrqa_Error",
    "_Z9accrqa_RPPiPdmiid15Accrqa_DistanceP12Accrqa_Error",
    "_Z33accrqa_RP_output_size_in_elementslll",
]


def c_abi_names():
    src = open(os.path.join(ROOT, "include", "accrqa_amd.h")).read()
    return sorted(set(re.findall(r"\b(accrqa_amd_\w+)\s*\(", src)))


def test_library_built():
    assert os.path.exists(LIB), "libaccrqa_amd.so missing: run __graft_entry__.build()"


def test_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    names = c_abi_names()
    assert len(names) >= 12
    for n in names + CXX_SYMBOLS:
        assert hasattr(lib, n), n


def test_dynamic_symbol_table_has_cxx_api():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], stdout=subprocess.PIPE, text=True).stdout
    for n in CXX_SYMBOLS:
        assert n in out, n


def test_output_size_helpers_without_gpu():
    lib = C.CDLL(LIB)
    f = getattr(lib, "_Z34accrqa_DET_output_size_in_elementsiiii")
    f.restype = C.c_int
    assert f(2, 3, 4, 5) == 2 * 3 * 4 * 5 * 5
    g = getattr(lib, "_Z33accrqa_RR_output_size_in_elementsiii")
    g.restype = C.c_int
    assert g(2, 3, 4) == 24


def test_python_layer_imports_and_raises_on_bad_dtype():
    import numpy as np

    import accrqa_amd
    with pytest.raises(accrqa_amd.AccrqaError):
        accrqa_amd.RR(np.zeros(10, dtype=np.int32), [1], [1], [0.5])


def test_argument_errors_before_device_use():
    """Count / NULL checks happen before any device call (reference order:
    counts -> 2, NULL arrays -> 4), so they are testable without a GPU."""
    import numpy as np

    import accrqa_amd
    L = accrqa_amd.load()
    x = np.ones(10, dtype=np.float32)
    out = np.zeros(5, dtype=np.float32)
    tau = np.array([1], dtype=np.int32)
    thr = np.array([0.5], dtype=np.float32)
    err = C.c_int(-1)
    L.accrqa_amd_RR_f32(out.ctypes.data, x.ctypes.data, 10, tau.ctypes.data, 0, tau.ctypes.data, 1,
                        thr.ctypes.data, 1, 1, C.byref(err))
    assert err.value == 2
    L.accrqa_amd_RR_f32(out.ctypes.data, None, 10, tau.ctypes.data, 1, tau.ctypes.data, 1, thr.ctypes.data, 1, 1,
                        C.byref(err))
    assert err.value == 4
