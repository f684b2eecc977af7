"""Self-checks of the kernels' arithmetic building blocks on the GPU.

The kernels take the square root of every flux and norm through sqrt_rn, a
correctly rounded float square root without the scaling and special-value
patches of sqrtf's full expansion (those lanes fall back to sqrtf).  Every
result of the solvers is pinned bit for bit to the oracle, which computes
sqrtf on the CPU; this test pins sqrt_rn to the device's sqrtf on all 2^32
float bit patterns, so no input the solvers can produce is left unchecked.
"""
import ctypes

import pytest

pytestmark = pytest.mark.gpu


def test_sqrt_rn_equals_sqrtf_on_every_float():
    from ccpi.filters.utils import require_gpu_lib
    lib = require_gpu_lib()
    f = lib.cilreg_check_sqrt_rn
    f.argtypes = [ctypes.POINTER(ctypes.c_ulonglong)] * 2
    f.restype = ctypes.c_int
    bad, seen = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    assert f(ctypes.byref(bad), ctypes.byref(seen)) == 0
    assert seen.value == 1 << 32
    assert bad.value == 0, f"{bad.value} floats whose sqrt_rn differs from sqrtf"
