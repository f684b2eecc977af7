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


@pytest.mark.parametrize("eps", [1e-8, 1e-12], ids=["rof", "llt"])
def test_rof_sqrt_d_equals_double_sqrt_on_every_float(eps):
    """The 3D D1 flux normaliser (float)sqrt((double)s + eps) of ROF (eps 1e-8)
    and LLT-ROF (1e-12) skips sqrt's scaling and special-value patches for
    in-range arguments; pinned to the device's sqrt for every non-negative
    float s."""
    from ccpi.filters.utils import require_gpu_lib
    lib = require_gpu_lib()
    f = lib.cilreg_check_rof_sqrt_d
    f.argtypes = [ctypes.c_double] + [ctypes.POINTER(ctypes.c_ulonglong)] * 2
    f.restype = ctypes.c_int
    bad, seen = ctypes.c_ulonglong(), ctypes.c_ulonglong()
    assert f(eps, ctypes.byref(bad), ctypes.byref(seen)) == 0
    assert seen.value == 1 << 31
    assert bad.value == 0, f"{bad.value} floats whose flux normaliser differs from the double sqrt"
