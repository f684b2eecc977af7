"""Loads the regulariser libraries from this package's directory.

Mirrors src/Python/ccpi/filters/utils.py of the reference: ``cilregcuda`` is
this package's MI355X library (libcilregcuda.so, built by
``make -C ccpi-regularisation-toolkit_amd``).  ``cilreg`` is the reference's
C-OMP CPU library; it is not part of this package and is loaded only when a
user has placed their own ``libcilreg.so`` beside this file (otherwise the
``device='cpu'`` functions raise).
"""
import ctypes
import os
import platform
import warnings

try:
    pre = {'Linux': 'lib', 'Windows': '', 'Darwin': 'lib'}[platform.system()]
except KeyError:
    raise ValueError(f"unsupported platform: {platform.system()}")
ext = {'Linux': '.so', 'Windows': '.dll', 'Darwin': '.dylib'}[platform.system()]

_here = os.path.dirname(os.path.abspath(__file__))
GPU_LIB_PATH = os.path.join(_here, f'{pre}cilregcuda{ext}')
CPU_LIB_PATH = os.path.join(_here, f'{pre}cilreg{ext}')


def _load(path):
    try:
        return ctypes.cdll.LoadLibrary(path)
    except OSError as exc:
        warnings.warn(f"{os.path.basename(path)} not loaded: {exc}", ImportWarning, stacklevel=2)
        return None


cilregcuda = _load(GPU_LIB_PATH) if os.path.exists(GPU_LIB_PATH) else None
cilreg = _load(CPU_LIB_PATH) if os.path.exists(CPU_LIB_PATH) else None


def require_gpu_lib():
    """The GPU library, or a RuntimeError that says how to build it."""
    if cilregcuda is None:
        raise RuntimeError(
            f"libcilregcuda.so is not built or failed to load ({GPU_LIB_PATH}); "
            "run `make -C ccpi-regularisation-toolkit_amd` (or __graft_entry__.build())")
    return cilregcuda


def last_error():
    lib = require_gpu_lib()
    lib.cilreg_last_error.restype = ctypes.c_char_p
    msg = lib.cilreg_last_error()
    return msg.decode() if msg else ""
