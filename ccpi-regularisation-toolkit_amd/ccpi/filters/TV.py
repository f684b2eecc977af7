"""ctypes wrappers of the TV-family regularisers.

Same function names, argument order and defaults as the reference's
src/Python/ccpi/filters/TV.py (GPU wrappers :477-771, CPU wrappers :8-292):
numpy arrays in (Z, Y, X) or (Y, X) order, ``dims = shape[::-1]`` (+1 for 2D),
``out`` and ``infovector`` optional and filled in place, the output returned.

Deliberate differences:
* inputs are converted to C-contiguous float32 (the reference passes the raw
  buffer of whatever it is given);
* a -1 from the library raises RuntimeError with the library's message (the
  reference ignores the return code and hands back zeros);
* TGV passes ``shape[::-1]`` like every other method (the reference reverses
  twice, TV.py:722-725, which its CUDA kernels then index transposed).

The ``*_CPU`` functions call the reference's own C-OMP library when the user
has put ``libcilreg.so`` beside this package; this package does not ship one.
"""
import ctypes

import numpy as np

from .utils import cilreg, cilregcuda, require_gpu_lib, last_error

_fp = ctypes.POINTER(ctypes.c_float)
_i = ctypes.c_int
_f = ctypes.c_float


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _dims(a):
    d = list(a.shape)[::-1]
    if a.ndim == 2:
        d.append(1)
    if len(d) != 3:
        raise ValueError(f"only 2D and 3D arrays are supported, got shape {a.shape}")
    return d


def _out(x, out, ones_init=False):
    if out is None:
        return np.zeros_like(x)
    if out.shape != x.shape or out.dtype != np.float32 or not out.flags.c_contiguous:
        raise ValueError("out must be a C-contiguous float32 array of the input's shape")
    return out


def _info(infovector):
    if infovector is None:
        return np.zeros((2,), dtype='float32')
    if infovector.dtype != np.float32 or infovector.size < 2 or not infovector.flags.c_contiguous:
        raise ValueError("infovector must be a contiguous float32 array of at least 2 elements")
    return infovector


def _p(a):
    return a.ctypes.data_as(_fp)


def _gpu_device(gpu_device):
    return 0 if gpu_device == 'gpu' else int(gpu_device)


def _check(rc, name):
    if rc != 0:
        raise RuntimeError(f"{name} failed: {last_error()}")


# ---------------------------------------------------------------------------
# GPU (MI355X) wrappers
# ---------------------------------------------------------------------------
def TV_ROF_GPU(inputData, regularisation_parameter, iterationsNumb, marching_step_parameter, tolerance_param,
               gpu_device, out=None, infovector=None):
    lib = require_gpu_lib()
    lib.TV_ROF_GPU_main.argtypes = [_fp, _fp, _fp, _f, _i, _f, _f, _i, _i, _i, _i]
    lib.TV_ROF_GPU_main.restype = _i
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    d = _dims(x)
    rc = lib.TV_ROF_GPU_main(_p(x), _p(out), _p(infovector), regularisation_parameter, iterationsNumb,
                             marching_step_parameter, tolerance_param, _gpu_device(gpu_device), *d)
    _check(rc, "TV_ROF_GPU_main")
    return out


def TV_FGP_GPU(inputData, lambdaPar, iterationsNumb, epsil, methodTV, nonneg, gpu_device, out=None, infovector=None):
    lib = require_gpu_lib()
    lib.TV_FGP_GPU_main.argtypes = [_fp, _fp, _fp, _f, _i, _f, _i, _i, _i, _i, _i, _i]
    lib.TV_FGP_GPU_main.restype = _i
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    d = _dims(x)
    rc = lib.TV_FGP_GPU_main(_p(x), _p(out), _p(infovector), lambdaPar, iterationsNumb, epsil, methodTV, nonneg,
                             _gpu_device(gpu_device), *d)
    _check(rc, "TV_FGP_GPU_main")
    return out


def PDTV_GPU(Input, lambdaPar, iter, epsil, lipschitz_const, methodTV, nonneg, gpu_device, Output=None,
             infovector=None):
    lib = require_gpu_lib()
    lib.TV_PD_GPU_main.argtypes = [_fp, _fp, _fp, _f, _i, _f, _f, _i, _i, _i, _i, _i, _i]
    lib.TV_PD_GPU_main.restype = _i
    x = _f32(Input)
    Output = _out(x, Output)
    infovector = _info(infovector)
    d = _dims(x)
    rc = lib.TV_PD_GPU_main(_p(x), _p(Output), _p(infovector), lambdaPar, iter, epsil, lipschitz_const, methodTV,
                            nonneg, _gpu_device(gpu_device), *d)
    _check(rc, "TV_PD_GPU_main")
    return Output


def SB_TV_GPU(inputData, lambdaPar, iterationsNumb, epsil, methodTV, gpu_device, out=None, infovector=None):
    lib = require_gpu_lib()
    lib.TV_SB_GPU_main.argtypes = [_fp, _fp, _fp, _f, _i, _f, _i, _i, _i, _i, _i]
    lib.TV_SB_GPU_main.restype = _i
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    d = _dims(x)
    rc = lib.TV_SB_GPU_main(_p(x), _p(out), _p(infovector), lambdaPar, iterationsNumb, epsil, methodTV,
                            _gpu_device(gpu_device), *d)
    _check(rc, "TV_SB_GPU_main")
    return out


def TGV_GPU(inputData, lambdaPar, alpha1, alpha0, iterationsNumb, L2, epsil, gpu_device, out=None, infovector=None):
    lib = require_gpu_lib()
    lib.TGV_GPU_main.argtypes = [_fp, _fp, _fp, _f, _f, _f, _i, _f, _f, _i, _i, _i, _i]
    lib.TGV_GPU_main.restype = _i
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    d = _dims(x)
    rc = lib.TGV_GPU_main(_p(x), _p(out), _p(infovector), lambdaPar, alpha1, alpha0, iterationsNumb, L2, epsil,
                          _gpu_device(gpu_device), *d)
    _check(rc, "TGV_GPU_main")
    return out


def LLT_ROF_GPU(inputData, lambdaROF, lambdaLLT, iterationsNumb, tau, epsil, gpu_device, out=None, infovector=None):
    lib = require_gpu_lib()
    lib.LLT_ROF_GPU_main.argtypes = [_fp, _fp, _fp, _f, _f, _i, _f, _f, _i, _i, _i, _i]
    lib.LLT_ROF_GPU_main.restype = _i
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    rc = lib.LLT_ROF_GPU_main(_p(x), _p(out), _p(infovector), lambdaROF, lambdaLLT, iterationsNumb, tau, epsil,
                              _gpu_device(gpu_device), *_dims(x))
    _check(rc, "LLT_ROF_GPU_main")
    return out


def _ref_image(x, inputRef):
    r = _f32(inputRef)
    if r.shape != x.shape:
        raise ValueError(f"inputRef shape {r.shape} differs from the input's {x.shape}")
    return r


def dTV_FGP_GPU(inputData, inputRef, lambdaPar, iterationsNumb, epsil, eta, methodTV, nonneg, gpu_device, out=None,
                infovector=None):
    lib = require_gpu_lib()
    lib.dTV_FGP_GPU_main.argtypes = [_fp, _fp, _fp, _fp, _f, _i, _f, _f, _i, _i, _i, _i, _i, _i]
    lib.dTV_FGP_GPU_main.restype = _i
    x = _f32(inputData)
    r = _ref_image(x, inputRef)
    out = _out(x, out)
    infovector = _info(infovector)
    d = _dims(x)
    rc = lib.dTV_FGP_GPU_main(_p(x), _p(r), _p(out), _p(infovector), lambdaPar, iterationsNumb, epsil, eta, methodTV,
                              nonneg, _gpu_device(gpu_device), *d)
    _check(rc, "dTV_FGP_GPU_main")
    return out


# ---------------------------------------------------------------------------
# CPU wrappers: the reference's own C-OMP library, when the user supplies it
# ---------------------------------------------------------------------------
def _require_cpu_lib():
    if cilreg is None:
        raise RuntimeError("device='cpu' needs the reference C-OMP library libcilreg.so beside this package; "
                           "this package provides the MI355X path only (device='gpu' or an integer)")
    return cilreg


def TV_ROF_CPU(inputData, regularisation_parameter, iterationsNumb, marching_step_parameter, tolerance_param,
               out=None, infovector=None):
    lib = _require_cpu_lib()
    lib.TV_ROF_CPU_main.argtypes = [_fp, _fp, _fp, _fp, _i, _i, _f, _f, _i, _i, _i]
    lib.TV_ROF_CPU_main.restype = _f
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    lam = ctypes.c_float(regularisation_parameter)
    lib.TV_ROF_CPU_main(_p(x), _p(out), _p(infovector), ctypes.byref(lam), 0, iterationsNumb,
                        marching_step_parameter, tolerance_param, *_dims(x))
    return out


def TV_FGP_CPU(inputData, lambdaPar, iterationsNumb, epsil, methodTV, nonneg, out=None, infovector=None):
    lib = _require_cpu_lib()
    lib.TV_FGP_CPU_main.argtypes = [_fp, _fp, _fp, _f, _i, _f, _i, _i, _i, _i, _i]
    lib.TV_FGP_CPU_main.restype = _f
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    lib.TV_FGP_CPU_main(_p(x), _p(out), _p(infovector), lambdaPar, iterationsNumb, epsil, methodTV, nonneg,
                        *_dims(x))
    return out


def PDTV_CPU(inputData, lambdaPar, iterationsNumb, epsil, lipschitz_const, methodTV, nonneg, out=None,
             infovector=None):
    lib = _require_cpu_lib()
    lib.PDTV_CPU_main.argtypes = [_fp, _fp, _fp, _f, _i, _f, _f, _i, _i, _i, _i, _i]
    lib.PDTV_CPU_main.restype = _f
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    lib.PDTV_CPU_main(_p(x), _p(out), _p(infovector), lambdaPar, iterationsNumb, epsil, lipschitz_const, methodTV,
                      nonneg, *_dims(x))
    return out


def SB_TV_CPU(inputData, mu, iter, epsil, methodTV, out=None, infovector=None):
    lib = _require_cpu_lib()
    lib.SB_TV_CPU_main.argtypes = [_fp, _fp, _fp, _f, _i, _f, _i, _i, _i, _i]
    lib.SB_TV_CPU_main.restype = _f
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    lib.SB_TV_CPU_main(_p(x), _p(out), _p(infovector), mu, iter, epsil, methodTV, *_dims(x))
    return out


def TGV_CPU(inputData, lambdaPar, alpha1, alpha0, iterationsNumb, L2, epsil, out=None, infovector=None):
    lib = _require_cpu_lib()
    lib.TGV_main.argtypes = [_fp, _fp, _fp, _f, _f, _f, _i, _f, _f, _i, _i, _i]
    lib.TGV_main.restype = _f
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    lib.TGV_main(_p(x), _p(out), _p(infovector), lambdaPar, alpha1, alpha0, iterationsNumb, L2, epsil, *_dims(x))
    return out


def dTV_FGP_CPU(inputData, inputRef, lambdaPar, iterationsNumb, epsil, eta, methodTV, nonneg, out=None,
                infovector=None):
    lib = _require_cpu_lib()
    lib.dTV_FGP_CPU_main.argtypes = [_fp, _fp, _fp, _fp, _f, _i, _f, _f, _i, _i, _i, _i, _i]
    lib.dTV_FGP_CPU_main.restype = _f
    x = _f32(inputData)
    r = _ref_image(x, inputRef)
    out = _out(x, out)
    infovector = _info(infovector)
    lib.dTV_FGP_CPU_main(_p(x), _p(r), _p(out), _p(infovector), lambdaPar, iterationsNumb, epsil, eta, methodTV,
                         nonneg, *_dims(x))
    return out


def LLT_ROF_CPU(inputData, lambdaROF, lambdaLLT, iterationsNumb, tau, epsil, out=None, infovector=None):
    lib = _require_cpu_lib()
    lib.LLT_ROF_CPU_main.argtypes = [_fp, _fp, _fp, _f, _f, _i, _f, _f, _i, _i, _i]
    lib.LLT_ROF_CPU_main.restype = _f
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    lib.LLT_ROF_CPU_main(_p(x), _p(out), _p(infovector), lambdaROF, lambdaLLT, iterationsNumb, tau, epsil, *_dims(x))
    return out
