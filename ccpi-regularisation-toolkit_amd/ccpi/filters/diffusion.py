"""ctypes wrappers of the diffusion regularisers: NDF and Diff4th.

Same function names, argument order and defaults as the reference's
src/Python/ccpi/filters/diffusion.py (CPU wrappers :9-87, GPU wrappers
:89-171), with the conventions of ``TV.py`` in this package (float32
C-contiguous inputs, RuntimeError on a -1 from the library).

NDF on the GPU computes the synchronous explicit step: every neighbour comes
from the previous iterate.  The reference's CPU and CUDA loops update the
image in place while other threads read it, so their results depend on the
thread schedule; on the reference's own test problem the two stay within the
1e-5 per-pixel tolerance its CPU-vs-GPU test applies.
"""
from .TV import _f, _fp, _i, _check, _dims, _f32, _gpu_device, _info, _out, _p, _require_cpu_lib
from .utils import require_gpu_lib


def NDF_GPU(inputData, lambdaPar, sigmaPar, iterationsNumb, tau, penaltytype, epsil, gpu_device, out=None,
            infovector=None):
    lib = require_gpu_lib()
    lib.NonlDiff_GPU_main.argtypes = [_fp, _fp, _fp, _f, _f, _i, _f, _i, _f, _i, _i, _i, _i]
    lib.NonlDiff_GPU_main.restype = _i
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    rc = lib.NonlDiff_GPU_main(_p(x), _p(out), _p(infovector), lambdaPar, sigmaPar, iterationsNumb, tau, penaltytype,
                               epsil, _gpu_device(gpu_device), *_dims(x))
    _check(rc, "NonlDiff_GPU_main")
    return out


def Diffus4th_GPU(inputData, lambdaPar, sigmaPar, iterationsNumb, tau, epsil, gpu_device, out=None, infovector=None):
    lib = require_gpu_lib()
    lib.Diffus4th_GPU_main.argtypes = [_fp, _fp, _fp, _f, _f, _i, _f, _f, _i, _i, _i, _i]
    lib.Diffus4th_GPU_main.restype = _i
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    rc = lib.Diffus4th_GPU_main(_p(x), _p(out), _p(infovector), lambdaPar, sigmaPar, iterationsNumb, tau, epsil,
                                _gpu_device(gpu_device), *_dims(x))
    _check(rc, "Diffus4th_GPU_main")
    return out


def NDF_CPU(inputData, lambdaPar, sigmaPar, iterationsNumb, tau, penaltytype, epsil, out=None, infovector=None):
    lib = _require_cpu_lib()
    lib.Diffusion_CPU_main.argtypes = [_fp, _fp, _fp, _f, _f, _i, _f, _i, _f, _i, _i, _i]
    lib.Diffusion_CPU_main.restype = _f
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    lib.Diffusion_CPU_main(_p(x), _p(out), _p(infovector), lambdaPar, sigmaPar, iterationsNumb, tau, penaltytype,
                           epsil, *_dims(x))
    return out


def Diffus4th_CPU(inputData, lambdaPar, sigmaPar, iterationsNumb, tau, epsil, out=None, infovector=None):
    lib = _require_cpu_lib()
    lib.Diffus4th_CPU_main.argtypes = [_fp, _fp, _fp, _f, _f, _i, _f, _f, _i, _i, _i]
    lib.Diffus4th_CPU_main.restype = _f
    x = _f32(inputData)
    out = _out(x, out)
    infovector = _info(infovector)
    lib.Diffus4th_CPU_main(_p(x), _p(out), _p(infovector), lambdaPar, sigmaPar, iterationsNumb, tau, epsil, *_dims(x))
    return out
