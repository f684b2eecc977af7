"""User API: ``ROF_TV, FGP_TV, PD_TV, SB_TV, TGV, FGP_dTV, LLT_ROF, NDF, Diff4th``
with ``device=`` dispatch.

Mirrors src/Python/ccpi/filters/regularisers.py:19-49 (FGP_dTV :44, LLT_ROF :42, NDF :46, Diff4th :47) of the reference:
``device='cpu'`` calls the C-OMP function, ``'gpu'`` or an integer device
index calls the MI355X function with ``gpu_device`` set, anything else raises
KeyError.  Positional argument order per method:

* ``ROF_TV(x, lambda, iterations, tau, tol)``
* ``FGP_TV(x, lambda, iterations, tol, methodTV, nonneg)``
* ``PD_TV(x, lambda, iterations, tol, lipschitz_const, methodTV, nonneg)``
* ``SB_TV(x, mu, iterations, tol, methodTV)``
* ``TGV(x, lambda, alpha1, alpha0, iterations, L2, tol)``
* ``FGP_dTV(x, x_ref, lambda, iterations, tol, eta, methodTV, nonneg)``
* ``LLT_ROF(x, lambdaROF, lambdaLLT, iterations, tau, tol)``
* ``NDF(x, lambda, sigma, iterations, tau, penalty_type, tol)``
* ``Diff4th(x, lambda, sigma, iterations, tau, tol)``

each also taking ``out=`` and ``infovector=`` (``Output=`` for PD on the GPU,
as in the reference, TV.py:567).
"""
import numbers
from functools import wraps

from .TV import TV_ROF_CPU, TV_ROF_GPU
from .TV import TV_FGP_CPU, TV_FGP_GPU
from .TV import PDTV_CPU, PDTV_GPU
from .TV import SB_TV_CPU, SB_TV_GPU
from .TV import TGV_CPU, TGV_GPU
from .TV import dTV_FGP_CPU, dTV_FGP_GPU
from .TV import LLT_ROF_CPU, LLT_ROF_GPU
from .diffusion import NDF_CPU, NDF_GPU
from .diffusion import Diffus4th_CPU, Diffus4th_GPU
from .utils import cilregcuda


def _gpu_index(device):
    """The GPU index a ``device`` argument selects, or None if it selects none:
    'gpu' is device 0 (even when the library did not load: the GPU function
    then reports that), an integer is that device provided the library is
    loaded (regularisers.py:19-24 of the reference)."""
    if device == 'gpu':
        return 0
    if isinstance(device, numbers.Integral) and cilregcuda is not None:
        return device
    return None


def create_wrapper(CPU_func, GPU_func):
    """The user function over a CPU and a GPU implementation: ``device='cpu'``
    (the default) runs CPU_func, a device _gpu_index accepts runs GPU_func with
    ``gpu_device`` set, any other value raises KeyError."""
    def dispatch(*args, device='cpu', **kwargs):
        if device == 'cpu':
            return CPU_func(*args, **kwargs)
        index = _gpu_index(device)
        if index is None:
            raise KeyError(f"{GPU_func.__name__}: device {device} not available")
        return GPU_func(*args, gpu_device=index, **kwargs)
    return wraps(CPU_func)(dispatch)


ROF_TV = create_wrapper(TV_ROF_CPU, TV_ROF_GPU)
FGP_TV = create_wrapper(TV_FGP_CPU, TV_FGP_GPU)
PD_TV = create_wrapper(PDTV_CPU, PDTV_GPU)
SB_TV = create_wrapper(SB_TV_CPU, SB_TV_GPU)
TGV = create_wrapper(TGV_CPU, TGV_GPU)
FGP_dTV = create_wrapper(dTV_FGP_CPU, dTV_FGP_GPU)
LLT_ROF = create_wrapper(LLT_ROF_CPU, LLT_ROF_GPU)
NDF = create_wrapper(NDF_CPU, NDF_GPU)
Diff4th = create_wrapper(Diffus4th_CPU, Diffus4th_GPU)
