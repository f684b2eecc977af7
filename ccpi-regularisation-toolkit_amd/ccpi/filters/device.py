"""Device-resident solves on torch tensors (no host copies).

The reference offers this only through CuPy for ROF/PD
(src/Python/ccpi/filters/regularisersCuPy.py:35-256).  Here every method takes
a float32 (or float64 for the ``*_f64`` solvers) CUDA tensor in (Z, Y, X) or
(Y, X) order and returns ``(out, infovector)``; the work is queued on the
tensor device's current torch stream and the call returns when it is done.

Options beyond the reference (see include/cilreg_ext.h):
``check_every`` (residual-check cadence, reference 5), ``precision``
('fp32', 'bf16', 'fp16' storage of the dual/auxiliary fields), ``tile_y`` and
``zchunk`` (tiling; 0 = automatic).
"""
import ctypes
import os

import numpy as np

from .utils import require_gpu_lib, last_error

PRECISIONS = {'fp32': 0, 'bf16': 1, 'fp16': 2}


class CilregOpts(ctypes.Structure):
    _fields_ = [("check_every", ctypes.c_int), ("dual_precision", ctypes.c_int), ("tile_y", ctypes.c_int),
                ("zchunk", ctypes.c_int), ("profile", ctypes.c_int), ("wide_index", ctypes.c_int),
                ("stream_launches", ctypes.c_int), ("temporal_depth", ctypes.c_int), ("autotune", ctypes.c_int),
                ("persistent", ctypes.c_int)]


def make_opts(check_every=5, precision='fp32', tile_y=0, zchunk=0, profile=False, wide_index=False,
              stream_launches=False, depth=0, autotune=False, persistent=False):
    if precision not in PRECISIONS:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}")
    return CilregOpts(int(check_every), PRECISIONS[precision], int(tile_y), int(zchunk), int(bool(profile)),
                      int(bool(wide_index)), int(bool(stream_launches)), int(depth), int(bool(autotune)),
                      int(bool(persistent)))


def last_profile():
    """(mean_ms, min_ms, count) of the per-iteration kernel time of the last
    solve on this thread run with ``profile=True``."""
    lib = require_gpu_lib()
    lib.cilreg_last_profile.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_int)]
    lib.cilreg_last_profile.restype = ctypes.c_int
    mean, mn, cnt = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    lib.cilreg_last_profile(ctypes.byref(mean), ctypes.byref(mn), ctypes.byref(cnt))
    return mean.value, mn.value, cnt.value


def release_cache():
    """Free the device workspaces the library keeps cached on the current device."""
    lib = require_gpu_lib()
    lib.cilreg_release_cache.restype = ctypes.c_int
    return lib.cilreg_release_cache()


def stream_bench(nbytes=1 << 30, reps=20):
    """Achievable HBM GB/s of the current device from the library's streaming
    kernels: {'copy', 'read', 'write', 'sum2'} (sum2: 2 reads + 1 write)."""
    lib = require_gpu_lib()
    lib.cilreg_stream_bench.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    lib.cilreg_stream_bench.restype = ctypes.c_int
    out = (ctypes.c_double * 4)()
    if lib.cilreg_stream_bench(int(nbytes), int(reps), out) != 0:
        raise RuntimeError(f"cilreg_stream_bench failed: {last_error()}")
    return dict(zip(("copy", "read", "write", "sum2"), list(out)))


def autotune(method, shape, method_tv=0, **kw):
    """The autotuner's choice for `method` ('rof', 'fgp', 'pd', 'tgv') on a
    volume of `shape` ((Z, Y, X) or (Y, X)) with the given options (e.g.
    precision): dict(tile_y, zchunk, depth, ms_per_iter)."""
    lib = require_gpu_lib()
    lib.cilreg_autotune.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(CilregOpts), ctypes.POINTER(CilregOpts),
                                                         ctypes.POINTER(ctypes.c_double)]
    lib.cilreg_autotune.restype = ctypes.c_int
    m = {'rof': 0, 'fgp': 1, 'pd': 2, 'tgv': 3}[method]
    d = list(shape)[::-1] + ([1] if len(shape) == 2 else [])
    base = make_opts(**kw)
    best = CilregOpts()
    ms = ctypes.c_double()
    if lib.cilreg_autotune(m, d[0], d[1], d[2], int(method_tv), ctypes.byref(base), ctypes.byref(best),
                           ctypes.byref(ms)) != 0:
        raise RuntimeError(f"cilreg_autotune failed: {last_error()}")
    return {"tile_y": best.tile_y, "zchunk": best.zchunk, "depth": best.temporal_depth, "ms_per_iter": ms.value}


def _tune_lib():
    lib = require_gpu_lib()
    for name in ("cilreg_tune_save", "cilreg_tune_load"):
        getattr(lib, name).argtypes = [ctypes.c_char_p]
        getattr(lib, name).restype = ctypes.c_int
    lib.cilreg_tune_entries.argtypes = []
    lib.cilreg_tune_entries.restype = ctypes.c_int
    lib.cilreg_tune_clear.argtypes = []
    lib.cilreg_tune_clear.restype = None
    return lib


def tune_save(path):
    """Write the autotuner's cached choices to `path`; the number written."""
    n = _tune_lib().cilreg_tune_save(os.fsencode(path))
    if n < 0:
        raise RuntimeError(f"cilreg_tune_save failed: {last_error()}")
    return n


def tune_load(path):
    """Merge an autotuner cache file into this process's choices; the number
    read.  A file with a malformed entry is rejected whole (RuntimeError)."""
    n = _tune_lib().cilreg_tune_load(os.fsencode(path))
    if n < 0:
        raise RuntimeError(f"cilreg_tune_load failed: {last_error()}")
    return n


def tune_entries():
    """Number of autotuner choices cached in this process."""
    return _tune_lib().cilreg_tune_entries()


def tune_clear():
    """Forget this process's autotuner choices (cache files are untouched)."""
    _tune_lib().cilreg_tune_clear()


_vp = ctypes.c_void_p
_i = ctypes.c_int
_f = ctypes.c_float
_d = ctypes.c_double
_SIGS = {
    "cilreg_rof_dev": [_vp, _vp, _vp, _f, _i, _f, _f, _i, _i, _i, _vp, _vp],
    "cilreg_fgp_dev": [_vp, _vp, _vp, _f, _i, _f, _i, _i, _i, _i, _i, _vp, _vp],
    "cilreg_pd_dev": [_vp, _vp, _vp, _f, _i, _f, _f, _i, _i, _i, _i, _i, _vp, _vp],
    "cilreg_sb_dev": [_vp, _vp, _vp, _f, _i, _f, _i, _i, _i, _i, _vp, _vp],
    "cilreg_tgv_dev": [_vp, _vp, _vp, _f, _f, _f, _i, _f, _f, _i, _i, _i, _vp, _vp],
    "cilreg_llt_dev": [_vp, _vp, _vp, _f, _f, _i, _f, _f, _i, _i, _i, _vp, _vp],
    "cilreg_ndf_dev": [_vp, _vp, _vp, _f, _f, _i, _f, _i, _f, _i, _i, _i, _vp, _vp],
    "cilreg_d4_dev": [_vp, _vp, _vp, _f, _f, _i, _f, _f, _i, _i, _i, _vp, _vp],
    "cilreg_dtv_dev": [_vp, _vp, _vp, _vp, _f, _i, _f, _f, _i, _i, _i, _i, _i, _vp, _vp],
    "cilreg_fgp_dev_f64": [_vp, _vp, _vp, _d, _i, _d, _i, _i, _i, _i, _i, _vp, _vp],
    "cilreg_pd_dev_f64": [_vp, _vp, _vp, _d, _i, _d, _d, _i, _i, _i, _i, _i, _vp, _vp],
    "cilreg_tgv_dev_f64": [_vp, _vp, _vp, _d, _d, _d, _i, _d, _d, _i, _i, _i, _vp, _vp],
    "cilreg_fgp_dev_ws": [_vp, _vp, _vp, _f, _i, _f, _i, _i, _i, _i, _i, _vp, _vp, _i, _vp, _vp],
    "cilreg_pd_dev_ws": [_vp, _vp, _vp, _f, _i, _f, _f, _i, _i, _i, _i, _i, _vp, _i, _vp, _vp],
    "cilreg_tgv_dev_ws": [_vp, _vp, _vp, _f, _f, _f, _i, _f, _f, _i, _i, _i, _vp, _i, _vp, _vp],
    "cilreg_sb_dev_ws": [_vp, _vp, _vp, _f, _i, _f, _i, _i, _i, _i, _vp, _i, _vp, _vp],
}


def _fn(name):
    lib = require_gpu_lib()
    f = getattr(lib, name)
    f.argtypes = _SIGS[name]
    f.restype = _i
    return f


def _dims(t):
    d = list(t.shape)[::-1]
    if t.dim() == 2:
        d.append(1)
    if len(d) != 3:
        raise ValueError(f"only 2D and 3D tensors are supported, got shape {tuple(t.shape)}")
    return d


def _prep(x, out, dtype):
    import torch
    if not (x.is_cuda and x.dtype == dtype and x.is_contiguous()):
        raise ValueError(f"input must be a contiguous {dtype} CUDA tensor")
    if out is None:
        out = torch.zeros_like(x)
    elif not (out.is_cuda and out.dtype == dtype and out.is_contiguous() and out.shape == x.shape
              and out.device == x.device):
        raise ValueError("out must be a contiguous tensor of the input's shape, dtype and device")
    return out


def _run(name, x, out, dtype, scalars, opts):
    import torch
    info = np.zeros(2, dtype=np.float64 if dtype == torch.float64 else np.float32)
    f = _fn(name)
    with torch.cuda.device(x.device):
        stream = torch.cuda.current_stream(x.device).cuda_stream
        rc = f(x.data_ptr(), out.data_ptr(), info.ctypes.data, *scalars, *_dims(x), ctypes.byref(opts),
               ctypes.c_void_p(stream))
    if rc != 0:
        raise RuntimeError(f"{name} failed: {last_error()}")
    return out, info


def rof(x, lam, iters, tau, tol=0.0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float32)
    return _run("cilreg_rof_dev", x, out, torch.float32, [lam, iters, tau, tol], make_opts(**kw))


def fgp(x, lam, iters, tol=0.0, method=0, nonneg=0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float32)
    return _run("cilreg_fgp_dev", x, out, torch.float32, [lam, iters, tol, method, nonneg], make_opts(**kw))


def pd(x, lam, iters, tol=0.0, L=12.0, method=0, nonneg=0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float32)
    return _run("cilreg_pd_dev", x, out, torch.float32, [lam, iters, tol, L, method, nonneg], make_opts(**kw))


def sb(x, mu, iters, tol=0.0, method=0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float32)
    return _run("cilreg_sb_dev", x, out, torch.float32, [mu, iters, tol, method], make_opts(**kw))


def tgv(x, lam, alpha1, alpha0, iters, L2=12.0, tol=0.0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float32)
    return _run("cilreg_tgv_dev", x, out, torch.float32, [lam, alpha1, alpha0, iters, L2, tol], make_opts(**kw))


def llt(x, lam_rof, lam_llt, iters, tau, tol=0.0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float32)
    return _run("cilreg_llt_dev", x, out, torch.float32, [lam_rof, lam_llt, iters, tau, tol], make_opts(**kw))


def ndf(x, lam, sigma, iters, tau, penalty=1, tol=0.0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float32)
    return _run("cilreg_ndf_dev", x, out, torch.float32, [lam, sigma, iters, tau, penalty, tol], make_opts(**kw))


def d4(x, lam, sigma, iters, tau, tol=0.0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float32)
    return _run("cilreg_d4_dev", x, out, torch.float32, [lam, sigma, iters, tau, tol], make_opts(**kw))


def dtv(x, ref, lam, iters, tol=0.0, eta=0.01, method=0, nonneg=0, out=None, **kw):
    """FGP-dTV guided by the reference image ``ref`` (a tensor of x's shape, dtype and device)."""
    import torch
    out = _prep(x, out, torch.float32)
    if not (ref.is_cuda and ref.dtype == torch.float32 and ref.is_contiguous() and ref.shape == x.shape
            and ref.device == x.device):
        raise ValueError("ref must be a contiguous float32 tensor of the input's shape and device")
    info = np.zeros(2, dtype=np.float32)
    f = _fn("cilreg_dtv_dev")
    with torch.cuda.device(x.device):
        rc = f(x.data_ptr(), ref.data_ptr(), out.data_ptr(), info.ctypes.data, lam, iters, tol, eta, method, nonneg,
               *_dims(x), ctypes.byref(make_opts(**kw)),
               ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream))
    if rc != 0:
        raise RuntimeError(f"cilreg_dtv_dev failed: {last_error()}")
    return out, info


def fgp_f64(x, lam, iters, tol=0.0, method=0, nonneg=0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float64)
    return _run("cilreg_fgp_dev_f64", x, out, torch.float64, [lam, iters, tol, method, nonneg], make_opts(**kw))


def pd_f64(x, lam, iters, tol=0.0, L=12.0, method=0, nonneg=0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float64)
    return _run("cilreg_pd_dev_f64", x, out, torch.float64, [lam, iters, tol, L, method, nonneg], make_opts(**kw))


def tgv_f64(x, lam, alpha1, alpha0, iters, L2=12.0, tol=0.0, out=None, **kw):
    import torch
    out = _prep(x, out, torch.float64)
    return _run("cilreg_tgv_dev_f64", x, out, torch.float64, [lam, alpha1, alpha0, iters, L2, tol],
                make_opts(**kw))


def tv_energy(u, u0, lam, type=1):
    """The ROF objective in FP64 (TV_energy2D/3D of the reference's utils.c:64-117
    summed in double): type 1 = sum 2 lam |grad+ u| + sum (u - u0)^2, type 2 =
    the gradient term alone.  u, u0: contiguous CUDA tensors of one shape,
    both float32 or both float64."""
    import torch
    if not (u.is_cuda and u0.is_cuda and u.shape == u0.shape and u.dtype == u0.dtype and u.is_contiguous()
            and u0.is_contiguous() and u.dtype in (torch.float32, torch.float64) and u.device == u0.device):
        raise ValueError("u and u0 must be contiguous CUDA tensors of one shape, device and dtype (float32/float64)")
    lib = require_gpu_lib()
    name = "cilreg_tv_energy_dev" if u.dtype == torch.float32 else "cilreg_tv_energy_dev_f64"
    f = getattr(lib, name)
    f.argtypes = [_vp, _vp, _d, _i, _i, _i, _i, ctypes.POINTER(ctypes.c_double), _vp]
    f.restype = _i
    e = ctypes.c_double()
    with torch.cuda.device(u.device):
        rc = f(u.data_ptr(), u0.data_ptr(), float(lam), int(type), *_dims(u), ctypes.byref(e),
               ctypes.c_void_p(torch.cuda.current_stream(u.device).cuda_stream))
    if rc != 0:
        raise RuntimeError(f"{name} failed: {last_error()}")
    return e.value


# ---------------------------------------------------------------------------
# warm start (include/cilreg_ext.h: cilreg_*_dev_ws)
# ---------------------------------------------------------------------------
STATE_CHANNELS = {"fgp": 6, "pd": 4, "tgv": 13, "sb": 7}


def new_state(method, x):
    """Zeroed iterate state for `method` ('fgp', 'pd', 'tgv') on volumes like x
    (FP32 channels of x's shape); for FGP also the host pair (t_k, next
    multiplier) as a float32 array."""
    import torch
    st = torch.zeros((STATE_CHANNELS[method],) + tuple(x.shape), dtype=torch.float32, device=x.device)
    if method == "fgp":
        return st, np.array([1.0, 0.0], dtype=np.float32)
    return st


def _check_state(st, method, x):
    import torch
    if not (st.is_cuda and st.dtype == torch.float32 and st.is_contiguous()
            and tuple(st.shape) == (STATE_CHANNELS[method],) + tuple(x.shape)):
        raise ValueError(f"state must be a contiguous float32 CUDA tensor of shape "
                         f"{(STATE_CHANNELS[method],) + tuple(x.shape)}")


def fgp_warm(x, lam, iters, state, tstate, warm=True, tol=0.0, method=0, nonneg=0, out=None, **kw):
    """FGP starting from (warm) and returning the iterate state in `state`
    (P, P_prev) and `tstate` (t_k, next multiplier; updated in place)."""
    import torch
    out = _prep(x, out, torch.float32)
    _check_state(state, "fgp", x)
    ts = np.ascontiguousarray(tstate, dtype=np.float32)
    info = np.zeros(2, dtype=np.float32)
    f = _fn("cilreg_fgp_dev_ws")
    with torch.cuda.device(x.device):
        rc = f(x.data_ptr(), out.data_ptr(), info.ctypes.data, lam, iters, tol, method, nonneg, *_dims(x),
               state.data_ptr(), ts.ctypes.data, int(bool(warm)), ctypes.byref(make_opts(**kw)),
               ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream))
    if rc != 0:
        raise RuntimeError(f"cilreg_fgp_dev_ws failed: {last_error()}")
    tstate[...] = ts
    return out, info


def pd_warm(x, lam, iters, state, warm=True, tol=0.0, L=12.0, method=0, nonneg=0, out=None, **kw):
    """PD starting from (warm) and returning the iterate state (u_bar, P) in `state`."""
    import torch
    out = _prep(x, out, torch.float32)
    _check_state(state, "pd", x)
    info = np.zeros(2, dtype=np.float32)
    f = _fn("cilreg_pd_dev_ws")
    with torch.cuda.device(x.device):
        rc = f(x.data_ptr(), out.data_ptr(), info.ctypes.data, lam, iters, tol, L, method, nonneg, *_dims(x),
               state.data_ptr(), int(bool(warm)), ctypes.byref(make_opts(**kw)),
               ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream))
    if rc != 0:
        raise RuntimeError(f"cilreg_pd_dev_ws failed: {last_error()}")
    return out, info


def tgv_warm(x, lam, alpha1, alpha0, iters, state, warm=True, L2=12.0, tol=0.0, out=None, **kw):
    """TGV starting from (warm) and returning the iterate state (u_bar, V, P, Q) in `state`."""
    import torch
    out = _prep(x, out, torch.float32)
    _check_state(state, "tgv", x)
    info = np.zeros(2, dtype=np.float32)
    f = _fn("cilreg_tgv_dev_ws")
    with torch.cuda.device(x.device):
        rc = f(x.data_ptr(), out.data_ptr(), info.ctypes.data, lam, alpha1, alpha0, iters, L2, tol, *_dims(x),
               state.data_ptr(), int(bool(warm)), ctypes.byref(make_opts(**kw)),
               ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream))
    if rc != 0:
        raise RuntimeError(f"cilreg_tgv_dev_ws failed: {last_error()}")
    return out, info


def sb_warm(x, mu, iters, state, warm=True, tol=0.0, method=0, out=None, **kw):
    """SB starting from (warm) and returning the iterate state (U, D, B) in `state`."""
    import torch
    out = _prep(x, out, torch.float32)
    _check_state(state, "sb", x)
    info = np.zeros(2, dtype=np.float32)
    f = _fn("cilreg_sb_dev_ws")
    with torch.cuda.device(x.device):
        rc = f(x.data_ptr(), out.data_ptr(), info.ctypes.data, mu, iters, tol, method, *_dims(x), state.data_ptr(),
               int(bool(warm)), ctypes.byref(make_opts(**kw)),
               ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream))
    if rc != 0:
        raise RuntimeError(f"cilreg_sb_dev_ws failed: {last_error()}")
    return out, info


def run(method, x, params, out=None, **kw):
    """Dispatch by method name with a parameter dict (keys of the oracle's run())."""
    p = params
    if method == "rof":
        return rof(x, p["lambda"], p["iters"], p["tau"], p["tol"], out=out, **kw)
    if method == "fgp":
        return fgp(x, p["lambda"], p["iters"], p["tol"], p["method"], p["nonneg"], out=out, **kw)
    if method == "pd":
        return pd(x, p["lambda"], p["iters"], p["tol"], p["L"], p["method"], p["nonneg"], out=out, **kw)
    if method == "sb":
        return sb(x, p["mu"], p["iters"], p["tol"], p["method"], out=out, **kw)
    if method == "tgv":
        return tgv(x, p["lambda"], p["alpha1"], p["alpha0"], p["iters"], p["L2"], p["tol"], out=out, **kw)
    if method == "llt":
        return llt(x, p["lambda_rof"], p["lambda_llt"], p["iters"], p["tau"], p["tol"], out=out, **kw)
    if method == "ndf":
        return ndf(x, p["lambda"], p["sigma"], p["iters"], p["tau"], p["penalty"], p["tol"], out=out, **kw)
    if method == "d4":
        return d4(x, p["lambda"], p["sigma"], p["iters"], p["tau"], p["tol"], out=out, **kw)
    if method == "dtv":   # p["ref"]: the reference image, a tensor of x's shape
        return dtv(x, p["ref"], p["lambda"], p["iters"], p["tol"], p["eta"], p["method"], p["nonneg"], out=out, **kw)
    raise ValueError(method)
