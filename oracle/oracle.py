"""ctypes bindings for the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg
import this module, and only as the checker.  The product package under
``ccpi-regularisation-toolkit_amd/`` never imports it.

Two libraries are bound with one wrapper surface:

* ``libcpuref.so`` -- this repository's C restatement (``cpu_ref.c``) in FP32
  (bit-exact with the reference C-OMP) and FP64 (precision study);
* ``_ref/libcilreg.so`` -- the reference C-OMP library compiled from its own
  sources by ``make -C oracle ref`` (present only where it was built).

Argument order and meaning follow the reference's Python wrappers
(src/Python/ccpi/filters/TV.py:8-250): numpy shape (Z, Y, X) or (Y, X) maps to
dims (dimX, dimY, dimZ) = shape[::-1] (+1 for 2D).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpuref.so")
REF_PATH = os.path.join(HERE, "_ref", "libcilreg.so")

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)


def build():
    """Compile the restatement (and the reference library when its sources exist)."""
    subprocess.check_call(["make", "-s", "-C", HERE])
    if os.path.isdir("/root/reference/src/Core/regularisers_CPU"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def _dims(a):
    d = list(a.shape)[::-1]
    if a.ndim == 2:
        d.append(1)
    return d


class _Lib:
    """Wrapper over one library; ``prec`` is 'f32' or 'f64' (f64 only for cpu_ref)."""

    def __init__(self, path, kind, prec="f32"):
        self.path = path
        self.kind = kind
        self.prec = prec
        self.lib = ctypes.CDLL(path)
        self.dtype = np.float32 if prec == "f32" else np.float64
        real = ctypes.c_float if prec == "f32" else ctypes.c_double
        self.real = real
        ptr = _f32p if prec == "f32" else _f64p
        self.ptr = ptr
        if kind == "cpuref":
            self.f_rof = getattr(self.lib, f"cr_rof_{prec}")
            self.f_fgp = getattr(self.lib, f"cr_fgp_{prec}")
            self.f_pd = getattr(self.lib, f"cr_pd_{prec}")
            self.f_sb = getattr(self.lib, f"cr_sb_{prec}")
            self.f_tgv = getattr(self.lib, f"cr_tgv_{prec}")
            self.f_dtv = getattr(self.lib, f"cr_dtv_{prec}")
            self.f_llt = getattr(self.lib, f"cr_llt_{prec}")
            self.f_ndf = getattr(self.lib, f"cr_ndf_{prec}")
            self.f_ndf_inplace = getattr(self.lib, f"cr_ndf_inplace_{prec}")
            self.f_d4 = getattr(self.lib, f"cr_d4_{prec}")
        else:  # the reference's own symbol names (src/Core/regularisers_CPU/*.h)
            self.f_rof = self.lib.TV_ROF_CPU_main
            self.f_fgp = self.lib.TV_FGP_CPU_main
            self.f_pd = self.lib.PDTV_CPU_main
            self.f_sb = self.lib.SB_TV_CPU_main
            self.f_tgv = self.lib.TGV_main
            self.f_dtv = self.lib.dTV_FGP_CPU_main
            self.f_llt = self.lib.LLT_ROF_CPU_main
            # the reference updates NDF in place inside its parallel loops;
            # run on one thread that is the serial in-place scheme
            self.f_ndf = None
            self.f_ndf_inplace = self.lib.Diffusion_CPU_main
            self.f_d4 = self.lib.Diffus4th_CPU_main
        i = ctypes.c_int
        self.f_rof.argtypes = [ptr, ptr, ptr, ptr, i, i, real, real, i, i, i]
        self.f_fgp.argtypes = [ptr, ptr, ptr, real, i, real, i, i, i, i, i]
        self.f_pd.argtypes = [ptr, ptr, ptr, real, i, real, real, i, i, i, i, i]
        self.f_sb.argtypes = [ptr, ptr, ptr, real, i, real, i, i, i, i]
        self.f_tgv.argtypes = [ptr, ptr, ptr, real, real, real, i, real, real, i, i, i]
        self.f_dtv.argtypes = [ptr, ptr, ptr, ptr, real, i, real, real, i, i, i, i, i]
        self.f_llt.argtypes = [ptr, ptr, ptr, real, real, i, real, real, i, i, i]
        for f in (self.f_ndf, self.f_ndf_inplace):
            if f is not None:
                f.argtypes = [ptr, ptr, ptr, real, real, i, real, i, real, i, i, i]
        self.f_d4.argtypes = [ptr, ptr, ptr, real, real, i, real, real, i, i, i]
        for f in (self.f_rof, self.f_fgp, self.f_pd, self.f_sb, self.f_tgv, self.f_dtv, self.f_llt, self.f_ndf,
                  self.f_ndf_inplace, self.f_d4):
            if f is not None:
                f.restype = real

    # -- helpers --------------------------------------------------------
    def _prep(self, x, out, zero_out=True):
        x = np.ascontiguousarray(x, dtype=self.dtype)
        if out is None:
            out = np.zeros_like(x)
        info = np.zeros(2, dtype=self.dtype)
        return x, out, info

    def _p(self, a):
        return a.ctypes.data_as(self.ptr)

    # -- solvers (argument order of ccpi.filters.regularisers) ------------
    def rof(self, x, lam, iters, tau, tol, out=None, lam_arr=None):
        x, out, info = self._prep(x, out)
        d = _dims(x)
        if lam_arr is None:
            lam_buf = np.array([lam], dtype=self.dtype)
            is_arr = 0
        else:
            lam_buf = np.ascontiguousarray(lam_arr, dtype=self.dtype)
            is_arr = 1
        self.f_rof(self._p(x), self._p(out), self._p(info), self._p(lam_buf), is_arr, iters, tau, tol, *d)
        return out, info

    def fgp(self, x, lam, iters, tol, method, nonneg, out=None):
        x, out, info = self._prep(x, out)
        d = _dims(x)
        self.f_fgp(self._p(x), self._p(out), self._p(info), lam, iters, tol, method, nonneg, *d)
        return out, info

    def pd(self, x, lam, iters, tol, L, method, nonneg, out=None):
        x, out, info = self._prep(x, out)
        d = _dims(x)
        self.f_pd(self._p(x), self._p(out), self._p(info), lam, iters, tol, L, method, nonneg, *d)
        return out, info

    def sb(self, x, mu, iters, tol, method, out=None):
        x, out, info = self._prep(x, out)
        d = _dims(x)
        self.f_sb(self._p(x), self._p(out), self._p(info), mu, iters, tol, method, *d)
        return out, info

    def tgv(self, x, lam, a1, a0, iters, L2, tol, out=None):
        x, out, info = self._prep(x, out)
        d = _dims(x)
        self.f_tgv(self._p(x), self._p(out), self._p(info), lam, a1, a0, iters, L2, tol, *d)
        return out, info

    def dtv(self, x, ref, lam, iters, tol, eta, method, nonneg, out=None):
        x, out, info = self._prep(x, out)
        r = np.ascontiguousarray(ref, dtype=self.dtype)
        if r.shape != x.shape:
            raise ValueError("reference image shape differs from the input's")
        d = _dims(x)
        self.f_dtv(self._p(x), self._p(r), self._p(out), self._p(info), lam, iters, tol, eta, method, nonneg, *d)
        return out, info

    def llt(self, x, lam_rof, lam_llt, iters, tau, tol, out=None):
        x, out, info = self._prep(x, out)
        self.f_llt(self._p(x), self._p(out), self._p(info), lam_rof, lam_llt, iters, tau, tol, *_dims(x))
        return out, info

    def ndf(self, x, lam, sigma, iters, tau, penalty, tol, out=None, in_place=False):
        """The synchronous explicit scheme; ``in_place=True`` runs the serial
        in-place order instead (the reference library's order on one thread)."""
        f = self.f_ndf_inplace if in_place else self.f_ndf
        if f is None:
            raise ValueError("the reference library has only the in-place scheme")
        x, out, info = self._prep(x, out)
        if in_place and self.kind == "ref":
            with _single_thread():
                f(self._p(x), self._p(out), self._p(info), lam, sigma, iters, tau, penalty, tol, *_dims(x))
        else:
            f(self._p(x), self._p(out), self._p(info), lam, sigma, iters, tau, penalty, tol, *_dims(x))
        return out, info

    def d4(self, x, lam, sigma, iters, tau, tol, out=None):
        x, out, info = self._prep(x, out)
        self.f_d4(self._p(x), self._p(out), self._p(info), lam, sigma, iters, tau, tol, *_dims(x))
        return out, info

    def run(self, method, x, params, out=None):
        """Dispatch by method name with a parameter dict (keys as in the
        reference Python API)."""
        p = params
        if method == "rof":
            return self.rof(x, p["lambda"], p["iters"], p["tau"], p["tol"], out=out)
        if method == "fgp":
            return self.fgp(x, p["lambda"], p["iters"], p["tol"], p["method"], p["nonneg"], out=out)
        if method == "pd":
            return self.pd(x, p["lambda"], p["iters"], p["tol"], p["L"], p["method"], p["nonneg"], out=out)
        if method == "sb":
            return self.sb(x, p["mu"], p["iters"], p["tol"], p["method"], out=out)
        if method == "tgv":
            return self.tgv(x, p["lambda"], p["alpha1"], p["alpha0"], p["iters"], p["L2"], p["tol"], out=out)
        if method == "llt":
            return self.llt(x, p["lambda_rof"], p["lambda_llt"], p["iters"], p["tau"], p["tol"], out=out)
        if method == "ndf":
            return self.ndf(x, p["lambda"], p["sigma"], p["iters"], p["tau"], p["penalty"], p["tol"], out=out,
                            in_place=p.get("in_place", False))
        if method == "d4":
            return self.d4(x, p["lambda"], p["sigma"], p["iters"], p["tau"], p["tol"], out=out)
        if method == "dtv":   # p["ref"]: the reference image, an array of x's shape
            return self.dtv(x, p["ref"], p["lambda"], p["iters"], p["tol"], p["eta"], p["method"], p["nonneg"],
                            out=out)
        raise ValueError(method)


_cache = {}


class _single_thread:
    """OpenMP on one thread for the duration of a call (libgomp's omp_set_num_threads)."""

    def __enter__(self):
        self.gomp = ctypes.CDLL("libgomp.so.1")
        self.prev = self.gomp.omp_get_max_threads()
        self.gomp.omp_set_num_threads(1)

    def __exit__(self, *exc):
        self.gomp.omp_set_num_threads(self.prev)


def cpu_ref(prec="f32"):
    """The C restatement; builds it on first use if the .so is missing."""
    key = ("cpuref", prec)
    if key not in _cache:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", HERE])
        _cache[key] = _Lib(LIB_PATH, "cpuref", prec)
    return _cache[key]


def reference():
    """The reference C-OMP library built from its sources, or None when absent."""
    key = ("ref", "f32")
    if key not in _cache:
        _cache[key] = _Lib(REF_PATH, "ref") if os.path.exists(REF_PATH) else None
    return _cache[key]


def tv_energy_f64(u, u0, lam, type=1):
    """The FP64 objective of cpu_ref.c (TV_energy2D/3D of utils.c:64-117 in double)."""
    lib = cpu_ref().lib
    f = lib.cr_tv_energy_f64
    f.argtypes = [_f32p, _f32p, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    f.restype = ctypes.c_double
    u = np.ascontiguousarray(u, dtype=np.float32)
    u0 = np.ascontiguousarray(u0, dtype=np.float32)
    d = _dims(u)
    return f(u.ctypes.data_as(_f32p), u0.ctypes.data_as(_f32p), lam, type, *d)


def reference_tv_energy(u, u0, lam, type=1):
    """TV_energy2D / TV_energy3D of the reference library (utils.c:64-117):
    a serial float sum.  None when the reference library is not built."""
    ref = reference()
    if ref is None:
        return None
    u = np.ascontiguousarray(u, dtype=np.float32)
    u0 = np.ascontiguousarray(u0, dtype=np.float32)
    e = np.zeros(1, dtype=np.float32)
    i = ctypes.c_int
    if u.ndim == 2:
        f = ref.lib.TV_energy2D
        f.argtypes = [_f32p, _f32p, _f32p, ctypes.c_float, i, i, i]
        f.restype = ctypes.c_float
        f(u.ctypes.data_as(_f32p), u0.ctypes.data_as(_f32p), e.ctypes.data_as(_f32p), lam, type, u.shape[1],
          u.shape[0])
    else:
        f = ref.lib.TV_energy3D
        f.argtypes = [_f32p, _f32p, _f32p, ctypes.c_float, i, i, i, i]
        f.restype = ctypes.c_float
        f(u.ctypes.data_as(_f32p), u0.ctypes.data_as(_f32p), e.ctypes.data_as(_f32p), lam, type, u.shape[2],
          u.shape[1], u.shape[0])
    return float(e[0])


class forced_stop:
    """Within the block the restatement's stopping rule fires at the check
    of iteration ``it`` and at no other (``cr_set_forced_stop``): the output of
    a solve that stopped there, whatever its residual sums decide."""

    def __init__(self, it, prec="f32"):
        self.it = int(it)
        self.f = getattr(cpu_ref(prec).lib, f"cr_set_forced_stop_{prec}")
        self.f.argtypes = [ctypes.c_int]
        self.f.restype = None

    def __enter__(self):
        self.f(self.it)
        return self

    def __exit__(self, *exc):
        self.f(-1)
