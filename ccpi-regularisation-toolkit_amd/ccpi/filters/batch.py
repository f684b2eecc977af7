"""Many volumes at once: the batch runner, the continuous-batching server and
the device phantom generator (cilreg_batch_run, cilreg_server_*,
cilreg_phantom_dev in include/cilreg_ext.h).

A job is a dict with the keys of ``device.run`` plus the method and the
volume:

    {"method": "fgp", "input": tensor (Z, Y, X) or (Y, X) float32 CUDA, or
     "shape": (Z, Y, X) / (Y, X) with "phantom": (n_shapes, sigma, seed),
     "lambda": .., "iters": .., "tol": .., ...method parameters}

Method parameters: rof -- lambda, tau; fgp -- lambda, method, nonneg;
pd -- lambda, L, method, nonneg; tgv -- lambda, alpha1, alpha0, L2.
Every result is bit-identical to solving the volume alone with the same
options (for a tolerance-stopped 3D job, the same ``zchunk``: the batch
marches 32 planes per workgroup unless ``zchunk`` is given).
"""
import ctypes

import numpy as np

from .device import make_opts, CilregOpts
from .utils import require_gpu_lib, last_error

METHODS = {"rof": 0, "fgp": 1, "pd": 2, "tgv": 3}


class CilregJob(ctypes.Structure):
    _fields_ = [("method", ctypes.c_int), ("dimX", ctypes.c_int), ("dimY", ctypes.c_int), ("dimZ", ctypes.c_int),
                ("iter", ctypes.c_int), ("epsil", ctypes.c_float), ("lambda_", ctypes.c_float),
                ("tau", ctypes.c_float), ("lipschitz", ctypes.c_float), ("alpha1", ctypes.c_float),
                ("alpha0", ctypes.c_float), ("L2", ctypes.c_float), ("methodTV", ctypes.c_int),
                ("nonneg", ctypes.c_int), ("d_input", ctypes.c_void_p), ("d_output", ctypes.c_void_p),
                ("phantom_shapes", ctypes.c_int), ("phantom_sigma", ctypes.c_float),
                ("phantom_seed", ctypes.c_ulonglong), ("infovector", ctypes.c_float * 2)]


def _lib():
    lib = require_gpu_lib()
    lib.cilreg_batch_run.argtypes = [ctypes.POINTER(CilregJob), ctypes.c_int, ctypes.POINTER(CilregOpts),
                                     ctypes.c_void_p]
    lib.cilreg_batch_run.restype = ctypes.c_int
    lib.cilreg_phantom_dev.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_float, ctypes.c_ulonglong, ctypes.c_void_p]
    lib.cilreg_phantom_dev.restype = ctypes.c_int
    lib.cilreg_server_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.POINTER(CilregOpts)]
    lib.cilreg_server_create.restype = ctypes.c_void_p
    lib.cilreg_server_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(CilregJob),
                                         ctypes.POINTER(ctypes.c_longlong)]
    lib.cilreg_server_submit.restype = ctypes.c_int
    lib.cilreg_server_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                       ctypes.POINTER(ctypes.c_int)]
    lib.cilreg_server_step.restype = ctypes.c_int
    lib.cilreg_server_poll.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.POINTER(ctypes.c_float),
                                       ctypes.c_void_p]
    lib.cilreg_server_poll.restype = ctypes.c_int
    lib.cilreg_server_status.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
    lib.cilreg_server_status.restype = ctypes.c_int
    lib.cilreg_server_cancel.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
    lib.cilreg_server_cancel.restype = ctypes.c_int
    lib.cilreg_server_destroy.argtypes = [ctypes.c_void_p]
    lib.cilreg_server_destroy.restype = None
    return lib


def _shape_dims(shape):
    shape = tuple(int(s) for s in shape)
    if len(shape) == 2:
        return shape[1], shape[0], 1
    if len(shape) == 3:
        return shape[2], shape[1], shape[0]
    raise ValueError(f"only 2D and 3D volumes are supported, got shape {shape}")


def job_struct(spec, output=None):
    """The C descriptor of a job dict (the tensors it names must stay alive)."""
    m = spec["method"]
    if m not in METHODS:
        raise ValueError(f"method must be one of {sorted(METHODS)}")
    j = CilregJob()
    j.method = METHODS[m]
    x = spec.get("input")
    if x is not None:
        import torch
        if not (x.is_cuda and x.dtype == torch.float32 and x.is_contiguous()):
            raise ValueError("input must be a contiguous float32 CUDA tensor")
        shape = tuple(x.shape)
        j.d_input = x.data_ptr()
    else:
        shape = tuple(spec["shape"])
        n_shapes, sigma, seed = spec["phantom"]
        j.phantom_shapes, j.phantom_sigma, j.phantom_seed = int(n_shapes), float(sigma), int(seed)
        if j.phantom_shapes <= 0:
            raise ValueError("a generated input needs phantom n_shapes > 0")
    j.dimX, j.dimY, j.dimZ = _shape_dims(shape)
    j.iter = int(spec["iters"])
    j.epsil = float(spec.get("tol", 0.0))
    j.lambda_ = float(spec.get("lambda", 0.0))
    j.tau = float(spec.get("tau", 0.0))
    j.lipschitz = float(spec.get("L", 8.0))
    j.alpha1 = float(spec.get("alpha1", 1.0))
    j.alpha0 = float(spec.get("alpha0", 2.0))
    j.L2 = float(spec.get("L2", 12.0))
    j.methodTV = int(spec.get("method_tv", 0))
    j.nonneg = int(spec.get("nonneg", 0))
    if output is not None:
        j.d_output = output.data_ptr()
    return j, shape


def _stream(device=None):
    import torch
    return torch.cuda.current_stream(device).cuda_stream


def phantom(shape, n_shapes=8, sigma=0.05, seed=0, device=None, out=None):
    """Seeded noisy phantom generated on the device (float32, shape (Z, Y, X) or (Y, X))."""
    import torch
    lib = _lib()
    if out is None:
        out = torch.empty(tuple(shape), dtype=torch.float32, device=device or "cuda")
    nx, ny, nz = _shape_dims(out.shape)
    with torch.cuda.device(out.device):
        rc = lib.cilreg_phantom_dev(out.data_ptr(), nx, ny, nz, int(n_shapes), float(sigma), int(seed),
                                    ctypes.c_void_p(_stream(out.device)))
    if rc != 0:
        raise RuntimeError(f"cilreg_phantom_dev failed: {last_error()}")
    return out


def batch_run(specs, outs=None, **kw):
    """Solve every job of `specs` with grouped kernels; returns [(out, infovector)].

    Options are those of device.make_opts (precision, check_every, zchunk,
    stream_launches=True to launch directly instead of one hipGraph)."""
    import torch
    lib = _lib()
    opts = make_opts(**kw)
    n = len(specs)
    arr = (CilregJob * max(n, 1))()
    results = []
    dev = None
    for i, sp in enumerate(specs):
        x = sp.get("input")
        shape = tuple(x.shape) if x is not None else tuple(sp["shape"])
        dev = x.device if x is not None else (dev or torch.device("cuda", torch.cuda.current_device()))
        o = outs[i] if outs is not None else torch.zeros(shape, dtype=torch.float32, device=dev)
        j, _ = job_struct(sp, o)
        arr[i] = j
        results.append(o)
    with torch.cuda.device(dev or torch.cuda.current_device()):
        rc = lib.cilreg_batch_run(arr, n, ctypes.byref(opts), ctypes.c_void_p(_stream()))
    if rc != 0:
        raise RuntimeError(f"cilreg_batch_run failed: {last_error()}")
    return [(results[i], np.array([arr[i].infovector[0], arr[i].infovector[1]], dtype=np.float32))
            for i in range(n)]


class Server:
    """Continuous-batching server: submit jobs at any time, advance all
    admitted jobs together with step(), collect finished ones with poll().

    Jobs wait until their device memory fits the HBM budget; results stay on
    the device until polled."""

    def __init__(self, gpu_device=0, hbm_budget=200e9, **kw):
        self._lib = _lib()
        self._opts = make_opts(**kw)
        self._h = self._lib.cilreg_server_create(int(gpu_device), float(hbm_budget), ctypes.byref(self._opts))
        if not self._h:
            raise RuntimeError(f"cilreg_server_create failed: {last_error()}")
        self.device = gpu_device
        self._keep = {}

    def submit(self, spec):
        """Queue a job; returns its ticket.  An input tensor must stay alive until polled."""
        j, shape = job_struct(spec)
        t = ctypes.c_longlong()
        if self._lib.cilreg_server_submit(self._h, ctypes.byref(j), ctypes.byref(t)) != 0:
            raise RuntimeError(f"cilreg_server_submit failed: {last_error()}")
        self._keep[t.value] = (spec.get("input"), shape)
        return t.value

    def step(self, levels=1):
        """Queue `levels` more iterations of every admitted job behind the step
        already queued, wait for that earlier step and retire what it finished
        (levels <= 0: wait for everything queued); returns (active, pending)."""
        a, p = ctypes.c_int(), ctypes.c_int()
        if self._lib.cilreg_server_step(self._h, int(levels), ctypes.byref(a), ctypes.byref(p)) != 0:
            raise RuntimeError(f"cilreg_server_step failed: {last_error()}")
        return a.value, p.value

    def status(self, ticket):
        """'pending', 'running', 'done' or 'failed'."""
        return {0: "pending", 1: "running", 2: "done"}.get(self._lib.cilreg_server_status(self._h, ticket), "failed")

    def poll(self, ticket, out=None, discard=False):
        """(out, infovector) once the job is done (its memory is then released), else None.

        out: a float32 CUDA tensor of the job's shape to receive the result
        (allocated when None); discard=True releases the job without copying
        the result (out is then None)."""
        import torch
        keep = self._keep.get(ticket)
        if keep is None:
            raise KeyError(ticket)
        st = self._lib.cilreg_server_status(self._h, ticket)
        if st in (0, 1):
            return None
        info = (ctypes.c_float * 2)()
        if out is None and not discard and st == 2:
            out = torch.empty(keep[1], dtype=torch.float32, device=torch.device("cuda", self.device))
        rc = self._lib.cilreg_server_poll(self._h, ticket, info, ctypes.c_void_p(out.data_ptr() if out is not None
                                                                                 else None))
        self._keep.pop(ticket, None)
        if rc < 0:
            raise RuntimeError(f"job {ticket} failed: {last_error()}")
        return out, np.array([info[0], info[1]], dtype=np.float32)

    def cancel(self, ticket):
        """Withdraw a job in any state; its device memory is released."""
        if ticket not in self._keep:
            raise KeyError(ticket)
        if self._lib.cilreg_server_cancel(self._h, int(ticket)) != 0:
            raise RuntimeError(f"cilreg_server_cancel failed: {last_error()}")
        self._keep.pop(ticket, None)

    def close(self):
        if self._h:
            self._lib.cilreg_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def job_voxels(spec):
    shape = tuple(spec["input"].shape) if spec.get("input") is not None else tuple(spec["shape"])
    return int(np.prod(shape))


def serve_stream(server, requests, levels=8, keep_results=False, clock=None, progress=None, progress_every=20.0):
    """Replay a timed request stream through `server` in real time.

    requests: list of (arrival_seconds, job dict), sorted by arrival.  Each
    request is submitted once the wall clock passes its arrival time, all
    admitted jobs advance `levels` iterations per server step, and finished
    jobs are collected after every step.  Returns a dict with per-request
    latency (completion - arrival, seconds), iterations, voxel counts, the
    makespan and the mean number of jobs on the device per step; with
    keep_results the outputs too (ticket order = request order).  progress:
    called as progress(done, total) every `progress_every` seconds."""
    import time
    clock = clock or time.perf_counter
    n = len(requests)
    lat = [None] * n
    iters = [0] * n
    outs = [None] * n if keep_results else None
    vox = [job_voxels(sp) for _, sp in requests]
    t0 = clock()
    nxt = 0
    live = {}
    steps = 0
    occupancy = 0
    done = 0
    t_report = t0
    while nxt < n or live:
        if progress is not None and clock() - t_report >= progress_every:
            t_report = clock()
            progress(done, n)
        now = clock() - t0
        while nxt < n and requests[nxt][0] <= now:
            live[server.submit(requests[nxt][1])] = nxt
            nxt += 1
        if not live:
            time.sleep(max(0.0, min(requests[nxt][0] - now, 0.01)))
            continue
        active, _ = server.step(levels)
        steps += 1
        occupancy += active
        for t in list(live):
            r = server.poll(t, discard=not keep_results)
            if r is None:
                continue
            i = live.pop(t)
            done += 1
            lat[i] = clock() - t0 - requests[i][0]
            iters[i] = int(r[1][0])
            if keep_results:
                outs[i] = r
    makespan = clock() - t0 - (requests[0][0] if n else 0.0)
    res = {"latency_s": lat, "iterations": iters, "voxels": vox, "makespan_s": makespan,
           "steps": steps, "mean_active_per_step": occupancy / steps if steps else 0.0}
    if keep_results:
        res["results"] = outs
    return res
