"""Seeded multi-volume workloads: the uneven batch and the served request
stream (synthetic inputs only; phantoms are generated on the device from
their seeds by ccpi.filters.batch.phantom or at admission by the server).

Each job / request is a dict in the form ccpi.filters.batch takes; the
helpers here also say how to solve one of them alone (``solve_alone``), the
plain form the batched and served results are compared with.
"""
import numpy as np


def batched_jobs(seed=4):
    """64 uneven ellipsoid-phantom volumes, each extent drawn from U{64..320},
    noise sigma U[0.02, 0.08]; even entries FGP-TV (lambda U[0.01, 0.05], 300
    iterations, isotropic), odd entries ROF-TV (lambda U[0.01, 0.05], time
    step 0.001, 1000 iterations); fixed iteration counts.  Returns
    [(shape (Z, Y, X), (n_shapes, sigma, seed), job dict)]."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(64):
        shape = tuple(int(v) for v in rng.integers(64, 321, size=3))
        sigma = float(rng.uniform(0.02, 0.08))
        lam = float(rng.uniform(0.01, 0.05))
        if i % 2 == 0:
            spec = {"method": "fgp", "iters": 300, "lambda": lam, "method_tv": 0, "nonneg": 0}
        else:
            spec = {"method": "rof", "iters": 1000, "lambda": lam, "tau": 0.001}
        jobs.append((shape, (20, sigma, 9000 + i), spec))
    return jobs


def request_stream(rate=200.0, seconds=60.0, seed=5):
    """Poisson arrivals (``rate`` per second over ``seconds``); each request
    ROF/FGP/PD/TGV uniformly, a 2D slice (square, side U{256..2048}) with
    p = 0.5 else a 3D cube (side U{64..384}), generated phantom with noise
    sigma U[0.02, 0.08], lambda U[0.01, 0.05], iterations U{100..500}, no
    tolerance.  Returns [(arrival_seconds, job dict)] sorted by arrival."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > seconds:
            break
        method = ["rof", "fgp", "pd", "tgv"][int(rng.integers(0, 4))]
        if rng.random() < 0.5:
            side = int(rng.integers(256, 2049))
            shape = (side, side)
        else:
            side = int(rng.integers(64, 385))
            shape = (side, side, side)
        sigma = float(rng.uniform(0.02, 0.08))
        spec = {"method": method, "shape": shape, "phantom": (20, sigma, 20000 + len(reqs)),
                "lambda": float(rng.uniform(0.01, 0.05)), "iters": int(rng.integers(100, 501)),
                "tau": 0.001, "L": 12.0, "alpha1": 1.0, "alpha0": 2.0, "L2": 12.0}
        reqs.append((t, spec))
    return reqs


def solve_alone(D, spec, x, **kw):
    """One job dict solved by itself with the single-volume device solver D
    (ccpi.filters.device) on input tensor x: (out, infovector)."""
    m = spec["method"]
    tol = spec.get("tol", 0.0)
    mtv, nn = spec.get("method_tv", 0), spec.get("nonneg", 0)
    if m == "rof":
        return D.rof(x, spec["lambda"], spec["iters"], spec["tau"], tol, **kw)
    if m == "fgp":
        return D.fgp(x, spec["lambda"], spec["iters"], tol, mtv, nn, **kw)
    if m == "pd":
        return D.pd(x, spec["lambda"], spec["iters"], tol, spec.get("L", 8.0), mtv, nn, **kw)
    if m == "tgv":
        return D.tgv(x, spec["lambda"], spec.get("alpha1", 1.0), spec.get("alpha0", 2.0), spec["iters"],
                     spec.get("L2", 12.0), tol, **kw)
    raise ValueError(m)
