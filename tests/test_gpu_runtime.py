"""Runtime behaviour of the drop-in boundary: concurrent callers, memory
ownership, failure handling and server cancellation.

* Two host threads calling TV_FGP_GPU_main / TGV_GPU_main on one device at
  once (ctypes drops the GIL; each thread has its own stream): both results
  bit-exact against the oracle.
* After a drop-in call the library holds no device memory (the reference
  allocates and frees everything inside each call, shared.h / *_core.cu).
* An allocation failure injected at each allocation a call makes returns -1
  with the allocation error and leaves no device memory held; the next call
  succeeds (the reference's error convention, regularisers_GPU/shared.h:15-38).
* Server: cancelling a waiting, a running and a finished job frees their
  memory and leaves the other jobs' results unchanged.
"""
import ctypes
import threading

import numpy as np
import pytest

import oracle
from conftest import bits_equal

pytestmark = pytest.mark.gpu


def _lib():
    from ccpi.filters.utils import require_gpu_lib
    lib = require_gpu_lib()
    lib.cilreg_pool_stats.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
    lib.cilreg_pool_stats.restype = ctypes.c_int
    lib.cilreg_inject_alloc_failure.argtypes = [ctypes.c_int]
    lib.cilreg_inject_alloc_failure.restype = None
    return lib


def pool_stats():
    lib = _lib()
    b, i = ctypes.c_size_t(), ctypes.c_size_t()
    assert lib.cilreg_pool_stats(ctypes.byref(b), ctypes.byref(i)) == 0
    return b.value, i.value


def _vol(shape, seed):
    return np.random.default_rng(seed).random(shape).astype(np.float32)


def test_two_threads_on_one_device():
    from ccpi.filters import regularisers as R
    xa, xb = _vol((40, 70, 130), 1), _vol((33, 65, 129), 2)
    res = {}

    def run_a():
        for _ in range(3):
            res["a"] = R.FGP_TV(xa, 0.03, 60, 0.0, 0, 1, device="gpu")

    def run_b():
        for _ in range(3):
            res["b"] = R.TGV(xb, 0.03, 1.0, 2.0, 40, 12.0, 0.0, device="gpu")

    ta, tb = threading.Thread(target=run_a), threading.Thread(target=run_b)
    ta.start()
    tb.start()
    ta.join()
    tb.join()
    ya, _ = oracle.cpu_ref().fgp(xa, 0.03, 60, 0.0, 0, 1)
    yb, _ = oracle.cpu_ref().tgv(xb, 0.03, 1.0, 2.0, 40, 12.0, 0.0)
    assert bits_equal(res["a"], ya)
    assert bits_equal(res["b"], yb)


def test_drop_in_call_holds_no_device_memory():
    from ccpi.filters import regularisers as R
    from ccpi.filters import device as D
    D.release_cache()
    x = _vol((20, 36, 70), 3)
    for fn, args in [(R.FGP_TV, (0.02, 10, 1e-4, 0, 1)), (R.ROF_TV, (0.02, 10, 0.001, 0.0)),
                     (R.TGV, (0.02, 1.0, 2.0, 10, 12.0, 0.0)), (R.SB_TV, (0.02, 10, 0.0, 0))]:
        fn(x, *args, device="gpu")
        assert pool_stats() == (0, 0), fn.__name__


@pytest.mark.parametrize("method", ["fgp", "tgv", "rof", "pd", "sb", "dtv", "ndf"])
def test_injected_allocation_failure(method):
    """Each allocation of the call fails in turn: -1 (RuntimeError in
    Python) with the allocation error, nothing held afterwards; once the
    injection point lies beyond the call's allocations it succeeds and equals
    the oracle."""
    from ccpi.filters import regularisers as R
    lib = _lib()
    x = _vol((18, 33, 67), 4)
    calls = {
        "fgp": (lambda: R.FGP_TV(x, 0.02, 20, 1e-4, 0, 1, device="gpu"),
                lambda: oracle.cpu_ref().fgp(x, 0.02, 20, 1e-4, 0, 1)[0]),
        "tgv": (lambda: R.TGV(x, 0.02, 1.0, 2.0, 20, 12.0, 1e-4, device="gpu"),
                lambda: oracle.cpu_ref().tgv(x, 0.02, 1.0, 2.0, 20, 12.0, 1e-4)[0]),
        "rof": (lambda: R.ROF_TV(x, 0.02, 20, 0.001, 1e-4, device="gpu"),
                lambda: oracle.cpu_ref().rof(x, 0.02, 20, 0.001, 1e-4)[0]),
        "pd": (lambda: R.PD_TV(x, 0.02, 20, 1e-4, 12.0, 0, 1, device="gpu"),
               lambda: oracle.cpu_ref().pd(x, 0.02, 20, 1e-4, 12.0, 0, 1)[0]),
        "sb": (lambda: R.SB_TV(x, 0.02, 20, 1e-4, 0, device="gpu"),
               lambda: oracle.cpu_ref().sb(x, 0.02, 20, 1e-4, 0)[0]),
        "dtv": (lambda: R.FGP_dTV(x, x, 0.02, 20, 1e-4, 0.01, 0, 1, device="gpu"),
                lambda: oracle.cpu_ref().dtv(x, x, 0.02, 20, 1e-4, 0.01, 0, 1)[0]),
        "ndf": (lambda: R.NDF(x, 0.02, 0.05, 20, 0.01, 1, 1e-4, device="gpu"),
                lambda: oracle.cpu_ref().ndf(x, 0.02, 0.05, 20, 0.01, 1, 1e-4)[0]),
    }
    run, ref = calls[method]
    failures = 0
    y = None
    for k in range(1, 12):
        lib.cilreg_inject_alloc_failure(k)
        try:
            y = run()
        except RuntimeError as e:
            assert "alloc" in str(e), str(e)
            assert pool_stats() == (0, 0)
            failures += 1
            continue
        finally:
            lib.cilreg_inject_alloc_failure(0)
        break
    assert failures >= 1, "no allocation failed"
    assert y is not None, "the call did not succeed once the injection lay beyond its allocations"
    assert pool_stats() == (0, 0)
    assert bits_equal(y, ref())


def test_server_cancel():
    import torch
    from ccpi.filters import batch as B, device as D
    specs = [{"method": "fgp", "shape": (24, 40, 70), "phantom": (6, 0.05, 100 + i), "lambda": 0.02, "iters": 40}
             for i in range(4)]
    with B.Server(0, hbm_budget=200e9) as srv:
        t = [srv.submit(sp) for sp in specs]
        srv.step(4)                       # all four admitted and running
        srv.cancel(t[1])                  # running
        with pytest.raises(KeyError):
            srv.cancel(t[1])
        while srv.status(t[0]) != "done":
            srv.step(8)
        srv.cancel(t[2])                  # finished with t[0] (same size and count), not polled
        late = srv.submit(dict(specs[3], phantom=(6, 0.05, 200)))
        srv.cancel(late)                  # before it ran a step
        while srv.status(t[3]) != "done":
            srv.step(8)
        y0, _ = srv.poll(t[0])
        y3, _ = srv.poll(t[3])
    for ticket_spec, y in ((specs[0], y0), (specs[3], y3)):
        x = B.phantom(ticket_spec["shape"], *ticket_spec["phantom"])
        r, _ = D.fgp(x, 0.02, 40, 0.0, 0, 0)
        assert torch.equal(y, r)
