"""The allocation-free FGP entry point (cilreg_fgp_dev_noalloc) against the
allocating one (cilreg_fgp_dev), through the C ABI.

* Two calls in a row, the second on a smaller input, each bit-identical to
  cilreg_fgp_dev (output and infovector): 3D with a tolerance (the stop rule
  without pinned polling), 2D, BF16 duals.
* No allocation: with the library's allocation-failure hook armed for its
  next allocation, the call still succeeds (it allocates nothing), and the
  hook is still armed afterwards (the next allocating call fails).
* A volume whose workspace exceeds the scratch array is refused with -1 and
  the reason.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_vp, _i, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
_SIG = [_vp, _vp, _vp, _f, _i, _f, _i, _i, _i, _i, _i, _vp, _vp]


def _lib():
    from ccpi.filters.utils import require_gpu_lib
    lib = require_gpu_lib()
    for name in ("cilreg_fgp_dev", "cilreg_fgp_dev_noalloc"):
        f = getattr(lib, name)
        f.argtypes = _SIG
        f.restype = _i
    lib.cilreg_inject_alloc_failure.argtypes = [_i]
    lib.cilreg_inject_alloc_failure.restype = None
    lib.cilreg_noalloc_scratch_bytes.restype = ctypes.c_size_t
    return lib


def _call(name, x, out, lam, iters, tol, method, nonneg, precision="fp32"):
    import torch
    from ccpi.filters.device import make_opts
    lib = _lib()
    info = np.zeros(2, dtype=np.float32)
    d = list(x.shape)[::-1] + ([1] if x.dim() == 2 else [])
    opts = make_opts(precision=precision)
    rc = getattr(lib, name)(x.data_ptr(), out.data_ptr(), info.ctypes.data, lam, iters, tol, method, nonneg, *d,
                            ctypes.byref(opts), ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, info


def _vol(shape, seed):
    import torch
    return torch.from_numpy(np.random.default_rng(seed).random(shape).astype(np.float32)).cuda()


CASES = [
    ("3d-tol", (40, 96, 130), (17, 33, 70), dict(lam=0.02, iters=60, tol=1e-4, method=0, nonneg=1)),
    ("2d", (700, 900), (129, 65), dict(lam=0.03, iters=41, tol=0.0, method=1, nonneg=0)),
    ("3d-bf16", (32, 80, 200), (9, 20, 66), dict(lam=0.02, iters=30, tol=0.0, method=0, nonneg=1, precision="bf16")),
]


@pytest.mark.parametrize("name,big,small,p", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_allocating_entry_point(name, big, small, p):
    import torch
    for shape, seed in ((big, 1), (small, 2)):   # two calls in a row, the second smaller
        x = _vol(shape, seed)
        out_ref = torch.zeros_like(x)   # with a tolerance the output's contents seed the first residual
        out = torch.zeros_like(x)
        rc_ref, info_ref = _call("cilreg_fgp_dev", x, out_ref, **p)
        rc, info = _call("cilreg_fgp_dev_noalloc", x, out, **p)
        assert rc_ref == 0 and rc == 0
        assert np.array_equal(info, info_ref), (info, info_ref)
        assert torch.equal(out, out_ref), f"{name} {shape}: max |diff| {(out - out_ref).abs().max().item():.3e}"


def test_noalloc_allocates_nothing():
    import torch
    from ccpi.filters.utils import last_error
    lib = _lib()
    x = _vol((24, 40, 70), 3)
    out_ref = torch.zeros_like(x)
    out = torch.zeros_like(x)
    assert _call("cilreg_fgp_dev", x, out_ref, 0.02, 20, 0.0, 0, 1)[0] == 0
    lib.cilreg_inject_alloc_failure(1)   # the next device allocation would fail
    try:
        rc, _ = _call("cilreg_fgp_dev_noalloc", x, out, 0.02, 20, 0.0, 0, 1)
        assert rc == 0, last_error()
        # still armed: an allocating call now fails on its first allocation
        rc2, _ = _call("cilreg_fgp_dev", x, torch.zeros_like(x), 0.02, 20, 0.0, 0, 1)
        assert rc2 == -1 and "alloc" in last_error()
    finally:
        lib.cilreg_inject_alloc_failure(0)
    assert torch.equal(out, out_ref)


def test_noalloc_rejects_too_large():
    import torch
    from ccpi.filters.utils import last_error
    lib = _lib()
    cap = lib.cilreg_noalloc_scratch_bytes()
    assert cap == 64 << 20
    n = cap // 56 + 4096   # FP32 3D needs 56 B per voxel
    x = torch.zeros((n // (64 * 64) + 1, 64, 64), dtype=torch.float32, device="cuda")
    rc, _ = _call("cilreg_fgp_dev_noalloc", x, torch.zeros_like(x), 0.02, 4, 0.0, 0, 1)
    assert rc == -1
    assert "scratch" in last_error()
