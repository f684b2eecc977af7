"""SB-TV: the fused pass (both sweeps, shrinkage and Bregman update in one
kernel) against the two-kernel path (``depth=1``), bit for bit.

The two-kernel path is pinned to the reference fixtures by
test_gpu_parity.py (which itself runs the fused pass by default); here every
tiling, z chunking (down to one plane per workgroup), boundary, extent that
falls back to the two-kernel path, TV type and tolerance stop must agree.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(20, 36, 70), (9, 13, 130), (2, 17, 65), (33, 64, 64), (5, 70, 3), (3, 9, 129), (4, 17, 40),
          (6, 10, 66), (1, 30, 50), (2, 2, 2)]


def _vol(shape, seed):
    from ccpi.filters import batch as B
    return B.phantom(shape, 6, 0.05, seed)


def _pair(x, **kw):
    from ccpi.filters import device as D
    a = D.sb(x, depth=1, **kw)
    b = D.sb(x, depth=0, **kw)
    return a, b


def _same(a, b, tag):
    (ya, ia), (yb, ib) = a, b
    assert np.array_equal(ia, ib), f"{tag}: infovector {ia} vs {ib}"
    assert bool((ya == yb).all()), f"{tag}: {int((ya != yb).sum())} elements differ"


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("zchunk", [0, 1, 3])
def test_sb_fused_matches_two_kernels(shape, zchunk):
    x = _vol(shape, 31)
    for tile_y in (4, 8):
        for iters in (1, 2, 5):
            for method in (0, 1):
                a, b = _pair(x, mu=0.05, iters=iters, method=method, zchunk=zchunk, tile_y=tile_y)
                _same(a, b, f"{shape} zchunk={zchunk} tile_y={tile_y} iters={iters} method={method}")


@pytest.mark.parametrize("shape", [(37, 53), (64, 9), (65, 9), (3, 130), (1, 70), (70, 1)],
                         ids=lambda s: "x".join(map(str, s)))
def test_sb_fused_matches_two_kernels_2d(shape):
    x = _vol(shape, 32)
    for iters in (1, 4):
        for method in (0, 1):
            a, b = _pair(x, mu=0.05, iters=iters, method=method)
            _same(a, b, f"2d {shape} iters={iters} method={method}")


@pytest.mark.parametrize("check_every", [1, 5])
def test_sb_fused_tolerance_stops(check_every):
    x = _vol((16, 48, 80), 33)
    stopped = False
    for tol in (1e-2, 3e-3, 1e-4):
        a, b = _pair(x, mu=0.05, iters=80, tol=tol, check_every=check_every)
        _same(a, b, f"tol={tol} check_every={check_every}")
        stopped |= int(a[1][0]) < 80
    assert stopped


def test_sb_fused_wide_index():
    x = _vol((12, 30, 66), 34)
    a, b = _pair(x, mu=0.05, iters=6, wide_index=True)
    _same(a, b, "wide index")
