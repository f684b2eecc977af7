"""Two-iteration FGP, PD, TGV and FGP-dTV passes (temporal blocking) against
one-iteration passes.

The FGP solver fuses iterations k and k+1 into one pass by default (3D, and
2D images below 2^20 pixels; depth=2 forces it for any 2D image); the pass
evaluates iteration k on a ring around its tile and every plane through a
four-stage z pipeline.  Each value is computed by the same expressions in
the same order as the one-iteration kernel, so the results and the
infovector must be bit-identical to depth 1 (itself pinned to the oracle by
test_gpu_parity.py) for every tiling, chunking (down to one plane per
workgroup), boundary, TV type, nonnegativity, storage precision, odd and even
iteration counts, and tolerance stops at every check cadence.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(20, 36, 70), (9, 13, 130), (2, 17, 65), (33, 64, 64), (5, 70, 3)]


def _vol(shape, seed):
    from ccpi.filters import batch as B
    return B.phantom(shape, 6, 0.05, seed)


def _pair(x, **kw):
    from ccpi.filters import device as D
    a = D.fgp(x, depth=1, **kw)
    b = D.fgp(x, depth=2, **kw)
    return a, b


def _same(a, b, tag):
    (ya, ia), (yb, ib) = a, b
    assert np.array_equal(ia, ib), f"{tag}: infovector {ia} vs {ib}"
    assert bool((ya == yb).all()), f"{tag}: {int((ya != yb).sum())} elements differ"


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("zchunk", [0, 1, 2, 3, 7])
def test_depth2_matches_depth1_tilings(shape, zchunk):
    x = _vol(shape, 11)
    for tile_y in (4, 8):
        for iters in (1, 2, 7, 10):
            a, b = _pair(x, lam=0.03, iters=iters, zchunk=zchunk, tile_y=tile_y, nonneg=1)
            _same(a, b, f"{shape} zchunk={zchunk} tile_y={tile_y} iters={iters}")


@pytest.mark.parametrize("method", [0, 1], ids=["iso", "aniso"])
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_depth2_matches_depth1_methods(method, precision):
    x = _vol((24, 40, 72), 12)
    for nonneg in (0, 1):
        a, b = _pair(x, lam=0.02, iters=25, method=method, nonneg=nonneg, precision=precision)
        _same(a, b, f"method={method} nonneg={nonneg} {precision}")


@pytest.mark.parametrize("check_every", [1, 2, 3, 5])
def test_depth2_matches_depth1_tolerance_stops(check_every):
    x = _vol((16, 48, 80), 13)
    stopped = False
    for tol in (1e-2, 3e-3, 1e-4):
        for zchunk in (0, 4):
            a, b = _pair(x, lam=0.03, iters=120, tol=tol, check_every=check_every, zchunk=zchunk)
            _same(a, b, f"tol={tol} check_every={check_every} zchunk={zchunk}")
            stopped |= int(a[1][0]) < 120
    assert stopped


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_depth2_matches_depth1_tolerance_stops_16bit(precision):
    """With 16-bit P the pass keeps its FP64 residual sums in LDS slots: the
    reported residual and the stop iteration must equal the one-iteration
    kernel's, 3D and 2D."""
    for shape, seed in (((16, 48, 80), 13), ((200, 240), 17)):
        x = _vol(shape, seed)
        stopped = False
        for tol in (1e-2, 3e-3, 1e-4):
            a, b = _pair(x, lam=0.03, iters=120, tol=tol, check_every=1, precision=precision)
            _same(a, b, f"{shape} {precision} tol={tol}")
            stopped |= int(a[1][0]) < 120
        assert stopped, f"{shape} {precision}: no tolerance stop exercised"


def test_depth2_wide_index():
    x = _vol((12, 30, 66), 14)
    a, b = _pair(x, lam=0.03, iters=9, wide_index=True)
    _same(a, b, "wide index")


# 2D: the same pipeline on one plane (P has two components, no z terms).
SHAPES_2D = [(70, 130), (1, 65), (33, 2), (257, 200), (3, 3), (9, 64), (1030, 1027)]


@pytest.mark.parametrize("shape", SHAPES_2D, ids=lambda s: "x".join(map(str, s)))
def test_depth2_matches_depth1_2d_tilings(shape):
    x = _vol(shape, 15)
    for tile_y in (4, 8):
        for iters in (1, 2, 7, 10):
            a, b = _pair(x, lam=0.03, iters=iters, tile_y=tile_y, nonneg=1)
            _same(a, b, f"{shape} tile_y={tile_y} iters={iters}")


@pytest.mark.parametrize("method", [0, 1], ids=["iso", "aniso"])
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_depth2_matches_depth1_2d_methods(method, precision):
    x = _vol((150, 190), 16)
    for nonneg in (0, 1):
        a, b = _pair(x, lam=0.02, iters=25, method=method, nonneg=nonneg, precision=precision)
        _same(a, b, f"2D method={method} nonneg={nonneg} {precision}")


@pytest.mark.parametrize("check_every", [1, 2, 3, 5])
def test_depth2_matches_depth1_2d_tolerance_stops(check_every):
    x = _vol((200, 240), 17)
    stopped = False
    for tol in (1e-2, 3e-3, 1e-4):
        a, b = _pair(x, lam=0.03, iters=120, tol=tol, check_every=check_every)
        _same(a, b, f"2D tol={tol} check_every={check_every}")
        stopped |= int(a[1][0]) < 120
    assert stopped


def test_depth2_2d_wide_index():
    x = _vol((90, 130), 18)
    a, b = _pair(x, lam=0.03, iters=9, wide_index=True)
    _same(a, b, "2D wide index")


# ---------------------------------------------------------------------------
# TGV: the two-iteration pass reads the state once and writes it once; the
# intermediate fields are rounded through the storage type as the
# one-iteration kernel stores and re-reads them, so depth 2 == depth 1 bit
# for bit, and the stop output (u_bar after the stop iteration, from the
# pass's side buffer when the stop falls on its first iteration) too.
# ---------------------------------------------------------------------------
def _tgv_pair(x, **kw):
    from ccpi.filters import device as D
    a = D.tgv(x, depth=1, **kw)
    b = D.tgv(x, depth=2, **kw)
    return a, b


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("zchunk", [0, 1, 2, 3, 7])
def test_tgv_depth2_matches_depth1_tilings(shape, zchunk):
    x = _vol(shape, 21)
    for tile_y in (4, 8):
        for iters in (1, 2, 7, 10):
            a, b = _tgv_pair(x, lam=0.03, alpha1=1.0, alpha0=2.0, iters=iters, zchunk=zchunk, tile_y=tile_y)
            _same(a, b, f"TGV {shape} zchunk={zchunk} tile_y={tile_y} iters={iters}")


@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_tgv_depth2_matches_depth1_precisions(precision):
    x = _vol((24, 40, 72), 22)
    a, b = _tgv_pair(x, lam=0.02, alpha1=1.0, alpha0=2.0, iters=25, precision=precision)
    _same(a, b, f"TGV {precision}")


@pytest.mark.parametrize("check_every", [1, 2, 3, 5])
def test_tgv_depth2_matches_depth1_tolerance_stops(check_every):
    x = _vol((16, 48, 80), 23)
    stopped = set()
    for tol in (3e-2, 1e-2, 3e-3):
        for zchunk in (0, 4):
            a, b = _tgv_pair(x, lam=0.03, alpha1=1.0, alpha0=2.0, iters=150, tol=tol, check_every=check_every,
                             zchunk=zchunk)
            _same(a, b, f"TGV tol={tol} check_every={check_every} zchunk={zchunk}")
            if int(a[1][0]) < 150:
                stopped.add(int(a[1][0]) % 2)
    assert stopped, "no tolerance stop exercised"


@pytest.mark.parametrize("precision", ["bf16", "fp16"])
def test_tgv_depth2_matches_depth1_tolerance_stops_16bit(precision):
    """16-bit duals on the ring-wave pass (residual sums and the tile point's
    z pipeline in LDS slots): stop iteration, reported residual and output
    equal the one-iteration kernel's."""
    x = _vol((16, 48, 80), 23)
    stopped = False
    for tol in (3e-2, 1e-2, 3e-3):
        a, b = _tgv_pair(x, lam=0.03, alpha1=1.0, alpha0=2.0, iters=150, tol=tol, check_every=1,
                         precision=precision)
        _same(a, b, f"TGV {precision} tol={tol}")
        stopped |= int(a[1][0]) < 150
    assert stopped, f"TGV {precision}: no tolerance stop exercised"


def test_tgv_depth2_wide_index():
    x = _vol((12, 30, 66), 24)
    a, b = _tgv_pair(x, lam=0.03, alpha1=1.0, alpha0=2.0, iters=9, wide_index=True)
    _same(a, b, "TGV wide index")


# ---------------------------------------------------------------------------
# PD: the two-iteration pass (state u_bar, P), including the mirrored forward
# difference at the last index of every axis and the un-extrapolated stop
# output of either iteration of a pass.
# ---------------------------------------------------------------------------
def _pd_pair(x, **kw):
    from ccpi.filters import device as D
    a = D.pd(x, depth=1, **kw)
    b = D.pd(x, depth=2, **kw)
    return a, b


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("zchunk", [0, 1, 2, 3, 7])
def test_pd_depth2_matches_depth1_tilings(shape, zchunk):
    x = _vol(shape, 31)
    for tile_y in (4, 8):
        for iters in (1, 2, 7, 10):
            a, b = _pd_pair(x, lam=0.03, iters=iters, L=12.0, zchunk=zchunk, tile_y=tile_y, nonneg=1)
            _same(a, b, f"PD {shape} zchunk={zchunk} tile_y={tile_y} iters={iters}")


@pytest.mark.parametrize("method", [0, 1], ids=["iso", "aniso"])
@pytest.mark.parametrize("precision", ["fp32", "bf16", "fp16"])
def test_pd_depth2_matches_depth1_methods(method, precision):
    x = _vol((24, 40, 72), 32)
    for nonneg in (0, 1):
        a, b = _pd_pair(x, lam=0.02, iters=25, L=8.0, method=method, nonneg=nonneg, precision=precision)
        _same(a, b, f"PD method={method} nonneg={nonneg} {precision}")


@pytest.mark.parametrize("check_every", [1, 2, 3, 5])
def test_pd_depth2_matches_depth1_tolerance_stops(check_every):
    x = _vol((16, 48, 80), 33)
    parities = set()
    for tol in (3e-2, 1e-2, 3e-3):
        for zchunk in (0, 4):
            a, b = _pd_pair(x, lam=0.03, iters=150, tol=tol, L=12.0, check_every=check_every, zchunk=zchunk)
            _same(a, b, f"PD tol={tol} check_every={check_every} zchunk={zchunk}")
            if int(a[1][0]) < 150:
                parities.add(int(a[1][0]) % 2)
    assert parities, "no tolerance stop exercised"


def test_pd_depth2_wide_index():
    x = _vol((12, 30, 66), 34)
    a, b = _pd_pair(x, lam=0.03, iters=9, wide_index=True)
    _same(a, b, "PD wide index")


# ---------------------------------------------------------------------------
# FGP-dTV: the FGP pipeline with dTV's projections onto the complement of the
# reference gradient B; depth 2 == depth 1 bit for bit.
# ---------------------------------------------------------------------------
def _dtv_pair(x, r, **kw):
    from ccpi.filters import device as D
    a = D.dtv(x, r, depth=1, **kw)
    b = D.dtv(x, r, depth=2, **kw)
    return a, b


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("zchunk", [0, 1, 3])
def test_dtv_depth2_matches_depth1_tilings(shape, zchunk):
    x = _vol(shape, 21)
    r = _vol(shape, 22)
    for tile_y in (4, 8):
        for iters in (1, 2, 7, 10):
            a, b = _dtv_pair(x, r, lam=0.03, iters=iters, eta=0.05, zchunk=zchunk, tile_y=tile_y, nonneg=1)
            _same(a, b, f"dtv {shape} zchunk={zchunk} tile_y={tile_y} iters={iters}")


@pytest.mark.parametrize("method", [0, 1], ids=["iso", "aniso"])
def test_dtv_depth2_matches_depth1_methods(method):
    x = _vol((24, 40, 72), 23)
    r = _vol((24, 40, 72), 24)
    for nonneg in (0, 1):
        for eta in (0.01, 0.3):
            a, b = _dtv_pair(x, r, lam=0.02, iters=25, eta=eta, method=method, nonneg=nonneg)
            _same(a, b, f"dtv method={method} nonneg={nonneg} eta={eta}")


@pytest.mark.parametrize("check_every", [1, 2, 5])
def test_dtv_depth2_matches_depth1_tolerance_stops(check_every):
    x = _vol((16, 48, 80), 25)
    r = _vol((16, 48, 80), 26)
    stopped = False
    for tol in (1e-2, 3e-3, 1e-4):
        for zchunk in (0, 4):
            a, b = _dtv_pair(x, r, lam=0.03, iters=120, tol=tol, eta=0.05, check_every=check_every, zchunk=zchunk)
            _same(a, b, f"dtv tol={tol} check_every={check_every} zchunk={zchunk}")
            stopped |= int(a[1][0]) < 120
    assert stopped


def test_dtv_depth2_wide_index():
    x = _vol((12, 30, 66), 27)
    r = _vol((12, 30, 66), 28)
    a, b = _dtv_pair(x, r, lam=0.03, iters=9, eta=0.05, wide_index=True)
    _same(a, b, "dtv wide index")


# A stop on the first iteration of a two-iteration pass returns a value the
# pass keeps on chip (FGP / FGP-dTV: D_k; PD: the un-extrapolated U_k; TGV:
# u_bar after iteration k), and a stop on the second one PD's U_{k+1}: the
# solvers recompute those after the passes.  Both positions must occur and
# match one iteration per pass.
@pytest.mark.parametrize("solver", ["fgp", "pd", "dtv", "tgv"])
def test_depth2_stop_on_either_iteration_of_a_pass(solver):
    from ccpi.filters import device as D
    x = _vol((16, 48, 80), 41)
    r = _vol((16, 48, 80), 42)
    parities = set()
    for ce in (1, 3):
        for tol in (3e-2, 1e-2, 3e-3, 1e-3, 3e-4):
            if solver == "fgp":
                kw = dict(lam=0.03, iters=200, tol=tol, check_every=ce)
                a, b = D.fgp(x, depth=1, **kw), D.fgp(x, depth=2, **kw)
            elif solver == "pd":
                kw = dict(lam=0.03, iters=200, tol=tol, L=12.0, check_every=ce)
                a, b = D.pd(x, depth=1, **kw), D.pd(x, depth=2, **kw)
            elif solver == "dtv":
                kw = dict(lam=0.03, iters=200, tol=tol, eta=0.05, check_every=ce)
                a, b = D.dtv(x, r, depth=1, **kw), D.dtv(x, r, depth=2, **kw)
            else:
                kw = dict(lam=0.03, alpha1=1.0, alpha0=2.0, iters=200, tol=tol, check_every=ce)
                a, b = D.tgv(x, depth=1, **kw), D.tgv(x, depth=2, **kw)
            _same(a, b, f"{solver} tol={tol} check_every={ce}")
            if int(a[1][0]) < 200:
                parities.add(int(a[1][0]) % 2)
    # TGV's second-iteration stop returns the pass's own output; the others
    # must see both positions
    need = {0} if solver == "tgv" else {0, 1}
    assert need <= parities, f"stop positions exercised: {parities}"


# ROF 2D images below 2^20 pixels run two iterations per launch by default
# (depth=2 forces it for any 2D image): iteration k's fluxes and u_{k+1} on the
# tile and a one-point ring, then iteration k+1 on the tile, from LDS.  A stop
# on the pass's first iteration is recomputed by one one-iteration launch.
ROF2D_SHAPES = [(1, 1), (2, 1), (1, 70), (70, 1), (3, 5), (8, 64), (9, 65), (7, 63), (17, 127), (130, 5),
                (512, 512), (700, 1000)]


@pytest.mark.parametrize("shape", ROF2D_SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_rof2d_depth2_matches_depth1(shape):
    import torch
    from ccpi.filters import device as D
    x = torch.from_numpy(np.random.default_rng(43).random(shape, dtype=np.float32)).cuda()
    for iters in (1, 2, 3, 8, 41):
        for lam in (0.02, 0.2):
            kw = dict(lam=lam, iters=iters, tau=0.001)
            _same(D.rof(x, depth=1, **kw), D.rof(x, depth=2, **kw), f"rof {shape} iters={iters} lam={lam}")
            _same(D.rof(x, depth=1, **kw), D.rof(x, **kw), f"rof default {shape} iters={iters} lam={lam}")


def test_rof2d_depth2_stop_on_either_iteration_of_a_pass():
    import torch
    from ccpi.filters import device as D
    x = torch.from_numpy(np.random.default_rng(44).random((96, 200), dtype=np.float32)).cuda()
    parities = set()
    for ce in (1, 3, 5):
        for tol in np.geomspace(3e-2, 1e-5, 24):
            kw = dict(lam=0.02, iters=300, tau=0.001, tol=float(tol), check_every=ce)
            a = D.rof(x, depth=1, **kw)
            _same(a, D.rof(x, depth=2, **kw), f"rof tol={tol} check_every={ce}")
            if int(a[1][0]) < 300:
                parities.add(int(a[1][0]) % 2)
    assert parities == {0, 1}, f"stop positions exercised: {parities}"
