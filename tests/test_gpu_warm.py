"""Warm start (cilreg_{fgp,pd,sb,tgv}_dev_ws): the iterate state in caller memory.

N iterations, then a warm call of M more from the returned state, must equal
one call of N + M bit for bit -- output and state -- for every method, 2D and
3D, odd and even splits (the two-iteration passes realign), and every dual
storage type (the state is FP32; 16-bit storage converts exactly both ways
because the stored values are 16-bit representable).  A cold call must equal
the plain solve.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(18, 40, 70), (90, 77)]


def _vol(shape, seed):
    from ccpi.filters import batch as B
    return B.phantom(shape, 5, 0.05, seed)


def _run(method, x, iters, state, tstate, warm, **kw):
    from ccpi.filters import device as D
    if method == "fgp":
        return D.fgp_warm(x, 0.03, iters, state, tstate, warm=warm, nonneg=1, **kw)
    if method == "pd":
        return D.pd_warm(x, 0.03, iters, state, warm=warm, L=12.0, **kw)
    if method == "sb":
        kw.pop("precision", None)   # SB stores its fields in FP32
        return D.sb_warm(x, 0.03, iters, state, warm=warm, **kw)
    return D.tgv_warm(x, 0.03, 1.0, 2.0, iters, state, warm=warm, **kw)


def _plain(method, x, iters, **kw):
    from ccpi.filters import device as D
    if method == "fgp":
        return D.fgp(x, 0.03, iters, nonneg=1, **kw)
    if method == "pd":
        return D.pd(x, 0.03, iters, L=12.0, **kw)
    if method == "sb":
        kw.pop("precision", None)
        return D.sb(x, 0.03, iters, **kw)
    return D.tgv(x, 0.03, 1.0, 2.0, iters, **kw)


@pytest.mark.parametrize("method", ["fgp", "pd", "tgv", "sb"])
@pytest.mark.parametrize("shape", SHAPES, ids=["3d", "2d"])
@pytest.mark.parametrize("precision", ["fp32", "bf16"])
def test_split_equals_single_call(method, shape, precision):
    from ccpi.filters import device as D
    x = _vol(shape, 50)
    for n, m in ((6, 7), (5, 8), (1, 12)):
        one = D.new_state(method, x)
        st1, ts1 = (one if method == "fgp" else (one, np.zeros(2, np.float32)))
        y1, i1 = _run(method, x, n + m, st1, ts1, False, precision=precision)
        two = D.new_state(method, x)
        st2, ts2 = (two if method == "fgp" else (two, np.zeros(2, np.float32)))
        _run(method, x, n, st2, ts2, False, precision=precision)
        y2, i2 = _run(method, x, m, st2, ts2, True, precision=precision)
        tag = f"{method} {shape} {precision} {n}+{m}"
        assert bool((y1 == y2).all()), tag
        assert bool((st1 == st2).all()), tag
        if method == "fgp":
            assert np.array_equal(ts1, ts2), tag
        ref, _ = _plain(method, x, n + m, precision=precision)
        assert bool((y1 == ref).all()), tag + " vs plain"


@pytest.mark.parametrize("method", ["fgp", "pd", "tgv", "sb"])
def test_tolerance_state_is_the_stop_iterations(method):
    # with a tolerance the state is that after the stop iteration: resuming
    # for M more equals a fixed-count solve of (stop + 1 + M) iterations
    from ccpi.filters import device as D
    x = _vol((16, 40, 72), 51)
    got = D.new_state(method, x)
    st, ts = (got if method == "fgp" else (got, np.zeros(2, np.float32)))
    _, info = _run(method, x, 200, st, ts, False, tol=2e-2, check_every=1)
    stop = int(info[0])
    assert stop < 200
    y_resumed, _ = _run(method, x, 5, st, ts, True)
    ref, _ = _plain(method, x, stop + 1 + 5)
    assert bool((y_resumed == ref).all())
