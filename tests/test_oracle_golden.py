"""Pins the CPU oracle to the reference C-OMP library.

Every fixture in tests/golden/ref_cases.npz was written by the reference
library compiled from its own sources (tests/golden/make_golden.py); the
oracle must reproduce output and infovector bit for bit, including early-stop
iteration counts and the serial-FP32 residual.  When the reference library is
built in this checkout (oracle/_ref/libcilreg.so), extra seeded cases are
compared against it directly.
"""
import zlib

import numpy as np
import pytest

from conftest import golden, case_id, bits_equal

import oracle

CASES = golden().cases


@pytest.mark.parametrize("case", CASES, ids=[case_id(c) for c in CASES])
def test_oracle_matches_reference_fixture(case):
    g = golden()
    x = g.input(case)
    out0 = g.out_init(case)
    p = g.params(case)
    if case["method"] == "ndf":
        p["in_place"] = True   # the fixture holds the reference's serial in-place order
    y, info = oracle.cpu_ref().run(case["method"], x, p, out=out0)
    y_ref, info_ref = g.expected(case)
    assert bits_equal(y, y_ref), f"max |diff| = {np.abs(y - y_ref).max()}"
    assert bits_equal(info, info_ref), f"infovector {info} vs reference {info_ref}"


def _ref_or_skip():
    ref = oracle.reference()
    if ref is None:
        pytest.skip("reference C-OMP library not built here (make -C oracle ref)")
    return ref


EXTRA = [
    ("llt", {"lambda_rof": 0.03, "lambda_llt": 0.01, "iters": 25, "tau": 0.004, "tol": 0.0}),
    ("llt", {"lambda_rof": 0.01, "lambda_llt": 0.03, "iters": 120, "tau": 0.01, "tol": 5e-3}),
    ("ndf", {"lambda": 0.03, "sigma": 0.1, "iters": 25, "tau": 0.02, "penalty": 3, "tol": 0.0, "in_place": True}),
    ("ndf", {"lambda": 0.03, "sigma": 0.2, "iters": 60, "tau": 0.02, "penalty": 5, "tol": 5e-3, "in_place": True}),
    ("d4", {"lambda": 1.0, "sigma": 0.05, "iters": 25, "tau": 0.001, "tol": 0.0}),
    ("rof", {"lambda": 0.03, "iters": 25, "tau": 0.002, "tol": 0.0}),
    ("fgp", {"lambda": 0.07, "iters": 25, "tol": 0.0, "method": 0, "nonneg": 1}),
    ("fgp", {"lambda": 0.01, "iters": 120, "tol": 5e-3, "method": 1, "nonneg": 0}),
    ("pd", {"lambda": 0.07, "iters": 25, "tol": 0.0, "L": 6.0, "method": 1, "nonneg": 1}),
    ("sb", {"mu": 0.08, "iters": 15, "tol": 0.0, "method": 1}),
    ("tgv", {"lambda": 0.1, "alpha1": 0.5, "alpha0": 1.5, "iters": 25, "L2": 10.0, "tol": 0.0}),
]


@pytest.mark.parametrize("shape", [(29, 31), (9, 10, 23), (3, 5, 7), (2, 3), (2, 2, 2)])
@pytest.mark.parametrize("method,params", EXTRA, ids=[f"{m}{i}" for i, (m, _) in enumerate(EXTRA)])
def test_oracle_matches_reference_library(shape, method, params):
    ref = _ref_or_skip()
    rng = np.random.default_rng(zlib.crc32(repr((shape, method, params)).encode()))
    x = (rng.random(shape) * 2 - 0.5).astype(np.float32)   # negative values exercise nonneg
    a, ia = oracle.cpu_ref().run(method, x, params)
    b, ib = ref.run(method, x, params)
    assert bits_equal(a, b)
    assert bits_equal(ia, ib)


DTV_EXTRA = [
    {"lambda": 0.04, "iters": 25, "tol": 0.0, "eta": 0.02, "method": 0, "nonneg": 1},
    {"lambda": 0.01, "iters": 120, "tol": 5e-3, "eta": 0.3, "method": 1, "nonneg": 0},
]


@pytest.mark.parametrize("shape", [(29, 31), (9, 10, 23), (3, 5, 7), (1, 9), (4, 1, 6)])
@pytest.mark.parametrize("params", DTV_EXTRA, ids=["dtv0", "dtv1"])
def test_oracle_dtv_matches_reference_library(shape, params):
    ref = _ref_or_skip()
    rng = np.random.default_rng(zlib.crc32(repr((shape, "dtv", params)).encode()))
    x = (rng.random(shape) * 2 - 0.5).astype(np.float32)
    r = rng.random(shape).astype(np.float32)
    p = dict(params, ref=r)
    a, ia = oracle.cpu_ref().run("dtv", x, p)
    b, ib = ref.run("dtv", x, p)
    assert bits_equal(a, b)
    assert bits_equal(ia, ib)


def test_rof_lambda_array_matches_reference():
    ref = _ref_or_skip()
    rng = np.random.default_rng(5)
    x = rng.random((12, 14, 9)).astype(np.float32)
    lam = (0.01 + 0.05 * rng.random(x.shape)).astype(np.float32)
    a, ia = oracle.cpu_ref().rof(x, 0.0, 30, 0.002, 0.0, lam_arr=lam)
    b, ib = ref.rof(x, 0.0, 30, 0.002, 0.0, lam_arr=lam)
    assert bits_equal(a, b) and bits_equal(ia, ib)
