"""Checks on the CPU oracle that the mathematics fixes, independent of the
reference fixtures.

* Scale equivariance, bit-exact for a factor of 2 (every scaling by 2 is exact
  in binary floating point and each solver's step sizes scale accordingly):
  FGP(2f, 2 lambda) = 2 FGP(f, lambda), PD likewise, SB(2f, 2 mu) = 2 SB(f, mu),
  TGV(2f; lambda, 2 alpha1, 2 alpha0) = 2 TGV(f; lambda, alpha1, alpha0).
  A dropped or misplaced step-size factor breaks these.
* Constant images are fixed points: exactly for ROF and FGP, to a few ulps for
  PD, SB and TGV (whose prox steps divide c (1 + t) by (1 + t)).
* TGV in 3D with a singleton y axis equals TGV in 2D on the (z, x) plane, bit
  for bit: every y-derivative is zero, so P2, V2, Q22, Q12, Q23 stay zero and
  the remaining terms are summed in the same order.  A transposed operand or a
  wrong component in the symmetric gradient breaks this.
* Zero iterations return the input (the FGP output buffer is left alone).
* FGP-dTV: scale equivariance in (f, lambda) as FGP; the reference image's
  normalised gradient is unchanged by scaling (ref, eta) together by 2, so the
  output is too (a misplaced eta or a wrong square-root argument breaks it);
  with a zero reference image the gradient field B is zero, no projection
  happens, and 2D dTV equals 2D FGP (same 1/(8 lambda) step) up to the
  rounding of t_k, which dTV takes through a double square root.
* NDF, synchronous explicit scheme: with sigma = 0 (linear diffusion) a
  product of cosines cos(pi a i/(nx-1)) cos(pi b j/(ny-1)) is an eigenvector
  of the mirrored Laplacian, so from u = f = v every iterate is s_n v with
  s_{n+1} = s_n + tau (lambda mu s_n - (s_n - 1)); the oracle must follow that
  closed form (a wrong boundary mirror, a swapped tau / lambda or an in-place
  update breaks it).  The threshold-linear penalty with a threshold above
  every difference is linear diffusion bit for bit.
* Diff4th: constant images are fixed points bit for bit; scaling f by 2 and
  sigma by 2 scales the output by 2 bit for bit over a few iterations (while
  no |grad|^2 reaches the fixed 1e-7 floor, which does not scale).
* LLT-ROF (on smooth ramps, whose differences stay far above the flux
  floors): with lambda_LLT = 0 it is ROF-TV with the floor 1e-12 instead of
  1e-8, within 1e-6 of the ROF oracle; scaling f and both lambdas by 2
  scales the output by 2 within 1e-6.
* An axis of extent 1 (where the reference's mirrored neighbour reads out of
  bounds) equals, bit for bit, the same axis at extent 2 with the data
  duplicated, where the reference's mirror is defined and lands on the
  duplicate: every difference along the axis is zero either way (ROF, PD,
  SB; checked against the reference library on the extent-2 input too).
* The FP64 build agrees with the FP32 build to FP32 accuracy.
"""
import numpy as np
import pytest

import oracle
from conftest import bits_equal

RNG = np.random.default_rng(7)
F2 = RNG.random((23, 31)).astype(np.float32)
F3 = RNG.random((7, 9, 13)).astype(np.float32)


@pytest.mark.parametrize("f", [F2, F3], ids=["2d", "3d"])
@pytest.mark.parametrize("method,nonneg", [(0, 0), (1, 1)])
def test_fgp_scale_equivariance(f, method, nonneg):
    lib = oracle.cpu_ref()
    a, _ = lib.fgp(f, 0.03, 40, 0.0, method, nonneg)
    b, _ = lib.fgp(2 * f, 0.06, 40, 0.0, method, nonneg)
    assert bits_equal(b, 2 * a)


@pytest.mark.parametrize("f", [F2, F3], ids=["2d", "3d"])
@pytest.mark.parametrize("method,nonneg", [(0, 1), (1, 0)])
def test_pd_scale_equivariance(f, method, nonneg):
    lib = oracle.cpu_ref()
    a, _ = lib.pd(f, 0.03, 40, 0.0, 8.0, method, nonneg)
    b, _ = lib.pd(2 * f, 0.06, 40, 0.0, 8.0, method, nonneg)
    assert bits_equal(b, 2 * a)


@pytest.mark.parametrize("f", [F2, F3], ids=["2d", "3d"])
@pytest.mark.parametrize("method", [0, 1])
def test_sb_scale_equivariance(f, method):
    lib = oracle.cpu_ref()
    a, _ = lib.sb(f, 0.05, 20, 0.0, method)
    b, _ = lib.sb(2 * f, 0.1, 20, 0.0, method)
    assert bits_equal(b, 2 * a)


@pytest.mark.parametrize("f", [F2, F3], ids=["2d", "3d"])
def test_tgv_scale_equivariance(f):
    lib = oracle.cpu_ref()
    a, _ = lib.tgv(f, 0.05, 1.0, 2.0, 40, 12.0, 0.0)
    b, _ = lib.tgv(2 * f, 0.05, 2.0, 4.0, 40, 12.0, 0.0)
    assert bits_equal(b, 2 * a)


@pytest.mark.parametrize("shape", [(17, 19), (5, 6, 7)], ids=["2d", "3d"])
def test_constant_image_is_fixed(shape):
    lib = oracle.cpu_ref()
    c = np.full(shape, 0.37, dtype=np.float32)
    y, _ = lib.rof(c, 0.02, 30, 0.001, 0.0)
    assert np.array_equal(y, c)
    y, _ = lib.fgp(c, 0.02, 30, 0.0, 0, 1)
    assert np.array_equal(y, c)
    for y in (lib.pd(c, 0.02, 30, 0.0, 12.0, 0, 0)[0], lib.sb(c, 0.02, 30, 0.0, 0)[0],
              lib.tgv(c, 0.02, 1.0, 2.0, 30, 12.0, 0.0)[0]):
        assert np.max(np.abs(y - c)) <= 8 * np.spacing(np.float32(0.37))


def test_tgv_singleton_axis_reduces_to_2d():
    lib = oracle.cpu_ref()
    f2 = RNG.random((11, 14)).astype(np.float32)           # (Y', X) plane
    f3 = np.ascontiguousarray(f2[:, None, :])                # (Z, 1, X) volume
    a, ia = lib.tgv(f2, 0.05, 1.0, 2.0, 60, 12.0, 0.0)
    b, ib = lib.tgv(f3, 0.05, 1.0, 2.0, 60, 12.0, 0.0)
    assert bits_equal(b[:, 0, :], a)


def test_zero_iterations_identity():
    lib = oracle.cpu_ref()
    for y in (lib.rof(F3, 0.02, 0, 0.001, 0.0)[0], lib.pd(F3, 0.02, 0, 0.0, 8.0, 0, 0)[0],
              lib.sb(F3, 0.02, 0, 0.0, 0)[0], lib.tgv(F3, 0.02, 1.0, 2.0, 0, 12.0, 0.0)[0]):
        assert bits_equal(y, F3)
    out = np.full_like(F3, 5.0)
    y, info = lib.fgp(F3, 0.02, 0, 1e-3, 0, 0, out=out)
    assert np.all(y == 5.0) and info[0] == 0


@pytest.mark.parametrize("method", ["rof", "fgp", "pd", "sb", "tgv"])
def test_fp64_build_agrees_with_fp32(method):
    params = {
        "rof": {"lambda": 0.02, "iters": 50, "tau": 0.001, "tol": 0.0},
        "fgp": {"lambda": 0.02, "iters": 50, "tol": 0.0, "method": 0, "nonneg": 1},
        "pd": {"lambda": 0.02, "iters": 50, "tol": 0.0, "L": 12.0, "method": 0, "nonneg": 1},
        "sb": {"mu": 0.02, "iters": 20, "tol": 0.0, "method": 0},
        "tgv": {"lambda": 0.02, "alpha1": 1.0, "alpha0": 2.0, "iters": 50, "L2": 12.0, "tol": 0.0},
    }[method]
    a, _ = oracle.cpu_ref("f32").run(method, F3, params)
    b, _ = oracle.cpu_ref("f64").run(method, F3.astype(np.float64), params)
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    assert err < 1e-5, err


ENERGY_SHAPES = [(37, 53), (11, 13, 17), (9, 16, 23)]


@pytest.mark.skipif(oracle.reference() is None, reason="reference library not built (make -C oracle ref)")
@pytest.mark.parametrize("shape", ENERGY_SHAPES, ids=["2d-37x53", "3d-11x13x17", "3d-9x16x23"])
@pytest.mark.parametrize("etype", [1, 2], ids=["grad+data", "grad"])
def test_tv_energy_matches_reference(shape, etype):
    """cr_tv_energy_f64 against the reference's TV_energy2D/3D (utils.c:64-117).

    The reference sums n positive terms serially in float; every term is
    itself within a few ulps (powf, sqrtf, the products), so its total is
    within (n + 8) * 2^-24 of the exact sum, relatively (first-order bound for
    a float sum of non-negative terms).  The FP64 sum is exact to ~1e-15.
    Non-cubic shapes make a transposed axis, a wrong boundary rule or a
    dropped factor miss by far more than that bound; type 2 pins the gradient
    term alone, type 1 adds the fidelity term with u0 independent of u."""
    u = RNG.random(shape).astype(np.float32)
    u0 = RNG.random(shape).astype(np.float32)
    lam = 0.07
    ours = oracle.tv_energy_f64(u, u0, lam, etype)
    ref = oracle.reference_tv_energy(u, u0, lam, etype)
    bound = (u.size + 8) * 2.0 ** -24
    assert abs(ours - ref) <= bound * abs(ours), (ours, ref, abs(ours - ref) / abs(ours), bound)


@pytest.mark.parametrize("shape", [(5, 7, 9), (1, 6, 4), (6, 11)])
def test_tv_energy_linear_ramp_closed_form(shape):
    """u = x + 2 y + 2 z (integers, so every difference is exact): the forward
    gradient is (1, 2, 2) except that the replicate boundary zeroes the
    component along each axis at its last index, so |grad u| is sqrt of the
    sum of the squares of the components still present; counting the points
    of each boundary class gives the gradient term in closed form."""
    import itertools
    dims = tuple(shape)[::-1] + ((1,) if len(shape) == 2 else ())   # (nx, ny, nz)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    u = (x + 2 * y + 2 * z).astype(np.float32).reshape(shape)
    comp = (1.0, 2.0, 2.0)
    grad = 0.0
    for last in itertools.product((False, True), repeat=3):   # at the last index along x, y, z?
        count = 1
        for n, at_last in zip(dims, last):
            count *= 1 if at_last else n - 1
        mag = np.sqrt(sum(c * c for c, at_last in zip(comp, last) if not at_last))
        grad += count * mag
    lam = 0.25
    assert oracle.tv_energy_f64(u, u, lam, 2) == pytest.approx(2 * lam * grad, rel=1e-13)
    u0 = u + np.float32(0.5)
    assert oracle.tv_energy_f64(u, u0, lam, 1) == pytest.approx(2 * lam * grad + 0.25 * u.size, rel=1e-13)


@pytest.mark.parametrize("f", [F2, F3], ids=["2d", "3d"])
@pytest.mark.parametrize("method,nonneg", [(0, 0), (1, 1)])
def test_dtv_scale_equivariance(f, method, nonneg):
    lib = oracle.cpu_ref()
    ref = np.flip(f).copy()
    a, ia = lib.dtv(f, ref, 0.03, 40, 0.0, 0.05, method, nonneg)
    b, ib = lib.dtv(2 * f, ref, 0.06, 40, 0.0, 0.05, method, nonneg)
    assert bits_equal(b, 2 * a)


@pytest.mark.parametrize("f", [F2, F3], ids=["2d", "3d"])
def test_dtv_reference_scale_invariance(f):
    lib = oracle.cpu_ref()
    ref = (np.roll(f, 3) * 1.7).astype(np.float32)
    a, _ = lib.dtv(f, ref, 0.03, 30, 0.0, 0.1, 0, 0)
    b, _ = lib.dtv(f, 2 * ref, 0.03, 30, 0.0, 0.2, 0, 0)
    assert bits_equal(a, b)
    c, _ = lib.dtv(f, ref, 0.03, 30, 0.0, 0.2, 0, 0)   # eta alone does matter
    assert not bits_equal(a, c)


@pytest.mark.parametrize("method", [0, 1])
def test_dtv_zero_reference_is_fgp_2d(method):
    lib = oracle.cpu_ref()
    a, _ = lib.dtv(F2, np.zeros_like(F2), 0.03, 60, 0.0, 0.1, method, 0)
    b, _ = lib.fgp(F2, 0.03, 60, 0.0, method, 0)
    err = np.linalg.norm(a.astype(np.float64) - b) / np.linalg.norm(b.astype(np.float64))
    assert err < 1e-6, err
    assert not np.allclose(a, F2, atol=1e-3)   # and it did smooth


def _eig2d(nx, ny, a, b):
    i = np.arange(nx)
    j = np.arange(ny)
    vx = np.cos(np.pi * a * i / (nx - 1))
    vy = np.cos(np.pi * b * j / (ny - 1))
    mu = -(2 - 2 * np.cos(np.pi * a / (nx - 1))) - (2 - 2 * np.cos(np.pi * b / (ny - 1)))
    return np.outer(vy, vx), mu


@pytest.mark.parametrize("a,b", [(1, 0), (2, 3), (5, 1)])
def test_ndf_linear_eigenvector_closed_form(a, b):
    lib = oracle.cpu_ref()
    v, mu = _eig2d(40, 33, a, b)
    f = v.astype(np.float32)
    lam, tau, n = 1.0, 0.1, 30
    u, _ = lib.ndf(f, lam, 0.0, n, tau, 1, 0.0)
    s = 1.0
    for _ in range(n):
        s = s + tau * (lam * mu * s - (s - 1.0))
    assert np.abs(u - s * v).max() < 2e-6, np.abs(u - s * v).max()
    # the serial in-place order is a different scheme: it leaves the closed form
    w, _ = lib.ndf(f, lam, 0.0, n, tau, 1, 0.0, in_place=True)
    assert np.abs(w - s * v).max() > 2e-5


@pytest.mark.parametrize("f", [F2, F3], ids=["2d", "3d"])
def test_ndf_threshold_linear_above_all_differences_is_linear(f):
    lib = oracle.cpu_ref()
    a, _ = lib.ndf(f, 0.03, 0.0, 30, 0.02, 1, 0.0)
    b, _ = lib.ndf(f, 0.03, 10.0, 30, 0.02, 4, 0.0)
    assert bits_equal(a, b)


@pytest.mark.parametrize("shape", [(17, 19), (5, 6, 7)], ids=["2d", "3d"])
def test_d4_constant_image_is_fixed(shape):
    lib = oracle.cpu_ref()
    x = np.full(shape, 0.37, dtype=np.float32)
    y, _ = lib.d4(x, 0.8, 0.02, 20, 1e-3, 0.0)
    assert bits_equal(y, x)


@pytest.mark.parametrize("f", [F2, F3], ids=["2d", "3d"])
def test_d4_scale_equivariance(f):
    # a few iterations: on the mirrored edges two central differences are 0,
    # and later the third can fall under the fixed 1e-7 floor in one run and
    # not in the scaled one
    lib = oracle.cpu_ref()
    a, _ = lib.d4(f, 1.0, 0.05, 5, 1e-3, 0.0)
    b, _ = lib.d4(2 * f, 1.0, 0.1, 5, 1e-3, 0.0)
    assert bits_equal(b, 2 * a)


def _ramp(shape):
    # differences bounded away from 0, so the 1e-12 / 1e-8 flux floors stay negligible
    g = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    v = sum((0.05 + 0.02 * d) * g[d] for d in range(len(shape))) + 0.01 * np.sin(0.7 * g[-1]) * np.cos(0.3 * g[0])
    return v.astype(np.float32)


@pytest.mark.parametrize("shape", [(23, 31), (7, 9, 13)], ids=["2d", "3d"])
def test_llt_without_second_order_term_is_rof(shape):
    lib = oracle.cpu_ref()
    f = _ramp(shape)
    a, _ = lib.llt(f, 0.03, 0.0, 40, 0.002, 0.0)
    b, _ = lib.rof(f, 0.03, 40, 0.002, 0.0)
    err = np.abs(a.astype(np.float64) - b).max()
    assert err < 1e-6, err
    assert np.abs(a - f).max() > 1e-3   # and it did move


@pytest.mark.parametrize("shape", [(23, 31), (7, 9, 13)], ids=["2d", "3d"])
def test_llt_scale_equivariance(shape):
    lib = oracle.cpu_ref()
    f = _ramp(shape)
    a, _ = lib.llt(f, 0.02, 0.01, 30, 0.005, 0.0)
    b, _ = lib.llt(2 * f, 0.04, 0.02, 30, 0.005, 0.0)
    err = np.abs(b.astype(np.float64) - 2 * a).max()
    assert err < 1e-6, err
    assert np.abs(a - f).max() > 1e-3


EXT1_CASES = [
    ("rof", {"lambda": 0.04, "iters": 20, "tau": 0.002, "tol": 0.0}),
    ("pd", {"lambda": 0.04, "iters": 20, "tol": 0.0, "L": 12.0, "method": 0, "nonneg": 1}),
    ("pd", {"lambda": 0.04, "iters": 20, "tol": 0.0, "L": 12.0, "method": 1, "nonneg": 0}),
    ("sb", {"mu": 0.05, "iters": 10, "tol": 0.0, "method": 0}),
    ("sb", {"mu": 0.05, "iters": 10, "tol": 0.0, "method": 1}),
]


@pytest.mark.parametrize("shape,axis", [((1, 70), 0), ((65, 1), 1), ((9, 1, 66), 1), ((7, 40, 1), 2)],
                         ids=["2d-y1", "2d-x1", "3d-y1", "3d-x1"])
@pytest.mark.parametrize("method,params", EXT1_CASES, ids=[f"{m}{i}" for i, (m, _) in enumerate(EXT1_CASES)])
def test_extent1_axis_equals_duplicated_extent2(shape, axis, method, params):
    x = (np.random.default_rng(11).random(shape) * 2 - 0.5).astype(np.float32)
    x2 = np.ascontiguousarray(np.concatenate([x, x], axis=axis))
    y1, i1 = oracle.cpu_ref().run(method, x, params)
    y2, i2 = oracle.cpu_ref().run(method, x2, params)
    assert np.isfinite(y1).all()
    for half in np.split(y2, 2, axis=axis):
        assert bits_equal(np.ascontiguousarray(half), y1)
    ref = oracle.reference()
    if ref is not None:
        y2r, _ = ref.run(method, x2, params)
        assert bits_equal(y2r, y2)


@pytest.mark.parametrize("method,params", [
    ("rof", {"lambda": 0.04, "tau": 0.002, "tol": 1e-30}),
    ("fgp", {"lambda": 0.04, "tol": 1e-30, "method": 0, "nonneg": 1}),
    ("tgv", {"lambda": 0.04, "alpha1": 1.0, "alpha0": 2.0, "L2": 12.0, "tol": 1e-30}),
], ids=["rof", "fgp", "tgv"])
def test_forced_stop_hook(method, params):
    """oracle.forced_stop(j) stops the restatement at the check of iteration
    j: ROF, FGP and TGV then return the iterate after j + 1 iterations, as the
    fixed-count solve of j + 1 iterations does, and report j; outside the
    block the rule is the reference's again (a tolerance of 1e-30 never
    fires)."""
    j = 10
    with oracle.forced_stop(j):
        y, info = oracle.cpu_ref().run(method, F3, dict(params, iters=40))
    assert int(info[0]) == j
    z, _ = oracle.cpu_ref().run(method, F3, dict(params, iters=j + 1, tol=0.0))
    assert np.array_equal(y.view(np.uint32), z.view(np.uint32))
    _, info2 = oracle.cpu_ref().run(method, F3, dict(params, iters=40))
    assert int(info2[0]) == 40
