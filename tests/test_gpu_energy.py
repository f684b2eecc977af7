"""The FP64 objective on the device (cilreg_tv_energy_dev / _f64) against the
oracle's cr_tv_energy_f64, which tests/test_oracle_invariants.py pins to the
reference's TV_energy2D/3D (utils.c:64-117) and to a counted closed form.

Both sum every term in double, in different orders, so they agree to a few
units of 1e-15 relatively; 1e-12 leaves room for the order alone, while a
dropped factor, a wrong boundary or a transposed axis misses by far more.
"""
import numpy as np
import pytest

SHAPES = [(37, 53), (1, 70), (11, 13, 17), (9, 1, 66), (64, 96, 130)]


@pytest.mark.gpu
@pytest.mark.parametrize("shape", SHAPES, ids=[str(s) for s in SHAPES])
@pytest.mark.parametrize("etype", [1, 2])
def test_device_energy_matches_oracle(shape, etype):
    import torch
    import oracle
    from ccpi.filters import device as D
    rng = np.random.default_rng(sum(shape) + etype)
    u = rng.random(shape).astype(np.float32)
    u0 = rng.random(shape).astype(np.float32)
    e_or = oracle.tv_energy_f64(u, u0, 0.07, etype)
    tu, tu0 = torch.from_numpy(u).cuda(), torch.from_numpy(u0).cuda()
    e = D.tv_energy(tu, tu0, 0.07, etype)
    assert abs(e - e_or) <= 1e-12 * abs(e_or), (e, e_or)
    # the FP64 entry point on the same values computes the same terms in the same order
    e64 = D.tv_energy(tu.double(), tu0.double(), 0.07, etype)
    assert e64 == e
    # deterministic: the same call gives the same bits
    assert D.tv_energy(tu, tu0, 0.07, etype) == e


@pytest.mark.gpu
def test_device_energy_matches_reference_library():
    """Directly against the reference's float TV_energy3D where its library is
    built: within its serial float summation bound (n + 8) * 2^-24."""
    import torch
    import oracle
    from ccpi.filters import device as D
    rng = np.random.default_rng(5)
    u = rng.random((11, 13, 17)).astype(np.float32)
    u0 = rng.random((11, 13, 17)).astype(np.float32)
    ref = oracle.reference_tv_energy(u, u0, 0.07, 1)
    if ref is None:
        pytest.skip("reference library not built")
    e = D.tv_energy(torch.from_numpy(u).cuda(), torch.from_numpy(u0).cuda(), 0.07, 1)
    assert abs(e - ref) <= (u.size + 8) * 2.0 ** -24 * abs(e)


@pytest.mark.gpu
def test_device_energy_rejects_bad_arguments():
    import torch
    from ccpi.filters import device as D
    u = torch.zeros((4, 5), device="cuda")
    with pytest.raises(RuntimeError):
        D.tv_energy(u, u, 0.1, 3)
    with pytest.raises(ValueError):
        D.tv_energy(u, u.double(), 0.1, 1)
