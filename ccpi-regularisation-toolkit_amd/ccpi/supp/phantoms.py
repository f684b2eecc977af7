"""Seeded synthetic phantoms for the benchmark and test workloads.

Every generator draws from ``numpy.random.Generator(PCG64(seed)).random()``
only, in a fixed order, and makes Gaussian noise with its own Box-Muller
transform, so a (generator, seed, shape) triple always gives the same volume
on one machine.  Coordinates are normalised to [0, 1) per axis with voxel
centres at (i + 0.5) / N.  Arrays are float32, C-contiguous, numpy axis order
(Z, Y, X) -- X is the fastest axis, as the regularisers expect.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "gaussian_noise",
    "shapes_2d",
    "ellipsoids_3d",
    "ramp_ellipsoids_3d",
    "noisy",
]


def _rng(seed):
    return np.random.Generator(np.random.PCG64(seed))


def gaussian_noise(rng, n, sigma):
    """n samples of N(0, sigma^2) by Box-Muller from uniform draws."""
    m = (n + 1) // 2
    u1 = rng.random(m)
    u2 = rng.random(m)
    u1 = np.where(u1 <= 0.0, np.finfo(np.float64).tiny, u1)
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.concatenate([r * np.cos(2.0 * np.pi * u2), r * np.sin(2.0 * np.pi * u2)])[:n]
    return (sigma * z).astype(np.float32)


def _axis_coords(n):
    return (np.arange(n, dtype=np.float64) + 0.5) / n


def _box(c, r, n):
    """Index range [lo, hi) of voxels whose centre lies in [c - r, c + r]."""
    lo = max(int(np.floor((c - r) * n - 0.5)), 0)
    hi = min(int(np.ceil((c + r) * n + 0.5)) + 1, n)
    return lo, max(hi, lo)


def shapes_2d(shape=(512, 512), n_shapes=20, seed=1):
    """Axis-aligned rectangles and ellipses (p = 1/2 each), centres U[0,1]^2,
    half-extents U[0.05, 0.3], intensities U[0,1], combined by max on a zero
    background."""
    rng = _rng(seed)
    ny, nx = shape
    img = np.zeros(shape, dtype=np.float32)
    ys, xs = _axis_coords(ny), _axis_coords(nx)
    for _ in range(n_shapes):
        is_rect = rng.random() < 0.5
        cy, cx = rng.random(), rng.random()
        ry, rx = 0.05 + 0.25 * rng.random(), 0.05 + 0.25 * rng.random()
        val = np.float32(rng.random())
        y0, y1 = _box(cy, ry, ny)
        x0, x1 = _box(cx, rx, nx)
        if y1 <= y0 or x1 <= x0:
            continue
        dy = (ys[y0:y1, None] - cy) / ry
        dx = (xs[None, x0:x1] - cx) / rx
        inside = (np.abs(dy) <= 1.0) & (np.abs(dx) <= 1.0) if is_rect else (dy * dy + dx * dx <= 1.0)
        sub = img[y0:y1, x0:x1]
        np.maximum(sub, np.where(inside, val, np.float32(0)), out=sub)
    return img


def ellipsoids_3d(shape=(256, 256, 256), n_ellipsoids=40, seed=2):
    """Axis-aligned ellipsoids: centres U[0,1]^3, semi-axes U[0.05, 0.3] per
    axis, intensities U[0,1], combined by max on a zero background."""
    rng = _rng(seed)
    nz, ny, nx = shape
    vol = np.zeros(shape, dtype=np.float32)
    zs, ys, xs = _axis_coords(nz), _axis_coords(ny), _axis_coords(nx)
    for _ in range(n_ellipsoids):
        cz, cy, cx = rng.random(), rng.random(), rng.random()
        rz, ry, rx = (0.05 + 0.25 * rng.random() for _ in range(3))
        val = np.float32(rng.random())
        z0, z1 = _box(cz, rz, nz)
        y0, y1 = _box(cy, ry, ny)
        x0, x1 = _box(cx, rx, nx)
        if z1 <= z0 or y1 <= y0 or x1 <= x0:
            continue
        q = (((zs[z0:z1, None, None] - cz) / rz) ** 2 + ((ys[None, y0:y1, None] - cy) / ry) ** 2
             + ((xs[None, None, x0:x1] - cx) / rx) ** 2)
        sub = vol[z0:z1, y0:y1, x0:x1]
        np.maximum(sub, np.where(q <= 1.0, val, np.float32(0)), out=sub)
    return vol


def ramp_ellipsoids_3d(shape=(192, 512, 512), n_ellipsoids=30, seed=3):
    """Piecewise-linear phantom: ellipsoids (centres U[0,1]^3, semi-axes
    U[0.05, 0.3]) each holding c + g . (x - x0) with c ~ U[0,1] and
    g ~ U[-1,1]/256 per axis in intensity per voxel; painted in order, later
    ellipsoids overwrite earlier ones."""
    rng = _rng(seed)
    nz, ny, nx = shape
    vol = np.zeros(shape, dtype=np.float32)
    zs, ys, xs = _axis_coords(nz), _axis_coords(ny), _axis_coords(nx)
    for _ in range(n_ellipsoids):
        cz, cy, cx = rng.random(), rng.random(), rng.random()
        rz, ry, rx = (0.05 + 0.25 * rng.random() for _ in range(3))
        c = rng.random()
        gz, gy, gx = ((2.0 * rng.random() - 1.0) / 256.0 for _ in range(3))
        z0, z1 = _box(cz, rz, nz)
        y0, y1 = _box(cy, ry, ny)
        x0, x1 = _box(cx, rx, nx)
        if z1 <= z0 or y1 <= y0 or x1 <= x0:
            continue
        q = (((zs[z0:z1, None, None] - cz) / rz) ** 2 + ((ys[None, y0:y1, None] - cy) / ry) ** 2
             + ((xs[None, None, x0:x1] - cx) / rx) ** 2)
        # gradient in intensity per voxel along each index axis, measured from the centre voxel
        ramp = (c + gz * (np.arange(z0, z1)[:, None, None] - cz * nz)
                + gy * (np.arange(y0, y1)[None, :, None] - cy * ny)
                + gx * (np.arange(x0, x1)[None, None, :] - cx * nx)).astype(np.float32)
        sub = vol[z0:z1, y0:y1, x0:x1]
        inside = q <= 1.0
        sub[inside] = ramp[inside]
    return vol


def noisy(clean, sigma, seed):
    """clean + N(0, sigma^2) noise drawn from its own seeded stream."""
    rng = _rng(seed)
    n = clean.size
    return (clean + gaussian_noise(rng, n, sigma).reshape(clean.shape)).astype(np.float32)
