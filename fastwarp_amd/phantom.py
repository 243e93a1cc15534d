"""Seeded synthetic phantoms of the benchmark workloads.

Fixed image: 40 random 3-D ellipsoids (semi-axes 4-20 voxels, random orientation, centres in the
central 80 % box) painted in order with labels 0-3, mapped to class intensities ({0, 0.3, 0.6, 1.0}
T1-like, {1.0, 0.7, 0.4, 0.1} T2-like), Gaussian-smoothed with sigma 1 voxel, plus N(0, 0.02)
noise. Moving image: the fixed image resampled through A(x + u(x)), where A is a random affine
about the grid centre (rotation <= 10 degrees per axis, translation <= 5 mm, scale 0.95-1.05) and
u is white noise Gaussian-smoothed with sigma 8 voxels and scaled to a maximum of `disp_max_mm`;
then clamped at zero with gamma 0.8. Grids are centred on the origin, identity direction.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy import ndimage

from ._abi import Affine, Grid

T1 = (0.0, 0.3, 0.6, 1.0)
T2 = (1.0, 0.7, 0.4, 0.1)


@dataclass
class Pair:
    grid: Grid
    fixed: list[np.ndarray]   # one flat x-fastest f64 array per channel
    moving: list[np.ndarray]
    affine: Affine            # the true A (moving-grid points to fixed-grid points)
    displacement: np.ndarray  # the true u, (n, 3) mm


def _rotation(rng: np.random.Generator) -> np.ndarray:
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
        ]
    )


def labels(shape_xyz, rng: np.random.Generator, count: int = 40) -> np.ndarray:
    nx, ny, nz = shape_xyz
    lab = np.zeros((nz, ny, nx), dtype=np.int8)
    lo = 0.1 * np.array([nx, ny, nz])
    hi = 0.9 * np.array([nx, ny, nz])
    for _ in range(count):
        c = rng.uniform(lo, hi)
        axes = rng.uniform(4.0, 20.0, size=3)
        rot = _rotation(rng)
        label = int(rng.integers(0, 4))
        r = int(np.ceil(axes.max())) + 1
        x0, x1 = max(int(c[0]) - r, 0), min(int(c[0]) + r + 1, nx)
        y0, y1 = max(int(c[1]) - r, 0), min(int(c[1]) + r + 1, ny)
        z0, z1 = max(int(c[2]) - r, 0), min(int(c[2]) + r + 1, nz)
        if x0 >= x1 or y0 >= y1 or z0 >= z1:
            continue
        zz, yy, xx = np.meshgrid(np.arange(z0, z1), np.arange(y0, y1), np.arange(x0, x1), indexing="ij")
        d = np.stack([xx - c[0], yy - c[1], zz - c[2]], axis=-1) @ rot  # body coordinates
        inside = np.sum((d / axes) ** 2, axis=-1) <= 1.0
        lab[z0:z1, y0:y1, x0:x1][inside] = label
    return lab


def make_pair(shape_xyz=(96, 112, 96), spacing: float = 2.0, seed: int = 1, disp_max_mm: float = 4.0,
              channels: int = 1, affine_only: bool = False) -> Pair:
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape_xyz
    origin = tuple(-(n - 1) * spacing / 2.0 for n in shape_xyz)
    grid = Grid.make(shape_xyz, (spacing,) * 3, origin)
    lab = labels(shape_xyz, rng)
    fixed_vols = []
    for c in range(channels):
        table = np.array(T1 if c == 0 else T2)
        v = ndimage.gaussian_filter(table[lab], 1.0, mode="nearest")
        v = v + rng.normal(0.0, 0.02, size=v.shape)
        fixed_vols.append(v)
    # the random affine about the grid centre (the origin)
    ang = np.deg2rad(rng.uniform(-10.0, 10.0, size=3))
    cx, cy, cz = np.cos(ang)
    sx, sy, sz = np.sin(ang)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    mat = rz @ rx @ ry @ np.diag(rng.uniform(0.95, 1.05, size=3))
    trans = rng.uniform(-5.0, 5.0, size=3)
    aff = Affine.make(mat, trans, (0.0, 0.0, 0.0))
    if affine_only or disp_max_mm <= 0:
        u = np.zeros((nz, ny, nx, 3))
    else:
        u = np.stack([ndimage.gaussian_filter(rng.normal(size=(nz, ny, nx)), 8.0, mode="nearest") for _ in range(3)], axis=-1)
        u *= disp_max_mm / np.max(np.linalg.norm(u, axis=-1))
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    pts = np.stack([xx, yy, zz], axis=-1) * spacing + np.array(origin)  # physical, (nz, ny, nx, 3)
    q = (pts + u) @ mat.T + trans
    idx = (q - np.array(origin)) / spacing  # continuous x, y, z index in the fixed grid
    coords = [idx[..., 2], idx[..., 1], idx[..., 0]]
    moving_vols = []
    for v in fixed_vols:
        m = ndimage.map_coordinates(v, coords, order=1, mode="constant", cval=0.0)
        moving_vols.append(np.maximum(m, 0.0) ** 0.8)
    flat = lambda a: np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
    return Pair(grid, [flat(v) for v in fixed_vols], [flat(v) for v in moving_vols], aff, u.reshape(-1, 3))


# The benchmark configurations: the 2 mm plain registration and the 1 mm full normalisation.
CONFIGS = {
    "2mm": dict(shape_xyz=(96, 112, 96), spacing=2.0, seed=1, disp_max_mm=4.0),
    "1mm": dict(shape_xyz=(182, 218, 182), spacing=1.0, seed=2, disp_max_mm=6.0),
}
