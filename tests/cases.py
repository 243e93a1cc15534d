"""Deterministic parity cases: odd, anisotropic and oblique grids with textured images and smooth
fields, in the spirit of the reference's known-answer cases (crates/fastwarp-gpu/src/parity.rs:56-240)."""
from __future__ import annotations

import numpy as np

from fastwarp_amd._abi import Grid


def rot_z(a: float) -> np.ndarray:
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1.0]])


def rot_x(a: float) -> np.ndarray:
    c, s = np.cos(a), np.sin(a)
    return np.array([[1.0, 0, 0], [0, c, -s], [0, s, c]])


def grid(size, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0), oblique: float = 0.0) -> Grid:
    d = rot_z(oblique) @ rot_x(0.6 * oblique) if oblique else np.eye(3)
    return Grid.make(size, spacing, origin, d)


def coords(g: Grid):
    nx, ny, nz = g.shape_xyz
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return i.reshape(-1).astype(float), j.reshape(-1).astype(float), k.reshape(-1).astype(float)


def image(g: Grid, seed: int, noise: float = 0.05, f32: bool = True) -> np.ndarray:
    """A smooth texture in about [0.05, 0.95] plus uniform noise, x-fastest."""
    i, j, k = coords(g)
    rng = np.random.default_rng(seed)
    s = (
        0.5
        + 0.22 * np.sin(0.43 * i + 0.21 * j + seed) * np.cos(0.35 * k + 0.17 * j)
        + 0.14 * np.sin(0.27 * i - 0.33 * k + 0.5 * seed)
    )
    v = np.clip(s + noise * (rng.random(i.size) - 0.5), 0.02, 0.98)
    return v.astype(np.float32).astype(np.float64) if f32 else v


def field(g: Grid, amp: float, seed: int, f32: bool = True) -> np.ndarray:
    """A smooth displacement field (mm) with a little noise, shape (n, 3)."""
    nx, ny, nz = g.shape_xyz
    i, j, k = coords(g)
    u, v, w = i / nx, j / ny, k / nz
    rng = np.random.default_rng(seed)
    nz_ = rng.random((i.size, 3)) - 0.5
    out = np.stack(
        [
            amp * (np.sin(6.0 * v + 2.0 * w) + 0.1 * nz_[:, 0]),
            amp * (0.7 * np.cos(5.0 * w + 3.0 * u) + 0.1 * nz_[:, 1]),
            amp * (0.5 * np.sin(4.0 * u - 3.0 * v) + 0.1 * nz_[:, 2]),
        ],
        axis=1,
    )
    return out.astype(np.float32).astype(np.float64) if f32 else out


def rel_err(got: np.ndarray, want: np.ndarray) -> float:
    """Largest absolute error over the largest reference magnitude (parity.rs `compare`)."""
    rng = max(float(np.max(np.abs(want))), 1e-12)
    return float(np.max(np.abs(np.asarray(got) - np.asarray(want)))) / rng


def interior(g: Grid, margin: int = 1) -> np.ndarray:
    nx, ny, nz = g.shape_xyz
    i, j, k = coords(g)
    return (i >= margin) & (j >= margin) & (k >= margin) & (i < nx - margin) & (j < ny - margin) & (k < nz - margin)
