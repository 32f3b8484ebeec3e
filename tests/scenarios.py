"""Seeded synthetic inputs shared by the tests and the benchmark (reference layout:
padded by g = 3, j fastest; numpy shape (nx + 6, ny + 6))."""
from __future__ import annotations

import numpy as np

G = 3


def centres(nx, ny, dx, dy, base=(0.0, 0.0), g=G):
    """Cell-centre coordinates x = base + (i + 0.5 - g) dx (reference
    tests/LargeTests/test_mac2d_large.cpp:473-474)."""
    x = base[0] + (np.arange(nx + 2 * g) + 0.5 - g) * dx
    y = base[1] + (np.arange(ny + 2 * g) + 0.5 - g) * dy
    return np.meshgrid(x, y, indexing="ij")


def faces_u(nx, ny, dx, dy, base=(0.0, 0.0), g=G):
    """West-face coordinates of each cell (u[idx(i, j)] sits at x = (i - g) dx)."""
    x = base[0] + (np.arange(nx + 2 * g) - g) * dx
    y = base[1] + (np.arange(ny + 2 * g) + 0.5 - g) * dy
    return np.meshgrid(x, y, indexing="ij")


def faces_v(nx, ny, dx, dy, base=(0.0, 0.0), g=G):
    x = base[0] + (np.arange(nx + 2 * g) + 0.5 - g) * dx
    y = base[1] + (np.arange(ny + 2 * g) - g) * dy
    return np.meshgrid(x, y, indexing="ij")


def circle_sdf(nx, ny, dx, dy, cx, cy, r, heavy_inside=True):
    X, Y = centres(nx, ny, dx, dy)
    d = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2) - r
    return np.ascontiguousarray(-d if heavy_inside else d)


def single_vortex(nx, ny, dx, dy, t=0.0, T=8.0):
    """u = -sin^2(pi x) sin(2 pi y) cos(pi t / T), v = sin^2(pi y) sin(2 pi x) cos(pi t / T)
    sampled on the MAC faces."""
    Xu, Yu = faces_u(nx, ny, dx, dy)
    Xv, Yv = faces_v(nx, ny, dx, dy)
    c = np.cos(np.pi * t / T)
    u = -np.sin(np.pi * Xu) ** 2 * np.sin(2 * np.pi * Yu) * c
    v = np.sin(np.pi * Yv) ** 2 * np.sin(2 * np.pi * Xv) * c
    return np.ascontiguousarray(u), np.ascontiguousarray(v)


def random_droplets(seed, nx, ny, dx, dy, n_range=(1, 20), r_range=(0.02, 0.1), lx=1.0, ly=1.0):
    """1-20 non-overlapping circles, radii U[0.02, 0.1], centres uniform; phi = max over
    circles of (r_k - |x - c_k|) (heavy phase inside)."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(n_range[0], n_range[1] + 1))
    circles = []
    tries = 0
    while len(circles) < n and tries < 10000:
        tries += 1
        r = rng.uniform(*r_range)
        c = rng.uniform([r, r], [lx - r, ly - r])
        if all(np.hypot(*(c - c2)) > r + r2 + 2 * max(dx, dy) for c2, r2 in circles):
            circles.append((c, r))
    X, Y = centres(nx, ny, dx, dy)
    phi = np.full(X.shape, -np.inf)
    for c, r in circles:
        phi = np.maximum(phi, r - np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2))
    return np.ascontiguousarray(phi), circles


def density_ratio(seed):
    """log-uniform in [10, 1000]"""
    rng = np.random.default_rng(seed + 7919)
    return float(10 ** rng.uniform(1.0, 3.0))


def smooth_velocity(seed, nx, ny, dx, dy, modes=(8, 16)):
    """Seeded smooth random velocity (sum of Fourier modes) on the MAC faces."""
    rng = np.random.default_rng(seed + 104729)
    m = int(rng.integers(modes[0], modes[1] + 1))
    Xu, Yu = faces_u(nx, ny, dx, dy)
    Xv, Yv = faces_v(nx, ny, dx, dy)
    u = np.zeros_like(Xu)
    v = np.zeros_like(Xv)
    for _ in range(m):
        kx, ky = rng.integers(1, 6, size=2)
        a, b = rng.standard_normal(2)
        px, py = rng.uniform(0, 2 * np.pi, size=2)
        u += a * np.sin(np.pi * kx * Xu + px) * np.cos(np.pi * ky * Yu + py)
        v += b * np.cos(np.pi * kx * Xv + px) * np.sin(np.pi * ky * Yv + py)
    return np.ascontiguousarray(u / m), np.ascontiguousarray(v / m)


def random_field(seed, shape, scale=1.0):
    return np.ascontiguousarray(np.random.default_rng(seed).standard_normal(shape) * scale)
