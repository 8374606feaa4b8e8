"""Seeded synthetic phantoms of the benchmark configurations (host-side, float64 CPU).

A water cylinder (1.0 g/cm^3, axis along z) holding 3-6 bone ellipsoids (density U[1.4, 1.9]) and 1-3 air
ellipsoids (0.001 g/cm^3), returned as two material channels (water-like, bone-like) of densities in g/cm^3 — the
``rho [M, R, C, S]`` input of the solver (the reference's WLAB phantom, citrus/utilities.py:104, plays the same role
for its clinical geometry). Also the scan helpers ``view_angles`` and ``helical_scan``.
"""

from __future__ import annotations

import math

import numpy as np
import torch
from torch import Tensor

from .scatter_solver.geometry import Grid


def ellipsoid_phantom(grid: Grid, seed: int) -> Tensor:
    """[2, R, C, S] densities (water channel, bone channel) of a seeded water cylinder with ellipsoid inserts."""
    gen = torch.Generator().manual_seed(seed)

    def u(a, b):
        return a + (b - a) * float(torch.rand((), generator=gen, dtype=torch.float64))

    pos = grid.center_coordinates(dtype=torch.float64)  # [R, C, S, 3]
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2] - grid.recon_slice_offset
    ex, ey, ez = grid.extent
    radius = 0.42 * min(ex, ey)
    half_len = 0.45 * ez
    water = ((x * x + y * y <= radius * radius) & (z.abs() <= half_len)).to(torch.float64)
    bone = torch.zeros_like(water)

    def ellipsoid():
        c = torch.tensor([u(-0.5, 0.5) * radius, u(-0.5, 0.5) * radius, u(-0.6, 0.6) * half_len])
        ax = torch.tensor([u(0.08, 0.22) * radius, u(0.08, 0.22) * radius, u(0.1, 0.3) * half_len])
        t = u(0.0, math.pi)
        ct, st = math.cos(t), math.sin(t)
        dx, dy, dz = x - c[0], y - c[1], z - c[2]
        rx, ry = ct * dx + st * dy, -st * dx + ct * dy
        return (rx / ax[0]) ** 2 + (ry / ax[1]) ** 2 + (dz / ax[2]) ** 2 <= 1.0

    for _ in range(int(torch.randint(3, 7, (), generator=gen))):
        inside = ellipsoid() & (water > 0)
        bone[inside] = u(1.4, 1.9)
        water[inside] = 0.0
    for _ in range(int(torch.randint(1, 4, (), generator=gen))):
        inside = ellipsoid() & ((water > 0) | (bone > 0))
        water[inside] = 0.001
        bone[inside] = 0.0
    return torch.stack([water, bone])


def coarsen(rho: Tensor, factor: int) -> Tensor:
    """Cell means over ``factor^3`` blocks of [M, R, C, S] (exact conservative resampling for aligned grids)."""
    M, R, C, S = rho.shape
    f = int(factor)
    return rho.reshape(M, R // f, f, C // f, f, S // f, f).mean(dim=(2, 4, 6))


def view_angles(num_views: int, start_view_angle: float = 0.0, end_view_angle: float = 2.0 * math.pi):
    """[V] float64 view angles ``start + (end - start) i / V`` (utilities.py:1119; the end is excluded)."""
    n = int(num_views)
    if n < 1:
        raise ValueError("num_views must be >= 1")
    if not float(end_view_angle) > float(start_view_angle):
        raise ValueError("end_view_angle must be larger than start_view_angle")
    return float(start_view_angle) + (float(end_view_angle) - float(start_view_angle)) * \
        np.arange(n, dtype=np.float64) / n


def helical_scan(geometry, angles, pitch_cm: float, center_cm: float = 0.0):
    """Geometry of a helical scan with constant pitch (utilities.py:1153): the z shift of view angle a is
    ``center_cm + pitch_cm (a - a_mid) / (2 pi)``, a_mid the middle of the angular range."""
    a = np.asarray(angles.detach().cpu().numpy() if isinstance(angles, Tensor) else angles,
                   dtype=np.float64).reshape(-1)
    middle = 0.5 * (float(a.min()) + float(a.max())) if a.size else 0.0
    return geometry.with_helical_z_shifts(a, float(center_cm) + float(pitch_cm) * (a - middle) / (2.0 * math.pi))
