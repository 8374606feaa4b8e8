"""Per-view transport grids cropped to the view's beam (helical scans and uneven batching).

The reference solves every view on one covering grid (``citrus/scatter_solver/scatter_solver.py:495-507``). In a
helical scan each view's cone beam reaches only a slab of slices, so its first-collision source, and with it every
collided order, vanishes outside that slab; a view can then be solved on the slab alone. ``beam_slice_range`` bounds
the slab conservatively: a ray from the source (at the view's z shift) to panel row v reaches
``|z - z_source| <= v_max rho / SDD`` at horizontal distance rho from the source (a flat panel sits at horizontal
distance >= SDD, a curved one at exactly SDD), and rho <= SID + r over a volume of horizontal half-diagonal r. The
range is widened to whole transport cells. Rays from the source to any point of the slab stay inside the beam, so the
uncollided fluence there does not depend on material outside it: the cropped solve equals the covering-grid solve of
the same densities with everything outside the slab removed.
"""

from __future__ import annotations

import math
from dataclasses import replace

from torch import Tensor

from .geometry import ConeGeometry, Grid


def beam_slice_range(geometry: ConeGeometry, angle: float, grid: Grid, multiple: int = 1) -> tuple[int, int]:
    """Slices [k0, k1) of ``grid`` the view's beam can reach, widened outward to multiples of ``multiple``."""
    dx, dy, dz = grid.deltas
    r = 0.5 * math.hypot(grid.num_cols * dx, grid.num_rows * dy)
    v_max = 0.5 * geometry.num_det_rows * geometry.delta_det_row + abs(geometry.det_row_offset)
    half = v_max * (geometry.source_iso_dist + r) / geometry.source_detector_dist
    zs = geometry.z_shift(angle)
    z0 = grid.lower_corner()[2]
    ns = grid.num_slices
    k0 = max(0, math.floor((zs - half - z0) / dz))
    k1 = min(ns, math.ceil((zs + half - z0) / dz))
    m = max(1, int(multiple))
    k0 = (k0 // m) * m
    k1 = min(ns, -(-k1 // m) * m)
    if k1 <= k0:
        raise ValueError(f"the beam of view {angle} misses the grid's slices")
    return k0, k1


def crop_slices(grid: Grid, rho: Tensor, k0: int, k1: int) -> tuple[Grid, Tensor]:
    """(grid of slices [k0, k1), the densities [M, R, C, k1 - k0] on it): same pitch, recentred in z."""
    ns = grid.num_slices
    if not 0 <= k0 < k1 <= ns:
        raise ValueError(f"slice range [{k0}, {k1}) outside [0, {ns})")
    dz = grid.deltas[2]
    sub = replace(grid, shape=(grid.num_rows, grid.num_cols, k1 - k0),
                  recon_slice_offset=grid.recon_slice_offset + dz * (0.5 * (k0 + k1) - 0.5 * ns))
    return sub, rho[..., k0:k1].contiguous()
