"""Uncollided fluence on the HIP kernels.

Keeps the interface of ``citrus/scatter_solver/uncollided.py``: ``present_materials`` (:27), ``uncollided_fluence``
(:99), ``uncollided_path_lengths`` (:137) and ``uncollided_from_path_lengths`` (:160):
``phi_UC,g(r) = b(Omega_UC(r)) / |r - r0|^2 sum_{E in g} w_E exp(-sum_m P_m(r) m_m(E) - mu_f(E) P_f)``. The upstream
density line integrals come from the projector's ``upstream_cumulative`` (one traversal launch), the rest is one
fused launch (``citrus_uncollided_fluence``).
"""

from __future__ import annotations

from typing import Optional, Sequence, Union

import torch
from torch import Tensor

from .. import _hip
from .cross_sections import MultigroupLibrary
from .geometry import ConeGeometry, Grid
from .ray_tracing import HipProjectors
from .source import SourceModel


def present_materials(rho: Tensor) -> list[int]:
    """Library indices of the materials with a nonzero density map (one host synchronization)."""
    return [int(m) for m in torch.nonzero(rho.reshape(rho.shape[0], -1).abs().amax(dim=1) > 0).flatten().tolist()]


def uncollided_path_lengths(projectors, rho: Tensor, grid: Grid, geometry: ConeGeometry, angle: float,
                            points: Optional[Grid] = None) -> Tensor:
    """[M', *points.shape] upstream density line integrals in g/cm^2 to the voxel centres of ``points``."""
    point_grid = grid if points is None else points
    if point_grid == grid:
        return projectors.upstream_cumulative(rho, grid, geometry, angle)
    r = point_grid.center_coordinates(rho.device, torch.float64).reshape(-1, 3)
    return projectors.upstream_to_points(rho, grid, geometry, angle, r).reshape(-1, *point_grid.shape)


def uncollided_from_path_lengths(library: MultigroupLibrary, path_lengths: Tensor,
                                 materials: Optional[Sequence[int]], points: Grid, geometry: ConeGeometry,
                                 angle: float, source: SourceModel) -> Tensor:
    """[G, *points.shape] uncollided fluence from the upstream line integrals (uncollided.py:160)."""
    device = path_lengths.device
    Mp = int(path_lengths.shape[0])
    fine_mu = library.fine_mu_t if materials is None else library.fine_mu_t[[int(m) for m in materials]]
    if int(fine_mu.shape[0]) != Mp:
        raise ValueError(f"path_lengths has {Mp} materials, expected {int(fine_mu.shape[0])}")
    paths = path_lengths.reshape(Mp, -1).contiguous()
    _hip.require_gpu_tensor(paths, "path_lengths")
    mu = fine_mu.to(device, torch.float32).contiguous()
    grp = library.fine_group.to(device, torch.int32).contiguous()
    w = library.fine_weight.to(device, torch.float32).contiguous()
    profile, thickness, mu_f = source.device_tables(geometry, device)
    values, ints, table = source.host_block(geometry, angle)
    x0, y0, z0 = points.lower_corner()
    dx, dy, dz = points.deltas
    import numpy as np

    lattice = np.array([x0 + 0.5 * dx, y0 + 0.5 * dy, z0 + 0.5 * dz, dx, dy, dz], dtype=np.float64)
    shape = np.array(points.shape, dtype=np.int32)
    G = library.num_groups
    out = torch.empty((G,) + tuple(points.shape), device=device, dtype=torch.float32)
    with torch.cuda.device(device):
        _hip.check(_hip.library().citrus_uncollided_fluence(
            paths.data_ptr(), Mp, lattice.ctypes.data, shape.ctypes.data, mu.data_ptr(), grp.data_ptr(), w.data_ptr(),
            int(mu.shape[1]), G, values.ctypes.data, ints.ctypes.data, None if table is None else table.ctypes.data,
            _hip.ptr(profile), _hip.ptr(thickness), _hip.ptr(mu_f), out.data_ptr(), _hip.stream_of(device)),
            "citrus_uncollided_fluence")
    return out


def uncollided_fluence(library: MultigroupLibrary, rho: Tensor, grid: Grid, geometry: ConeGeometry, angle: float,
                       source: SourceModel, projectors=None, points: Optional[Grid] = None,
                       materials: Union[Sequence[int], str, None] = "auto") -> Tensor:
    """[G, *points.shape] uncollided scalar fluence at the voxel centres of ``points`` (uncollided.py:99)."""
    if int(rho.shape[0]) != library.num_materials:
        raise ValueError(f"rho has {int(rho.shape[0])} materials, the library {library.num_materials}")
    projectors = HipProjectors() if projectors is None else projectors
    point_grid = grid if points is None else points
    present = present_materials(rho) if isinstance(materials, str) else (None if materials is None
                                                                        else [int(m) for m in materials])
    if present is not None and not present:  # vacuum: trace one all-zero map (zero path, no attenuation)
        present = [0]
    traced = rho if present is None else rho.index_select(0, torch.tensor(present, device=rho.device))
    paths = uncollided_path_lengths(projectors, traced.contiguous(), grid, geometry, angle, point_grid)
    return uncollided_from_path_lengths(library, paths, present, point_grid, geometry, angle, source)
