"""Voxel grids of the solver (host-side setup).

Mirrors ``Grid`` of ``citrus/scatter_solver/geometry.py:44-230``: volumes are [rows, cols, slices] with columns along
x, rows along y and slices along z, lengths in cm, the grid centred on the rotation axis.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace

import torch
from torch import Tensor


@dataclass(frozen=True)
class Grid:
    """Uniform cuboid voxel grid centred on the rotation axis (geometry.py:44).

    Attributes:
        shape (tuple[int, int, int]): (rows, cols, slices), the extents along (y, x, z) in voxels.
        delta_voxel (float): pitch along x in cm.
        voxel_row_aspect (float): pitch along y over ``delta_voxel``. Defaults to 1.0.
        voxel_slice_aspect (float): pitch along z over ``delta_voxel``. Defaults to 1.0.
        recon_slice_offset (float): z of the grid centre in cm. Defaults to 0.0.
    """

    shape: tuple[int, int, int]
    delta_voxel: float
    voxel_row_aspect: float = 1.0
    voxel_slice_aspect: float = 1.0
    recon_slice_offset: float = 0.0

    @property
    def num_rows(self) -> int:
        return self.shape[0]

    @property
    def num_cols(self) -> int:
        return self.shape[1]

    @property
    def num_slices(self) -> int:
        return self.shape[2]

    @property
    def num_voxels(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]

    @property
    def deltas(self) -> tuple[float, float, float]:
        """Pitches (dx, dy, dz) in cm."""
        d = self.delta_voxel
        return (d, self.voxel_row_aspect * d, self.voxel_slice_aspect * d)

    @property
    def voxel_volume(self) -> float:
        dx, dy, dz = self.deltas
        return dx * dy * dz

    @property
    def extent(self) -> tuple[float, float, float]:
        """Side lengths (x, y, z) of the box in cm."""
        dx, dy, dz = self.deltas
        return (self.num_cols * dx, self.num_rows * dy, self.num_slices * dz)

    def axis_coordinates(self, device=None, dtype=torch.float32) -> tuple[Tensor, Tensor, Tensor]:
        """Voxel-centre coordinates x [cols], y [rows], z [slices] in cm (geometry.py:102)."""
        dx, dy, dz = self.deltas
        nr, nc, ns = self.shape
        x = dx * (torch.arange(nc, device=device, dtype=dtype) - (nc - 1) / 2)
        y = dy * (torch.arange(nr, device=device, dtype=dtype) - (nr - 1) / 2)
        z = dz * (torch.arange(ns, device=device, dtype=dtype) - (ns - 1) / 2) + self.recon_slice_offset
        return x, y, z

    def center_coordinates(self, device=None, dtype=torch.float32) -> Tensor:
        """[rows, cols, slices, 3] voxel-centre positions (x, y, z) in cm."""
        x, y, z = self.axis_coordinates(device, dtype)
        return torch.stack([x[None, :, None].expand(self.shape), y[:, None, None].expand(self.shape),
                            z[None, None, :].expand(self.shape)], dim=-1)

    def lower_corner(self) -> tuple[float, float, float]:
        """(x, y, z) of the minimum corner in cm."""
        dx, dy, dz = self.deltas
        nr, nc, ns = self.shape
        return (-nc * dx / 2, -nr * dy / 2, -ns * dz / 2 + self.recon_slice_offset)

    def axis_edges(self, device=None, dtype=torch.float64) -> tuple[Tensor, Tensor, Tensor]:
        """Voxel boundaries x [cols+1], y [rows+1], z [slices+1] in cm."""
        x0, y0, z0 = self.lower_corner()
        dx, dy, dz = self.deltas
        nr, nc, ns = self.shape
        return (x0 + dx * torch.arange(nc + 1, device=device, dtype=dtype),
                y0 + dy * torch.arange(nr + 1, device=device, dtype=dtype),
                z0 + dz * torch.arange(ns + 1, device=device, dtype=dtype))

    @staticmethod
    def covering(grid: "Grid", cell_cm: float) -> "Grid":
        """Grid of cubic cells of edge ``cell_cm`` covering the box of ``grid``, centred on it (geometry.py:146)."""
        c = float(cell_cm)
        if not c > 0:
            raise ValueError(f"cell size must be > 0, got {cell_cm}")

        def count(e: float) -> int:
            r = e / c
            k = round(r)
            return max(1, int(k) if abs(r - k) <= 1e-9 * max(1.0, r) else int(math.ceil(r)))

        ex, ey, ez = grid.extent
        return Grid(shape=(count(ey), count(ex), count(ez)), delta_voxel=c, recon_slice_offset=grid.recon_slice_offset)

    def refine(self, factor: int) -> "Grid":
        """The same box with ``factor`` voxels per voxel along each axis (geometry.py:172)."""
        f = int(factor)
        if f < 1:
            raise ValueError(f"refinement factor must be >= 1, got {factor}")
        if f == 1:
            return self
        nr, nc, ns = self.shape
        return replace(self, shape=(nr * f, nc * f, ns * f), delta_voxel=self.delta_voxel / f)
