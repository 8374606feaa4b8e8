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


@dataclass(frozen=True)
class ConeGeometry:
    """Cone-beam geometry with a flat or cylindrical panel, circular or helical views (geometry.py:331-765).

    In the view frame the source sits at ``(0, +source_iso_dist, z_shift)``, u along +x_view (channels), v along +z
    (rows); the flat panel lies in the plane ``y_view = source_iso_dist - source_detector_dist`` and the curved panel on
    the cylinder of radius ``source_detector_dist`` about the vertical axis through the source. A view at ``angle``
    rotates the view frame into the recon frame by ``(x, y) = (c x_v + s y_v, -s x_v + c y_v)``.
    """

    source_iso_dist: float
    source_detector_dist: float
    num_det_rows: int
    num_det_channels: int
    delta_det_row: float
    delta_det_channel: float
    det_row_offset: float = 0.0
    det_channel_offset: float = 0.0
    use_curved_detector: bool = False
    helical_view_angles: tuple[float, ...] = ()
    helical_z_shifts: tuple[float, ...] = ()

    def __post_init__(self) -> None:
        angles = tuple(float(a) for a in self.helical_view_angles)
        shifts = tuple(float(z) for z in self.helical_z_shifts)
        if len(angles) != len(shifts):
            raise ValueError("helical_view_angles and helical_z_shifts must have the same length")
        steps = [b - a for a, b in zip(angles, angles[1:])]
        if steps and not (all(s > 0 for s in steps) or all(s < 0 for s in steps)):
            raise ValueError("helical_view_angles must be strictly monotone")
        if steps and steps[0] < 0:
            angles, shifts = angles[::-1], shifts[::-1]
        object.__setattr__(self, "helical_view_angles", angles)
        object.__setattr__(self, "helical_z_shifts", shifts)
        object.__setattr__(self, "use_curved_detector", bool(self.use_curved_detector))

    def with_helical_z_shifts(self, angles, z_shifts) -> "ConeGeometry":
        return replace(self, helical_view_angles=tuple(float(a) for a in angles),
                       helical_z_shifts=tuple(float(z) for z in z_shifts))

    @property
    def is_parallel(self) -> bool:
        return False

    @property
    def is_helical(self) -> bool:
        return len(self.helical_z_shifts) > 0 and max(self.helical_z_shifts) > min(self.helical_z_shifts)

    def z_shift(self, angle: float) -> float:
        """z shift of source and panel, linear in the helical table, constant outside it."""
        angles, shifts = self.helical_view_angles, self.helical_z_shifts
        if not angles:
            return 0.0
        a = float(angle)
        if a <= angles[0]:
            return shifts[0]
        if a >= angles[-1]:
            return shifts[-1]
        import bisect

        j = bisect.bisect_right(angles, a) - 1
        t = (a - angles[j]) / (angles[j + 1] - angles[j])
        return shifts[j] + t * (shifts[j + 1] - shifts[j])

    @property
    def magnification(self) -> float:
        return self.source_detector_dist / self.source_iso_dist

    @property
    def panel_y_view(self) -> float:
        return self.source_iso_dist - self.source_detector_dist

    @property
    def image_shape(self) -> tuple[int, int]:
        return (self.num_det_rows, self.num_det_channels)

    @property
    def pixel_area(self) -> float:
        return self.delta_det_row * self.delta_det_channel

    def pixel_uv(self, device=None, dtype=torch.float32) -> tuple[Tensor, Tensor]:
        """u [channels] and v [rows] of the pixel centres in cm."""
        nc, nr = self.num_det_channels, self.num_det_rows
        u = (torch.arange(nc, device=device, dtype=dtype) - (nc - 1) / 2) * self.delta_det_channel \
            - self.det_channel_offset
        v = (torch.arange(nr, device=device, dtype=dtype) - (nr - 1) / 2) * self.delta_det_row - self.det_row_offset
        return u, v

    def source_position(self, angle: float) -> tuple[float, float, float]:
        c, s = math.cos(angle), math.sin(angle)
        return (s * self.source_iso_dist, c * self.source_iso_dist, self.z_shift(angle))

    def panel_center(self, angle: float) -> tuple[float, float, float]:
        c, s = math.cos(angle), math.sin(angle)
        yv = self.panel_y_view
        return (s * yv, c * yv, self.z_shift(angle))

    def panel_axes(self, angle: float):
        """(u_hat, v_hat, n_hat) of the panel at its centre in the recon frame; n_hat points toward the source."""
        c, s = math.cos(angle), math.sin(angle)
        return (c, -s, 0.0), (0.0, 0.0, 1.0), (s, c, 0.0)

    def panel_points(self, u: Tensor, v: Tensor, angle: float) -> Tensor:
        """Recon-frame positions of panel points (u, v); on a curved panel the point at azimuth u/SDD."""
        device, dtype = u.device, u.dtype
        u, v = torch.broadcast_tensors(u, v.to(device, dtype))
        uh, vh, nh = (torch.tensor(a, device=device, dtype=dtype) for a in self.panel_axes(angle))
        pc = torch.tensor(self.panel_center(angle), device=device, dtype=dtype)
        if not self.use_curved_detector:
            return pc + u[..., None] * uh + v[..., None] * vh
        R = float(self.source_detector_dist)
        gamma = u / R
        return (pc + (R * torch.sin(gamma))[..., None] * uh + (2.0 * R * torch.sin(0.5 * gamma) ** 2)[..., None] * nh
                + (v * torch.cos(gamma))[..., None] * vh)

    def panel_half_extent(self) -> tuple[float, float]:
        """Half widths (u, v) of the active panel area, edge to edge, in cm (geometry.py:482)."""
        return (self.num_det_channels * self.delta_det_channel / 2, self.num_det_rows * self.delta_det_row / 2)

    def local_panel_axes(self, points: Tensor, angle: float) -> tuple[Tensor, Tensor, Tensor]:
        """(u_hat, v_hat, n_hat) [..., 3] of the panel surface at panel points, n_hat toward the source
        (geometry.py:624): the view's axes on a flat panel; on a curved one the horizontal normal toward the cylinder
        axis through the source, the vertical v_hat and ``u_hat = n_hat x v_hat``."""
        device, dtype = points.device, points.dtype
        u_hat, v_hat, n_hat = (torch.tensor(a, device=device, dtype=dtype).expand_as(points)
                               for a in self.panel_axes(angle))
        if not self.use_curved_detector:
            return u_hat, v_hat, n_hat
        to_axis = torch.tensor(self.source_position(angle), device=device, dtype=dtype) - points
        to_axis = to_axis - (to_axis * v_hat).sum(-1, keepdim=True) * v_hat
        n_local = to_axis / torch.linalg.norm(to_axis, dim=-1, keepdim=True)
        return torch.linalg.cross(n_local, v_hat, dim=-1), v_hat, n_local

    def pixel_centers(self, angle: float, device=None, dtype=torch.float32) -> Tensor:
        """[rows, channels, 3] pixel-centre positions in cm."""
        u, v = self.pixel_uv(device, dtype)
        return self.panel_points(u[None, :], v[:, None], angle)

    def panel_distance_along(self, points: Tensor, omega: Tensor, angle: float) -> Tensor:
        """Signed distance t with ``points + t omega`` on the panel surface (larger root on the cylinder)."""
        if self.use_curved_detector:
            sp = torch.tensor(self.source_position(angle), device=points.device, dtype=points.dtype)
            d = (points - sp)[..., :2]
            w = omega.to(points.dtype)[..., :2]
            a = (w * w).sum(-1)
            b = (d * w).sum(-1)
            c = (d * d).sum(-1) - float(self.source_detector_dist) ** 2
            a, b, c = torch.broadcast_tensors(a, b, c)
            disc = b * b - a * c
            root = torch.sqrt(disc.clamp_min(0.0))
            safe_a = torch.where(a > 0, a, torch.ones_like(a))
            pos = b >= 0
            den = torch.where(pos, b + root, safe_a)
            num = torch.where(pos, -c, root - b)
            t = num / torch.where(den != 0, den, torch.ones_like(den))
            return torch.where((a > 0) & (disc >= 0) & (den != 0), t, torch.full_like(t, math.inf))
        n = torch.tensor(self.panel_axes(angle)[2], device=points.device, dtype=points.dtype)
        pc = torch.tensor(self.panel_center(angle), device=points.device, dtype=points.dtype)
        return ((pc - points) * n).sum(-1) / (omega * n).sum(-1)
