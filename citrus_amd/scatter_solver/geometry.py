"""Voxel grids, conservative resampling and the cone-beam scan geometry (host-side setup).

Written from the frame conventions of ``docs/source/conventions.rst`` ("Frames and tensor layouts", "Grids"); the
class and method names, fields and argument meanings are those of ``citrus/scatter_solver/geometry.py`` (``Grid``
:44, ``resample_conservative`` :304, ``ConeGeometry`` :331), which callers observe.

* Volumes are ``[rows, cols, slices]``: column j along +x, row i along +y, slice k along +z; voxel centres
  ``x = dx (j - (C-1)/2)``, ``y = dy (i - (R-1)/2)``, ``z = dz (k - (S-1)/2) + recon_slice_offset``.
* View frame of view angle a: ``x_v = cos(a) x - sin(a) y``, ``y_v = sin(a) x + cos(a) y``, ``z_v = z``. The source
  sits at ``(0, SID, z_shift(a))``, rays travel toward -y_v, the flat panel is the plane ``y_v = SID - SDD`` with u
  along +x_v and v along +z; a curved panel is the cylinder of radius SDD about the vertical line through the source,
  pixel (u, v) at azimuth ``gamma = u / SDD``: ``(SDD sin gamma, SID - SDD cos gamma, z_shift + v cos gamma)``.
  Every position is formed in the view frame and rotated into the recon frame by one helper (``_view_to_recon``).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np
import torch
from torch import Tensor

_COUNT_TOLERANCE = 1e-9  # relative slack of extent / cell before a partial cell is counted
_OVERLAP_SNAP = 1e-12    # overlap fractions this close to 0 or 1 are taken as exactly 0 or 1


def _centred_axis(n: int, pitch: float, offset: float, device, dtype) -> Tensor:
    """``offset + pitch (i - (n-1)/2)`` for i < n, formed in float64 and cast once."""
    i = torch.arange(n, dtype=torch.float64)
    return (offset + pitch * (i - 0.5 * (n - 1))).to(device=device, dtype=dtype)


@dataclass(frozen=True)
class Grid:
    """Uniform voxel grid centred on the rotation axis.

    Attributes:
        shape: (rows, cols, slices), the voxel counts along (y, x, z).
        delta_voxel: pitch along x in cm.
        voxel_row_aspect: pitch along y over ``delta_voxel``. Defaults to 1.0.
        voxel_slice_aspect: pitch along z over ``delta_voxel``. Defaults to 1.0.
        recon_slice_offset: z of the grid centre in cm. Defaults to 0.0.
    """

    shape: tuple[int, int, int]
    delta_voxel: float
    voxel_row_aspect: float = 1.0
    voxel_slice_aspect: float = 1.0
    recon_slice_offset: float = 0.0

    num_rows = property(lambda self: int(self.shape[0]))
    num_cols = property(lambda self: int(self.shape[1]))
    num_slices = property(lambda self: int(self.shape[2]))
    num_voxels = property(lambda self: int(np.prod(self.shape)))

    @property
    def deltas(self) -> tuple[float, float, float]:
        """Pitches (dx, dy, dz) in cm."""
        return (self.delta_voxel, self.delta_voxel * self.voxel_row_aspect, self.delta_voxel * self.voxel_slice_aspect)

    @property
    def voxel_volume(self) -> float:
        return float(np.prod(self.deltas))

    @property
    def extent(self) -> tuple[float, float, float]:
        """Side lengths (x, y, z) of the box in cm."""
        counts = (self.num_cols, self.num_rows, self.num_slices)
        return tuple(n * d for n, d in zip(counts, self.deltas))

    def _axes(self):
        """((count, pitch, centre) of x, y, z)."""
        dx, dy, dz = self.deltas
        return ((self.num_cols, dx, 0.0), (self.num_rows, dy, 0.0), (self.num_slices, dz, self.recon_slice_offset))

    def axis_coordinates(self, device=None, dtype=torch.float32) -> tuple[Tensor, Tensor, Tensor]:
        """Voxel-centre coordinates x [cols], y [rows], z [slices] in cm."""
        return tuple(_centred_axis(n, d, c, device, dtype) for n, d, c in self._axes())

    def center_coordinates(self, device=None, dtype=torch.float32) -> Tensor:
        """[rows, cols, slices, 3] voxel-centre positions (x, y, z) in cm."""
        x, y, z = self.axis_coordinates(device, dtype)
        yy, xx, zz = torch.meshgrid(y, x, z, indexing="ij")
        return torch.stack([xx, yy, zz], dim=-1)

    def lower_corner(self) -> tuple[float, float, float]:
        """(x, y, z) of the minimum corner in cm."""
        return tuple(c - 0.5 * n * d for n, d, c in self._axes())

    def axis_edges(self, device=None, dtype=torch.float64) -> tuple[Tensor, Tensor, Tensor]:
        """Voxel boundaries x [cols+1], y [rows+1], z [slices+1] in cm."""
        return tuple(_centred_axis(n + 1, d, c, device, dtype) for n, d, c in self._axes())

    @staticmethod
    def covering(grid: "Grid", cell_cm: float) -> "Grid":
        """Cubic cells of edge ``cell_cm`` over the box of ``grid``, ``ceil(extent / cell)`` per axis (a ratio within
        a relative 1e-9 of an integer counts as that integer), centred on the same box."""
        c = float(cell_cm)
        if not c > 0:
            raise ValueError(f"cell size must be > 0, got {cell_cm}")
        counts = []
        for e in grid.extent:
            r = e / c
            counts.append(max(1, math.ceil(r - _COUNT_TOLERANCE * max(1.0, r))))
        nx, ny, nz = counts
        return Grid(shape=(ny, nx, nz), delta_voxel=c, recon_slice_offset=grid.recon_slice_offset)

    def refine(self, factor: int) -> "Grid":
        """The same box with ``factor`` voxels per voxel along each axis."""
        f = int(factor)
        if f < 1:
            raise ValueError(f"refinement factor must be >= 1, got {factor}")
        if f == 1:
            return self
        return replace(self, shape=tuple(int(n) * f for n in self.shape), delta_voxel=self.delta_voxel / f)


# ---- conservative resampling ------------------------------------------------------------------------------------

def overlap_matrix(edges_from: Tensor, edges_to: Tensor) -> Tensor:
    """[n_to, n_from] float64 share of target cell j covered by source cell i (overlap length / target length), so
    that ``W @ f`` is the exact average over each target cell of a piecewise-constant field that is zero outside the
    source partition (geometry.py:252). Shares within 1e-12 of 0 or 1 are snapped to 0 or 1."""
    a = edges_from.detach().to("cpu", torch.float64)
    b = edges_to.detach().to("cpu", torch.float64)
    left = torch.maximum(b[:-1].unsqueeze(1), a[:-1].unsqueeze(0))
    right = torch.minimum(b[1:].unsqueeze(1), a[1:].unsqueeze(0))
    share = torch.clamp(right - left, min=0.0) / torch.diff(b).unsqueeze(1)
    share[share < _OVERLAP_SNAP] = 0.0
    share[(share - 1.0).abs() < _OVERLAP_SNAP] = 1.0
    return share


def resampling_bands(grid_from: Grid, grid_to: Grid) -> tuple:
    """Per axis (rows, cols, slices): (first source index [n_to] int32, count [n_to] int32, shares [n_to, width]
    float32) of the overlap matrix, whose nonzero shares in each row are contiguous (both partitions ascend)."""
    xf, yf, zf = grid_from.axis_edges()
    xt, yt, zt = grid_to.axis_edges()
    out = []
    for ef, et in ((yf, yt), (xf, xt), (zf, zt)):
        W = overlap_matrix(ef, et)
        nz = W > 0
        count = nz.sum(1)
        first = torch.where(count > 0, nz.to(torch.int64).argmax(1), torch.zeros_like(count))
        width = max(1, int(count.max()) if count.numel() else 1)
        idx = (first.unsqueeze(1) + torch.arange(width).unsqueeze(0)).clamp(max=W.shape[1] - 1)
        shares = torch.gather(W, 1, idx) * (torch.arange(width).unsqueeze(0) < count.unsqueeze(1))
        out.append((first.to(torch.int32), count.to(torch.int32), shares.to(torch.float32)))
    return tuple(out)


def resample_conservative(field: Tensor, grid_from: Grid, grid_to: Grid) -> Tensor:
    """[..., R', C', S'] exact cell averages on ``grid_to`` of the voxel-wise constant ``field [..., R, C, S]`` on
    ``grid_from`` (geometry.py:304): parts of a target cell outside the source box count as zero, so every channel's
    integral is conserved when ``grid_to`` covers ``grid_from``. On the GPU one HIP kernel
    (``citrus_resample_conservative``) sums each target cell's overlapping source voxels with the product of the three
    per-axis shares; ``field`` itself is returned when the grids are equal."""
    if tuple(field.shape[-3:]) != tuple(grid_from.shape):
        raise ValueError(f"field spatial shape {tuple(field.shape[-3:])} != grid shape {grid_from.shape}")
    if grid_from == grid_to:
        return field
    from .. import _hip

    lead = tuple(field.shape[:-3])
    channels = int(np.prod(lead)) if lead else 1
    src = field.reshape(channels, *grid_from.shape).to(torch.float32).contiguous()
    _hip.require_gpu_tensor(src, "field")
    bands = [tuple(t.to(src.device) for t in band) for band in resampling_bands(grid_from, grid_to)]
    out = torch.empty(channels, *grid_to.shape, device=src.device, dtype=torch.float32)
    lib = _hip.library()
    args = []
    for first, count, shares in bands:
        args += [first.data_ptr(), count.data_ptr(), shares.data_ptr(), int(shares.shape[1])]
    with torch.cuda.device(src.device):
        _hip.check(lib.citrus_resample_conservative(src.data_ptr(), channels, *grid_from.shape, *grid_to.shape,
                                                    *args, out.data_ptr(), _hip.stream_of(src.device)),
                   "citrus_resample_conservative")
    return out.reshape(*lead, *grid_to.shape)


# ---- cone-beam geometry -------------------------------------------------------------------------------------------

def _view_to_recon(p_view: Tensor, angle: float) -> Tensor:
    """[..., 3] recon-frame vectors of view-frame vectors: the inverse of ``x_v = c x - s y``, ``y_v = s x + c y``."""
    c, s = math.cos(angle), math.sin(angle)
    xv, yv, zv = p_view.unbind(-1)
    return torch.stack([c * xv + s * yv, c * yv - s * xv, zv], dim=-1)


def _triple(t: Tensor) -> tuple[float, float, float]:
    return tuple(float(v) for v in t.tolist())


@dataclass(frozen=True)
class ConeGeometry:
    """Cone-beam scan with a flat or cylindrical panel and circular or helical views.

    Attributes (the reference's fields and meanings): source-isocentre and source-detector distances (cm), detector
    rows and channels with their pitches and offsets (cm), ``use_curved_detector``, and the helical table
    ``helical_view_angles`` / ``helical_z_shifts`` (strictly monotone angles, stored ascending; the shift is linear in
    the angle between entries and constant beyond the ends).
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
        a = np.asarray(self.helical_view_angles, dtype=np.float64).reshape(-1)
        z = np.asarray(self.helical_z_shifts, dtype=np.float64).reshape(-1)
        if a.shape != z.shape:
            raise ValueError("helical_view_angles and helical_z_shifts must have the same length")
        step = np.diff(a)
        if step.size and not (np.all(step > 0) or np.all(step < 0)):
            raise ValueError("helical_view_angles must be strictly monotone")
        if step.size and step[0] < 0:
            a, z = a[::-1], z[::-1]
        object.__setattr__(self, "helical_view_angles", tuple(a.tolist()))
        object.__setattr__(self, "helical_z_shifts", tuple(z.tolist()))
        object.__setattr__(self, "use_curved_detector", bool(self.use_curved_detector))

    def with_helical_z_shifts(self, angles, z_shifts) -> "ConeGeometry":
        return replace(self, helical_view_angles=tuple(float(v) for v in angles),
                       helical_z_shifts=tuple(float(v) for v in z_shifts))

    @property
    def is_parallel(self) -> bool:
        return False

    @property
    def is_helical(self) -> bool:
        return len(set(self.helical_z_shifts)) > 1

    def z_shift(self, angle: float) -> float:
        """z shift of the source and the panel at a view angle (the table interpolated, clamped at its ends)."""
        if not self.helical_view_angles:
            return 0.0
        return float(np.interp(float(angle), self.helical_view_angles, self.helical_z_shifts))

    @property
    def magnification(self) -> float:
        return self.source_detector_dist / self.source_iso_dist

    @property
    def panel_y_view(self) -> float:
        """y_v of the flat panel plane (and of the curved panel's centre line)."""
        return self.source_iso_dist - self.source_detector_dist

    @property
    def image_shape(self) -> tuple[int, int]:
        return (self.num_det_rows, self.num_det_channels)

    @property
    def pixel_area(self) -> float:
        return self.delta_det_row * self.delta_det_channel

    def pixel_uv(self, device=None, dtype=torch.float32) -> tuple[Tensor, Tensor]:
        """u [channels] and v [rows] of the pixel centres in cm."""
        u = _centred_axis(self.num_det_channels, self.delta_det_channel, -self.det_channel_offset, device, dtype)
        v = _centred_axis(self.num_det_rows, self.delta_det_row, -self.det_row_offset, device, dtype)
        return u, v

    def _view_point(self, xv: float, yv: float, angle: float) -> tuple[float, float, float]:
        p = torch.tensor([xv, yv, self.z_shift(angle)], dtype=torch.float64)
        return _triple(_view_to_recon(p, angle))

    def source_position(self, angle: float) -> tuple[float, float, float]:
        return self._view_point(0.0, self.source_iso_dist, angle)

    def panel_center(self, angle: float) -> tuple[float, float, float]:
        return self._view_point(0.0, self.panel_y_view, angle)

    def panel_axes(self, angle: float):
        """(u_hat, v_hat, n_hat) of the panel at its centre in the recon frame: +x_v, +z and +y_v (toward the
        source)."""
        axes = _view_to_recon(torch.eye(3, dtype=torch.float64)[[0, 2, 1]], angle)
        return tuple(_triple(a) for a in axes)

    def panel_points(self, u: Tensor, v: Tensor, angle: float) -> Tensor:
        """[..., 3] recon-frame positions of panel points (u, v)."""
        device, dtype = u.device, u.dtype
        u64, v64 = torch.broadcast_tensors(u.to(torch.float64), v.to(u.device, torch.float64))
        zs = self.z_shift(angle)
        if self.use_curved_detector:
            R = float(self.source_detector_dist)
            gamma = u64 / R
            view = torch.stack([R * torch.sin(gamma), self.source_iso_dist - R * torch.cos(gamma),
                                zs + v64 * torch.cos(gamma)], dim=-1)
        else:
            view = torch.stack([u64, torch.full_like(u64, self.panel_y_view), zs + v64], dim=-1)
        return _view_to_recon(view, angle).to(device=device, dtype=dtype)

    def panel_half_extent(self) -> tuple[float, float]:
        """Half widths (u, v) of the active panel area, edge to edge, in cm."""
        return (0.5 * self.num_det_channels * self.delta_det_channel, 0.5 * self.num_det_rows * self.delta_det_row)

    def local_panel_axes(self, points: Tensor, angle: float) -> tuple[Tensor, Tensor, Tensor]:
        """(u_hat, v_hat, n_hat) [..., 3] of the panel surface at panel points, n_hat toward the source side: the
        centre axes on a flat panel; on a curved panel the point's azimuth ``gamma = atan2(x_v, SID - y_v)`` gives
        the horizontal normal ``(-sin gamma, cos gamma, 0)`` toward the cylinder axis and ``u_hat = n_hat x v_hat =
        (cos gamma, sin gamma, 0)`` in the view frame."""
        device, dtype = points.device, points.dtype
        if not self.use_curved_detector:
            return tuple(torch.tensor(a, device=device, dtype=dtype).expand_as(points)
                         for a in self.panel_axes(angle))
        c, s = math.cos(angle), math.sin(angle)
        p = points.to(torch.float64)
        xv = c * p[..., 0] - s * p[..., 1]
        yv = s * p[..., 0] + c * p[..., 1]
        gamma = torch.atan2(xv, self.source_iso_dist - yv)
        zero, one = torch.zeros_like(gamma), torch.ones_like(gamma)
        n_view = torch.stack([-torch.sin(gamma), torch.cos(gamma), zero], dim=-1)
        u_view = torch.stack([torch.cos(gamma), torch.sin(gamma), zero], dim=-1)
        v_view = torch.stack([zero, zero, one], dim=-1)
        return tuple(_view_to_recon(a, angle).to(device=device, dtype=dtype) for a in (u_view, v_view, n_view))

    def pixel_centers(self, angle: float, device=None, dtype=torch.float32) -> Tensor:
        """[rows, channels, 3] pixel-centre positions in cm."""
        u, v = self.pixel_uv(device, torch.float64)
        return self.panel_points(u[None, :], v[:, None], angle).to(dtype)

    def panel_distance_along(self, points: Tensor, omega: Tensor, angle: float) -> Tensor:
        """Signed distance t with ``points + t omega`` on the panel surface (``inf`` where the line misses it). On a
        curved panel t is the far root of ``|(points + t omega - source)_xy|^2 = SDD^2``, taken from the pair
        ``(q / a, c / q)`` with ``q = -(b + sign(b) sqrt(b^2 - a c))`` so that neither root cancels."""
        if not self.use_curved_detector:
            n = torch.tensor(self.panel_axes(angle)[2], device=points.device, dtype=points.dtype)
            pc = torch.tensor(self.panel_center(angle), device=points.device, dtype=points.dtype)
            return ((pc - points) * n).sum(-1) / (omega * n).sum(-1)
        src = torch.tensor(self.source_position(angle), device=points.device, dtype=points.dtype)
        d = (points - src)[..., :2]
        w = omega.to(points.dtype)[..., :2]
        a = (w * w).sum(-1)
        b = (d * w).sum(-1)
        c = (d * d).sum(-1) - float(self.source_detector_dist) ** 2
        a, b, c = torch.broadcast_tensors(a, b, c)
        disc = b * b - a * c
        q = -(b + torch.where(b >= 0, 1.0, -1.0) * torch.sqrt(disc.clamp_min(0.0)))
        ok = (a > 0) & (disc >= 0) & (q != 0)
        q_safe = torch.where(q != 0, q, torch.ones_like(q))
        a_safe = torch.where(a > 0, a, torch.ones_like(a))
        far = torch.maximum(q_safe / a_safe, c / q_safe)
        return torch.where(ok, far, torch.full_like(far, math.inf))
