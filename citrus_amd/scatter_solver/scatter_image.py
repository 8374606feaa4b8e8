"""Collapsed-cone scatter image and primary image of one view on the HIP kernels.

Keeps the interface of ``citrus/scatter_solver/scatter_image.py``: ``covering_polar_angle`` (:59),
``panel_direction_set`` (:98), ``DirectionSetContext`` / ``make_direction_context`` (:138, :542),
``scatter_image_from_source`` (:737), ``supervoxel_subpoint_emission`` (:829), ``box_matrix`` / ``box_blur_*``
(:897-1007), ``panel_height_field`` (:1010), the finite-cone footprint (:1033), ``combined_scatter_image`` (:1141),
``primary_path_lengths`` / ``primary_from_path_lengths`` / ``primary_image`` (:1368-1506) and ``scatter_fraction``
(:1509). Per chunk of cone directions the image is:

1. the optical depth from the transport cells to the panel (``citrus_ray_downstream``);
2. the multiple-collision source ``S_MC = Y(w) T`` of the cells, one GEMM;
3. the attenuated total emission at the quadrature points, first-collision source included, with the sums of its
   mean panel height (``citrus_attenuated_emission``);
4. oblique line integrals through the pixel centres (``citrus_ray_oblique``);
5. the footprint box blur (``citrus_box_matrix`` and two batched GEMMs);
6. accumulation with the signal weights ``w_k cos(theta0) D_g(theta0)`` (no grid).

With ``signal_only`` on a flat panel the groups are summed with the signal weights inside the emission kernel, so a
single channel is traced. Anti-scatter grids whose transmission depends on the angles are outside this build.
"""

from __future__ import annotations

import contextlib
import ctypes
import math
from dataclasses import dataclass, field, replace
from typing import Callable, Optional, Sequence, Union

import torch
from torch import Tensor

from .. import _hip
from .angular_quadrature import Quadrature, hemisphere_quadrature, real_spherical_harmonics
from .cross_sections import MultigroupLibrary, group_transmission, mu_t_volume
from .detector import DetectorResponse, GridResponse, detector_angle
from .geometry import ConeGeometry, Grid
from .source import SourceModel
from .uncollided import _device_index, _fc_args, present_materials

SourceFunction = Callable[[Tensor], Tensor]


# ---- direction sets ---------------------------------------------------------------------------------------------

def covering_polar_angle(grid: Grid, geometry: ConeGeometry, angle: float, margin: float = 1e-6) -> float:
    """Smallest polar angle about the inward panel normal containing every direction from the grid box to the panel
    (scatter_image.py:59): the maximum over pairs of a box corner and a panel corner (a curved panel's top and bottom
    edges sampled four times per channel), pi/2 when a corner is not strictly on the source side."""
    lo, ext = grid.lower_corner(), grid.extent
    corners = torch.tensor([[lo[0] + i * ext[0], lo[1] + j * ext[1], lo[2] + k * ext[2]]
                            for i in (0, 1) for j in (0, 1) for k in (0, 1)], dtype=torch.float64)
    u_hat, v_hat, n_hat = (torch.tensor(a, dtype=torch.float64) for a in geometry.panel_axes(angle))
    pc = torch.tensor(geometry.panel_center(angle), dtype=torch.float64)
    hu, hv = geometry.panel_half_extent()
    cu, cv = -geometry.det_channel_offset, -geometry.det_row_offset
    if geometry.use_curved_detector:
        eu = cu + torch.linspace(-hu, hu, 4 * geometry.num_det_channels + 1, dtype=torch.float64)
        ev = cv + torch.tensor([-hv, hv], dtype=torch.float64)
        panel = geometry.panel_points(eu[:, None], ev[None, :], angle).reshape(-1, 3)
    else:
        panel = torch.stack([pc + (cu + su * hu) * u_hat + (cv + sv * hv) * v_hat for su in (-1, 1) for sv in (-1, 1)])
    d = panel[:, None, :] - corners[None, :, :]
    normal = (d * (-n_hat)).sum(-1)
    if bool((normal <= 0).any()):
        return math.pi / 2
    lateral = torch.linalg.norm(d - normal[..., None] * (-n_hat), dim=-1)
    return float(min(math.pi / 2, float(torch.atan2(lateral, normal).max()) + margin))


def panel_direction_set(geometry: ConeGeometry, angle: float, config_or_counts, *,
                        max_polar_angle: Union[float, Grid] = math.pi / 2, device=None,
                        dtype: torch.dtype = torch.float32) -> Quadrature:
    """Collapsed-cone direction set about the inward panel normal with its footprint cell widths (scatter_image.py:98).
    ``config_or_counts`` is (num_polar, num_azimuthal) or an object with ``cone_num_polar`` / ``cone_num_azimuthal``."""
    if hasattr(config_or_counts, "cone_num_polar"):
        P, A = config_or_counts.cone_num_polar, config_or_counts.cone_num_azimuthal
    else:
        P, A = (int(c) for c in config_or_counts)
    if isinstance(max_polar_angle, Grid):
        max_polar_angle = covering_polar_angle(max_polar_angle, geometry, angle)
    u_hat, v_hat, n_hat = geometry.panel_axes(angle)
    return hemisphere_quadrature([-c for c in n_hat], P, A, max_polar_angle, device=device, dtype=dtype,
                                 tangent_axes=(u_hat, v_hat))


# ---- per-(view, direction set) context -----------------------------------------------------------------------------

@dataclass(eq=False)
class DirectionSetContext:
    """The emission-independent part of one view's collapsed-cone images (scatter_image.py:138).

    K panel-facing directions, G groups, Rd x Cd pixels. ``signal_weight`` is [K, G] (flat panel: the same at every
    pixel) or [K, G, Cd] (curved panel: the incidence angle varies along the channels); ``planar_weight`` [K, Cd] on a
    curved panel. ``grid_weight`` [G] is the transmission of an angle-independent grid (ones with ``pixel_grid``).
    ``pixel_grid`` is a grid whose transmission depends on the photon's angles: the signal weights then exclude it,
    and each chunk's images are weighted by its transmission at every (direction, group, pixel) after the footprint
    blur (``pixel_signal_weight``; the reference's per-row, per-channel and per-pixel septa factors,
    scatter_image.py:373-390 and :606-650, are this product, applied there between or after the blurs, with which a
    per-row or per-channel factor commutes)."""

    quadrature: Quadrature
    geometry: ConeGeometry
    angle: float
    detector: DetectorResponse
    grid_response: GridResponse
    group_energy: Tensor
    omega: Tensor
    weight: Tensor
    cos_theta0: Tensor
    detector_weight: Tensor
    grid_weight: Tensor
    signal_weight: Tensor
    separable: bool
    chunk: int
    chunk_bounds: tuple
    omega_chunks: tuple
    weight_chunks: tuple
    image_shape: tuple
    cell_width: Optional[Tensor] = None
    planar_weight: Optional[Tensor] = None
    signal_group_weight: Optional[Tensor] = None
    pixel_grid: Optional[GridResponse] = None
    _cache: dict = field(default_factory=dict, repr=False)

    @property
    def num_directions(self) -> int:
        return int(self.omega.shape[0])

    def pixel_signal_weight(self, chunk_index: int) -> Tensor:
        """Signal weights of the chunk's directions at the pixels: [k, G, Rd, Cd] with ``pixel_grid`` (the signal
        weights times the grid's transmission at each direction's grid angles at each pixel, ``transmission_at`` in
        float64), else ``signal_weight`` shaped to broadcast against [k, G, Rd, Cd] images."""
        start, stop = self.chunk_bounds[chunk_index]
        sw = self.signal_weight[start:stop]
        sw = sw[:, :, None, None] if sw.dim() == 2 else sw[:, :, None, :]
        if self.pixel_grid is None:
            return sw
        key = ("pixels_f64",)
        if key not in self._cache:
            self._cache[key] = self.geometry.pixel_centers(self.angle, self.device, torch.float64)
        omega = self.omega_chunks_f64()[chunk_index]
        t = self.pixel_grid.transmission_at(self._cache[("group_energy_f64",)], omega, self._cache[key],
                                            self.geometry, self.angle)  # [G, k, Rd, Cd]
        return sw * t.permute(1, 0, 2, 3).to(sw.dtype)

    def omega_chunks_f64(self) -> tuple:
        """The direction chunks in float64 (the ray kernels' directions), formed once per context."""
        key = ("omega_chunks_f64",)
        if key not in self._cache:
            self._cache[key] = tuple(c.to(torch.float64).contiguous() for c in self.omega_chunks)
        return self._cache[key]

    @property
    def num_groups(self) -> int:
        return int(self.group_energy.numel())

    @property
    def device(self) -> torch.device:
        return self.omega.device

    @property
    def dtype(self) -> torch.dtype:
        return self.omega.dtype

    def check(self, geometry, angle, quadrature, detector, grid_response, num_groups, device, dtype) -> None:
        """Raise ``ValueError`` unless the context was built for these image arguments (host-side)."""
        bad = []
        if quadrature is not self.quadrature:
            bad.append("direction set (quadrature)")
        if geometry != self.geometry:
            bad.append("geometry")
        if abs(float(angle) - self.angle) > 1e-12:
            bad.append(f"angle ({angle} vs {self.angle})")
        if detector is not self.detector:
            bad.append("detector response")
        if grid_response is not self.grid_response:
            bad.append("grid response")
        if int(num_groups) != self.num_groups:
            bad.append(f"group count ({num_groups} vs {self.num_groups})")
        if torch.device(device) != self.device:
            bad.append(f"device ({device} vs {self.device})")
        if dtype != self.dtype:
            bad.append(f"dtype ({dtype} vs {self.dtype})")
        if bad:
            raise ValueError("DirectionSetContext does not match the image arguments: " + ", ".join(bad))

    def spherical_harmonics(self, L: int) -> tuple:
        """Per chunk [k, (L+1)^2] real spherical harmonics of the directions (cached per order)."""
        key = ("sh", int(L))
        if key not in self._cache:
            Y = real_spherical_harmonics(int(L), self.omega.double()).to(self.dtype)
            self._cache[key] = tuple(Y[s:e].contiguous() for s, e in self.chunk_bounds)
        return self._cache[key]

    def new_accumulator(self) -> Tensor:
        G, (Rd, Cd) = self.num_groups, self.image_shape
        return torch.zeros(3, G, Rd, Cd, device=self.device, dtype=self.dtype)

    def accumulate(self, accumulator: Tensor, image: Tensor, chunk_index: int) -> None:
        """Add the fluence, planar fluence and signal of one chunk's images [k, G, Rd, Cd]."""
        start, stop = self.chunk_bounds[chunk_index]
        w = self.weight_chunks[chunk_index]
        accumulator[0].add_(torch.einsum("kgrc,k->grc", image, w))
        if self.planar_weight is not None:
            accumulator[1].add_(torch.einsum("kgrc,kc->grc", image, self.planar_weight[start:stop]))
        else:
            accumulator[1].add_(torch.einsum("kgrc,k->grc", image, w * self.cos_theta0[start:stop]))
        if self.pixel_grid is not None:
            accumulator[2].add_((image * self.pixel_signal_weight(chunk_index)).sum(0))
        elif self.planar_weight is not None:
            accumulator[2].add_(torch.einsum("kgrc,kgc->grc", image, self.signal_weight[start:stop]))
        else:
            accumulator[2].add_(torch.einsum("kgrc,kg->grc", image, self.signal_weight[start:stop]))

    def finalize(self, accumulator: Tensor) -> dict:
        fl, pl, sg = (accumulator[j].contiguous() for j in range(3))
        return {"fluence": fl, "planar_fluence": pl, "signal_per_group": sg, "signal": sg.sum(0)}


def make_direction_context(quadrature: Quadrature, geometry: ConeGeometry, angle: float, detector: DetectorResponse,
                           grid_response: GridResponse, library: Optional[MultigroupLibrary] = None, device=None,
                           dtype: Optional[torch.dtype] = None, *, group_energy: Optional[Tensor] = None,
                           chunk: int = 8, factorize: bool = True) -> DirectionSetContext:
    """Precompute the emission-independent part of one view's images (scatter_image.py:542): the panel-facing
    directions (``w . -n_hat > 0`` in the image dtype), their weights, ``cos(theta0)``, the detector response at
    ``theta0`` and the signal weights ``w cos(theta0) D_g(theta0) G_g``; on a curved panel the incidence angle of
    each direction at each channel's local normal. ``signal_group_weight`` [K, G] holds the weights by which the
    groups of one direction collapse into one signal image: ``w cos(theta0) D_g G_g`` on a flat panel and, on a
    curved panel with an angle-independent detector, ``D_g G_g`` (the channel factor ``planar_weight`` is applied
    after the footprint blur)."""
    if group_energy is None:
        if library is None:
            raise ValueError("pass library or group_energy")
        group_energy = library.group_energy
    device = torch.device(device) if device is not None else quadrature.omega.device
    if dtype is None:
        dtype = quadrature.omega.dtype if quadrature.omega.dtype.is_floating_point else torch.float32
    f64 = torch.float64
    energy = group_energy.detach().to(device, f64).reshape(-1)
    G = int(energy.numel())
    _, _, n_hat = geometry.panel_axes(angle)
    n_in = -torch.tensor(n_hat, device=device, dtype=f64)
    keep = torch.nonzero((quadrature.omega.detach().to(device, dtype) * n_in.to(dtype)).sum(-1) > 0).reshape(-1)
    omega = quadrature.omega.detach().to(device, f64)[keep]
    weight = quadrature.weight.detach().to(device, f64)[keep]
    cos0 = (omega * n_in).sum(-1)
    K = int(omega.shape[0])
    pixels = geometry.pixel_centers(angle, device, f64)
    Rd, Cd = int(pixels.shape[0]), int(pixels.shape[1])
    # a grid whose transmission depends on the photon's angles weights every pixel of the images (pixel_grid);
    # otherwise its transmission at the group energies is part of the signal weights
    pixel_grid = grid_response if grid_response.depends_on_angles else None
    grid_t = (torch.ones(G, device=device, dtype=f64) if pixel_grid is not None
              else grid_response.at_energy(energy, device, f64))  # [G]
    theta0 = detector_angle(omega, geometry, angle)
    det = detector.at(energy, theta0) if K else torch.zeros(G, 0, device=device, dtype=f64)  # [G, K]
    planar = None
    group_weight = None
    if geometry.use_curved_detector:
        normals = geometry.local_panel_axes(pixels[0], angle)[2]  # [Cd, 3]
        cos_c = (omega @ normals.neg().t()).clamp(0.0, 1.0)  # [K, Cd]
        planar = weight[:, None] * cos_c
        signal = (planar[:, None, :] * detector.at(energy, torch.arccos(cos_c)).permute(1, 0, 2)
                  * grid_t[None, :, None])  # [K, G, Cd]
        if detector.angle_independent and pixel_grid is None:  # signal = planar[k, c] D_g G_g: the groups
            group_weight = (det.t() * grid_t[None, :]).to(dtype).contiguous()  # collapse before the image [K, G]
    else:
        signal = (weight * cos0)[:, None] * det.t() * grid_t[None, :]  # [K, G]
        if pixel_grid is None:
            group_weight = signal.to(dtype).contiguous()
    chunk = max(1, int(chunk))
    bounds = tuple((s, min(s + chunk, K)) for s in range(0, K, chunk))
    om = omega.to(dtype).contiguous()
    w = weight.to(dtype).contiguous()
    cw = getattr(quadrature, "cell_width", None)
    cw = None if cw is None else cw.detach().to(device, dtype)[keep].contiguous()
    context = DirectionSetContext(
        quadrature=quadrature, geometry=geometry, angle=float(angle), detector=detector, grid_response=grid_response,
        group_energy=energy.to(dtype), omega=om, weight=w, cos_theta0=cos0.to(dtype),
        detector_weight=det.t().to(dtype).contiguous(), grid_weight=grid_t.to(dtype),
        signal_weight=signal.to(dtype).contiguous(), separable=not geometry.use_curved_detector, chunk=chunk,
        chunk_bounds=bounds, omega_chunks=tuple(om[s:e] for s, e in bounds),
        weight_chunks=tuple(w[s:e] for s, e in bounds), image_shape=(Rd, Cd), cell_width=cw,
        planar_weight=None if planar is None else planar.to(dtype).contiguous(), signal_group_weight=group_weight,
        pixel_grid=pixel_grid)
    if pixel_grid is not None:
        context._cache[("group_energy_f64",)] = energy
        context._cache[("omega_chunks_f64",)] = tuple(omega[s:e].contiguous() for s, e in bounds)
    return context


def _resolve_context(context, quadrature, geometry, angle, detector, grid_response, library, mu_t, chunk,
                     group_energy) -> DirectionSetContext:
    G = int(mu_t.shape[0])
    if context is None:
        context = make_direction_context(quadrature, geometry, angle, detector, grid_response, library, mu_t.device,
                                         mu_t.dtype, group_energy=group_energy, chunk=chunk)
        if context.num_groups != G:
            raise ValueError(f"group_energy has {context.num_groups} entries, mu_t has {G} groups")
        return context
    context.check(geometry, angle, quadrature, detector, grid_response, G, mu_t.device, mu_t.dtype)
    return context


# ---- finite-cone footprint --------------------------------------------------------------------------------------

def box_matrix(N: int, w: Tensor, dtype: Optional[torch.dtype] = None) -> Tensor:
    """[k, N, N] footprint filter matrices ``B[c, j] = |[j, j+1] n W_c| / |W_c|`` (scatter_image.py:952), one
    ``citrus_box_matrix`` launch."""
    widths = w.reshape(-1).to(torch.float32).contiguous()
    _hip.require_gpu_tensor(widths, "widths")
    k = int(widths.numel())
    out = torch.empty(k, int(N), int(N), device=w.device, dtype=torch.float32)
    with torch.cuda.device(w.device):
        _hip.check(_hip.library().citrus_box_matrix(widths.data_ptr(), k, int(N), out.data_ptr(),
                                                    _hip.stream_of(w.device)), "citrus_box_matrix")
    return out if dtype in (None, torch.float32) else out.to(dtype)


def box_blur_channels(image: Tensor, width_channels: Tensor) -> Tensor:
    """[k, G, Rd, Cd] images blurred along the channels (scatter_image.py:922)."""
    k, G, Rd, Cd = image.shape
    bt = box_matrix(Cd, width_channels, image.dtype).transpose(1, 2).contiguous()  # [k, Cd, Cd]
    return torch.bmm(image.reshape(k, G * Rd, Cd), bt).view(k, G, Rd, Cd)


def box_blur_rows(image: Tensor, width_rows: Tensor) -> Tensor:
    """[k, G, Rd, Cd] images blurred along the rows (scatter_image.py:937)."""
    Rd = int(image.shape[2])
    return torch.matmul(box_matrix(Rd, width_rows, image.dtype)[:, None], image.contiguous())


def box_blur_rows_cols(image: Tensor, width_channels: Tensor, width_rows: Tensor) -> Tensor:
    """Separable box blur with per-image fractional widths in pixels, the window clipped to the panel and
    renormalized (scatter_image.py:897)."""
    return box_blur_rows(box_blur_channels(image, width_channels), width_rows)


def panel_height_field(grid: Grid, geometry: ConeGeometry, angle: float, device=None,
                       dtype: torch.dtype = torch.float32) -> Tensor:
    """[R, C, S] normal distance ``(r - panel_center) . n_hat`` of the voxel centres (scatter_image.py:1010)."""
    _, _, n_hat = geometry.panel_axes(angle)
    r = grid.center_coordinates(device, torch.float64)
    pc = torch.tensor(geometry.panel_center(angle), device=device, dtype=torch.float64)
    return ((r - pc) @ torch.tensor(n_hat, device=device, dtype=torch.float64)).to(dtype)


class _Footprint:
    """Finite-cone footprint of one image (scatter_image.py:1033): the single-ray image of direction k is box-blurred
    with ``hbar_k (W_alpha_k, W_beta_k)`` in pixels, ``hbar_k`` the emission-weighted mean panel height."""

    def __init__(self, grid: Grid, geometry: ConeGeometry, angle: float, context: DirectionSetContext) -> None:
        if context.cell_width is None:
            raise ValueError("the finite-cone footprint needs quadrature.cell_width (panel_direction_set fills it)")
        self.context = context
        key = ("footprint", grid)
        cached = context._cache.get(key)
        if cached is None:
            device, dtype = context.device, context.dtype
            scale = torch.tensor([1.0 / geometry.delta_det_channel, 1.0 / geometry.delta_det_row], device=device,
                                 dtype=dtype)
            heights = panel_height_field(grid, geometry, angle, device, dtype).reshape(-1).contiguous()
            _, _, n_hat = geometry.panel_axes(angle)
            pc = geometry.panel_center(angle)
            lo, ext = grid.lower_corner(), grid.extent
            hs = [sum((lo[a] + c[a] * ext[a] - pc[a]) * n_hat[a] for a in range(3))
                  for c in [(i, j, k) for i in (0, 1) for j in (0, 1) for k in (0, 1)]]
            cached = (scale, heights, 0.5 * (min(hs) + max(hs)))
            context._cache[key] = cached
        self.scale, self.heights, self.center_height = cached
        # the same float32 scales on the host, for the filter kernel's arguments (no device read)
        self.scale_host = (1.0 / geometry.delta_det_channel, 1.0 / geometry.delta_det_row)

    def mean_height_from_sums(self, sums: Tensor) -> Tensor:
        """[k] mean heights from the [k, 2] sums (sum e, sum e h); the box-centre height where the emission vanishes."""
        num, den = sums[:, 1], sums[:, 0]
        tiny = torch.finfo(den.dtype).tiny
        return torch.where(den > 0, num / den.clamp_min(tiny), torch.full_like(den, self.center_height))

    def mean_height(self, emission: Tensor) -> Tensor:
        k = int(emission.shape[0])
        total = emission.sum(dim=1).reshape(k, -1)
        return self.mean_height_from_sums(torch.stack([total.sum(-1), total @ self.heights], dim=-1))

    def widths(self, chunk_index: int, mean_heights: Tensor) -> Tensor:
        """[k, 2] box widths in pixels along the channels and the rows."""
        start, stop = self.context.chunk_bounds[chunk_index]
        return mean_heights[:, None] * self.context.cell_width[start:stop] * self.scale

    def blur(self, image: Tensor, chunk_index: int, mean_heights: Tensor) -> Tensor:
        w = self.widths(chunk_index, mean_heights)
        return box_blur_rows_cols(image, w[:, 0], w[:, 1])

    def filters(self, chunk_index: int, partial: Tensor, channel_weight: Optional[Tensor] = None) -> tuple:
        """(transposed channel filters [k, Cd, Cd], row filters [k, Rd, Rd]) of a chunk straight from the emission
        kernel's per-block sums ``partial [k, blocks, 2]``: ``mean_height_from_sums``, ``widths`` and ``box_matrix``
        in one ``citrus_footprint_filters`` call. ``channel_weight [k, Cd]`` scales the blurred images' channels
        (folded into the channel filters)."""
        start, stop = self.context.chunk_bounds[chunk_index]
        k, nb = int(partial.shape[0]), int(partial.shape[1])
        cw = self.context.cell_width[start:stop].contiguous()
        _hip.require_gpu_tensor(partial, "partial")
        _hip.require_gpu_tensor(cw, "cell_width")
        Rd, Cd = self.context.image_shape
        if channel_weight is not None:
            _hip.require_gpu_tensor(channel_weight, "channel_weight")
            if tuple(channel_weight.shape) != (k, Cd):
                raise ValueError(f"channel_weight must be [{k}, {Cd}], got {tuple(channel_weight.shape)}")
        dev = partial.device
        heights = torch.empty(k, device=dev, dtype=torch.float32)
        bct = torch.empty(k, Cd, Cd, device=dev, dtype=torch.float32)
        br = torch.empty(k, Rd, Rd, device=dev, dtype=torch.float32)
        with torch.cuda.device(dev):
            _hip.check(_hip.library().citrus_footprint_filters(
                partial.data_ptr(), k, nb, float(self.center_height), cw.data_ptr(), self.scale_host[0],
                self.scale_host[1], Cd, Rd, _hip.ptr(channel_weight), heights.data_ptr(), bct.data_ptr(),
                br.data_ptr(), _hip.stream_of(dev)), "citrus_footprint_filters")
        return bct, br

    def blur_accumulate(self, acc: Tensor, image: Tensor, chunk_index: int, partial: Tensor,
                        channel_weight: Optional[Tensor] = None) -> None:
        """acc [Rd, Cd] += sum_k (row blur of (channel blur of image_k)) for the single-channel images [k, 1, Rd, Cd]
        of chunk ``chunk_index``, the footprint widths from the emission kernel's per-block sums ``partial``: the
        filters of ``filters`` applied window by window (``citrus_footprint_blur``: two banded kernels, no dense
        filter matrices, each pixel summed over the chunk's directions by one thread)."""
        start, stop = self.context.chunk_bounds[chunk_index]
        k, G, Rd, Cd = image.shape
        if G != 1:
            raise ValueError("blur_accumulate takes single-channel images")
        if tuple(acc.shape) != (Rd, Cd) or not acc.is_contiguous() or acc.dtype != torch.float32:
            raise ValueError(f"acc must be a contiguous float32 [{Rd}, {Cd}] tensor")
        nb = int(partial.shape[1])
        if int(partial.shape[0]) != k:
            raise ValueError(f"partial has {int(partial.shape[0])} directions, the images {k}")
        cw = self.context.cell_width[start:stop].contiguous()
        for t, name in ((partial, "partial"), (cw, "cell_width"), (image, "image"), (acc, "acc")):
            _hip.require_gpu_tensor(t, name)
        if channel_weight is not None:
            _hip.require_gpu_tensor(channel_weight, "channel_weight")
            if tuple(channel_weight.shape) != (k, Cd) or not channel_weight.is_contiguous():
                raise ValueError(f"channel_weight must be contiguous [{k}, {Cd}], got {tuple(channel_weight.shape)}")
        dev = partial.device
        image = image.contiguous()
        heights = torch.empty(k, device=dev, dtype=torch.float32)
        work = torch.empty(k, Rd, Cd, device=dev, dtype=torch.float32)
        with torch.cuda.device(dev):
            _hip.check(_hip.library().citrus_footprint_blur(
                partial.data_ptr(), k, nb, float(self.center_height), cw.data_ptr(), self.scale_host[0],
                self.scale_host[1], Cd, Rd, _hip.ptr(channel_weight), image.data_ptr(), heights.data_ptr(),
                work.data_ptr(), acc.data_ptr(), _hip.stream_of(dev)), "citrus_footprint_blur")

    def accumulate_blurred(self, acc: Tensor, image: Tensor, channel_filters_t: Tensor, row_filters: Tensor) -> None:
        """acc [Rd, Cd] += sum_k (row blur of (channel blur of image_k)) for single-channel images [k, 1, Rd, Cd]:
        two batched products and one accumulating product with a ones vector (the sum over k)."""
        k, G, Rd, Cd = image.shape
        if G != 1:
            raise ValueError("accumulate_blurred takes single-channel images")
        x = torch.bmm(image.reshape(k, Rd, Cd), channel_filters_t)
        y = torch.bmm(row_filters, x)
        key = ("ones", k, str(acc.device))
        ones = self.context._cache.get(key)
        if ones is None:
            ones = self.context._cache[key] = torch.ones(1, k, device=acc.device, dtype=torch.float32)
        acc.view(1, Rd * Cd).addmm_(ones, y.view(k, Rd * Cd))

    @staticmethod
    def blur_with(image: Tensor, channel_filters_t: Tensor, row_filters: Tensor) -> Tensor:
        """``box_blur_rows_cols`` with the filters of ``filters``."""
        k, G, Rd, Cd = image.shape
        x = torch.bmm(image.reshape(k, G * Rd, Cd), channel_filters_t).view(k, G, Rd, Cd)
        return torch.matmul(row_filters[:, None], x.contiguous())


# ---- emission ---------------------------------------------------------------------------------------------------

def _subcell_offsets(omega: Tensor, grid_quadrature: Grid, f: int) -> Tensor:
    """[K, f, f, f] offsets ``w . (r_j - r_cell)`` of the f^3 points of a cell (axes rows, cols, slices)."""
    dx, dy, dz = grid_quadrature.deltas
    c = torch.arange(f, device=omega.device, dtype=omega.dtype) - (f - 1) / 2.0
    w = omega.reshape(-1, 3)
    return (w[:, 1, None, None, None] * (dy * c)[None, :, None, None]
            + w[:, 0, None, None, None] * (dx * c)[None, None, :, None]
            + w[:, 2, None, None, None] * (dz * c)[None, None, None, :])


def supervoxel_subpoint_emission(source_quadrature: Tensor, tau_coarse: Tensor, mu_t_coarse: Tensor, omegas: Tensor,
                                 grid_quadrature: Grid, q: int, *, offsets: Optional[Tensor] = None) -> Tensor:
    """[k, G, Rq, Cq, Sq] attenuated emission ``S_j exp(-tau_j)``, ``tau_j = max(0, tau_cell - mu_cell
    (w . (r_j - r_cell)))`` (scatter_image.py:829), for a source already on the quadrature grid."""
    f = int(q)
    k, G = source_quadrature.shape[:2]
    Rc, Cc, Sc = (n // f for n in grid_quadrature.shape)
    if tuple(tau_coarse.shape) != (k, G, Rc, Cc, Sc):
        raise ValueError(f"tau_coarse {tuple(tau_coarse.shape)} != {(k, G, Rc, Cc, Sc)}")
    src = source_quadrature.to(tau_coarse.device, tau_coarse.dtype)
    if f == 1:
        return torch.exp(-tau_coarse) * src
    off = (_subcell_offsets(omegas.to(tau_coarse.device, tau_coarse.dtype).reshape(k, 3), grid_quadrature, f)
           if offsets is None else offsets.to(tau_coarse.device, tau_coarse.dtype))
    off = off[:, None, None, :, None, :, None, :]
    mu = mu_t_coarse.to(tau_coarse.device, tau_coarse.dtype)[None, :, :, None, :, None, :, None]
    tau = (tau_coarse[:, :, :, None, :, None, :, None] - mu * off).clamp(min=0.0)
    return (torch.exp(-tau) * src.reshape(k, G, Rc, f, Cc, f, Sc, f)).reshape(src.shape)


def _attenuated_emission(library, rho, phi_uncollided, grid_q, geometry, angle, directions, tau, mu_c, multiple, f,
                         heights, signal_weight, tau_mix=None, collapsed=False):
    """One ``citrus_attenuated_emission`` launch: (emission [k, G or 1, Rq, Cq, Sq], per-block sums
    [k, blocks, 2] of (sum_g e, h sum_g e), to be summed over the blocks). ``tau`` is
    [k, G, cells], or [k, M, cells] of traced densities whose group optical depths are ``tau_mix [G, M]`` mixes.
    ``collapsed``: ``multiple`` is the [k, 2, cells] output of ``multiple_emission`` (f = 1 with signal weights)."""
    if collapsed and (f != 1 or signal_weight is None):
        raise ValueError("the collapsed multiple-collision emission needs one point per cell and signal weights")
    a, keep, _ = _fc_args(library, rho, phi_uncollided, grid_q, geometry, angle, directions)
    k, G, P = a.K, library.num_groups, grid_q.num_voxels
    device = phi_uncollided.device
    tau = tau.to(device, torch.float32).contiguous()
    mu_c = mu_c.to(device, torch.float32).contiguous()
    multiple = multiple.to(device, torch.float32).contiguous()
    sw = None if signal_weight is None else signal_weight.to(device, torch.float32).contiguous()
    out = torch.empty((k, 1 if sw is not None else G) + tuple(grid_q.shape), device=device, dtype=torch.float32)
    nb = -(-P // _hip.EMISSION_BLOCK)
    partial = torch.empty(k, nb, 2, device=device, dtype=torch.float32)
    h = None if heights is None else heights.to(device, torch.float32).contiguous()
    channels = G
    if tau_mix is not None:
        tau_mix = tau_mix.to(device, torch.float32).contiguous()
        channels = int(tau_mix.shape[1])
        if tuple(tau_mix.shape) != (G, channels):
            raise ValueError(f"tau_mix must be [{G}, channels], got {tuple(tau_mix.shape)}")
    if int(tau.shape[1]) != channels:
        raise ValueError(f"tau has {int(tau.shape[1])} channels, expected {channels}")
    with torch.cuda.device(device):
        if collapsed:
            _hip.check(_hip.library().citrus_attenuated_emission_collapsed(
                ctypes.byref(a), tau.data_ptr(), channels, _hip.ptr(tau_mix), multiple.data_ptr(), _hip.ptr(h),
                sw.data_ptr(), out.data_ptr(), partial.data_ptr(), _hip.stream_of(device)),
                "citrus_attenuated_emission_collapsed")
        else:
            _hip.check(_hip.library().citrus_attenuated_emission(
                ctypes.byref(a), tau.data_ptr(), channels, _hip.ptr(tau_mix), mu_c.data_ptr(), multiple.data_ptr(),
                int(f), _hip.ptr(h), _hip.ptr(sw), out.data_ptr(), partial.data_ptr(), _hip.stream_of(device)),
                "citrus_attenuated_emission")
    del keep
    return out, partial


# ---- images -----------------------------------------------------------------------------------------------------

def scatter_image_from_source(source_function: SourceFunction, grid: Grid, geometry: ConeGeometry, angle: float,
                              quadrature: Quadrature, mu_t: Tensor, projectors, detector: DetectorResponse,
                              grid_response: GridResponse, library: Optional[MultigroupLibrary] = None,
                              chunk: int = 8, group_energy: Optional[Tensor] = None,
                              context: Optional[DirectionSetContext] = None, *, footprint: bool = True) -> dict:
    """Collapsed-cone image of an arbitrary emission ``source_function(directions [k, 3]) -> [k, G, R, C, S]``
    constant per voxel, attenuated from the voxel centre to the panel (scatter_image.py:737)."""
    context = _resolve_context(context, quadrature, geometry, angle, detector, grid_response, library, mu_t, chunk,
                               group_energy)
    G = int(mu_t.shape[0])
    fp = _Footprint(grid, geometry, angle, context) if footprint else None
    acc = context.new_accumulator()
    for i, directions in enumerate(context.omega_chunks):
        k = int(directions.shape[0])
        source = source_function(directions)
        if tuple(source.shape) != (k, G, *grid.shape):
            raise ValueError(f"source_function returned {tuple(source.shape)}, expected {(k, G, *grid.shape)}")
        tau = projectors.downstream_cumulative(mu_t, grid, geometry, angle, directions)
        emission = torch.exp(-tau) * source.to(mu_t.device, mu_t.dtype)
        image = projectors.oblique_project(emission.contiguous(), grid, geometry, angle, directions)
        if fp is not None:
            image = fp.blur(image, i, fp.mean_height(emission))
        context.accumulate(acc, image, i)
    return context.finalize(acc)


def sh_expand(Y: Tensor, B: Tensor, out: Optional[Tensor] = None) -> Tensor:
    """``Y [K, nsh] @ B [nsh, ncol]`` on the HIP kernel ``citrus_sh_expand`` (float32 on the matrix cores, four
    moments per step): the multiple-collision source of a chunk of cone directions from moment-major cell means
    (scatter_image.py:1246)."""
    _hip.require_gpu_tensor(Y, "Y")
    _hip.require_gpu_tensor(B, "B")
    K, nsh = (int(v) for v in Y.shape)
    if B.dim() != 2 or int(B.shape[0]) != nsh:
        raise ValueError(f"B must be [{nsh}, ncol], got {tuple(B.shape)}")
    ncol = int(B.shape[1])
    if out is None:
        out = torch.empty(K, ncol, device=B.device, dtype=torch.float32)
    lib = _hip.library()
    with torch.cuda.device(B.device):
        _hip.check(lib.citrus_sh_expand(Y.data_ptr(), K, nsh, B.data_ptr(), ncol, out.data_ptr(),
                                        _hip.stream_of(B.device)), "citrus_sh_expand")
    return out


def multiple_emission(Y: Tensor, means: Tensor, tau: Tensor, signal_weight: Tensor,
                      tau_mix: Optional[Tensor] = None) -> Tensor:
    """[K, 2, N] multiple-collision emission of a direction chunk at one point per cell, attenuated and summed over
    the groups on the HIP kernel ``citrus_multiple_emission``: ``[k, 0]`` weighted by ``signal_weight [K, G]``,
    ``[k, 1]`` unweighted. ``Y [K, nsh]`` harmonics, ``means [nsh, G, N]`` moment-major cell means, ``tau`` the
    chunk's optical depths [K, G, N], or [K, M, N] traced channels (M <= 2) mixed by ``tau_mix [G, M]``
    (scatter_image.py:1246-1262, uncollided_kernels.py:1138)."""
    for name, x in (("Y", Y), ("means", means), ("tau", tau), ("signal_weight", signal_weight)):
        _hip.require_gpu_tensor(x, name)
    K, nsh = (int(v) for v in Y.shape)
    if means.dim() != 3 or int(means.shape[0]) != nsh:
        raise ValueError(f"means must be [{nsh}, G, N], got {tuple(means.shape)}")
    G, N = int(means.shape[1]), int(means.shape[2])
    channels = G
    if tau_mix is not None:
        _hip.require_gpu_tensor(tau_mix, "tau_mix")
        channels = int(tau_mix.shape[1])
        if tuple(tau_mix.shape) != (G, channels) or channels > 2:
            raise ValueError(f"tau_mix must be [{G}, 1 or 2], got {tuple(tau_mix.shape)}")
    if tuple(tau.shape) != (K, channels, N):
        raise ValueError(f"tau must be [{K}, {channels}, {N}], got {tuple(tau.shape)}")
    if tuple(signal_weight.shape) != (K, G):
        raise ValueError(f"signal_weight must be [{K}, {G}], got {tuple(signal_weight.shape)}")
    out = torch.empty(K, 2, N, device=means.device, dtype=torch.float32)
    lib = _hip.library()
    with torch.cuda.device(means.device):
        _hip.check(lib.citrus_multiple_emission(Y.data_ptr(), K, nsh, means.data_ptr(), G, N, tau.data_ptr(),
                                                channels, _hip.ptr(tau_mix), signal_weight.data_ptr(),
                                                out.data_ptr(), _hip.stream_of(means.device)),
                   "citrus_multiple_emission")
    return out


def combined_scatter_image(library: MultigroupLibrary, rho_quadrature: Tensor, phi_uncollided: Tensor,
                           grid_quadrature: Grid, q: int, scatter_source_moments: Tensor, L: int,
                           geometry: ConeGeometry, angle: float, quadrature: Quadrature,
                           mu_t_coarse: Optional[Tensor], projectors, detector: DetectorResponse,
                           grid_response: GridResponse, context: Optional[DirectionSetContext] = None,
                           chunk: int = 8, *, first_collision_chunk: int = 4, footprint: bool = True, pairs=None,
                           grid_coarse: Optional[Grid] = None, signal_only: bool = False,
                           density_coarse: Optional[Tensor] = None, source_means: Optional[Tensor] = None) -> dict:
    """Total scatter image (first order plus multiple collision) of one view in one collapsed-cone pass
    (scatter_image.py:1141). ``scatter_source_moments`` is [G, (L+1)^2, Rc, Cc, Sc] cell means of the multiple-
    collision source moments T, or [G, (L+1)^2, 4, Rc, Cc, Sc] whose means are used. With ``density_coarse``
    [M, Rc, Cc, Sc] (the densities whose ``library.mu_t`` mix is ``mu_t_coarse``) the downstream optical depths come
    from the M density channels traced along one shared crossing sequence per direction. ``source_means``
    [(L+1)^2, G, Rc, Cc, Sc] (contiguous) are the same cell means already moment-major, the operand of the
    multiple-collision expansion (``scatter_moments_views(cell_means=True)`` writes them), which then needs no
    transposition."""
    device = phi_uncollided.device
    rho = rho_quadrature.to(device=device, dtype=torch.float32)
    f = int(q)
    grid_coarse = grid_coarse or Grid(tuple(n // f for n in grid_quadrature.shape), grid_quadrature.delta_voxel * f,
                                      grid_quadrature.voxel_row_aspect, grid_quadrature.voxel_slice_aspect,
                                      grid_quadrature.recon_slice_offset)
    if any(n % f for n in grid_quadrature.shape) or tuple(grid_coarse.shape) != tuple(n // f for n in
                                                                                      grid_quadrature.shape):
        raise ValueError(f"grid_coarse {grid_coarse.shape} is not grid_quadrature {grid_quadrature.shape} / q = {f}")
    if mu_t_coarse is None:
        mu_f = mu_t_volume(library, rho)
        mu_t_coarse = mu_f.reshape(mu_f.shape[0], grid_coarse.shape[0], f, grid_coarse.shape[1], f,
                                   grid_coarse.shape[2], f).mean(dim=(2, 4, 6))
    mu_t_coarse = mu_t_coarse.to(device, torch.float32).contiguous()
    G = int(mu_t_coarse.shape[0])
    T = scatter_source_moments
    if T.dim() == 6:
        T = T[:, :, 0]
    if T.dim() != 5 or int(T.shape[0]) != G or tuple(T.shape[2:]) != tuple(grid_coarse.shape):
        raise ValueError(f"scatter_source_moments must be [G={G}, N_sh, *{grid_coarse.shape}], got "
                         f"{tuple(scatter_source_moments.shape)}")
    if int(T.shape[1]) != (L + 1) ** 2:
        raise ValueError(f"scatter_source_moments has {int(T.shape[1])} moments, expected {(L + 1) ** 2}")
    context = _resolve_context(context, quadrature, geometry, angle, detector, grid_response, library, mu_t_coarse,
                               chunk, None)
    mix = None
    if (density_coarse is not None and int(density_coarse.shape[0]) <= 8
            and hasattr(projectors, "downstream_cumulative_mixed") and hasattr(projectors, "crossing_sequences")):
        density_coarse = density_coarse.to(device, torch.float32).contiguous()
        if tuple(density_coarse.shape) != (library.num_materials,) + tuple(grid_coarse.shape):
            raise ValueError(f"density_coarse must be [{library.num_materials}, *{grid_coarse.shape}], got "
                             f"{tuple(density_coarse.shape)}")
        mix = _cached(("mu_t_mix", str(device)), (library,),
                      lambda: library.mu_t.to(device, torch.float32).t().contiguous())  # [G, M]
    harmonics = context.spherical_harmonics(L)
    if mix is not None:
        key = ("crossing_sequences", tuple(grid_coarse.shape), tuple(grid_coarse.deltas))
        if key not in context._cache:
            context._cache[key] = projectors.crossing_sequences(grid_coarse, context.omega)
        sequences = context._cache[key]
    nsh = int(T.shape[1])
    if source_means is not None:
        if tuple(source_means.shape) != (nsh, G) + tuple(grid_coarse.shape) or not source_means.is_contiguous():
            raise ValueError(f"source_means must be contiguous [{nsh}, {G}, *{grid_coarse.shape}]")
        by_cell = source_means.reshape(nsh, -1)
    else:
        by_cell = T.to(device, torch.float32).transpose(0, 1).reshape(nsh, -1).contiguous()  # [N_sh, G cells]
    Rc, Cc, Sc = grid_coarse.shape
    fp = _Footprint(grid_quadrature, geometry, angle, context) if footprint else None
    group_weight = context.signal_group_weight if signal_only else None
    Rd, Cd = context.image_shape
    acc = torch.zeros(context.image_shape, device=device, dtype=torch.float32) if signal_only else \
        context.new_accumulator()
    # the chunks trace the one density field: one channel-last copy of it for all of them
    hold = (projectors.holding(density_coarse, grid_coarse) if mix is not None and hasattr(projectors, "holding")
            else contextlib.nullcontext())
    with hold:
        for i, directions in enumerate(context.omega_chunks):
            k = int(directions.shape[0])
            start, stop = context.chunk_bounds[i]
            rays = context.omega_chunks_f64()[i]  # the same directions as the ray kernels take them
            if mix is not None:  # density line integrals; the emission kernel mixes them into group optical depths
                tau = projectors.downstream_cumulative_mixed(density_coarse, None, grid_coarse, geometry, angle,
                                                             rays, sequences[start:stop])
            else:
                tau = projectors.downstream_cumulative(mu_t_coarse, grid_coarse, geometry, angle, rays)
            weight = None if group_weight is None else group_weight[start:stop].to(device, torch.float32).contiguous()
            collapsed = (f == 1 and weight is not None and (mix is None or int(mix.shape[1]) <= 2)
                         and (3 * G + 1) * Rc * Cc * Sc * 4 < 2 ** 32)
            if collapsed:  # signal-only at one point per cell: expansion, attenuation and group sum in one pass
                tau_c = tau.reshape(k, -1, Rc * Cc * Sc).to(device, torch.float32).contiguous()
                multiple = multiple_emission(harmonics[i].contiguous(), by_cell.view(nsh, G, -1), tau_c, weight, mix)
            else:
                multiple = sh_expand(harmonics[i], by_cell).view(k, G, Rc, Cc, Sc)
            emission, partial = _attenuated_emission(
                library, rho, phi_uncollided, grid_quadrature, geometry, angle, directions, tau, mu_t_coarse,
                multiple, f, None if fp is None else fp.heights, weight, mix, collapsed=collapsed)
            del tau, multiple
            image = projectors.oblique_project(emission, grid_quadrature, geometry, angle, rays)
            del emission
            if fp is not None and group_weight is not None:  # blur, channel weights and the sum over k, banded
                pw = None if context.planar_weight is None else context.planar_weight[start:stop].contiguous()
                fp.blur_accumulate(acc, image, i, partial, pw)
                del image
                continue
            if fp is not None:
                image = fp.blur_with(image, *fp.filters(i, partial))
            if group_weight is not None and context.planar_weight is None:
                acc.add_(image.sum(dim=(0, 1)))
            elif group_weight is not None:  # curved panel, angle-independent detector: w cos(theta_c) per channel
                acc.add_(torch.einsum("kgrc,kc->rc", image, context.planar_weight[start:stop]))
            elif signal_only and context.pixel_grid is not None:  # an angle-dependent grid: weights per pixel
                acc.add_((image * context.pixel_signal_weight(i)).sum(dim=(0, 1)))
            elif signal_only:  # curved panel: the signal weights vary along the channels
                acc.add_(torch.einsum("kgrc,kgc->rc", image, context.signal_weight[start:stop]))
            else:
                context.accumulate(acc, image, i)
            del image
    if signal_only:
        return {"signal": acc}
    return context.finalize(acc)


# ---- primary ----------------------------------------------------------------------------------------------------

def primary_subpixel_geometry(geometry: ConeGeometry, samples_per_pixel_axis: int) -> ConeGeometry:
    """The panel with n times finer pixels on which the n^2 sub-rays of every pixel are traced (:1330)."""
    n = int(samples_per_pixel_axis)
    if n < 1:
        raise ValueError("samples_per_pixel_axis must be >= 1")
    if n == 1:
        return geometry
    return replace(geometry, num_det_rows=geometry.num_det_rows * n, num_det_channels=geometry.num_det_channels * n,
                   delta_det_row=geometry.delta_det_row / n, delta_det_channel=geometry.delta_det_channel / n)


def primary_path_lengths(projectors, rho: Tensor, grid: Grid, geometry: ConeGeometry, angle: float,
                         samples_per_pixel_axis: Optional[int] = None) -> Tensor:
    """[M', Rd n, Cd n] density line integrals from the focal spot to the sub-pixel centres (:1368)."""
    n = int(samples_per_pixel_axis if samples_per_pixel_axis is not None
            else getattr(projectors, "samples_per_pixel_axis", 1) or 1)
    return projectors.primary_project(rho, grid, primary_subpixel_geometry(geometry, n), angle,
                                      samples_per_pixel_axis=1)


def _bilinear_at(table: Tensor, u: Tensor, v: Tensor, geometry: ConeGeometry) -> Tensor:
    """Bilinear value of a [rows, channels] panel table at panel coordinates (u, v), clamped to the pixel centres."""
    rows, cols = table.shape

    def bracket(x, n):
        x = x.clamp(0.0, n - 1.0)
        i0 = x.floor().long().clamp(0, max(n - 2, 0))
        i1 = (i0 + 1).clamp(max=n - 1)
        return i0, i1, (x - i0).clamp(0.0, 1.0)

    c0, c1, wc = bracket((u + geometry.det_channel_offset) / geometry.delta_det_channel + (cols - 1) / 2.0, cols)
    r0, r1, wr = bracket((v + geometry.det_row_offset) / geometry.delta_det_row + (rows - 1) / 2.0, rows)
    top = table[r0, c0] + (table[r0, c1] - table[r0, c0]) * wc
    bot = table[r1, c0] + (table[r1, c1] - table[r1, c0]) * wc
    return top + (bot - top) * wr


def primary_from_path_lengths(library: MultigroupLibrary, path: Tensor, materials: Optional[Sequence[int]],
                              geometry: ConeGeometry, angle: float, source: SourceModel, detector: DetectorResponse,
                              grid_response: GridResponse) -> dict:
    """Primary image of one view, exact in energy, from [M', Rd n, Cd n] line integrals (:1419): fine-energy
    transmission (with the filter path ``rho_f t(u, v)`` at the sub-pixel centres) averaged over the sub-pixels,
    times ``b / r^2`` at the pixel centres (the field weight is 1 on the panel, times the emission profile), then
    ``cos(theta0)``, the grid and the detector response at the incidence angle."""
    device = path.device
    Rd, Cd = geometry.num_det_rows, geometry.num_det_channels
    n = int(path.shape[1]) // Rd if path.dim() == 3 else 0
    if n < 1 or tuple(path.shape[1:]) != (Rd * n, Cd * n):
        raise ValueError(f"path must be [M', {Rd} n, {Cd} n] (n >= 1), got {tuple(path.shape)}")
    fpath = fmu = None
    if source.has_filter:
        fpath = _cached(("filter_path", n, str(device)), (source, geometry),
                        lambda: _filter_path(source, geometry, n, device))
        fmu = source.filter_attenuation_table
    T = group_transmission(library, path.to(torch.float32).contiguous(), materials=materials, filter_path=fpath,
                           filter_attenuation=fmu)  # [G, Rd n, Cd n]
    if n > 1:
        T = T.reshape(-1, Rd, n, Cd, n).mean(dim=(2, 4))
    inv_r2, cos0, grid_t, resp = _cached(("primary_view", float(angle), str(device)),
                                         (library, geometry, source, detector, grid_response),
                                         lambda: _primary_view_factors(library, geometry, angle, source, detector,
                                                                       grid_response, device))
    fluence = inv_r2[None] * T
    planar = fluence * cos0
    signal = planar * grid_t * resp
    return {"fluence": fluence, "planar_fluence": planar, "signal_per_group": signal, "signal": signal.sum(0)}


# per-view factors that depend only on the geometry, the angle and the models: recomputing them each step costs
# about a hundred small launches per view; a bounded cache keeps the last views' factors (8 MB per view at
# 16 groups x 256 x 256 pixels)
_VIEW_CACHE: "dict" = {}
_VIEW_CACHE_LIMIT = 128


def _cached(key, owners: tuple, make):
    hit = _VIEW_CACHE.get(key)
    if hit is not None and len(hit[0]) == len(owners) and all(a is b for a, b in zip(hit[0], owners)):
        return hit[1]
    value = make()
    if len(_VIEW_CACHE) >= _VIEW_CACHE_LIMIT:
        _VIEW_CACHE.pop(next(iter(_VIEW_CACHE)))
    _VIEW_CACHE[key] = (owners, value)
    return value


def _filter_path(source: SourceModel, geometry: ConeGeometry, n: int, device) -> Tensor:
    """[Rd n, Cd n] filter (bowtie) path rho_f t(u, v) at the sub-pixel centres, float32."""
    Rd, Cd = geometry.num_det_rows, geometry.num_det_channels
    sub = primary_subpixel_geometry(geometry, n)
    us, vs = sub.pixel_uv(device, torch.float64)
    thick = source.device_tables(geometry, device)[1].double()
    return (_bilinear_at(thick, us[None, :].expand(Rd * n, Cd * n), vs[:, None].expand(Rd * n, Cd * n),
                         geometry) * source.filter_density).float()


def _primary_view_factors(library, geometry, angle, source, detector, grid_response, device) -> tuple:
    """(b / r^2 [Rd, Cd], cos theta0 [Rd, Cd], grid transmission [G, Rd, Cd], detector response [G, Rd, Cd]) at the
    pixel centres of one view, float32 from float64 forms."""
    Rd, Cd = geometry.num_det_rows, geometry.num_det_channels
    pixels = geometry.pixel_centers(angle, device, torch.float64)
    r0 = torch.tensor(geometry.source_position(angle), device=device, dtype=torch.float64)
    d = pixels - r0
    r2 = (d * d).sum(-1)
    b = torch.ones_like(r2)
    if source.emission_profile is not None:
        b = source.emission_profile.to(device, torch.float64).reshape(Rd, Cd)
    omega_p = d / r2.sqrt()[..., None]
    theta0 = detector_angle(omega_p, geometry, angle, pixels)
    energy = library.group_energy.to(device, torch.float64)
    return ((b / r2).float(), torch.cos(theta0).float(),
            grid_response.transmission_at(energy, omega_p, pixels, geometry, angle).float(),
            detector.at(energy, theta0).float())


def primary_image(library: MultigroupLibrary, rho: Tensor, grid: Grid, geometry: ConeGeometry, angle: float,
                  source: SourceModel, projectors, detector: DetectorResponse, grid_response: GridResponse,
                  samples_per_pixel_axis: Optional[int] = None, materials: Optional[Sequence[int]] = None) -> dict:
    """Primary image of one view from the density maps of the materials present (:1474); ``materials`` names them
    when the caller already knows (else one host synchronization finds them)."""
    if int(rho.shape[0]) != library.num_materials:
        raise ValueError(f"rho has {int(rho.shape[0])} materials, the library {library.num_materials}")
    present = (list(materials) if materials is not None else present_materials(rho)) or [0]
    if present == list(range(library.num_materials)) and rho.is_contiguous():
        sel = rho
    else:
        sel = rho.index_select(0, _device_index(present, rho.device)).contiguous()
    path = primary_path_lengths(projectors, sel, grid, geometry, angle, samples_per_pixel_axis)
    return primary_from_path_lengths(library, path, present, geometry, angle, source, detector, grid_response)


def scatter_fraction(scatter_signal: Tensor, primary_signal: Tensor) -> Tensor:
    """``S / (P + S)`` elementwise, 0 where ``P + S = 0`` (:1509)."""
    den = primary_signal + scatter_signal
    safe = torch.where(den > 0, den, torch.ones_like(den))
    return torch.where(den > 0, scatter_signal / safe, torch.zeros_like(den))


def scatter_to_primary_ratio(scatter_signal: Tensor, primary_signal: Tensor) -> Tensor:
    """``S / P`` elementwise, 0 where ``P = 0`` (the SPR of citrus.py:1969)."""
    safe = torch.where(primary_signal > 0, primary_signal, torch.ones_like(primary_signal))
    return torch.where(primary_signal > 0, scatter_signal / safe, torch.zeros_like(primary_signal))
