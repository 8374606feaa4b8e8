"""Exact line integrals of voxel fields and the ``Projectors`` interface, on the HIP traversal kernels.

Keeps the interface of ``citrus/scatter_solver/ray_tracing.py``: the ``Projectors`` protocol of five operations
(:27-115), ``line_integrals`` (:205) and ``SiddonProjectors`` (:291) with ``upstream_cumulative`` (:405),
``upstream_to_points`` (:426), ``primary_project`` (:447), ``downstream_cumulative`` (:501) and ``oblique_project``
(:548). Every operation runs one HIP launch that generates its rays on the device (``citrus_ray_*``), with the
intersection setup in float64 and compensated float32 sums. There is no CPU path: float32 GPU tensors only.
"""

from __future__ import annotations

import contextlib
import math
from typing import Optional, Protocol

import numpy as np
import torch
from torch import Tensor

from .. import _hip
from .geometry import ConeGeometry, Grid

__all__ = ["Projectors", "SiddonProjectors", "HipProjectors", "line_integrals", "trace_voxel_sequences"]


class Projectors(Protocol):
    """The five ray-tracing operations of the scatter solver, one view angle, batched over a leading group axis."""

    def upstream_cumulative(self, mu: Tensor, grid: Grid, geometry: ConeGeometry, angle: float) -> Tensor: ...

    def upstream_to_points(self, mu: Tensor, grid: Grid, geometry: ConeGeometry, angle: float,
                           points: Tensor) -> Tensor: ...

    def primary_project(self, mu: Tensor, grid: Grid, geometry: ConeGeometry, angle: float) -> Tensor: ...

    def downstream_cumulative(self, mu: Tensor, grid: Grid, geometry: ConeGeometry, angle: float,
                              omegas: Tensor) -> Tensor: ...

    def oblique_project(self, field: Tensor, grid: Grid, geometry: ConeGeometry, angle: float, omegas: Tensor,
                        out: Optional[Tensor] = None) -> Tensor: ...


def _box(grid: Grid) -> np.ndarray:
    return np.array(list(grid.lower_corner()) + list(grid.deltas), dtype=np.float64)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _panel(geometry: ConeGeometry, angle: float) -> tuple[np.ndarray, np.ndarray]:
    uh, vh, nh = geometry.panel_axes(angle)
    values = (list(geometry.source_position(angle)) + list(geometry.panel_center(angle)) + list(uh) + list(vh)
              + list(nh) + [geometry.source_detector_dist, geometry.delta_det_row, geometry.delta_det_channel,
                            geometry.det_row_offset, geometry.det_channel_offset])
    ints = [int(bool(geometry.use_curved_detector)), geometry.num_det_rows, geometry.num_det_channels]
    return np.array(values, dtype=np.float64), np.array(ints, dtype=np.int32)


def _channel_last(volume: Tensor, grid: Grid) -> Tensor:
    """[G, R, C, S] float32 GPU field -> contiguous [R, C, S, G]."""
    if volume.dim() != 4 or tuple(volume.shape[1:]) != tuple(grid.shape):
        raise ValueError(f"field must be [G, {grid.shape}], got {tuple(volume.shape)}")
    if volume.device.type != "cuda" or volume.dtype != torch.float32:
        raise ValueError(f"the HIP traversal needs float32 GPU fields, got {volume.dtype} on {volume.device}")
    if volume.shape[0] > 16:
        raise ValueError("at most 16 channels per traversal")
    return volume.permute(1, 2, 3, 0).contiguous()


def line_integrals(volume: Tensor, starts: Tensor, ends: Tensor, grid: Grid) -> Tensor:
    """Exact integrals of ``volume [G, R, C, S]`` along the segments ``starts -> ends`` ([N, 3], cm): [G, N]."""
    G = int(volume.shape[0])
    vol = _channel_last(volume, grid)
    p0 = starts.to(device=volume.device, dtype=torch.float64).reshape(-1, 3).contiguous()
    p1 = ends.to(device=volume.device, dtype=torch.float64).reshape(-1, 3).contiguous()
    shared = p0.shape[0] == 1 and p1.shape[0] != 1
    if not shared and p0.shape != p1.shape:
        raise ValueError("starts and ends must have the same number of rays")
    N = int(p1.shape[0])
    out = torch.empty(G, N, device=volume.device, dtype=torch.float32)
    if N == 0:
        return out
    R, C, S = grid.shape
    box = _box(grid)
    with torch.cuda.device(volume.device):
        _hip.check(_hip.library().citrus_ray_segments(vol.data_ptr(), R, C, S, G, _ptr(box), p0.data_ptr(),
                                                      p1.data_ptr(), int(shared), N, out.data_ptr(),
                                                      _hip.stream_of(volume.device)), "citrus_ray_segments")
    return out


def trace_voxel_sequences(starts: Tensor, ends: Tensor, grid: Grid, max_len: Optional[int] = None
                          ) -> tuple[Tensor, Tensor, Tensor]:
    """Voxel sequence of every segment on the device: (flat index [N, K] int32 (-1 padded), parameter length
    [N, K] float32, count [N] int32), K = max_len (default: the longest possible traversal)."""
    device = starts.device
    p0 = starts.to(torch.float64).reshape(-1, 3).contiguous()
    p1 = ends.to(device=device, dtype=torch.float64).reshape(-1, 3).contiguous()
    N = int(p0.shape[0])
    R, C, S = grid.shape
    K = int(max_len) if max_len is not None else R + C + S + 3
    idx = torch.full((N, K), -1, device=device, dtype=torch.int32)
    seg = torch.zeros(N, K, device=device, dtype=torch.float32)
    count = torch.zeros(N, device=device, dtype=torch.int32)
    box = _box(grid)
    with torch.cuda.device(device):
        _hip.check(_hip.library().citrus_ray_trace_indices(R, C, S, _ptr(box), p0.data_ptr(), p1.data_ptr(), N, K,
                                                           idx.data_ptr(), seg.data_ptr(), count.data_ptr(),
                                                           _hip.stream_of(device)), "citrus_ray_trace_indices")
    return idx, seg, count


class HipProjectors:
    """``Projectors`` on the HIP traversal kernels (the drop-in for ``SiddonProjectors``, ray_tracing.py:291).

    Args:
        samples_per_pixel_axis: default sub-rays per pixel axis of ``primary_project``. Defaults to 1.
        backend: "auto" or "hip" (both run the HIP kernels; there is no other path).
    """

    def __init__(self, samples_per_pixel_axis: int = 1, backend: str = "auto", **_ignored):
        if samples_per_pixel_axis < 1:
            raise ValueError("samples_per_pixel_axis must be >= 1")
        if backend in ("torch", "triton"):
            raise ValueError(f"backend {backend!r} does not exist in this implementation (use 'auto' or 'hip')")
        if backend not in ("auto", "hip"):
            raise ValueError("backend must be 'auto' or 'hip'")
        self.samples_per_pixel_axis = int(samples_per_pixel_axis)
        self.backend = backend
        self._held = None  # (field, its channel-last copy) while ``holding`` it

    @contextlib.contextmanager
    def holding(self, field: Tensor, grid: Grid):
        """Within the block the traversals of ``field`` (the same tensor object) reuse one channel-last copy of it
        instead of forming one per call: for the direction chunks of one image, which trace one density field
        several times and do not change it. The copy is dropped on leaving the block."""
        previous = self._held
        self._held = (field, _channel_last(field, grid))
        try:
            yield self
        finally:
            self._held = previous

    def _channel_last_of(self, field: Tensor, grid: Grid) -> Tensor:
        held = self._held
        if held is not None and held[0] is field:
            return held[1]
        return _channel_last(field, grid)

    def line_integrals(self, volume: Tensor, starts: Tensor, ends: Tensor, grid: Grid) -> Tensor:
        return line_integrals(volume, starts, ends, grid)

    def upstream_cumulative(self, mu: Tensor, grid: Grid, geometry: ConeGeometry, angle: float) -> Tensor:
        """[G, R, C, S] optical depth from the focal spot to every voxel centre."""
        G = int(mu.shape[0])
        vol = _channel_last(mu, grid)
        out = torch.empty(mu.shape, device=mu.device, dtype=torch.float32)
        R, C, S = grid.shape
        x0, y0, z0 = grid.lower_corner()
        dx, dy, dz = grid.deltas
        origin = np.array([x0 + 0.5 * dx, y0 + 0.5 * dy, z0 + 0.5 * dz], dtype=np.float64)
        delta = np.array([dx, dy, dz], dtype=np.float64)
        shape = np.array([R, C, S], dtype=np.int32)
        point = np.array(geometry.source_position(angle), dtype=np.float64)
        with torch.cuda.device(mu.device):
            _hip.check(_hip.library().citrus_ray_lattice(vol.data_ptr(), R, C, S, G, _ptr(_box(grid)), _ptr(point),
                                                         _ptr(origin), _ptr(delta), _ptr(shape), out.data_ptr(),
                                                         _hip.stream_of(mu.device)), "citrus_ray_lattice")
        return out

    def upstream_to_points(self, mu: Tensor, grid: Grid, geometry: ConeGeometry, angle: float,
                           points: Tensor) -> Tensor:
        """[G, N] optical depth from the focal spot to each point."""
        source = torch.tensor(geometry.source_position(angle), dtype=torch.float64, device=mu.device)[None]
        return line_integrals(mu, source, points.to(mu.device), grid)

    def primary_project(self, mu: Tensor, grid: Grid, geometry: ConeGeometry, angle: float,
                        samples_per_pixel_axis: Optional[int] = None) -> Tensor:
        """[G, rows, channels] optical depth from the focal spot to every pixel (sub-ray average for n > 1)."""
        n = self.samples_per_pixel_axis if samples_per_pixel_axis is None else int(samples_per_pixel_axis)
        G = int(mu.shape[0])
        vol = _channel_last(mu, grid)
        out = torch.empty(G, geometry.num_det_rows, geometry.num_det_channels, device=mu.device,
                          dtype=torch.float32)
        panel, ints = _panel(geometry, angle)
        R, C, S = grid.shape
        with torch.cuda.device(mu.device):
            _hip.check(_hip.library().citrus_ray_primary(vol.data_ptr(), R, C, S, G, _ptr(_box(grid)), _ptr(panel),
                                                         _ptr(ints), n, out.data_ptr(), _hip.stream_of(mu.device)),
                       "citrus_ray_primary")
        return out

    def downstream_cumulative(self, mu: Tensor, grid: Grid, geometry: ConeGeometry, angle: float,
                              omegas: Tensor) -> Tensor:
        """[K, G, R, C, S] optical depth from every voxel centre to the panel along each direction; +inf where
        ``omega . (-n_hat) <= 0``."""
        G = int(mu.shape[0])
        vol = _channel_last(mu, grid)
        w = omegas.to(device=mu.device, dtype=torch.float64).reshape(-1, 3).contiguous()
        K = int(w.shape[0])
        out = torch.empty((K,) + tuple(mu.shape), device=mu.device, dtype=torch.float32)
        panel, ints = _panel(geometry, angle)
        R, C, S = grid.shape
        with torch.cuda.device(mu.device):
            _hip.check(_hip.library().citrus_ray_downstream(vol.data_ptr(), R, C, S, G, _ptr(_box(grid)),
                                                            _ptr(panel), _ptr(ints), w.data_ptr(), K, out.data_ptr(),
                                                            _hip.stream_of(mu.device)), "citrus_ray_downstream")
        return out

    def crossing_sequences(self, grid: Grid, omegas: Tensor) -> Tensor:
        """[K, bytes] device records of the voxel crossings of every direction from a voxel centre
        (``citrus_ray_crossing_sequences``); they depend on the directions and the voxel sizes only."""
        w = omegas.to(dtype=torch.float64).reshape(-1, 3).contiguous()
        if w.device.type != "cuda":
            raise ValueError("crossing sequences are built on the GPU: pass device directions")
        R, C, S = grid.shape
        lib = _hip.library()
        nbytes = int(lib.citrus_ray_crossing_sequence_bytes(R, C, S))
        if nbytes < 0:
            raise ValueError(f"crossing sequences need 1..511 voxels per axis, got {grid.shape}")
        K = int(w.shape[0])
        records = torch.empty(K, nbytes, device=w.device, dtype=torch.uint8)
        with torch.cuda.device(w.device):
            _hip.check(lib.citrus_ray_crossing_sequences(R, C, S, _ptr(_box(grid)), w.data_ptr(), K,
                                                         records.data_ptr(), _hip.stream_of(w.device)),
                       "citrus_ray_crossing_sequences")
        return records

    def downstream_cumulative_mixed(self, density: Tensor, mix: Optional[Tensor], grid: Grid, geometry: ConeGeometry,
                                    angle: float, omegas: Tensor, sequences: Optional[Tensor] = None) -> Tensor:
        """``downstream_cumulative`` of the fields ``sum_c mix[g, c] density[c]`` ([K, G, R, C, S]; ``mix`` [G, M],
        or None for the densities' own integrals) along one shared crossing sequence per direction
        (``citrus_ray_downstream_shared``): only the M density channels are traced. ``sequences`` are the
        ``crossing_sequences`` records of these directions on this grid (built here when None)."""
        M = int(density.shape[0])
        vol = self._channel_last_of(density, grid)
        w = omegas.to(device=density.device, dtype=torch.float64).reshape(-1, 3).contiguous()
        K = int(w.shape[0])
        R, C, S = grid.shape
        if mix is not None:
            mix = mix.to(device=density.device, dtype=torch.float32).contiguous()
            if mix.dim() != 2 or int(mix.shape[1]) != M:
                raise ValueError(f"mix must be [G, {M}], got {tuple(mix.shape)}")
        nout = M if mix is None else int(mix.shape[0])
        out = torch.empty((K, nout, R, C, S), device=density.device, dtype=torch.float32)
        if K == 0:
            return out
        if sequences is None:
            sequences = self.crossing_sequences(grid, w)
        nbytes = int(_hip.library().citrus_ray_crossing_sequence_bytes(R, C, S))
        if (tuple(sequences.shape) != (K, nbytes) or sequences.dtype != torch.uint8 or not sequences.is_contiguous()
                or sequences.device != density.device):
            raise ValueError(f"sequences must be contiguous uint8 [{K}, {nbytes}] on {density.device}")
        panel, ints = _panel(geometry, angle)
        with torch.cuda.device(density.device):
            _hip.check(_hip.library().citrus_ray_downstream_shared(
                vol.data_ptr(), R, C, S, M, _ptr(_box(grid)), _ptr(panel), _ptr(ints), w.data_ptr(), K,
                sequences.data_ptr(), _hip.ptr(mix), nout, out.data_ptr(), _hip.stream_of(density.device)),
                "citrus_ray_downstream_shared")
        return out

    def oblique_project(self, field: Tensor, grid: Grid, geometry: ConeGeometry, angle: float, omegas: Tensor,
                        out: Optional[Tensor] = None) -> Tensor:
        """[K, G, rows, channels] line integrals of the field along each direction through every pixel centre."""
        w = omegas.to(device=field.device, dtype=torch.float64).reshape(-1, 3).contiguous()
        K = int(w.shape[0])
        if field.dim() == 4:
            vol = _channel_last(field, grid)
            G, k_stride = int(field.shape[0]), 0
        elif field.dim() == 5 and field.shape[0] == K:
            G = int(field.shape[1])
            if tuple(field.shape[2:]) != tuple(grid.shape):
                raise ValueError(f"field spatial shape {tuple(field.shape[2:])} != grid shape {grid.shape}")
            if field.device.type != "cuda" or field.dtype != torch.float32:
                raise ValueError("the HIP traversal needs float32 GPU fields")
            vol = field.permute(0, 2, 3, 4, 1).contiguous()
            k_stride = grid.num_voxels * G
        else:
            raise ValueError(f"field must be [G, ...] or [K={K}, G, ...], got {tuple(field.shape)}")
        Rd, Cd = geometry.num_det_rows, geometry.num_det_channels
        if out is None:
            out = torch.empty(K, G, Rd, Cd, device=field.device, dtype=torch.float32)
        if out.shape != (K, G, Rd, Cd) or out.stride(3) != 1 or out.stride(2) != Cd:
            raise ValueError("out must be [K, G, rows, channels] with contiguous rows")
        panel, ints = _panel(geometry, angle)
        R, C, S = grid.shape
        with torch.cuda.device(field.device):
            _hip.check(_hip.library().citrus_ray_oblique(vol.data_ptr(), k_stride, R, C, S, G, _ptr(_box(grid)),
                                                         _ptr(panel), _ptr(ints), w.data_ptr(), K, out.data_ptr(),
                                                         out.stride(0), out.stride(1),
                                                         _hip.stream_of(field.device)), "citrus_ray_oblique")
        return out


SiddonProjectors = HipProjectors


# The reference's name of its default projector class (ray_tracing.py:291); callers that construct
# ``SiddonProjectors()`` get the HIP traversal.
SiddonProjectors = HipProjectors
