"""Uncollided fluence on the HIP kernels.

Keeps the interface of ``citrus/scatter_solver/uncollided.py``: ``present_materials`` (:27), ``uncollided_fluence``
(:99), ``uncollided_path_lengths`` (:137) and ``uncollided_from_path_lengths`` (:160):
``phi_UC,g(r) = b(Omega_UC(r)) / |r - r0|^2 sum_{E in g} w_E exp(-sum_m P_m(r) m_m(E) - mu_f(E) P_f)``. The upstream
density line integrals come from the projector's ``upstream_cumulative`` (one traversal launch), the rest is one
fused launch (``citrus_uncollided_fluence``).
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence, Union

import numpy as np
import torch
from torch import Tensor

from .. import _hip
from .cross_sections import MultigroupLibrary, fine_device_tables
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
    mu, grp, w = fine_device_tables(library, device, materials)
    if int(mu.shape[0]) != Mp:
        raise ValueError(f"path_lengths has {Mp} materials, expected {int(mu.shape[0])}")
    paths = path_lengths.reshape(Mp, -1).contiguous()
    _hip.require_gpu_tensor(paths, "path_lengths")
    profile, thickness, mu_f = source.device_tables(geometry, device)
    values, ints, table = source.host_block(geometry, angle)
    x0, y0, z0 = points.lower_corner()
    dx, dy, dz = points.deltas
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
    traced = rho if present is None else rho.index_select(0, _device_index(present, rho.device))
    paths = uncollided_path_lengths(projectors, traced.contiguous(), grid, geometry, angle, point_grid)
    return uncollided_from_path_lengths(library, paths, present, point_grid, geometry, angle, source)


# ==================================================================================================================
# first-collision source at the sweep directions (uncollided.py:200-656 of the reference)
# ==================================================================================================================

def primary_directions_to_voxels(grid: Grid, geometry: ConeGeometry, angle: float, device=None,
                                 dtype=torch.float32) -> Tensor:
    """[rows, cols, slices, 3] unit vectors ``(r - r0)/|r - r0|`` from the focal spot to the voxel centres
    (uncollided.py:74-96), formed in float64."""
    r = grid.center_coordinates(device, torch.float64)
    r0 = _device_vector(geometry.source_position(angle), device, torch.float64)
    d = r - r0
    return (d / torch.linalg.norm(d, dim=-1, keepdim=True)).to(dtype)


_FC_TABLES: dict = {}


def _fc_tables(library: MultigroupLibrary, device: torch.device):
    """Device tables of the first-collision kernels, cached per library and device: the node-major, downscatter-masked
    eta table [n_eta][M][G][GP] (GP from the kernel's group count), mu_s_l [M][G][G][L+1] and the adaptive map."""
    key = (id(library), str(device))
    hit = _FC_TABLES.get(key)
    if hit is not None and hit[0] is library:
        return hit[1]
    M, G = library.num_materials, library.num_groups
    gp = 4 if G <= 4 else 8 if G <= 8 else 16 if G <= 16 else 32
    down = torch.triu(torch.ones(G, G, dtype=torch.float64))  # [g_in, g_out], 1 where g_out >= g_in
    masked = library.mu_s_eta * down[None, :, :, None]  # [M, Gi, Go, N]
    table = torch.zeros(library.mu_s_eta.shape[-1], M, G, gp, dtype=torch.float64)
    table[..., :G] = masked.permute(3, 0, 1, 2)
    tables = (table.to(device, torch.float32).contiguous(),
              library.mu_s_l.to(device, torch.float32).contiguous(),
              None if library.eta_map is None else library.eta_map.to(device, torch.float32).contiguous())
    _FC_TABLES[key] = (library, tables)
    return tables


_SLOT_MATERIALS: dict = {}
_INDEX: dict = {}


_VECTORS: dict = {}


def _device_vector(values, device, dtype) -> Tensor:
    """A small constant vector on the device, kept per (values, device, dtype) in a bounded table."""
    key = (tuple(float(v) for v in values), str(device), dtype)
    hit = _VECTORS.get(key)
    if hit is None:
        if len(_VECTORS) >= 4096:
            _VECTORS.clear()
        hit = _VECTORS[key] = torch.tensor(key[0], device=device, dtype=dtype)
    return hit


def _device_index(values, device) -> Tensor:
    """A small int64 index tensor on the device, kept per (values, device): no host-to-device copy per call."""
    key = (tuple(int(v) for v in values), str(device))
    hit = _INDEX.get(key)
    if hit is None:
        hit = _INDEX[key] = torch.tensor(key[0], device=device, dtype=torch.int64)
    return hit


def _material_slots(rho: Tensor):
    """(slot material int32 [n][P], slot density float32 [n][P]) with n = the largest count of nonzero materials at
    one point (at least 1, at most 4): every material when M <= 4, else each point's nonzero ones (the reference's
    MaterialPairs, cross_sections.py:656, generalised)."""
    M = int(rho.shape[0])
    flat = rho.reshape(M, -1).to(torch.float32)
    if M <= 4:
        key = (M, int(flat.shape[1]), str(rho.device))
        mats = _SLOT_MATERIALS.get(key)
        if mats is None:  # every material in its own slot: the same table for every call on this grid size
            mats = torch.arange(M, device=rho.device, dtype=torch.int32)[:, None].expand(M, flat.shape[1]).contiguous()
            _SLOT_MATERIALS[key] = mats
        return mats, flat.contiguous()
    n = int((flat != 0).sum(0).max())
    if n > 4:
        raise ValueError(f"{n} nonzero materials at one point; the first-collision kernels take at most 4")
    n = max(n, 1)
    dens, idx = torch.topk(flat.abs(), n, dim=0)
    dens = torch.gather(flat, 0, idx)
    return idx.to(torch.int32).contiguous(), dens.contiguous()


def _fc_args(library, rho, phi_uncollided, grid: Grid, geometry, angle, omegas):
    device = phi_uncollided.device
    _hip.require_gpu_tensor(phi_uncollided, "phi_uncollided")
    if tuple(phi_uncollided.shape[1:]) != tuple(grid.shape) or tuple(rho.shape[1:]) != tuple(grid.shape):
        raise ValueError(f"rho {tuple(rho.shape)} and phi_uncollided {tuple(phi_uncollided.shape)} must lie on "
                         f"the grid {grid.shape}")
    if int(phi_uncollided.shape[0]) != library.num_groups or int(rho.shape[0]) != library.num_materials:
        raise ValueError("rho / phi_uncollided disagree with the library's materials / groups")
    table, mu_s_l, emap = _fc_tables(library, device)
    mats, dens = _material_slots(rho.to(device))
    om = omegas.reshape(-1, 3).to(device, torch.float32).contiguous()
    a = _hip.FcArgs()
    a.M, a.G, a.n_eta = library.num_materials, library.num_groups, int(library.mu_s_eta.shape[-1])
    a.n_map = 0 if emap is None else int(emap.numel())
    a.table, a.eta_map = table.data_ptr(), _hip.ptr(emap)
    a.n_slots, a.slot_material, a.slot_density = int(mats.shape[0]), mats.data_ptr(), dens.data_ptr()
    a.phi_uc = phi_uncollided.data_ptr()
    a.R, a.C, a.S = grid.shape
    x, y, z = grid.axis_coordinates(dtype=torch.float64)
    a.lattice[:] = [float(x[0]), float(y[0]), float(z[0]), *grid.deltas]
    a.focal[:] = list(geometry.source_position(angle))
    a.parallel = 0
    a.omegas, a.K = om.data_ptr(), int(om.shape[0])
    keep = (table, mu_s_l, emap, mats, dens, om)  # alive until the launch is queued
    return a, keep, mu_s_l


def first_collision_source_at_directions(library: MultigroupLibrary, rho: Tensor, phi_uncollided: Tensor, grid: Grid,
                                         geometry: ConeGeometry, angle: float, omegas: Tensor, chunk: int = 4,
                                         backend: str = "auto", pairs=None) -> Tensor:
    """[K, G, rows, cols, slices] first-collision source at the voxel centres of ``grid`` for each direction
    (uncollided.py:200), one HIP launch (``citrus_fc_source``). ``chunk`` and ``pairs`` are accepted for the
    reference's signature; the kernel takes every direction at once and forms material slots itself."""
    if backend not in ("auto", "hip"):
        raise ValueError(f"backend must be 'auto' or 'hip', got {backend!r}")
    a, keep, _ = _fc_args(library, rho, phi_uncollided, grid, geometry, angle, omegas)
    out = torch.empty((a.K, library.num_groups) + tuple(grid.shape), device=phi_uncollided.device,
                      dtype=torch.float32)
    with torch.cuda.device(phi_uncollided.device):
        _hip.check(_hip.library().citrus_fc_source(ctypes.byref(a), out.data_ptr(),
                                                   _hip.stream_of(phi_uncollided.device)), "citrus_fc_source")
    del keep
    return out


def first_collision_exact_moments(library: MultigroupLibrary, rho: Tensor, phi_uncollided: Tensor,
                                  l_max: int = 1) -> Tensor:
    """[l_max+1, G, R, C, S] ``M_l[g] = sum_{g' <= g} sum_m rho_m mu_s_l^{g'->g} phi_UC,g'`` (uncollided.py:357)."""
    if library.legendre_order < l_max:
        raise ValueError(f"library Legendre order {library.legendre_order} < {l_max}")
    device, dtype = phi_uncollided.device, phi_uncollided.dtype
    G = phi_uncollided.shape[0]
    transfer = library.mu_s_l[..., : l_max + 1].to(device=device, dtype=dtype)
    down = torch.triu(torch.ones(G, G, device=device, dtype=dtype))
    return torch.einsum("mabl,ab,m...,a...->lb...", transfer, down, rho.to(device=device, dtype=dtype),
                        phi_uncollided)


@dataclass(frozen=True, eq=False)
class FirstCollisionSamplingCorrection:
    """Per-point, per-group correction ``S <- scale S (1 + beta omega . Omega_UC)`` (uncollided.py:387-433).

    Attributes: scale [G, R, C, S], beta [G, R, C, S] in [-1, 1], omega_uncollided [R, C, S, 3]."""

    scale: Tensor
    beta: Tensor
    omega_uncollided: Tensor

    def apply(self, source: Tensor, omegas: Tensor) -> Tensor:
        """[K, G, R, C, S] corrected source from the samples at ``omegas`` [K, 3]."""
        directions = omegas.to(device=source.device, dtype=source.dtype).reshape(-1, 3)
        eta = torch.einsum("rcsx,kx->krcs", self.omega_uncollided.to(source.dtype), directions)
        return source * self.scale * (1.0 + self.beta[None] * eta[:, None])


def _fc_sums_raw(library, rho, phi_uncollided, grid, geometry, angle, omegas, weights,
                 samples: Optional[Tensor] = None, blocked_d8: int = 0, factors=None) -> Tensor:
    """[5, G, R, C, S] float32 I0, I1, I2, E0, E1 from one launch of ``citrus_fc_sums`` (which also stores the
    uncorrected samples [G, R, C, S, K] into ``samples`` when given), or with ``blocked_d8`` (K = 8 blocked_d8
    octant-major sweep directions) of ``citrus_fc_sums_blocked``, the samples [G, blocks 8 64 dp + pad] in the
    sweep's block order (each group's region followed by ``FC_GROUP_PAD`` unused floats); ``factors`` = (scale, beta
    [G, blocks 64 + pad], u [blocks 64, 4]) then receive the correction factors and primary directions per cell in
    block order."""
    if library.legendre_order < 1:
        raise ValueError("the sampling correction needs a library of Legendre order >= 1")
    a, keep, mu_s_l = _fc_args(library, rho, phi_uncollided, grid, geometry, angle, omegas)
    w = weights.reshape(-1).to(phi_uncollided.device, torch.float32).contiguous()
    if int(w.numel()) != a.K:
        raise ValueError("weights and omegas disagree in length")
    out = torch.empty((5, library.num_groups) + tuple(grid.shape), device=phi_uncollided.device, dtype=torch.float32)
    if samples is not None:
        if blocked_d8:
            want = (library.num_groups, fc_blocks(grid.shape) * 8 * 64 * (-(-blocked_d8 // 4) * 4) + _hip.FC_GROUP_PAD)
        else:
            want = (library.num_groups,) + tuple(grid.shape) + (a.K,)
        if (tuple(samples.shape) != want or samples.dtype != torch.float32 or not samples.is_contiguous()
                or samples.device != phi_uncollided.device):
            raise ValueError(f"samples must be a contiguous float32 {want} tensor on {phi_uncollided.device}")
    with torch.cuda.device(phi_uncollided.device):
        if blocked_d8:
            if samples is None:
                raise ValueError("the block order needs a samples tensor")
            fs, fb, fu = (None, None, None) if factors is None else factors
            if factors is not None:
                cells = fc_blocks(grid.shape) * 64
                for t, shp, name in ((fs, (library.num_groups, cells + _hip.FC_GROUP_PAD), "scale"),
                                     (fb, (library.num_groups, cells + _hip.FC_GROUP_PAD), "beta"),
                                     (fu, (cells, 4), "u")):
                    if (tuple(t.shape) != shp or t.dtype != torch.float32 or not t.is_contiguous()
                            or t.device != phi_uncollided.device):
                        raise ValueError(f"{name} must be a contiguous float32 {shp} tensor on {phi_uncollided.device}")
            _hip.check(_hip.library().citrus_fc_sums_blocked(
                ctypes.byref(a), w.data_ptr(), mu_s_l.data_ptr(), int(mu_s_l.shape[-1]), out.data_ptr(),
                samples.data_ptr(), int(blocked_d8), _hip.ptr(fs), _hip.ptr(fb), _hip.ptr(fu),
                _hip.stream_of(phi_uncollided.device)), "citrus_fc_sums_blocked")
        else:
            _hip.check(_hip.library().citrus_fc_sums(ctypes.byref(a), w.data_ptr(), mu_s_l.data_ptr(),
                                                     int(mu_s_l.shape[-1]), out.data_ptr(), _hip.ptr(samples),
                                                     _hip.stream_of(phi_uncollided.device)), "citrus_fc_sums")
    del keep
    return out


def first_collision_quadrature_sums(library: MultigroupLibrary, rho: Tensor, phi_uncollided: Tensor, grid: Grid,
                                    geometry: ConeGeometry, angle: float, omegas: Tensor, weights: Tensor,
                                    omega_uncollided: Optional[Tensor] = None, chunk: int = 4, batch: int = 8,
                                    pairs=None) -> Tensor:
    """[3, G, R, C, S] float64 ``I_j = sum_k w_k (omega_k . Omega_UC)^j S_k`` (uncollided.py:538), accumulated in
    float32 over all directions in one kernel like the reference's fused path (uncollided_kernels.py:637)."""
    return _fc_sums_raw(library, rho, phi_uncollided, grid, geometry, angle, omegas, weights)[:3].double()


def first_collision_correction_factors(I: Tensor, E: Tensor) -> tuple[Tensor, Tensor]:
    """(scale, beta) in float64 from the sums and exact moments (uncollided.py:611-640): ``c = E1/E0``,
    ``beta = (c I0 - I1)/(I2 - c I1)`` clamped to [-1, 1] (0 where E0 <= 0 or the denominator is below 1e-12 |I0|),
    ``scale = E0/(I0 + beta I1)`` (1 where that denominator is not positive)."""
    I = I.double()
    E = E.to(device=I.device, dtype=torch.float64)
    I0, I1, I2 = I[0], I[1], I[2]
    E0, E1 = E[0], E[1]
    pos = E0 > 0
    c = torch.where(pos, E1 / torch.where(pos, E0, torch.ones_like(E0)), torch.zeros_like(E0))
    den = I2 - c * I1
    ok = pos & (den.abs() > 1e-12 * I0.abs())
    beta = torch.where(ok, (c * I0 - I1) / torch.where(ok, den, torch.ones_like(den)), torch.zeros_like(den))
    beta = beta.clamp(-1.0, 1.0)
    d0 = I0 + beta * I1
    scale = torch.where(d0 > 0, E0 / torch.where(d0 > 0, d0, torch.ones_like(d0)), torch.ones_like(d0))
    return scale, beta


def first_collision_sampling_correction(library: MultigroupLibrary, rho: Tensor, phi_uncollided: Tensor, grid: Grid,
                                        geometry: ConeGeometry, angle: float, omegas: Tensor, weights: Tensor,
                                        chunk: int = 4, batch: int = 8,
                                        pairs=None) -> FirstCollisionSamplingCorrection:
    """Correction that makes the direction quadrature integrate the sampled source exactly for l = 0, 1
    (uncollided.py:436-461): sums and exact moments from one ``citrus_fc_sums`` launch (float32), factors in
    float64, returned in float32."""
    sums = _fc_sums_raw(library, rho, phi_uncollided, grid, geometry, angle, omegas, weights).double()
    scale, beta = first_collision_correction_factors(sums[:3], sums[3:])
    ou = primary_directions_to_voxels(grid, geometry, angle, phi_uncollided.device, torch.float32)
    return FirstCollisionSamplingCorrection(scale=scale.float().contiguous(), beta=beta.float().contiguous(),
                                            omega_uncollided=ou)


def first_collision_corrected_samples(library: MultigroupLibrary, rho: Tensor, phi_uncollided: Tensor, grid: Grid,
                                      geometry: ConeGeometry, angle: float, omegas: Tensor, weights: Tensor,
                                      out: Optional[Tensor] = None
                                      ) -> tuple[Tensor, FirstCollisionSamplingCorrection]:
    """Corrected first-collision source at every point and sweep direction, group-major [G, R, C, S, K], with its
    correction, from one evaluation of the source: ``citrus_fc_sums`` stores the samples while it sums them, the
    factors follow in float64 (uncollided.py:436-461, :611-640) and ``citrus_fc_correct`` applies them in place.
    With one point per cell this is ``first_collision_moments_at_directions(..., layout="group_major_means")``
    bit for bit, at the cost of one source evaluation instead of two."""
    K = int(omegas.reshape(-1, 3).shape[0])
    shape = (library.num_groups,) + tuple(grid.shape) + (K,)
    if out is None:
        out = torch.empty(shape, device=phi_uncollided.device, dtype=torch.float32)
    sums = _fc_sums_raw(library, rho, phi_uncollided, grid, geometry, angle, omegas, weights, samples=out).double()
    scale, beta = first_collision_correction_factors(sums[:3], sums[3:])
    ou = primary_directions_to_voxels(grid, geometry, angle, phi_uncollided.device, torch.float32)
    correction = FirstCollisionSamplingCorrection(scale=scale.float().contiguous(), beta=beta.float().contiguous(),
                                                  omega_uncollided=ou)
    a, keep, _ = _fc_args(library, rho, phi_uncollided, grid, geometry, angle, omegas)
    with torch.cuda.device(phi_uncollided.device):
        _hip.check(_hip.library().citrus_fc_correct(ctypes.byref(a), correction.scale.data_ptr(),
                                                    correction.beta.data_ptr(), out.data_ptr(),
                                                    _hip.stream_of(phi_uncollided.device)), "citrus_fc_correct")
    del keep
    return out, correction


def fc_blocks(shape) -> int:
    """Blocks of 4^3 cells of the sweep's block order over a (rows, cols, slices) grid (partial blocks at the edges)."""
    return int(np.prod([-(-int(n) // 4) for n in shape]))


def to_block_order(x: Tensor, shape) -> Tensor:
    """[*lead, R, C, S, *tail] -> [*lead, blocks, 64, *tail] in the sweep's block order: blocks of 4^3 cells (row
    block, column block, slice block) row-major, the cells of a block (row, column, slice) row-major; cells past the
    grid (partial blocks) zero. The spatial dimensions are the first run of ``shape`` in ``x.shape``."""
    R, C, S = (int(n) for n in shape)
    dims = list(x.shape)
    s0 = next((i for i in range(len(dims) - 2) if dims[i:i + 3] == [R, C, S]), None)
    if s0 is None:
        raise ValueError(f"no spatial dimensions {(R, C, S)} in {tuple(x.shape)}")
    lead, tail = tuple(dims[:s0]), tuple(dims[s0 + 3:])
    Rp, Cp, Sp = -(-R // 4) * 4, -(-C // 4) * 4, -(-S // 4) * 4
    if (Rp, Cp, Sp) != (R, C, S):
        pad = torch.zeros(lead + (Rp, Cp, Sp) + tail, device=x.device, dtype=x.dtype)
        pad[(slice(None),) * s0 + (slice(0, R), slice(0, C), slice(0, S))] = x
        x = pad
    y = x.reshape(lead + (Rp // 4, 4, Cp // 4, 4, Sp // 4, 4) + tail)
    nl = len(lead)
    perm = list(range(nl)) + [nl + 0, nl + 2, nl + 4, nl + 1, nl + 3, nl + 5] + [nl + 6 + i for i in range(len(tail))]
    return y.permute(perm).reshape(lead + ((Rp // 4) * (Cp // 4) * (Sp // 4), 64) + tail)


def from_block_order(x: Tensor, shape, lead_dims: int) -> Tensor:
    """Inverse of ``to_block_order``: [*lead, blocks, 64, *tail] -> [*lead, R, C, S, *tail]."""
    R, C, S = (int(n) for n in shape)
    nbr, nbc, nbs = -(-R // 4), -(-C // 4), -(-S // 4)
    lead, tail = tuple(x.shape[:lead_dims]), tuple(x.shape[lead_dims + 2:])
    nl = lead_dims
    y = x.reshape(lead + (nbr, nbc, nbs, 4, 4, 4) + tail)
    perm = list(range(nl)) + [nl + 0, nl + 3, nl + 1, nl + 4, nl + 2, nl + 5] + [nl + 6 + i for i in range(len(tail))]
    y = y.permute(perm).reshape(lead + (nbr * 4, nbc * 4, nbs * 4) + tail)
    return y[(slice(None),) * nl + (slice(0, R), slice(0, C), slice(0, S))]


def octant_task_slots() -> np.ndarray:
    """[8, 64]: where the sample of a block's cell lc (row, column, slice row-major within the 4^3 block) sits in
    octant o's run of the block order, i.e. the cell's place in that octant's sweep task: by diagonal plane
    a + b + c, then a, then b, where (a, b, c) are the cell's (column, row, slice) counted from the octant's upstream
    corner (o & 4 reverses the columns, o & 2 the rows, o & 1 the slices)."""
    order = sorted((a + b + c, a, b, c) for a in range(4) for b in range(4) for c in range(4))
    slot_of = {abc[1:]: n for n, abc in enumerate(order)}
    out = np.empty((8, 64), dtype=np.int64)
    for o in range(8):
        for lc in range(64):
            ly, lx, lz = lc // 16, (lc // 4) % 4, lc % 4
            a = 3 - lx if o & 4 else lx
            b = 3 - ly if o & 2 else ly
            c = 3 - lz if o & 1 else lz
            out[o, lc] = slot_of[(a, b, c)]
    return out


@dataclass(eq=False)
class DeferredFirstCollision:
    """The first-collision cell means of V views at every sweep direction (one quadrature point per cell), stored
    before their sampling correction in the sweep's block order, with what the correction needs: the persistent sweep
    applies it as it reads them (source layout 3), so the corrected source is never written; ``apply_`` forms the
    corrected group-major means for any other consumer. Either way each value is ``mean scale (1 + beta eta)`` with
    the operations of ``citrus_fc_correct`` (uncollided.py:420-433), so the two are equal bit for bit.

    Block order (``to_block_order``'s blocks): blocks of 4^3 cells, per block and octant the 64 cells' D8 directions
    padded to dp = D8 rounded up to 4, so that one sweep task (block, octant, group) reads one contiguous run, the
    cells of octant o's run in o's task order (``octant_task_slots``: the cells of one plane of the task adjacent);
    each group's region is followed by ``FC_GROUP_PAD`` unused floats, which skew the groups' and views' regions
    across HBM channels. The factors and directions per cell are in ``to_block_order``'s order.

    Attributes: shape (R, C, S); samples [V, G, blocks * 8 * 64 * dp + pad] (``block_samples`` views the blocks);
    scale and beta [V, G, blocks * 64 + pad] float32; u [V, blocks * 64, 4] the primary direction of each cell
    (``citrus_fc_primary_directions``); omega [M, 4] the reordered sweep directions; ``applied`` once ``apply_`` has
    formed the corrected means (kept in ``corrected``)."""

    shape: tuple
    samples: Tensor
    scale: Tensor
    beta: Tensor
    u: Tensor
    omega: Tensor
    applied: bool = False
    corrected: Optional[Tensor] = None

    @staticmethod
    def empty(V: int, G: int, shape, M: int, device) -> "DeferredFirstCollision":
        if M % 8:
            raise ValueError(f"the sweep's directions come in 8 octants, got {M}")
        nb, dp, pad = fc_blocks(shape), -(-(M // 8) // 4) * 4, _hip.FC_GROUP_PAD
        f = dict(device=device, dtype=torch.float32)
        return DeferredFirstCollision(shape=tuple(int(n) for n in shape),
                                      samples=torch.empty((V, G, nb * 8 * 64 * dp + pad), **f),
                                      scale=torch.empty(V, G, nb * 64 + pad, **f),
                                      beta=torch.empty(V, G, nb * 64 + pad, **f),
                                      u=torch.empty(V, nb * 64, 4, **f), omega=torch.empty(M, 4, **f))

    def block_samples(self) -> Tensor:
        """The samples as [V, G, blocks, 8, 64, dp] (a view without the pads)."""
        V, G = int(self.samples.shape[0]), int(self.samples.shape[1])
        nb, dp = fc_blocks(self.shape), -(-(int(self.omega.shape[0]) // 8) // 4) * 4
        return self.samples[..., :nb * 8 * 64 * dp].reshape(V, G, nb, 8, 64, dp)

    def cell_samples(self) -> Tensor:
        """The samples as [V, G, blocks, 8, 64, dp] with the cells of every run in ``to_block_order``'s order (row,
        column, slice row-major within the block)."""
        x = self.block_samples()
        slots = torch.as_tensor(octant_task_slots(), device=x.device)
        return x[:, :, :, torch.arange(8, device=x.device)[:, None], slots]

    def apply_(self) -> Tensor:
        """The group-major cell means [V, G, R, C, S, M] of the corrected source (formed once): the block order
        undone, then the correction of ``citrus_fc_correct_means``."""
        if not self.applied:
            V, G = int(self.samples.shape[0]), int(self.samples.shape[1])
            M = int(self.omega.shape[0])
            D8 = M // 8
            R, C, S = self.shape
            N = R * C * S
            nb = fc_blocks(self.shape)
            means = from_block_order(self.cell_samples()[..., :D8].permute(0, 1, 2, 4, 3, 5), self.shape, 2)
            means = means.reshape(V, G, R, C, S, M).contiguous()
            scale = from_block_order(self.scale[..., :nb * 64].reshape(V, G, nb, 64), self.shape, 2).contiguous()
            beta = from_block_order(self.beta[..., :nb * 64].reshape(V, G, nb, 64), self.shape, 2).contiguous()
            u = from_block_order(self.u.reshape(V, -1, 64, 4), self.shape, 1).contiguous()
            with torch.cuda.device(means.device):
                _hip.check(_hip.library().citrus_fc_correct_means(
                    means.data_ptr(), scale.data_ptr(), beta.data_ptr(), u.data_ptr(), self.omega.data_ptr(), V, G,
                    N, M, _hip.stream_of(means.device)), "citrus_fc_correct_means")
            self.corrected = means
            self.applied = True
        return self.corrected


def first_collision_deferred_samples(library: MultigroupLibrary, rho: Tensor, phi_uncollided: Tensor, grid: Grid,
                                     geometry: ConeGeometry, angle: float, omegas: Tensor, weights: Tensor,
                                     out: DeferredFirstCollision, view: int, return_correction: bool = True
                                     ) -> Optional[FirstCollisionSamplingCorrection]:
    """``first_collision_corrected_samples`` of view ``view`` of ``out`` without applying the correction: the
    uncorrected samples, the correction factors and the cells' primary directions go to ``out`` in block order, all
    from the one launch of ``citrus_fc_sums_blocked`` (one point per cell: ``grid`` is the transport grid). Returns
    the correction in the grid's layout (or None with ``return_correction=False``)."""
    K = int(omegas.reshape(-1, 3).shape[0])
    if tuple(grid.shape) != tuple(out.shape) or K != int(out.omega.shape[0]):
        raise ValueError(f"grid {tuple(grid.shape)} / {K} directions do not match the deferred source "
                         f"{tuple(out.shape)} / {int(out.omega.shape[0])}")
    _fc_sums_raw(library, rho, phi_uncollided, grid, geometry, angle, omegas, weights, samples=out.samples[view],
                 blocked_d8=K // 8, factors=(out.scale[view], out.beta[view], out.u[view]))
    out.omega[:, :3].copy_(omegas.reshape(-1, 3))
    out.omega[:, 3].zero_()
    out.applied = False
    out.corrected = None
    if not return_correction:
        return None
    nb = fc_blocks(grid.shape)
    scale = from_block_order(out.scale[view, :, :nb * 64].reshape(-1, nb, 64), grid.shape, 1).contiguous()
    beta = from_block_order(out.beta[view, :, :nb * 64].reshape(-1, nb, 64), grid.shape, 1).contiguous()
    u = from_block_order(out.u[view].reshape(nb, 64, 4), grid.shape, 0)
    return FirstCollisionSamplingCorrection(scale=scale, beta=beta, omega_uncollided=u[..., :3].contiguous())


_FC_LAYOUTS = {"moment_major": 0, "cell_major": 1, "group_major": 2, "group_major_means": 3}


def first_collision_moments_at_directions(library: MultigroupLibrary, rho: Tensor, phi_uncollided: Tensor, grid: Grid,
                                          geometry: ConeGeometry, angle: float, omegas: Tensor,
                                          correction: Optional[FirstCollisionSamplingCorrection], factor: int,
                                          chunk: int = 4, pairs=None, out: Optional[Tensor] = None,
                                          cell_major: bool = False, layout: Optional[str] = None) -> Tensor:
    """LD moments of the corrected first-collision source on the cells of the ``factor``-coarsened grid
    (uncollided.py:464-535), one ``citrus_fc_moments`` launch: [K, G, 4, Rc, Cc, Sc], or cell-major
    [K, Rc, Cc, Sc, G, 4] (the sweep's source chunk) with ``cell_major``; ``layout="group_major"`` gives
    [G, Rc, Cc, Sc, K, 4] and, for ``factor`` 1 (where the slopes vanish), ``layout="group_major_means"`` the cell
    means [G, Rc, Cc, Sc, K]. ``correction`` None leaves the samples uncorrected."""
    f = int(factor)
    if f < 1 or any(n % f for n in grid.shape):
        raise ValueError(f"grid shape {grid.shape} not divisible by the factor {factor}")
    name = layout or ("cell_major" if cell_major else "moment_major")
    if name not in _FC_LAYOUTS:
        raise ValueError(f"layout must be one of {sorted(_FC_LAYOUTS)}, got {layout!r}")
    a, keep, _ = _fc_args(library, rho, phi_uncollided, grid, geometry, angle, omegas)
    G, K = library.num_groups, a.K
    Rc, Cc, Sc = (n // f for n in grid.shape)
    shape = {"moment_major": (K, G, 4, Rc, Cc, Sc), "cell_major": (K, Rc, Cc, Sc, G, 4),
             "group_major": (G, Rc, Cc, Sc, K, 4), "group_major_means": (G, Rc, Cc, Sc, K)}[name]
    if name == "group_major_means" and f != 1:
        raise ValueError("layout 'group_major_means' needs factor 1 (the LD slopes vanish only there)")
    if out is None:
        out = torch.empty(shape, device=phi_uncollided.device, dtype=torch.float32)
    elif tuple(out.shape) != shape or not out.is_contiguous() or out.dtype != torch.float32:
        raise ValueError(f"out must be a contiguous float32 tensor of shape {shape}")
    scale = beta = None
    if correction is not None:
        scale = correction.scale.to(phi_uncollided.device, torch.float32).contiguous()
        beta = correction.beta.to(phi_uncollided.device, torch.float32).contiguous()
    with torch.cuda.device(phi_uncollided.device):
        _hip.check(_hip.library().citrus_fc_moments(ctypes.byref(a), _hip.ptr(scale), _hip.ptr(beta), f,
                                                    _FC_LAYOUTS[name], out.data_ptr(),
                                                    _hip.stream_of(phi_uncollided.device)), "citrus_fc_moments")
    del keep
    return out


def restrict_to_linear_discontinuous_moments(field: Tensor, grid: Grid, factor: int) -> Tensor:
    """[..., 4, R/f, C/f, S/f] LD moments [0, x, y, z] of a field on the refined grid (uncollided.py:311-352):
    sub-point coordinates ``u_j = (2j + 1)/f - 1``, mean and ``mean(S u_a)/mean(u_a^2)``."""
    f = int(factor)
    if any(n % f for n in grid.shape):
        raise ValueError(f"grid shape {grid.shape} not divisible by the factor {factor}")
    if tuple(field.shape[-3:]) != tuple(grid.shape):
        raise ValueError(f"field spatial shape {tuple(field.shape[-3:])} != grid shape {grid.shape}")
    lead = field.shape[:-3]
    R, C, S = (n // f for n in grid.shape)
    x = field.reshape(*lead, R, f, C, f, S, f)
    nl = len(lead)
    ar, ac, as_ = nl + 1, nl + 3, nl + 5
    mean = x.mean(dim=(ar, ac, as_))
    if f == 1:
        zero = torch.zeros_like(mean)
        return torch.stack([mean, zero, zero, zero], dim=-4)
    u = (2 * torch.arange(f, device=field.device, dtype=field.dtype) + 1) / f - 1
    scale = 1.0 / (u * u).mean()

    def slope(axis: int) -> Tensor:
        shape = [1] * x.dim()
        shape[axis] = f
        return (x * u.reshape(shape)).mean(dim=(ar, ac, as_)) * scale

    return torch.stack([mean, slope(ac), slope(ar), slope(as_)], dim=-4)
