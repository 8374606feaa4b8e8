"""Per-view scatter solver on the HIP kernels.

Keeps the interface of ``ScatterSolver`` and ``MaterialFields`` (``citrus/scatter_solver/scatter_solver.py:155-1000``)
for the operators of the per-view transport path: the transport and quadrature grids, the sweep and cone direction
sets, the density fields, the uncollided fluence, the first-collision moment source with its sampling correction, the
scattering-moment source, the transport sweep, source iteration, the multiple-collision source moments, the scatter
image and the primary. Densities reach the transport and quadrature grids by conservative resampling (block means
when the grids nest, the overlap kernel otherwise); library building from photon data is outside this build.
"""

from __future__ import annotations

import time
from contextlib import contextmanager
from dataclasses import dataclass, field, replace
from typing import Optional, Sequence

import numpy as np

import torch
from torch import Tensor

from .angular_quadrature import Quadrature, level_symmetric_quadrature, product_quadrature
from .config import SolverConfig
from .cross_sections import MultigroupLibrary, mu_t_volume
from .detector import DetectorResponse, GridResponse, ideal_counting_response, no_grid
from .geometry import ConeGeometry, Grid, resample_conservative
from .ray_tracing import HipProjectors
from .scatter_image import (combined_scatter_image, make_direction_context, panel_direction_set, primary_image,
                            scatter_fraction)
from .source import SourceModel
from .source_iteration import (MaterialTransfer, SourceIterationResult, scatter_moments, scatter_source_function,
                               source_iteration, summed_source_function)
from .transport_sweep import SweepContext, make_sweep_context, sweep
from .uncollided import (first_collision_moments_at_directions, first_collision_sampling_correction,
                         present_materials, uncollided_fluence, first_collision_corrected_samples)


def coarsen_density(rho_fine: Tensor, grid_fine: Grid, grid_coarse: Grid) -> Tensor:
    """[M, Rc, Cc, Sc] cell-averaged densities of nested grids: each coarse cell is the mean of its f^3 fine voxels,
    the conservative resampling (geometry.py:304) in the case where every coarse cell holds whole fine voxels.
    Raises ``ValueError`` for grids that do not nest by one integer factor along every axis."""
    if grid_coarse == grid_fine:
        return rho_fine
    fs = [c // n if n and c % n == 0 else 0 for c, n in zip(grid_fine.shape, grid_coarse.shape)]
    f = fs[0]
    nested = f >= 1 and all(x == f for x in fs) and all(
        abs(dc - f * df) <= 1e-9 * f * df for dc, df in zip(grid_coarse.deltas, grid_fine.deltas))
    if not nested or abs(grid_coarse.recon_slice_offset - grid_fine.recon_slice_offset) > 1e-9:
        raise ValueError(f"grid {grid_coarse.shape} does not nest the fine grid {grid_fine.shape} by an integer factor")
    M = int(rho_fine.shape[0])
    Rc, Cc, Sc = grid_coarse.shape
    return rho_fine.reshape(M, Rc, f, Cc, f, Sc, f).mean(dim=(2, 4, 6))


def resample_density(rho_fine: Tensor, grid_fine: Grid, grid_to: Grid) -> Tensor:
    """Densities of ``grid_fine`` on ``grid_to`` by conservative resampling: block means when the grids nest by an
    integer factor, otherwise the overlap-weighted HIP kernel of ``resample_conservative``."""
    try:
        return coarsen_density(rho_fine, grid_fine, grid_to)
    except ValueError:
        return resample_conservative(rho_fine, grid_fine, grid_to)


@dataclass
class MaterialFields:
    """Density fields of a volume on the fine, transport and quadrature grids (scatter_solver.py:155)."""

    grid_fine: Grid
    grid_coarse: Grid
    grid_quadrature: Grid
    q: int
    rho_fine: Tensor
    rho_coarse: Tensor
    rho_quadrature: Tensor
    mu_t_coarse: Tensor
    transfer_coarse: MaterialTransfer
    traced_materials: Optional[tuple] = None
    fine_materials: Optional[tuple] = None

    @property
    def num_materials(self) -> int:
        return int(self.rho_coarse.shape[0])

    @property
    def mu_s_l_coarse(self) -> MaterialTransfer:
        """The transfer of the transport cells in the factored form ``rho_m sigma_m`` (scatter_moments accepts it)."""
        return self.transfer_coarse


_IMAGE_KEYS = ("fluence", "planar_fluence", "signal_per_group", "signal")


@dataclass
class ViewResult:
    """Primary and scatter images of one view from ``ScatterSolver.solve_view`` (scatter_solver.py:237-312): image
    dicts (``fluence``, ``planar_fluence``, ``signal_per_group`` [G, Rd, Cd] and ``signal`` [Rd, Cd], or the signal
    alone for a signal-only solve), the scatter signal and fraction, the source-iteration record (its ``phi_total``
    empty unless kept), the collided moments when kept, and seconds and peak bytes per stage."""

    angle: float
    primary: dict
    scatter: dict
    scatter_signal: Tensor
    scatter_fraction: Tensor
    source_iteration_result: SourceIterationResult
    phi_total: Optional[Tensor] = None
    timings: dict = field(default_factory=dict)
    peak_memory: dict = field(default_factory=dict)
    peak_memory_bytes: Optional[int] = None

    @property
    def scatter_fluence(self) -> Tensor:
        return self.scatter["fluence"]

    @property
    def scatter_planar_fluence(self) -> Tensor:
        return self.scatter["planar_fluence"]

    def to_numpy_dict(self) -> dict:
        """Flat dict of numpy arrays for ``np.savez`` with the reference's keys (scatter_solver.py:282-312)."""
        def host(t: Tensor) -> np.ndarray:
            return t.detach().cpu().numpy()

        out = {"angle": np.asarray(self.angle)}
        for name, images in (("primary", self.primary), ("scatter", self.scatter)):
            out.update({f"{name}_{k}": host(images[k]) for k in _IMAGE_KEYS if k in images})
        out["scatter_signal"] = host(self.scatter_signal)
        out["scatter_fraction"] = host(self.scatter_fraction)
        r = self.source_iteration_result
        out["source_iteration_history"] = np.asarray(r.history, dtype=np.float64)
        out["source_iteration_phi_norms"] = np.asarray(r.phi_norms, dtype=np.float64)
        out["source_iteration_num_iterations"] = np.asarray(r.num_iterations)
        out["source_iteration_converged"] = np.asarray(r.converged)
        out.update({f"timing_{k}": np.asarray(v) for k, v in self.timings.items()})
        out["peak_memory_bytes"] = np.asarray(-1 if self.peak_memory_bytes is None else self.peak_memory_bytes)
        if self.phi_total is not None:
            out["phi_total"] = host(self.phi_total)
        return out


class ScatterSolver:
    """Per-view scatter solver (scatter_solver.py:314) on one GPU, float32."""

    def __init__(self, library: MultigroupLibrary, config: SolverConfig, geometry: ConeGeometry, projectors=None,
                 detector: Optional[DetectorResponse] = None, grid_response: Optional[GridResponse] = None,
                 source: Optional[SourceModel] = None, device=None, dtype: torch.dtype = torch.float32) -> None:
        if dtype != torch.float32 or getattr(config, "dtype", "float32") != "float32":
            raise ValueError("the HIP path computes in float32 (dtype torch.float32, config.dtype 'float32')")
        self.library = library
        self.config = config.validate()
        self.geometry = geometry
        self.projectors = projectors if projectors is not None else HipProjectors()
        self.detector = detector if detector is not None else ideal_counting_response()
        self.grid_response = grid_response if grid_response is not None else no_grid()
        self.source = source if source is not None else SourceModel()
        self.device = torch.device(device if device is not None else "cuda")
        self.dtype = dtype
        self.L = int(config.legendre_order)
        if library.legendre_order < self.L:
            raise ValueError(f"library Legendre order {library.legendre_order} < {self.L}")
        self._transport_grids: dict = {}
        self._sweep_contexts: dict = {}
        self._cone_quadratures: dict = {}
        self._direction_contexts: dict = {}
        self._sigma = library.mu_s_l[..., : self.L + 1].to(self.device, torch.float32).contiguous()

    # ---- grids and direction sets ----------------------------------------------------------------------------
    def responses(self) -> tuple:
        return self.detector, self.grid_response

    @property
    def num_groups(self) -> int:
        return self.library.num_groups

    def quadrature_points(self, grid_fine: Grid) -> int:
        return self.config.quadrature_points(max(grid_fine.deltas))

    def transport_grid(self, grid_fine: Grid) -> Grid:
        g = self._transport_grids.get(grid_fine)
        if g is None:
            g = self._transport_grids[grid_fine] = Grid.covering(grid_fine, self.config.transport_cell_cm)
        return g

    def quadrature_grid(self, grid_fine: Grid) -> Grid:
        return self.transport_grid(grid_fine).refine(self.quadrature_points(grid_fine))

    def sweep_quadrature(self) -> Quadrature:
        c = self.config
        if c.sweep_quadrature == "level_symmetric":
            return level_symmetric_quadrature(c.sn_order)
        return product_quadrature(c.num_polar, c.num_azimuthal)

    def sweep_context(self, grid_coarse: Grid) -> SweepContext:
        ctx = self._sweep_contexts.get(grid_coarse)
        if ctx is None:
            ctx = self._sweep_contexts[grid_coarse] = make_sweep_context(grid_coarse, self.sweep_quadrature(), self.L,
                                                                         device=self.device)
        return ctx

    def cone_quadrature(self, grid: Grid, angle: float) -> Quadrature:
        key = (grid, float(angle))
        q = self._cone_quadratures.get(key)
        if q is None:
            q = self._cone_quadratures[key] = panel_direction_set(
                self.geometry, angle, self.config, max_polar_angle=self.transport_grid(grid), device=self.device,
                dtype=self.dtype)
        return q

    def direction_context(self, fields: MaterialFields, angle: float, quadrature: Optional[Quadrature] = None):
        quadrature = self.cone_quadrature(fields.grid_fine, angle) if quadrature is None else quadrature
        key = (id(quadrature), float(angle))
        hit = self._direction_contexts.get(key)
        if hit is not None and hit.quadrature is quadrature:
            return hit
        ctx = make_direction_context(quadrature, self.geometry, angle, self.detector, self.grid_response,
                                     self.library, self.device, self.dtype, chunk=self.config.direction_chunk)
        self._direction_contexts[key] = ctx
        return ctx

    def sweep_direction_chunk(self, fields: MaterialFields) -> int:
        """Directions per octant per sweep chunk: all of them when the cell-major source chunk
        [8 dc, N, G, 4] fits ``config.sweep_buffer_bytes``, else the largest that fits (at least 1)."""
        ctx = self.sweep_context(fields.grid_coarse)
        per_dir = 8 * ctx.num_cells * self.num_groups * 4 * 4
        return max(1, min(ctx.directions_per_octant, int(self.config.sweep_buffer_bytes // per_dir), 32))

    # ---- fields ----------------------------------------------------------------------------------------------
    def material_fields(self, rho_fine: Tensor, grid_fine: Grid) -> MaterialFields:
        """Density fields of fine densities on the transport and quadrature grids (scatter_solver.py:658)."""
        if tuple(rho_fine.shape[1:]) != tuple(grid_fine.shape):
            raise ValueError(f"rho spatial shape {tuple(rho_fine.shape[1:])} != grid shape {grid_fine.shape}")
        if int(rho_fine.shape[0]) != self.library.num_materials:
            raise ValueError(f"rho has {int(rho_fine.shape[0])} materials, the library {self.library.num_materials}")
        rho = rho_fine.to(self.device, torch.float32).contiguous()
        coarse = self.transport_grid(grid_fine)
        q = self.quadrature_points(grid_fine)
        gq = coarse.refine(q)
        rho_c = resample_density(rho, grid_fine, coarse).contiguous()
        rho_q = rho_c if gq == coarse else resample_density(rho, grid_fine, gq).contiguous()
        traced = present_materials(rho_q)
        return MaterialFields(grid_fine=grid_fine, grid_coarse=coarse, grid_quadrature=gq, q=q, rho_fine=rho,
                              rho_coarse=rho_c, rho_quadrature=rho_q,
                              mu_t_coarse=mu_t_volume(self.library, rho_c).contiguous(),
                              transfer_coarse=MaterialTransfer(rho_c, self._sigma),
                              traced_materials=tuple(traced) if traced else (0,),
                              fine_materials=tuple(present_materials(rho)))

    def material_fields_from_attenuation(self, X: Tensor, grid_fine: Grid, density_model) -> MaterialFields:
        """Fields of the densities ``density_model(X)`` of an attenuation volume X [R, C, S] (1/cm) on
        ``grid_fine`` (scatter_solver.py:692): the material maps come from the density model's HIP kernel in one pass
        and are kept whole on the fine grid (an MI355X holds them; the reference keeps them only under a budget),
        then resampled as ``material_fields`` does."""
        if tuple(X.shape) != tuple(grid_fine.shape):
            raise ValueError(f"X shape {tuple(X.shape)} != grid shape {grid_fine.shape}")
        X = X.to(self.device, torch.float32).contiguous()
        return self.material_fields(density_model(X), grid_fine)

    # ---- operators -------------------------------------------------------------------------------------------
    def uncollided_fluence(self, fields: MaterialFields, angle: float) -> Tensor:
        """[G, Rq, Cq, Sq] uncollided fluence at the quadrature points (scatter_solver.py:745)."""
        t = self.config.uncollided_trace_grid
        rho, grid = ((fields.rho_fine, fields.grid_fine) if t == "fine" else
                     (fields.rho_quadrature, fields.grid_quadrature) if t == "quadrature" else
                     (fields.rho_coarse, fields.grid_coarse))
        return uncollided_fluence(self.library, rho, grid, self.geometry, angle, self.source, self.projectors,
                                  points=fields.grid_quadrature, materials=fields.traced_materials)

    def first_collision_sampling_correction(self, fields: MaterialFields, angle: float, phi_uncollided: Tensor,
                                            sweep_context: Optional[SweepContext] = None):
        ctx = self.sweep_context(fields.grid_coarse) if sweep_context is None else sweep_context
        return first_collision_sampling_correction(self.library, fields.rho_quadrature, phi_uncollided,
                                                   fields.grid_quadrature, self.geometry, angle,
                                                   ctx.quadrature.omega, ctx.quadrature.weight)

    def first_collision_group_major(self, fields: MaterialFields, angle: float, phi_uncollided: Tensor,
                                    sweep_context: Optional[SweepContext] = None, out: Optional[Tensor] = None):
        """(source, correction): the corrected, restricted first-collision source of every sweep direction
        (reordered), group-major [G, Rc, Cc, Sc, M, 4], or its cell means [G, Rc, Cc, Sc, M] when every cell holds
        one quadrature point (the slopes vanish; then the samples come from the same source evaluation as the
        correction's sums)."""
        ctx = self.sweep_context(fields.grid_coarse) if sweep_context is None else sweep_context
        omega = ctx.quadrature.omega.to(self.device, torch.float32)
        if fields.q == 1:
            return first_collision_corrected_samples(self.library, fields.rho_quadrature, phi_uncollided,
                                                     fields.grid_quadrature, self.geometry, angle, omega,
                                                     ctx.quadrature.weight, out=out)
        correction = self.first_collision_sampling_correction(fields, angle, phi_uncollided, ctx)
        source = first_collision_moments_at_directions(self.library, fields.rho_quadrature, phi_uncollided,
                                                       fields.grid_quadrature, self.geometry, angle, omega,
                                                       correction, fields.q, out=out, layout="group_major")
        return source, correction

    def first_collision_moment_source(self, fields: MaterialFields, angle: float, phi_uncollided: Tensor,
                                      sweep_context: Optional[SweepContext] = None):
        """Cell-major sweep source of the corrected, restricted first-collision source (scatter_solver.py:774); its
        ``hip_group_major`` attribute hands every direction at once to the sweep kernel."""
        ctx = self.sweep_context(fields.grid_coarse) if sweep_context is None else sweep_context
        omega = ctx.quadrature.omega.to(self.device, torch.float32)
        cache: dict = {}
        if fields.q == 1:
            cache["all"], correction = self.first_collision_group_major(fields, angle, phi_uncollided, ctx)
        else:
            correction = self.first_collision_sampling_correction(fields, angle, phi_uncollided, ctx)

        def source_function(direction_index: Tensor, out: Optional[Tensor] = None) -> Tensor:
            directions = omega.index_select(0, direction_index.to(omega.device))
            return first_collision_moments_at_directions(self.library, fields.rho_quadrature, phi_uncollided,
                                                         fields.grid_quadrature, self.geometry, angle, directions,
                                                         correction, fields.q, out=out, cell_major=True)

        def group_major() -> Tensor:
            """[G, Rc, Cc, Sc, M, 4] moments of every sweep direction (reordered), one launch, kept for the view;
            with one quadrature point per cell (zero slopes) the cell means [G, Rc, Cc, Sc, M]."""
            if cache.get("all") is None:
                cache["all"] = first_collision_moments_at_directions(
                    self.library, fields.rho_quadrature, phi_uncollided, fields.grid_quadrature, self.geometry, angle,
                    omega, correction, fields.q, layout="group_major_means" if fields.q == 1 else "group_major")
            return cache["all"]

        source_function.cell_major = True
        source_function.hip_group_major = group_major
        return source_function

    def scatter_moment_source(self, fields: MaterialFields, phi: Tensor,
                              sweep_context: Optional[SweepContext] = None):
        """Scattering source of collided moments, ``T = scatter_moments(mu_s_l, phi)`` (scatter_solver.py:850)."""
        ctx = self.sweep_context(fields.grid_coarse) if sweep_context is None else sweep_context
        return scatter_source_function(ctx, scatter_moments(fields.mu_s_l_coarse, phi))

    def transport_sweep(self, fields: MaterialFields, source_function) -> Tensor:
        return sweep(self.sweep_context(fields.grid_coarse), fields.mu_t_coarse, source_function,
                     direction_chunk=self.sweep_direction_chunk(fields))

    def run_source_iteration(self, fields: MaterialFields, first_collision_source_function):
        c = self.config
        return source_iteration(self.sweep_context(fields.grid_coarse), fields.mu_t_coarse, fields.mu_s_l_coarse,
                                first_collision_source_function, tolerance=c.source_iteration_tolerance,
                                max_iterations=c.max_source_iterations,
                                direction_chunk=self.sweep_direction_chunk(fields))

    def multiple_collision_source_moments(self, fields: MaterialFields, phi_collided: Tensor) -> Tensor:
        """[G, (L+1)^2, Rc, Cc, Sc] cell-mean moments T of the multiple-collision source (scatter_solver.py:899)."""
        return scatter_moments(fields.mu_s_l_coarse, phi_collided)[:, :, 0]

    def scatter_image(self, fields: MaterialFields, angle: float, phi_uncollided: Tensor,
                      scatter_source_moments: Tensor, quadrature: Optional[Quadrature] = None, context=None,
                      signal_only: bool = False, source_means: Optional[Tensor] = None) -> dict:
        quadrature = self.cone_quadrature(fields.grid_fine, angle) if quadrature is None else quadrature
        context = self.direction_context(fields, angle, quadrature) if context is None else context
        return combined_scatter_image(self.library, fields.rho_quadrature, phi_uncollided, fields.grid_quadrature,
                                      fields.q, scatter_source_moments, self.L, self.geometry, angle, quadrature,
                                      fields.mu_t_coarse, self.projectors, self.detector, self.grid_response,
                                      context, chunk=self.config.direction_chunk, grid_coarse=fields.grid_coarse,
                                      signal_only=signal_only, density_coarse=fields.rho_coarse,
                                      source_means=source_means)

    def primary(self, fields: MaterialFields, angle: float) -> dict:
        return primary_image(self.library, fields.rho_fine, fields.grid_fine, self.geometry, angle, self.source,
                             self.projectors, self.detector, self.grid_response, materials=fields.fine_materials)

    # ---- whole views ------------------------------------------------------------------------------------------
    @contextmanager
    def _stage(self, name: str, timings: dict, peaks: dict):
        """Seconds (device work included) and peak allocated bytes of one stage (scatter_solver.py:1020-1036)."""
        torch.cuda.synchronize(self.device)
        torch.cuda.reset_peak_memory_stats(self.device)
        t0 = time.perf_counter()
        yield
        torch.cuda.synchronize(self.device)
        peaks[name] = int(torch.cuda.max_memory_allocated(self.device))
        timings[name] = time.perf_counter() - t0

    def solve_view(self, rho_fine: Tensor, grid_fine: Grid, angle: float, keep_phi_total: bool = False,
                   signal_only: bool = False) -> ViewResult:
        """One view: fields, uncollided fluence, source iteration, scatter image, primary and scatter fraction
        (scatter_solver.py:1038-1062)."""
        timings: dict = {}
        peaks: dict = {}
        start = time.perf_counter()
        with self._stage("materials", timings, peaks):
            fields = self.material_fields(rho_fine, grid_fine)
        return self._solve_view(fields, angle, keep_phi_total, timings, peaks, start, signal_only)

    def solve_view_from_fields(self, fields: MaterialFields, angle: float, keep_phi_total: bool = False,
                               signal_only: bool = False) -> ViewResult:
        """One view from prebuilt fields, which every view of a volume can share (scatter_solver.py:1063-1077)."""
        return self._solve_view(fields, angle, keep_phi_total, {}, {}, time.perf_counter(), signal_only)

    def solve_views(self, rho_fine: Tensor, grid_fine: Grid, angles: Sequence[float], keep_phi_total: bool = False,
                    signal_only: bool = False) -> list:
        """Every view independently with ``solve_view`` (scatter_solver.py:1125-1140)."""
        return [self.solve_view(rho_fine, grid_fine, float(a), keep_phi_total=keep_phi_total, signal_only=signal_only)
                for a in angles]

    def _solve_view(self, fields: MaterialFields, angle: float, keep_phi_total: bool, timings: dict, peaks: dict,
                    start_time: float, signal_only: bool = False) -> ViewResult:
        """The stages of one view after its fields (scatter_solver.py:1079-1124): the source iteration from the
        corrected first-collision source, the multiple-collision source moments of the summed collided fluence, the
        collapsed-cone scatter image over the view's cone set and the primary."""
        with self._stage("sweep_context", timings, peaks):
            sweep_context = self.sweep_context(fields.grid_coarse)
        with self._stage("uncollided", timings, peaks):
            phi_uncollided = self.uncollided_fluence(fields, angle)
        with self._stage("source_iteration", timings, peaks):
            result = self.run_source_iteration(
                fields, self.first_collision_moment_source(fields, angle, phi_uncollided, sweep_context))
        with self._stage("multiple_collision_source", timings, peaks):
            source_moments = self.multiple_collision_source_moments(fields, result.phi_total)
            phi_total = result.phi_total if keep_phi_total else None
            result = replace(result, phi_total=result.phi_total if keep_phi_total
                             else torch.empty(0, device=self.device))
        with self._stage("cone_directions", timings, peaks):
            quadrature = self.cone_quadrature(fields.grid_fine, angle)
            context = self.direction_context(fields, angle, quadrature)
        with self._stage("scatter_image", timings, peaks):
            scatter = self.scatter_image(fields, angle, phi_uncollided, source_moments, quadrature, context,
                                         signal_only=signal_only)
        del phi_uncollided, source_moments
        with self._stage("primary", timings, peaks):
            primary = self.primary(fields, angle)
        signal = scatter["signal"]
        timings["total"] = time.perf_counter() - start_time
        return ViewResult(angle=float(angle), primary=primary, scatter=scatter, scatter_signal=signal,
                          scatter_fraction=scatter_fraction(signal, primary["signal"]),
                          source_iteration_result=result, phi_total=phi_total, timings=timings, peak_memory=peaks,
                          peak_memory_bytes=max(peaks.values()) if peaks else None)
