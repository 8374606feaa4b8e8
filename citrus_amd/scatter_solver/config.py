"""Solver configuration of the transport hot path.

The fields of ``SolverConfig`` (``citrus/scatter_solver/config.py:60-160``) that the per-view path reads, with the
reference's names and defaults: angular discretization (``num_polar`` x ``num_azimuthal`` product set, Legendre order),
transport cell and quadrature-point spacing, uncollided trace grid, source-iteration stopping rule and the cone set of
the scatter image. New here: ``sweep_quadrature = "level_symmetric"`` with ``sn_order`` selects an S_N level-symmetric
sweep set (S2-S12) instead of the product set. Library building (group edges, photon data) is outside this build.
"""

from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class SolverConfig:
    legendre_order: int = 7
    num_polar: int = 8
    num_azimuthal: int = 16
    sweep_quadrature: str = "product"
    sn_order: int = 8
    transport_cell_cm: float = 1.0
    quadrature_point_spacing_cm: float = 0.5
    uncollided_trace_grid: str = "quadrature"
    source_iteration_tolerance: float = 1.0e-3
    max_source_iterations: int = 30
    cone_num_polar: int = 16
    cone_num_azimuthal: int = 32
    direction_chunk: int = 8
    sweep_buffer_bytes: int = 16 << 30

    def validate(self) -> "SolverConfig":
        """Raise ``ValueError`` on an out-of-range field; returns the configuration."""
        if not 0 <= self.legendre_order <= 7:
            raise ValueError("legendre_order must be in 0..7")
        if self.sweep_quadrature not in ("product", "level_symmetric"):
            raise ValueError("sweep_quadrature must be 'product' or 'level_symmetric'")
        if self.transport_cell_cm <= 0 or self.quadrature_point_spacing_cm <= 0:
            raise ValueError("transport_cell_cm and quadrature_point_spacing_cm must be > 0")
        if self.uncollided_trace_grid not in ("coarse", "quadrature", "fine"):
            raise ValueError("uncollided_trace_grid must be 'coarse', 'quadrature' or 'fine'")
        if self.source_iteration_tolerance <= 0 or self.max_source_iterations < 1:
            raise ValueError("source_iteration_tolerance must be > 0 and max_source_iterations >= 1")
        if self.direction_chunk < 1 or self.cone_num_polar < 1 or self.cone_num_azimuthal < 1:
            raise ValueError("direction_chunk and the cone counts must be >= 1")
        return self

    def quadrature_points(self, fine_voxel_cm: float) -> int:
        """Points per axis per transport cell (config.py:130-144):
        ``q = min(max(1, round(cell / spacing)), max(1, round(cell / fine_voxel_cm)))`` with Python's round."""
        cell = float(self.transport_cell_cm)
        target = max(1, round(cell / float(self.quadrature_point_spacing_cm)))
        return min(target, max(1, round(cell / float(fine_voxel_cm))))
