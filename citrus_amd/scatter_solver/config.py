"""Solver configuration of the transport hot path.

``SolverConfig`` carries every field of the reference's ``SolverConfig`` (``citrus/scatter_solver/config.py:66-101``)
with its name, default and validation rule, so a configuration written for the reference constructs here unchanged.
The library-building fields (``num_groups``, ``group_edges_kev``, ``group_edges_method``, ``grouping_object_size_cm``,
``grouping_medium``, ``energy_min_kev``, ``num_eta_table``, ``angular_table_grid``, ``materials``) are validated and
kept for the caller; they describe the library the caller hands to ``ScatterSolver`` (the XCOM library builder is
outside this build). New here: ``sweep_quadrature = "level_symmetric"`` with ``sn_order`` selects an S_N
level-symmetric sweep set (S2-S12) instead of the product set, and ``sweep_buffer_bytes`` bounds the sweep's source
chunk.
"""

from __future__ import annotations

from dataclasses import asdict, dataclass
from typing import Optional, Sequence, Union

# the rules of group_edges_method and the node placements of angular_table_grid (config.py:16-17)
GROUP_EDGES_METHODS = ("scattered_fluence", "equal_fluence", "log")
ANGULAR_TABLE_GRIDS = ("adaptive", "uniform")
# the seven default reference materials in library order (cross_sections.py:211)
MATERIALS = ("air", "lung_icrp", "adipose_icrp", "water", "muscle_icru44", "soft_tissue_icru44",
             "bone_cortical_icrp")


@dataclass(frozen=True)
class SolverConfig:
    # energy (the library's construction; config.py:67-75)
    num_groups: int = 6
    group_edges_kev: Optional[Sequence[float]] = None
    group_edges_method: str = "scattered_fluence"
    grouping_object_size_cm: float = 20.0
    grouping_medium: Union[str, tuple] = "auto"
    energy_min_kev: float = 10.0
    legendre_order: int = 7
    num_eta_table: int = 2001
    angular_table_grid: str = "adaptive"
    # materials
    materials: tuple = MATERIALS
    # angle
    num_polar: int = 8
    num_azimuthal: int = 16
    sweep_quadrature: str = "product"
    sn_order: int = 8
    # space
    transport_cell_cm: float = 1.0
    quadrature_point_spacing_cm: float = 0.5
    uncollided_trace_grid: str = "quadrature"
    # source iteration
    source_iteration_tolerance: float = 1.0e-3
    max_source_iterations: int = 30
    # scatter image
    cone_num_polar: int = 16
    cone_num_azimuthal: int = 32
    direction_chunk: int = 8
    # detector and grid responses
    group_averaged_responses: bool = False
    # numerics
    dtype: str = "float32"
    sweep_buffer_bytes: int = 16 << 30

    def __post_init__(self) -> None:
        if not isinstance(self.materials, str) and isinstance(self.materials, Sequence):
            object.__setattr__(self, "materials", tuple(self.materials))
        if self.group_edges_kev is not None and not isinstance(self.group_edges_kev, tuple):
            try:
                object.__setattr__(self, "group_edges_kev", tuple(float(e) for e in self.group_edges_kev))
            except (TypeError, ValueError):
                pass  # validate() reports it
        medium = self.grouping_medium
        if not isinstance(medium, str):
            try:  # a mapping or (name, density) pairs -> sorted pairs, hashable and order independent
                items = medium.items() if hasattr(medium, "items") else medium
                object.__setattr__(self, "grouping_medium", tuple(sorted((str(n), float(d)) for n, d in items)))
            except (TypeError, ValueError):
                pass

    def to_dict(self) -> dict:
        return asdict(self)

    def validate(self) -> "SolverConfig":
        """Raise ``ValueError`` on the first field outside its allowed values (the rules of config.py:177-222);
        returns the configuration."""
        L = self.legendre_order
        if not 0 <= L <= 7:
            raise ValueError("legendre_order must be in 0..7 (the sweep kernels hold up to 64 moments)")
        if self.sweep_quadrature not in ("product", "level_symmetric"):
            raise ValueError("sweep_quadrature must be 'product' or 'level_symmetric'")
        if self.sweep_quadrature == "product":
            if self.num_polar < L + 1:
                raise ValueError(f"num_polar={self.num_polar} must be >= legendre_order+1={L + 1}")
            if self.num_azimuthal < 2 * L + 1:
                raise ValueError(f"num_azimuthal={self.num_azimuthal} must be >= 2*legendre_order+1={2 * L + 1}")
        if self.num_groups < 1:
            raise ValueError("num_groups must be >= 1")
        if self.group_edges_kev is not None:
            try:
                edges = [float(e) for e in self.group_edges_kev]
            except (TypeError, ValueError):
                raise ValueError(f"group_edges_kev must be numbers, got {self.group_edges_kev!r}") from None
            if len(edges) != self.num_groups + 1:
                raise ValueError("group_edges_kev must have num_groups+1 entries")
            if any(not a > b for a, b in zip(edges, edges[1:])):
                raise ValueError("group_edges_kev must be strictly descending")
        if self.group_edges_method not in GROUP_EDGES_METHODS:
            raise ValueError(f"group_edges_method must be one of {GROUP_EDGES_METHODS}, "
                             f"got {self.group_edges_method!r}")
        if not self.grouping_object_size_cm > 0:
            raise ValueError(f"grouping_object_size_cm must be > 0, got {self.grouping_object_size_cm}")
        self._validate_grouping_medium()
        if self.num_eta_table < 2:
            raise ValueError(f"num_eta_table must be >= 2, got {self.num_eta_table}")
        if self.angular_table_grid not in ANGULAR_TABLE_GRIDS:
            raise ValueError(f"angular_table_grid must be one of {ANGULAR_TABLE_GRIDS}, "
                             f"got {self.angular_table_grid!r}")
        names = self.material_names()
        if len(names) < 2:
            raise ValueError(f"materials needs at least two reference materials, got {names}")
        if len(set(names)) != len(names):
            raise ValueError(f"materials has duplicates: {names}")
        if not self.transport_cell_cm > 0:
            raise ValueError(f"transport_cell_cm must be > 0, got {self.transport_cell_cm}")
        if not self.quadrature_point_spacing_cm > 0:
            raise ValueError(f"quadrature_point_spacing_cm must be > 0, got {self.quadrature_point_spacing_cm}")
        if self.uncollided_trace_grid not in ("coarse", "quadrature", "fine"):
            raise ValueError(f"uncollided_trace_grid must be 'coarse', 'quadrature' or 'fine', "
                             f"got {self.uncollided_trace_grid!r}")
        if self.max_source_iterations < 1 or self.source_iteration_tolerance <= 0:
            raise ValueError("need max_source_iterations >= 1 and source_iteration_tolerance > 0")
        if self.cone_num_polar < 1 or self.cone_num_azimuthal < 1:
            raise ValueError("need cone_num_polar >= 1 and cone_num_azimuthal >= 1")
        if self.direction_chunk < 1:
            raise ValueError("direction_chunk must be >= 1")
        if not isinstance(self.group_averaged_responses, bool):
            raise ValueError(f"group_averaged_responses must be True or False, got {self.group_averaged_responses!r}")
        if self.dtype not in ("float32", "float64"):
            raise ValueError(f"dtype must be 'float32' or 'float64', got {self.dtype!r}")
        return self

    def material_names(self) -> tuple:
        """Names of the reference materials: strings as given, or the ``name`` attribute of material objects."""
        if isinstance(self.materials, str):
            raise ValueError(f"materials must be a sequence of names, got the string {self.materials!r}")
        out = []
        for m in self.materials:
            name = m if isinstance(m, str) else getattr(m, "name", None)
            if not isinstance(name, str) or not name:
                raise ValueError(f"materials entries must be names or objects with a name, got {m!r}")
            out.append(name)
        return tuple(out)

    def _validate_grouping_medium(self) -> None:
        """``grouping_medium`` is "auto", one of the reference materials, or (reference material, density) pairs
        with densities >= 0 and at least one > 0 (config.py:146-175, restricted to the reference materials: the
        element and compound catalogue belongs to the library builder)."""
        medium = self.grouping_medium
        reference = self.material_names() if not isinstance(self.materials, str) else ()
        if isinstance(medium, str):
            if medium == "auto":
                return
            names = (medium,)
        else:
            try:
                names = tuple(str(n) for n, _ in medium)
                densities = [float(d) for _, d in medium]
            except (TypeError, ValueError):
                raise ValueError(f"grouping_medium must be 'auto', a material name or a mapping "
                                 f"{{reference material name: density}}, got {medium!r}") from None
            if any(not d >= 0 for d in densities) or not any(d > 0 for d in densities):
                raise ValueError(f"grouping_medium densities must be >= 0 with at least one > 0, got {medium!r}")
        outside = [n for n in names if n not in reference]
        if outside:
            raise ValueError(f"grouping_medium material(s) {outside} are not among the reference materials "
                             f"{reference}")

    def quadrature_points(self, fine_voxel_cm: float) -> int:
        """Points per axis per transport cell (config.py:130-144):
        ``q = min(max(1, round(cell / spacing)), max(1, round(cell / fine_voxel_cm)))`` with Python's round."""
        cell = float(self.transport_cell_cm)
        target = max(1, round(cell / float(self.quadrature_point_spacing_cm)))
        return min(target, max(1, round(cell / float(fine_voxel_cm))))
