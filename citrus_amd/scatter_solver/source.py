"""X-ray source model: collimated isotropic point source with emission profile, off-focal emission and beam filter.

Keeps the attributes of ``SourceModel`` (``citrus/scatter_solver/source.py:55-130``): the angular weight ``b(Omega)``
is 1 where the ray from the focal spot hits the collimated field (the panel widened by the field margins), the
off-focal fraction of the distance past the field edge outside it, times the emission profile at the ray's panel point;
a beam filter adds the path ``rho_f t(Omega)`` through it to every transmission from the focal spot. The weights are
evaluated on the device inside the uncollided-fluence kernel. Offline there is no photon database, so the filter's
fine mass attenuation is given as a table on the library's fine energies (``filter_attenuation_table``) instead of
being looked up by material name.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch
from torch import Tensor

from .geometry import ConeGeometry


@dataclass(frozen=True, eq=False)
class SourceModel:
    """Source model (source.py:55). ``filter_thickness_cm`` is [channels] or [rows, channels] in cm;
    ``filter_density`` in g/cm^3; ``filter_attenuation_table`` [E] in cm^2/g at the library's fine energies."""

    emission_profile: Optional[Tensor] = None
    field_margin_channels_cm: float = 0.0
    field_margin_rows_cm: float = 0.0
    off_focal_distance_cm: Optional[tuple] = None
    off_focal_fraction: Optional[tuple] = None
    filter_thickness_cm: Optional[Tensor] = None
    filter_density: float = 0.0
    filter_attenuation_table: Optional[Tensor] = None
    _cache: dict = field(default_factory=dict, init=False, repr=False)

    def __post_init__(self):
        if (self.off_focal_distance_cm is None) != (self.off_focal_fraction is None):
            raise ValueError("off_focal_distance_cm and off_focal_fraction must be given together")
        if self.off_focal_distance_cm is not None:
            d = tuple(float(v) for v in self.off_focal_distance_cm)
            f = tuple(float(v) for v in self.off_focal_fraction)
            if len(d) != len(f) or not d or len(d) > 16 or any(b <= a for a, b in zip(d, d[1:])):
                raise ValueError("off-focal table: 1..16 ascending distances with one fraction each")
            object.__setattr__(self, "off_focal_distance_cm", d)
            object.__setattr__(self, "off_focal_fraction", f)
        if (self.filter_thickness_cm is None) != (self.filter_attenuation_table is None):
            raise ValueError("a beam filter needs filter_thickness_cm and filter_attenuation_table")

    @property
    def has_filter(self) -> bool:
        return self.filter_thickness_cm is not None

    def _device(self, name: str, value: Tensor, geometry: ConeGeometry, device) -> Tensor:
        key = (name, str(device))
        t = self._cache.get(key)
        if t is None:
            t = torch.as_tensor(value, dtype=torch.float32)
            if name == "thickness":
                t = t.reshape(-1, geometry.num_det_channels).expand(geometry.num_det_rows, geometry.num_det_channels)
            t = t.contiguous().to(device)
            self._cache[key] = t
        return t

    def device_tables(self, geometry: ConeGeometry, device):
        """(profile, thickness, filter attenuation) float32 device tables, or None each."""
        profile = None if self.emission_profile is None else self._device("profile", self.emission_profile, geometry,
                                                                          device)
        thickness = None if self.filter_thickness_cm is None else self._device("thickness", self.filter_thickness_cm,
                                                                               geometry, device)
        mu_f = None if self.filter_attenuation_table is None else self._device("mu_f", self.filter_attenuation_table,
                                                                               geometry, device)
        return profile, thickness, mu_f

    def host_block(self, geometry: ConeGeometry, angle: float) -> tuple[np.ndarray, np.ndarray, Optional[np.ndarray]]:
        """(23 doubles, 4 ints, off-focal table) parameter blocks of ``citrus_uncollided_fluence``."""
        uh, vh, nh = geometry.panel_axes(angle)
        values = (list(geometry.source_position(angle)) + list(geometry.panel_center(angle)) + list(uh) + list(vh)
                  + list(nh) + [geometry.source_detector_dist, geometry.delta_det_row, geometry.delta_det_channel,
                                geometry.det_row_offset, geometry.det_channel_offset, self.field_margin_rows_cm,
                                self.field_margin_channels_cm, self.filter_density])
        k = 0 if self.off_focal_distance_cm is None else len(self.off_focal_distance_cm)
        ints = [int(bool(geometry.use_curved_detector)), geometry.num_det_rows, geometry.num_det_channels, k]
        table = None if k == 0 else np.array(list(self.off_focal_distance_cm) + list(self.off_focal_fraction))
        return np.array(values, dtype=np.float64), np.array(ints, dtype=np.int32), table


def bowtie_thickness(geometry: ConeGeometry, max_cm: float = 3.0) -> Tensor:
    """[channels] bowtie thickness growing with the square of the channel position, 0 on the central ray and
    ``max_cm`` at the outermost channel (the shape of ``bowtie_emission_profile``, source.py:513)."""
    u, _ = geometry.pixel_uv(dtype=torch.float64)
    return max_cm * (u / u.abs().max()) ** 2
