"""Detector and anti-scatter-grid responses of the signal weights (host-side tables).

Keeps the interface of ``citrus/scatter_solver/detector.py``: ``DetectorResponse`` (:197, a table D(E, theta0) in
keV deposited per incident photon, bilinear in energy and angle, clamped to the end nodes, |theta| used),
``ideal_counting_response`` (:294, D = 1), ``energy_weighted_response`` (:305, D = E), an absorber response with the
oblique path ``t / cos(theta)`` like ``csi_response`` (:258; the absorber's mass attenuation is passed in, since the
photon database is not available offline), ``GridResponse`` with ``no_grid`` (:657, :758) and ``detector_angle``
(:1054). Septa models of real anti-scatter grids are outside the transport hot path.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import torch
from torch import Tensor


def _bracket(nodes: Tensor, x: Tensor) -> tuple[Tensor, Tensor, Tensor]:
    """(i0, i1, w1) of the nodes bracketing each query, clamped to the end nodes (detector.py:147)."""
    n = nodes.numel()
    if n == 1:
        z = torch.zeros_like(x, dtype=torch.long)
        return z, z, torch.zeros_like(x)
    xc = x.clamp(nodes[0], nodes[-1])
    i1 = torch.searchsorted(nodes, xc.contiguous(), right=True).clamp(1, n - 1)
    i0 = i1 - 1
    w1 = ((xc - nodes[i0]) / (nodes[i1] - nodes[i0])).clamp(0.0, 1.0)
    return i0, i1, w1


@dataclass(frozen=True, eq=False)
class DetectorResponse:
    """D(E, theta0) [E, T] in keV per incident photon on ascending energy [E] (keV) and angle [T] (rad) nodes."""

    energy_kev: Tensor
    theta_rad: Tensor
    response: Tensor
    name: str = "detector"

    def __post_init__(self) -> None:
        if tuple(self.response.shape) != (self.energy_kev.numel(), self.theta_rad.numel()):
            raise ValueError(f"response must be [E, T] = [{self.energy_kev.numel()}, {self.theta_rad.numel()}], got "
                             f"{tuple(self.response.shape)}")

    @property
    def angle_independent(self) -> bool:
        """True when every energy row of the table is constant in the angle."""
        return bool((self.response == self.response[:, :1]).all())

    def at(self, group_energy: Tensor, theta: Tensor) -> Tensor:
        """[G, ...] response at the group energies and the angles ``theta [...]``, in the dtype of ``theta``."""
        device, dtype = theta.device, theta.dtype
        en, tn, table = (t.to(device, dtype) for t in (self.energy_kev, self.theta_rad, self.response))
        e = group_energy.to(device, dtype).reshape(-1)
        ie0, ie1, we = _bracket(en, e)
        it0, it1, wt = _bracket(tn, theta.abs())
        at_e = table[ie0] + (table[ie1] - table[ie0]) * we[:, None]  # [G, T]
        a, b = at_e[:, it0], at_e[:, it1]
        return a + (b - a) * wt


def ideal_counting_response() -> DetectorResponse:
    """D = 1 per photon: the signal is the planar photon fluence (detector.py:294)."""
    e = torch.tensor([1.0, 1000.0], dtype=torch.float64)
    t = torch.tensor([0.0, math.pi / 2], dtype=torch.float64)
    return DetectorResponse(e, t, torch.ones(2, 2, dtype=torch.float64), name="ideal counting")


def energy_weighted_response() -> DetectorResponse:
    """D = E: every photon deposits its energy (detector.py:305)."""
    e = torch.tensor([1.0, 1000.0], dtype=torch.float64)
    t = torch.tensor([0.0, math.pi / 2], dtype=torch.float64)
    return DetectorResponse(e, t, e[:, None].expand(2, 2).clone(), name="energy weighted")


def absorber_response(thickness_cm: float, density: float, mass_attenuation: Callable[[np.ndarray], np.ndarray],
                      energies: Optional[np.ndarray] = None, thetas: Optional[np.ndarray] = None,
                      name: str = "absorber") -> DetectorResponse:
    """Energy-integrating absorber with the oblique path ``t / cos(theta)`` (the model of csi_response,
    detector.py:258): ``D(E, theta) = E (1 - exp(-mu(E) rho t / cos(theta)))``, ``mass_attenuation`` in cm^2/g."""
    if thickness_cm < 0:
        raise ValueError("thickness must be non-negative")
    en = np.sort(np.asarray(np.linspace(10.0, 150.0, 281) if energies is None else energies,
                            dtype=np.float64).reshape(-1))
    th = np.sort(np.asarray(np.radians(np.linspace(0.0, 89.5, 181)) if thetas is None else thetas,
                            dtype=np.float64).reshape(-1))
    mu = np.asarray(mass_attenuation(en), dtype=np.float64)
    path = thickness_cm / np.cos(np.clip(th, 0.0, math.pi / 2 - 1e-9))
    resp = en[:, None] * (-np.expm1(-mu[:, None] * density * path[None, :]))
    return DetectorResponse(torch.as_tensor(en), torch.as_tensor(th), torch.as_tensor(resp), name=name)


@dataclass(frozen=True, eq=False)
class GridResponse:
    """Anti-scatter-grid transmission table [E, T1, T2] on energy and grid-angle nodes (detector.py:657); only the
    angle-independent table of ``no_grid`` is evaluated by this build."""

    energy_kev: Tensor
    theta1_rad: Tensor
    theta2_rad: Tensor
    transmission: Tensor
    name: str = "grid"

    @property
    def depends_on_angles(self) -> bool:
        t = self.transmission
        return not bool((t == t[:, :1, :1]).all())

    def at_energy(self, group_energy: Tensor, device, dtype) -> Tensor:
        """[G] transmission of an angle-independent grid at the group energies."""
        if self.depends_on_angles:
            raise NotImplementedError("septa whose transmission depends on the angles are not supported")
        en = self.energy_kev.to(device, dtype)
        table = self.transmission[:, 0, 0].to(device, dtype)
        i0, i1, w = _bracket(en, group_energy.to(device, dtype).reshape(-1))
        return table[i0] + (table[i1] - table[i0]) * w

    def transmission_at(self, group_energy: Tensor, omega: Tensor, pixel_positions: Tensor, geometry,
                        angle: float) -> Tensor:
        """[G, ...pixels] transmission of photons along ``omega`` at the pixels (angle-independent grids)."""
        t = self.at_energy(group_energy, pixel_positions.device, pixel_positions.dtype)
        return t.reshape(-1, *([1] * (pixel_positions.dim() - 1))).expand(-1, *pixel_positions.shape[:-1])


def no_grid() -> GridResponse:
    """No grid: transmission 1 at every energy and angle (detector.py:758)."""
    e = torch.tensor([1.0, 1000.0], dtype=torch.float64)
    t = torch.tensor([0.0, math.pi / 2], dtype=torch.float64)
    return GridResponse(e, t, t.clone(), torch.ones(2, 2, 2, dtype=torch.float64), name="no grid")


def detector_angle(omega: Tensor, geometry, angle: float, pixel_positions: Optional[Tensor] = None) -> Tensor:
    """[...] angle theta0 between photon directions and the inward panel normal ``-n_hat`` (detector.py:1054), with
    the local normal at ``pixel_positions`` on a curved panel."""
    _, _, n_hat = geometry.panel_axes(angle)
    n = torch.tensor(n_hat, device=omega.device, dtype=omega.dtype)
    if geometry.use_curved_detector and pixel_positions is not None:
        n = geometry.local_panel_axes(pixel_positions.to(omega.device, omega.dtype), angle)[2]
    c = (omega * (-n)).sum(-1) / torch.linalg.norm(omega, dim=-1)
    return torch.arccos(c.clamp(-1.0, 1.0))
