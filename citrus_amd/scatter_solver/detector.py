"""Detector and anti-scatter-grid responses of the signal weights (host-side tables).

Keeps the interface of ``citrus/scatter_solver/detector.py``: ``DetectorResponse`` (:197, a table D(E, theta0) in
keV deposited per incident photon, bilinear in energy and angle, clamped to the end nodes, |theta| used),
``ideal_counting_response`` (:294, D = 1), ``energy_weighted_response`` (:305, D = E), ``csi_response`` (:258, a CsI
scintillator with the oblique path ``t / cos(theta)`` and the coarse XCOM mass-attenuation table of CsI, :58-76) and
the general ``absorber_response`` behind it, ``csi_mass_attenuation`` / ``lead_mass_attenuation`` (:120-144, log-log
interpolation of the two embedded tables), ``GridResponse`` (:657: trilinear in energy and the grid angles, focused
or unfocused, septa along the channels or the rows, crossed septa) with ``no_grid`` (:758), ``grid_angles`` (:1077),
``default_focal_distance`` (:1039), ``focused_1d_grid`` (:779, the septum-shadow table of the reference's signature)
and the closed-form direct transmission of a 1-D grid (``analytic_1d_grid``, the direct part of grid_model.py:270),
and ``detector_angle`` (:1054).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np
import torch
from torch import Tensor


RHO_CSI = 4.51  # g/cm^3 (detector.py:52)
RHO_LEAD = 11.35  # g/cm^3 (detector.py:53)
_EDGE_STEP_KEV = 1e-3  # the two energies of an absorption-edge pair are this far apart (detector.py:54)

# Total mass attenuation with coherent scattering (cm^2/g) of the coarse NIST XCOM tables the reference embeds
# (detector.py:58-76 and :78-93). Data, kept as published: CsI with the iodine (33.17 keV) and caesium (35.98 keV) K
# edges, lead with its L3 (13.04 keV) and K (88.0 keV) edges.
CSI_TABLE = np.array([(10.0, 168.0), (15.0, 57.5), (20.0, 26.6), (25.0, 14.6), (30.0, 8.9), (33.17, 6.6),
                      (33.17 + _EDGE_STEP_KEV, 21.6), (35.98, 17.5), (35.98 + _EDGE_STEP_KEV, 31.3), (40.0, 23.5),
                      (50.0, 12.9), (60.0, 7.95), (80.0, 3.75), (100.0, 2.07), (150.0, 0.72)])
LEAD_TABLE = np.array([(10.0, 130.6), (13.04, 67.0), (13.04 + _EDGE_STEP_KEV, 157.0), (15.0, 111.6), (20.0, 86.4),
                       (30.0, 30.3), (40.0, 14.4), (50.0, 8.04), (60.0, 5.02), (80.0, 2.42), (88.0, 1.91),
                       (88.0 + _EDGE_STEP_KEV, 7.68), (100.0, 5.55), (150.0, 2.01)])


def _as_numpy(x) -> np.ndarray:
    return np.asarray(x.detach().cpu().numpy() if isinstance(x, Tensor) else x, dtype=np.float64)


def log_log_table(table: np.ndarray, energy_kev) -> np.ndarray:
    """Mass attenuation of an ascending ``[N, 2]`` (keV, cm^2/g) table at positive energies: a straight line in
    (log E, log mu) between neighbouring nodes, and the end segments' lines continued beyond the table
    (detector.py:96-118)."""
    e = _as_numpy(energy_kev)
    if np.any(e <= 0):
        raise ValueError("energies must be positive")
    le, lm = np.log(table[:, 0]), np.log(table[:, 1])
    x = np.log(e)
    seg = np.clip(np.searchsorted(le, x, side="right") - 1, 0, len(le) - 2)  # segment, end segments extended
    inside = (x >= le[0]) & (x <= le[-1])
    y_line = lm[seg] + (lm[seg + 1] - lm[seg]) / (le[seg + 1] - le[seg]) * (x - le[seg])
    y = np.where(inside, np.interp(x, le, lm), y_line)
    return np.exp(y)


def csi_mass_attenuation(energy_kev) -> np.ndarray:
    """Total mass attenuation of CsI (cm^2/g) from the embedded table (detector.py:121)."""
    return log_log_table(CSI_TABLE, energy_kev)


def lead_mass_attenuation(energy_kev) -> np.ndarray:
    """Total mass attenuation of lead (cm^2/g) from the embedded table (detector.py:134)."""
    return log_log_table(LEAD_TABLE, energy_kev)


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
                      energies=None, thetas=None, name: str = "absorber") -> DetectorResponse:
    """Energy-integrating absorber with the oblique path ``t / cos(theta)``: every interacting photon deposits its
    energy, ``D(E, theta) = E (1 - exp(-mu(E) rho t / cos(theta)))`` keV per incident photon, ``mass_attenuation`` in
    cm^2/g of keV (numpy in and out). Table nodes: 10-150 keV every 0.5 keV and 181 angles from 0 to 89.5 degrees
    unless given (detector.py:258-291)."""
    if thickness_cm < 0:
        raise ValueError("thickness must be non-negative")
    en = np.sort(_as_numpy(np.linspace(10.0, 150.0, 281) if energies is None else energies).reshape(-1))
    th = np.sort(_as_numpy(np.radians(np.linspace(0.0, 89.5, 181)) if thetas is None else thetas).reshape(-1))
    mu = np.asarray(mass_attenuation(en), dtype=np.float64)
    path = thickness_cm / np.cos(np.clip(th, 0.0, math.pi / 2 - 1e-9))
    resp = en[:, None] * (-np.expm1(-mu[:, None] * density * path[None, :]))
    return DetectorResponse(torch.as_tensor(en), torch.as_tensor(th), torch.as_tensor(resp), name=name)


def csi_response(thickness_cm: float = 0.06, density: float = RHO_CSI, energies=None,
                 mass_attenuation: Optional[Callable[[np.ndarray], np.ndarray]] = None,
                 thetas=None) -> DetectorResponse:
    """CsI scintillator of thickness t (default 0.6 mm) with the oblique path model (detector.py:258), its mass
    attenuation from the embedded table unless given."""
    return absorber_response(thickness_cm, density, csi_mass_attenuation if mass_attenuation is None
                             else mass_attenuation, energies, thetas,
                             name=f"CsI {thickness_cm * 10:.2f} mm (1/cos model)")


SEPTA_DIRECTIONS = ("channels", "rows")


@dataclass(frozen=True, eq=False)
class GridResponse:
    """Anti-scatter-grid transmission table G(E, theta1, theta2) (detector.py:657): trilinear in energy and the
    grid angles of ``grid_angles``, clamped to the end nodes, |theta| used. ``focal_distance_cm`` sets the focusing
    direction (None: ``default_focal_distance``; ``math.inf``: unfocused); ``septa_along`` is "channels" (lamellae
    along u, repeating along the rows) or "rows"; ``crossed`` is a second set of septa in the other direction, whose
    transmission multiplies this one."""

    energy_kev: Tensor
    theta1_rad: Tensor
    theta2_rad: Tensor
    transmission: Tensor
    name: str = "grid"
    focal_distance_cm: Optional[float] = None
    septa_along: str = "channels"
    crossed: Optional["GridResponse"] = None

    def __post_init__(self) -> None:
        shape = (self.energy_kev.numel(), self.theta1_rad.numel(), self.theta2_rad.numel())
        if tuple(self.transmission.shape) != shape:
            raise ValueError(f"transmission must be [E, T1, T2] = {list(shape)}, got {tuple(self.transmission.shape)}")
        if self.septa_along not in SEPTA_DIRECTIONS:
            raise ValueError(f"septa_along must be one of {SEPTA_DIRECTIONS}, got {self.septa_along!r}")
        if self.crossed is not None and (self.crossed.septa_along == self.septa_along
                                         or self.crossed.crossed is not None):
            raise ValueError("crossed septa must run in the other direction and have no crossed septa of their own")

    @property
    def depends_on_angles(self) -> bool:
        t = self.transmission
        return self.crossed is not None or not bool((t == t[:, :1, :1]).all())

    def at_energy(self, group_energy: Tensor, device, dtype) -> Tensor:
        """[G] transmission of an angle-independent grid at the group energies."""
        if self.depends_on_angles:
            raise NotImplementedError("this grid's transmission depends on the photon's angles: use transmission_at")
        en = self.energy_kev.to(device, dtype)
        table = self.transmission[:, 0, 0].to(device, dtype)
        i0, i1, w = _bracket(en, group_energy.to(device, dtype).reshape(-1))
        return table[i0] + (table[i1] - table[i0]) * w

    def at(self, group_energy: Tensor, theta1: Tensor, theta2: Tensor) -> Tensor:
        """[G, ...] table interpolated trilinearly at the group energies and the grid angles (|theta| used),
        on the device and in the dtype of ``theta1`` (detector.py:729)."""
        device, dtype = theta1.device, theta1.dtype
        theta1, theta2 = torch.broadcast_tensors(theta1, theta2.to(device, dtype))
        en, t1, t2 = (x.to(device, dtype) for x in (self.energy_kev, self.theta1_rad, self.theta2_rad))
        table = self.transmission.to(device, dtype)
        ie0, ie1, we = _bracket(en, group_energy.to(device, dtype).reshape(-1))
        at_e = table[ie0] + (table[ie1] - table[ie0]) * we[:, None, None]  # [G, T1, T2]
        i10, i11, w1 = _bracket(t1, theta1.abs())
        i20, i21, w2 = _bracket(t2, theta2.abs())
        lo = at_e[:, i10, i20] + (at_e[:, i10, i21] - at_e[:, i10, i20]) * w2
        hi = at_e[:, i11, i20] + (at_e[:, i11, i21] - at_e[:, i11, i20]) * w2
        return lo + (hi - lo) * w1

    def transmission_at(self, group_energy: Tensor, omega: Tensor, pixel_positions: Tensor, geometry,
                        angle: float) -> Tensor:
        """[G, ...] transmission of photons along ``omega`` arriving at the pixels (detector.py:704): the table at
        their grid angles, times the crossed septa's."""
        if not self.depends_on_angles:
            t = self.at_energy(group_energy, pixel_positions.device, pixel_positions.dtype)
            if omega.dim() == 2 and pixel_positions.dim() == 3:  # the shapes of grid_angles
                shape = (int(omega.shape[0]),) + tuple(pixel_positions.shape[:-1])
            else:
                shape = tuple(torch.broadcast_shapes(omega.shape[:-1], pixel_positions.shape[:-1]))
            return t.reshape(-1, *([1] * len(shape))).expand(-1, *shape)
        theta1, theta2 = grid_angles(omega, pixel_positions, geometry, angle, self.focal_distance_cm,
                                     self.septa_along)
        out = self.at(group_energy, theta1, theta2)
        if self.crossed is not None:
            out = out * self.crossed.transmission_at(group_energy, omega, pixel_positions, geometry, angle)
        return out


def no_grid() -> GridResponse:
    """No grid: transmission 1 at every energy and angle (detector.py:758)."""
    e = torch.tensor([1.0, 1000.0], dtype=torch.float64)
    t = torch.tensor([0.0, math.pi / 2], dtype=torch.float64)
    return GridResponse(e, t, t.clone(), torch.ones(2, 2, 2, dtype=torch.float64), name="no grid")


def default_focal_distance(geometry) -> float:
    """Focal distance of a grid built for a geometry (detector.py:1039): the source-detector distance of a cone
    beam, infinite (unfocused) for a parallel beam."""
    return math.inf if geometry.is_parallel else float(geometry.source_detector_dist)


def grid_angles(omega: Tensor, pixel_positions: Tensor, geometry, angle: float,
                focal_distance: Optional[float] = None, septa_along: str = "channels") -> tuple[Tensor, Tensor]:
    """(theta1, theta2) grid angles of photons along ``omega`` arriving at pixels (detector.py:1077).

    With the panel axes (u, v, n) at the pixel (n toward the source; on a curved panel the surface's axes at each
    pixel) and a = -omega: theta2 = atan2(a.u, a.n) and theta1 = atan2(a.v, a.n) - atan2(f.v, f.n), f the direction
    from the pixel to the focal line (f = n for an unfocused grid, theta1 then measured from the normal). Septa along
    the channels focus on the line parallel to u at ``focal_distance`` along the normal from the panel centre (on a
    curved panel: from the central row of each column); septa along the rows swap u and v and focus along each
    pixel's normal. ``omega [K, 3]`` with ``pixel_positions [rows, channels, 3]`` gives [K, rows, channels]."""
    if septa_along not in SEPTA_DIRECTIONS:
        raise ValueError(f"septa_along must be one of {SEPTA_DIRECTIONS}, got {septa_along!r}")
    device, dtype = pixel_positions.device, pixel_positions.dtype
    omega = omega.to(device, dtype)
    fd = default_focal_distance(geometry) if focal_distance is None else float(focal_distance)
    if omega.dim() == 2 and pixel_positions.dim() == 3:
        a, pixels = -omega[:, None, None, :], pixel_positions[None]
    else:
        a, pixels = -omega, pixel_positions
    pc = torch.tensor(geometry.panel_center(angle), device=device, dtype=dtype)
    if geometry.use_curved_detector:
        u_hat, v_hat, n_hat = geometry.local_panel_axes(pixels.expand(*torch.broadcast_shapes(a.shape, pixels.shape)),
                                                        angle)
    else:
        u_hat, v_hat, n_hat = (torch.tensor(x, device=device, dtype=dtype) for x in geometry.panel_axes(angle))
    if math.isinf(fd):
        f = None
    elif not geometry.use_curved_detector:
        f = pc + fd * n_hat - pixels
    elif septa_along == "channels":  # the focal line at the central row of each column
        f = -((pixels - pc) * v_hat).sum(-1, keepdim=True) * v_hat + fd * n_hat
    else:
        f = fd * n_hat + torch.zeros_like(pixels)
    if septa_along == "rows":
        u_hat, v_hat = v_hat, u_hat
    a_n = (a * n_hat).sum(-1)
    theta1 = torch.atan2((a * v_hat).sum(-1), a_n)
    if f is not None:
        theta1 = theta1 - torch.atan2((f * v_hat).sum(-1), (f * n_hat).sum(-1))
    theta2 = torch.atan2((a * u_hat).sum(-1), a_n)
    return torch.broadcast_tensors(theta1, theta2)


def _strip_measure(x: Tensor, pitch: float, thickness: float) -> Tensor:
    """Length of the strips [k p, k p + t) between 0 and x across them (negative for x < 0)."""
    k = torch.floor(x / pitch)
    return k * thickness + torch.clamp(x - k * pitch, 0.0, thickness)


def grid_direct_transmission(pitch: float, thickness: float, height: float, mu_strip: Tensor, mu_interspace: Tensor,
                             theta1: Tensor, theta2: Tensor) -> Tensor:
    """[E, T1, T2] float64 transmission of the photons that cross a 1-D grid without interacting, averaged over the
    entry offset across one pitch (the direct part of grid_model.py:270).

    A photon at grid angles (theta1, theta2) crosses the grid height h over the path L = h sqrt(1 + tan^2 theta1 +
    tan^2 theta2) and moves D = h tan|theta1| across the strips; entering at offset x it spends the share
    s(x) = (C(x + D) - C(x)) / D of the path in the strips (C the strip measure). s is linear between the offsets where
    x or x + D meets a strip face, so the mean of exp(-(mu_s - mu_i) L s(x)) over x is a sum of closed-form integrals
    of exponentials of linear functions; with D = 0 it is ((p - t) + t exp(-(mu_s - mu_i) L)) / p. Both parts carry
    the interspace's exp(-mu_i L)."""
    if not (pitch > 0 and 0 <= thickness <= pitch and height > 0):
        raise ValueError("need pitch > 0, 0 <= thickness <= pitch and height > 0")
    f64 = torch.float64
    ms = torch.as_tensor(mu_strip, dtype=f64).reshape(-1)
    mi = torch.as_tensor(mu_interspace, dtype=f64).reshape(-1)
    t1 = torch.as_tensor(theta1, dtype=f64).reshape(-1).abs().clamp(max=math.pi / 2 - 1e-9)
    t2 = torch.as_tensor(theta2, dtype=f64).reshape(-1).abs().clamp(max=math.pi / 2 - 1e-9)
    p, t, h = float(pitch), float(thickness), float(height)
    L = h * torch.sqrt(1.0 + torch.tan(t1)[:, None] ** 2 + torch.tan(t2)[None, :] ** 2)  # [T1, T2]
    D = h * torch.tan(t1)  # [T1]
    moving = D > 1e-12 * p
    Ds = torch.where(moving, D, torch.ones_like(D))
    # offsets in [0, p) where the share s is not linear: the entry or the exit meets a strip face
    knots = torch.stack([torch.zeros_like(D), torch.full_like(D, t), torch.remainder(-D, p),
                         torch.remainder(t - D, p), torch.full_like(D, p)], -1).sort(-1).values  # [T1, 5]
    share = (_strip_measure(knots + D[:, None], p, t) - _strip_measure(knots, p, t)) / Ds[:, None]  # [T1, 5]
    c = (ms - mi)[:, None, None, None] * L[None, :, :, None]  # [E, T1, T2, 1]
    e = c * share[None, :, None, :]  # exponent at the knots [E, T1, T2, 5]
    e0, e1 = e[..., :-1], e[..., 1:]
    d = (e1 - e0).abs()
    mean = torch.exp(-torch.minimum(e0, e1)) * torch.where(d > 1e-12, -torch.expm1(-d) / d.clamp(min=1e-300),
                                                          torch.ones_like(d))
    widths = (knots[:, 1:] - knots[:, :-1])[None, :, None, :]
    tilted = (mean * widths).sum(-1) / p
    normal = ((p - t) + t * torch.exp(-c[..., 0])) / p
    across = torch.where(moving[None, :, None], tilted, normal)
    return across * torch.exp(-mi[:, None, None] * L[None])


def analytic_1d_grid(ratio: float, lamella_thickness_cm: float, lamella_height_cm: float,
                     strip_attenuation: Callable[[np.ndarray], np.ndarray],
                     interspace_attenuation: Callable[[np.ndarray], np.ndarray], *,
                     energy_kev=None, theta1_rad=None, theta2_rad=None, focal_distance_cm: Optional[float] = None,
                     septa_along: str = "channels", name: str = "analytic 1-D grid") -> GridResponse:
    """A focused 1-D grid of lamellae of thickness t and height h with the interspace width h / ratio (pitch
    t + h / ratio), tabulated from its direct transmission through strips and interspace (``grid_direct_transmission``,
    the direct part of the reference's analytic model, grid_model.py:270). The attenuations (1/cm, of energy in keV)
    are passed in; the septa's own scatter and fluorescence (the secondary part of grid_model.py) are not modelled."""
    if ratio <= 0:
        raise ValueError("ratio must be positive")
    h, tl = float(lamella_height_cm), float(lamella_thickness_cm)
    pitch = tl + h / float(ratio)
    en = _as_numpy(np.linspace(10.0, 150.0, 141) if energy_kev is None else energy_kev).reshape(-1)
    t1 = _as_numpy(np.radians(np.linspace(0.0, 30.0, 241)) if theta1_rad is None else theta1_rad).reshape(-1)
    t2 = _as_numpy(np.radians(np.linspace(0.0, 60.0, 61)) if theta2_rad is None else theta2_rad).reshape(-1)
    table = grid_direct_transmission(pitch, tl, h, torch.as_tensor(strip_attenuation(en)),
                                     torch.as_tensor(interspace_attenuation(en)), torch.as_tensor(t1),
                                     torch.as_tensor(t2))
    return GridResponse(torch.as_tensor(en), torch.as_tensor(t1), torch.as_tensor(t2), table, name=name,
                        focal_distance_cm=focal_distance_cm, septa_along=septa_along)


def _grid_theta1_nodes() -> np.ndarray:
    """theta1 nodes (rad): 401 from 0 to 10 degrees, where a lamella's shadow opens, then 160 up to 89.5 degrees
    (detector.py:767-776)."""
    return np.radians(np.unique(np.concatenate([np.linspace(0.0, 10.0, 401), np.linspace(10.0, 89.5, 160)])))


def septum_shadow_transmission(pitch_cm: float, thickness_cm: float, height_cm: float, mu_lead_cm: np.ndarray,
                               theta1: np.ndarray) -> np.ndarray:
    """[E, T1] transmission of the geometric lamella-shadow model (detector.py:779-829): a photon tilted by theta1 from
    the focusing direction passes the open part of a period, ``open = max(0, (p - t) - h |tan theta1|) / p``, or
    crosses lead along the mean path ``L = (t / p) h / cos(theta1) / (1 - open)`` (the lead volume per period seen
    along the tilted track, spread over the shadowed share), so ``G = open + (1 - open) exp(-mu L)``; ``mu_lead_cm``
    [E] is the linear attenuation of the lamellae in 1/cm."""
    p, t, h = float(pitch_cm), float(thickness_cm), float(height_cm)
    th = np.clip(np.asarray(theta1, dtype=np.float64), -math.pi / 2 + 1e-9, math.pi / 2 - 1e-9)
    open_share = np.maximum(0.0, (p - t) - h * np.abs(np.tan(th))) / p
    shadow = 1.0 - open_share
    safe = np.where(shadow > 0, shadow, 1.0)
    lead_path = np.where(shadow > 0, (t / p) * h / np.cos(th) / safe, 0.0)
    mu = np.asarray(mu_lead_cm, dtype=np.float64).reshape(-1, 1)
    return open_share[None, :] + shadow[None, :] * np.exp(-mu * lead_path[None, :])


def focused_1d_grid(ratio: float = 10.0, lamella_thickness_cm: float = 0.0036, lamella_height_cm: float = 0.15,
                    pitch_cm: float = 1.0 / 60.0, focal_distance_cm: float = 150.0,
                    mass_attenuation: Optional[Callable[[np.ndarray], np.ndarray]] = None,
                    lead_density: float = RHO_LEAD, energies=None, theta1=None) -> GridResponse:
    """A focused 1-D lead grid in the geometric lamella-shadow model, with the reference's signature and defaults
    (detector.py:779): lamellae of thickness t and height h at pitch p, the transmission
    ``septum_shadow_transmission`` of lead (``lead_mass_attenuation`` x ``lead_density`` unless ``mass_attenuation``
    is given), independent of theta2. ``ratio`` only names the grid; t, h and p set its geometry."""
    p, t, h = float(pitch_cm), float(lamella_thickness_cm), float(lamella_height_cm)
    if not (0.0 <= t < p) or h < 0:
        raise ValueError("need 0 <= lamella thickness < pitch and height >= 0")
    en = np.sort(_as_numpy(np.linspace(10.0, 150.0, 281) if energies is None else energies).reshape(-1))
    t1 = np.sort(_as_numpy(_grid_theta1_nodes() if theta1 is None else theta1).reshape(-1))
    mu = np.asarray((lead_mass_attenuation if mass_attenuation is None else mass_attenuation)(en), dtype=np.float64)
    table = septum_shadow_transmission(p, t, h, mu * float(lead_density), t1)
    t2 = np.array([0.0, math.pi / 2])
    return GridResponse(torch.as_tensor(en), torch.as_tensor(t1), torch.as_tensor(t2),
                        torch.as_tensor(np.repeat(table[:, :, None], 2, axis=2)),
                        focal_distance_cm=float(focal_distance_cm),
                        name=f"focused 1D grid (nominal {ratio:g}:1, h/(p-t) = {h / (p - t):.1f}), {1 / p:.0f} lp/cm "
                             "(geometric model)")


def detector_angle(omega: Tensor, geometry, angle: float, pixel_positions: Optional[Tensor] = None) -> Tensor:
    """[...] angle theta0 between photon directions and the inward panel normal ``-n_hat`` (detector.py:1054), with
    the local normal at ``pixel_positions`` on a curved panel."""
    _, _, n_hat = geometry.panel_axes(angle)
    n = torch.tensor(n_hat, device=omega.device, dtype=omega.dtype)
    if geometry.use_curved_detector and pixel_positions is not None:
        n = geometry.local_panel_axes(pixel_positions.to(omega.device, omega.dtype), angle)[2]
    c = (omega * (-n)).sum(-1) / torch.linalg.norm(omega, dim=-1)
    return torch.arccos(c.clamp(-1.0, 1.0))
