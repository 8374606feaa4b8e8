"""Direction sets, real spherical harmonics and the moment transforms of the sweep (host-side setup).

Keeps the interface of ``citrus/scatter_solver/angular_quadrature.py``: ``Quadrature`` (:63), ``octant_of`` (:47),
``product_quadrature`` (:187), ``legendre_polynomials`` (:430), ``real_spherical_harmonics`` (:475, orthonormal,
no Condon-Shortley phase, flat index ``n = l*l + k + l``), ``SphericalHarmonicBasis`` / ``spherical_harmonic_basis``
(:528-567), ``moments_from_ordinates`` / ``ordinates_from_moments`` (:570-593), ``scatter_source_moments`` (:600) and
``scatter_source_at_directions`` (:680). New here: ``level_symmetric_quadrature`` (LQ_N sets, which the reference does
not have). All tables are computed in float64 once per context; the per-cell work runs in the HIP kernels.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Union

import numpy as np
import torch
from torch import Tensor

FOUR_PI = 4.0 * math.pi
_DeviceLike = Union[torch.device, str, None]


def octant_of(omega: Tensor) -> Tensor:
    """Octant code ``(w_x < 0) * 4 + (w_y < 0) * 2 + (w_z < 0)``; a zero component counts as non-negative."""
    neg = (omega < 0).to(torch.int64)
    return neg[..., 0] * 4 + neg[..., 1] * 2 + neg[..., 2]


@dataclass(frozen=True, eq=False)
class Quadrature:
    """Directions ``omega [M, 3]`` with weights ``weight [M]`` in sr, octant codes and optional cone cell widths."""

    omega: Tensor
    weight: Tensor
    octant: Tensor
    cell_width: Optional[Tensor] = None

    @staticmethod
    def from_directions(omega: Tensor, weight: Tensor) -> "Quadrature":
        if omega.dim() != 2 or omega.shape[1] != 3:
            raise ValueError(f"omega must be [M, 3], got {tuple(omega.shape)}")
        if weight.shape != omega.shape[:1]:
            raise ValueError(f"weight must be [M], got {tuple(weight.shape)}")
        return Quadrature(omega=omega, weight=weight, octant=octant_of(omega))

    @property
    def num_directions(self) -> int:
        return int(self.omega.shape[0])

    @property
    def device(self) -> torch.device:
        return self.omega.device

    @property
    def dtype(self) -> torch.dtype:
        return self.omega.dtype

    def octant_counts(self) -> Tensor:
        return torch.bincount(self.octant, minlength=8)

    def by_octant(self) -> tuple[Tensor, Tensor]:
        """Stable permutation listing octant 0 first, and the per-octant counts."""
        return torch.sort(self.octant, stable=True).indices, self.octant_counts()

    @property
    def directions_per_octant(self) -> Optional[int]:
        counts = self.octant_counts().tolist()
        return int(counts[0]) if all(c == counts[0] for c in counts) else None

    def reordered(self, perm: Tensor) -> "Quadrature":
        cw = None if self.cell_width is None else self.cell_width[perm.to(self.cell_width.device)]
        return Quadrature(omega=self.omega[perm], weight=self.weight[perm], octant=self.octant[perm], cell_width=cw)

    def to(self, device: _DeviceLike = None, dtype: Optional[torch.dtype] = None) -> "Quadrature":
        kw = {} if device is None else {"device": device}
        conv = (lambda t: t.to(dtype=dtype, **kw)) if dtype is not None else (lambda t: t.to(**kw))
        cw = None if self.cell_width is None else conv(self.cell_width)
        return Quadrature(omega=conv(self.omega), weight=conv(self.weight), octant=self.octant.to(**kw), cell_width=cw)


def product_quadrature(num_polar: int, num_azimuthal: int, device: _DeviceLike = None,
                       dtype: torch.dtype = torch.float64) -> Quadrature:
    """Gauss-Legendre in cos(theta) times equal azimuths ``(j + 1/2) 2 pi / A``; weights sum to 4 pi; m = i*A + j."""
    if num_azimuthal < 1:
        raise ValueError("num_azimuthal must be >= 1")
    if num_polar < 1:
        raise ValueError("number of Gauss-Legendre points must be >= 1")
    nodes, gw = np.polynomial.legendre.leggauss(num_polar)
    mu = torch.as_tensor(nodes, dtype=torch.float64)
    step = 2.0 * math.pi / num_azimuthal
    phi = (torch.arange(num_azimuthal, dtype=torch.float64) + 0.5) * step
    st = torch.sqrt((1.0 - mu * mu).clamp_min(0.0))
    omega = torch.stack([st[:, None] * torch.cos(phi)[None, :], st[:, None] * torch.sin(phi)[None, :],
                         mu[:, None].expand(num_polar, num_azimuthal)], dim=-1).reshape(-1, 3)
    weight = (torch.as_tensor(gw, dtype=torch.float64)[:, None] * step).expand(num_polar, num_azimuthal).reshape(-1)
    return Quadrature.from_directions(omega.to(device=device, dtype=dtype), weight.to(device=device, dtype=dtype))


# LQ_N level-symmetric sets: first cosine and the point weights (per octant, summing to 1) of each weight class,
# from the published table (Lewis & Miller, Computational Methods of Neutron Transport, 1984, table 4-1).
_LQN_TABLE = {
    2: (0.5773503, (1.0,)),
    4: (0.3500212, (0.3333333,)),
    6: (0.2666355, (0.1761263, 0.1572071)),
    8: (0.2182179, (0.1209877, 0.0907407, 0.0925926)),
    12: (0.1672126, (0.0707626, 0.0558811, 0.0373377, 0.0502819, 0.0258513)),
}
# weight class of a level triple, keyed by the triple sorted in descending order
_LQN_CLASS = {
    2: {(1, 1, 1): 0}, 4: {(2, 1, 1): 0}, 6: {(3, 1, 1): 0, (2, 2, 1): 1},
    8: {(4, 1, 1): 0, (3, 2, 1): 1, (2, 2, 2): 2},
    12: {(6, 1, 1): 0, (5, 2, 1): 1, (4, 3, 1): 2, (4, 2, 2): 3, (3, 3, 2): 4},
}


def level_symmetric_quadrature(order: int, device: _DeviceLike = None, dtype: torch.dtype = torch.float64
                               ) -> Quadrature:
    """Level-symmetric S_N set (N = 2, 4, 6, 8, 12) with N(N+2) directions and weights summing to 4 pi.

    The cosines are ``mu_i^2 = mu_1^2 + (i - 1) * 2 (1 - 3 mu_1^2) / (N - 2)``; a direction of octant 0 is
    ``(mu_i, mu_j, mu_k)`` with ``i + j + k = N/2 + 2``, listed with i then j descending; the other octants are its
    reflections, in octant-code order, so ``directions_per_octant = N(N+2)/8`` (S8: 10, S12: 21).
    """
    if order not in _LQN_TABLE:
        raise ValueError(f"level-symmetric order must be one of {sorted(_LQN_TABLE)}, got {order}")
    mu1, class_weights = _LQN_TABLE[order]
    half = order // 2
    if half == 1:
        cosines = [mu1]
    else:
        step = 2.0 * (1.0 - 3.0 * mu1 * mu1) / (order - 2)
        cosines = [math.sqrt(mu1 * mu1 + step * i) for i in range(half)]
    total = half + 2
    points, weights = [], []
    for i in range(half, 0, -1):
        for j in range(half, 0, -1):
            k = total - i - j
            if k < 1:
                continue
            points.append((cosines[i - 1], cosines[j - 1], cosines[k - 1]))
            weights.append(class_weights[_LQN_CLASS[order][tuple(sorted((i, j, k), reverse=True))]])
    base = torch.tensor(points, dtype=torch.float64)
    base = base / torch.linalg.norm(base, dim=1, keepdim=True)
    w = torch.tensor(weights, dtype=torch.float64)
    w = w / w.sum() * (FOUR_PI / 8.0)
    signs = torch.tensor([[-1.0 if o & 4 else 1.0, -1.0 if o & 2 else 1.0, -1.0 if o & 1 else 1.0] for o in range(8)],
                         dtype=torch.float64)
    omega = (signs[:, None, :] * base[None]).reshape(-1, 3)
    weight = w.repeat(8)
    return Quadrature.from_directions(omega.to(device=device, dtype=dtype), weight.to(device=device, dtype=dtype))


# ---- cone sets of the collapsed-cone scatter image (angular_quadrature.py:229-420) ---------------------------------

def _unit_vector(v) -> Tensor:
    t = torch.as_tensor(v, dtype=torch.float64).detach().to("cpu").reshape(-1)
    if t.numel() != 3:
        raise ValueError("axis must have 3 components")
    n = torch.linalg.norm(t)
    if not bool(n > 0):
        raise ValueError("axis must be non-zero")
    return t / n


def _local_frame(a: Tensor) -> tuple[Tensor, Tensor, Tensor]:
    """Right-handed frame (e1, e2, a) about a unit axis: the helper is the coordinate axis least aligned with a (lowest
    index on ties), ``e1 = normalize(helper - (helper . a) a)``, ``e2 = a x e1`` (angular_quadrature.py:248)."""
    helper = torch.zeros(3, dtype=torch.float64)
    helper[int(torch.argmin(a.abs()))] = 1.0
    e1 = helper - (helper @ a) * a
    e1 = e1 / torch.linalg.norm(e1)
    return e1, torch.linalg.cross(a, e1), a


def _cone_polar(num_polar: int, max_polar_angle: float) -> tuple[Tensor, Tensor]:
    """(cos theta, weights) [P] of the Gauss-Legendre nodes on [cos(max), 1], near-axis first."""
    cos_max = math.cos(max_polar_angle)
    x, w = np.polynomial.legendre.leggauss(num_polar)
    half = 0.5 * (1.0 - cos_max)
    return (torch.as_tensor(cos_max + half * (x + 1.0), dtype=torch.float64).flip(0),
            torch.as_tensor(w * half, dtype=torch.float64).flip(0))


def _cone_cell_widths(a: Tensor, u: Tensor, v: Tensor, num_polar: int, num_azimuthal: int, max_polar_angle: float,
                      subsamples: int) -> Tensor:
    """[M, 2] gnomonic widths of every cone cell (angular_quadrature.py:375): node (i, j) owns the cell between the
    cumulative-weight edges in cos(theta) and the azimuths [j dphi, (j+1) dphi]; the width is sqrt(12) times the
    standard deviation of ``(d.u / d.a, d.v / d.a)`` over an s x s equal-area lattice of the cell."""
    s = int(subsamples)
    frac = (torch.arange(s, dtype=torch.float64) + 0.5) / s
    cos_max = math.cos(max_polar_angle)
    _, pw = _cone_polar(num_polar, max_polar_angle)
    edges = torch.cat([torch.ones(1, dtype=torch.float64), 1.0 - torch.cumsum(pw, 0)])
    edges[-1] = cos_max
    step = 2.0 * math.pi / num_azimuthal
    e1, e2, _ = _local_frame(a)
    cs = edges[1:, None] + (edges[:-1] - edges[1:])[:, None] * frac[None, :]  # [P, s]
    az = (torch.arange(num_azimuthal, dtype=torch.float64)[:, None] + frac[None, :]) * step  # [A, s]
    sn = torch.sqrt((1.0 - cs * cs).clamp_min(0.0))
    lateral = torch.cos(az)[..., None] * e1 + torch.sin(az)[..., None] * e2  # [A, s, 3]
    d = sn[:, None, :, None, None] * lateral[None, :, None, :, :] + cs[:, None, :, None, None] * a
    nc = (d @ a).clamp_min(1e-6)
    alpha, beta = (d @ u) / nc, (d @ v) / nc
    n = float(s * s)

    def width(x: Tensor) -> Tensor:
        mean = x.sum(dim=(2, 3)) / n
        var = ((x * x).sum(dim=(2, 3)) / n - mean * mean).clamp_min(0.0)
        return torch.sqrt(12.0 * var).reshape(-1)

    return torch.stack([width(alpha), width(beta)], dim=-1)


def hemisphere_quadrature(axis, num_polar: int, num_azimuthal: int, max_polar_angle: float = math.pi / 2,
                          device: _DeviceLike = None, dtype: torch.dtype = torch.float64, *,
                          tangent_axes: Optional[tuple] = None, subsamples: int = 16) -> Quadrature:
    """Product set on the cone of half-angle ``max_polar_angle`` about ``axis`` (angular_quadrature.py:309):
    Gauss-Legendre in cos(theta) on [cos(max), 1] times equal azimuths ``(j + 1/2) 2 pi / A`` in ``_local_frame``,
    flat index ``i A + j`` with i = 0 nearest the axis; the weights sum to ``2 pi (1 - cos(max))``. With
    ``tangent_axes = (u, v)`` the gnomonic cell widths of the finite-cone footprint are filled."""
    if not (0.0 < max_polar_angle <= math.pi):
        raise ValueError("max_polar_angle must be in (0, pi]")
    if num_polar < 1 or num_azimuthal < 1:
        raise ValueError("num_polar and num_azimuthal must be >= 1")
    if device is None and isinstance(axis, Tensor):
        device = axis.device
    a = _unit_vector(axis)
    cos_t, pw = _cone_polar(num_polar, max_polar_angle)
    step = 2.0 * math.pi / num_azimuthal
    phi = (torch.arange(num_azimuthal, dtype=torch.float64) + 0.5) * step
    e1, e2, _ = _local_frame(a)
    sin_t = torch.sqrt((1.0 - cos_t * cos_t).clamp_min(0.0))
    local = ((sin_t[:, None] * torch.cos(phi)[None, :])[..., None] * e1
             + (sin_t[:, None] * torch.sin(phi)[None, :])[..., None] * e2 + cos_t[:, None, None] * a)
    omega = local.reshape(-1, 3)
    omega = omega / torch.linalg.norm(omega, dim=-1, keepdim=True)
    weight = (pw[:, None] * step).expand(num_polar, num_azimuthal).reshape(-1)
    q = Quadrature.from_directions(omega.to(device=device, dtype=dtype), weight.to(device=device, dtype=dtype))
    if tangent_axes is None:
        return q
    u = _unit_vector(tangent_axes[0])
    u = u - (u @ a) * a
    u = u / torch.linalg.norm(u)
    v = _unit_vector(tangent_axes[1])
    v = v - (v @ a) * a - (v @ u) * u
    v = v / torch.linalg.norm(v)
    width = _cone_cell_widths(a, u, v, num_polar, num_azimuthal, max_polar_angle, subsamples)
    return Quadrature(omega=q.omega, weight=q.weight, octant=q.octant, cell_width=width.to(device=device, dtype=dtype))


def legendre_polynomials(L: int, x: Tensor) -> Tensor:
    """P_0..P_L at x by Bonnet's recurrence, [..., L+1]."""
    if L < 0:
        raise ValueError("L must be >= 0")
    out = [torch.ones_like(x)]
    if L >= 1:
        out.append(x.clone())
    for l in range(1, L):
        out.append(((2 * l + 1) * x * out[l] - l * out[l - 1]) / (l + 1))
    return torch.stack(out, dim=-1)


def num_spherical_harmonics(L: int) -> int:
    return (L + 1) * (L + 1)


def spherical_harmonic_degree_index(L: int, device: _DeviceLike = None) -> Tensor:
    """Degree l of every flat index n = l*l + k + l."""
    return torch.cat([torch.full((2 * l + 1,), l, dtype=torch.int64) for l in range(L + 1)]).to(device)


def real_spherical_harmonics(L: int, omega: Tensor) -> Tensor:
    """Real orthonormal SH at unit directions, polar axis +z, no Condon-Shortley phase, [..., (L+1)^2].

    ``Y_{l,k} = sqrt(2) K_l^|k| D_l^|k|(z) Re/Im((x + i y)^|k|)`` for k > 0 / k < 0 and ``K_l^0 P_l(z)`` for k = 0,
    with ``D_l^m = d^m P_l / dz^m`` (recurrence ``(l-m+1) D_{l+1}^m = (2l+1) z D_l^m - (l+m) D_{l-1}^m``,
    ``D_m^m = (2m-1)!!``) and ``K_l^m = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!)``.
    """
    if L < 0:
        raise ValueError("L must be >= 0")
    x, y, z = omega[..., 0], omega[..., 1], omega[..., 2]
    re = [torch.ones_like(z)]
    im = [torch.zeros_like(z)]
    for m in range(1, L + 1):
        re.append(x * re[m - 1] - y * im[m - 1])
        im.append(x * im[m - 1] + y * re[m - 1])
    D = {}
    for m in range(L + 1):
        D[(m, m)] = torch.full_like(z, float(math.prod(range(1, 2 * m, 2))) if m > 0 else 1.0)
        if m + 1 <= L:
            D[(m + 1, m)] = (2 * m + 1) * z * D[(m, m)]
        for l in range(m + 1, L):
            D[(l + 1, m)] = ((2 * l + 1) * z * D[(l, m)] - (l + m) * D[(l - 1, m)]) / (l - m + 1)
    cols = []
    for l in range(L + 1):
        for k in range(-l, l + 1):
            m = abs(k)
            K = math.sqrt((2 * l + 1) / FOUR_PI * math.factorial(l - m) / math.factorial(l + m))
            if k == 0:
                cols.append(K * D[(l, 0)])
            else:
                cols.append(math.sqrt(2.0) * K * D[(l, m)] * (re[m] if k > 0 else im[m]))
    return torch.stack(cols, dim=-1)


@dataclass(frozen=True, eq=False)
class SphericalHarmonicBasis:
    """SH tables on a quadrature: ``Y [M, N]``, ``Y_weighted = w_m Y [M, N]``, ``degree_of_moment [N]``."""

    L: int
    Y: Tensor
    Y_weighted: Tensor
    degree_of_moment: Tensor

    @property
    def num_moments(self) -> int:
        return int(self.Y.shape[1])

    @property
    def num_directions(self) -> int:
        return int(self.Y.shape[0])


def spherical_harmonic_basis(L: int, quadrature: Quadrature) -> SphericalHarmonicBasis:
    Y = real_spherical_harmonics(L, quadrature.omega)
    return SphericalHarmonicBasis(L=L, Y=Y, Y_weighted=quadrature.weight[:, None] * Y,
                                  degree_of_moment=spherical_harmonic_degree_index(L, quadrature.omega.device))


def moments_from_ordinates(psi: Tensor, basis: SphericalHarmonicBasis) -> Tensor:
    """``phi_n = sum_m w_m Y_n(Omega_m) psi_m``: [..., M] -> [..., N]."""
    return psi @ basis.Y_weighted.to(psi)


def ordinates_from_moments(phi: Tensor, basis: SphericalHarmonicBasis) -> Tensor:
    """``psi_m = sum_n phi_n Y_n(Omega_m)``: [..., N] -> [..., M]."""
    return phi @ basis.Y.to(phi).T


def scatter_source_at_directions(T: Tensor, Y_selected: Tensor) -> Tensor:
    """``S[k, g, ...] = sum_n Y_selected[k, n] T[g, n, ...]``."""
    if T.shape[1] != Y_selected.shape[1]:
        raise ValueError(f"T has {T.shape[1]} moments, Y_selected has {Y_selected.shape[1]}")
    return torch.tensordot(Y_selected.to(T), T, dims=([1], [1]))
