"""Oracle first-collision source, sampling correction and LD restriction (numpy float64). Test infrastructure only.

Restates:

- ``angular_table_position`` / ``angular_table_nodes`` (``citrus/scatter_solver/cross_sections.py:905-973``): the
  interpolation position of a cosine eta in a transfer table of N nodes, uniform in eta or through the adaptive
  map ``u(q)`` at ``q = sin(theta/2)``;
- ``transfer_eta_at`` (``cross_sections.py:976-995``): linear interpolation of ``mu_s_eta`` between nodes;
- ``primary_directions_to_voxels`` (``citrus/scatter_solver/uncollided.py:74-96``);
- ``first_collision_source_at_directions`` (``uncollided.py:200-270``): ``S_FC,g(r, w) = sum_{g' <= g} sum_m
  rho_m(r) mu_s,eta^{g'->g}(Omega_UC(r) . w) phi_UC,g'(r)``;
- ``first_collision_exact_moments`` (``uncollided.py:357-384``), ``first_collision_quadrature_sums`` (``:538-608``),
  ``first_collision_correction_factors`` (``:611-640``) and ``FirstCollisionSamplingCorrection.apply``
  (``:420-433``);
- ``restrict_to_linear_discontinuous_moments`` (``uncollided.py:311-352``).

Pinned in tests/test_oracle.py by: hand evaluation with ``np.interp`` (the reference's test_uncollided.py:181-213);
the closed form of the reference's synthetic kernel ``a (1 + b eta) / 4 pi``, which a uniform table interpolates
exactly, and its exact Legendre moments ``a`` and ``a b / 3`` against a brute-force sphere quadrature; table positions
landing on the nodes of an adaptive table; exact reproduction of linear fields by the restriction and conservation
of the fine-grid integral (test_uncollided.py:217-238); and the l = 0 / l = 1 conservation identities of the
corrected samples (test_scatter_solver.py:282-323).
"""

from __future__ import annotations

import numpy as np


def synthetic_angular_map(num_knots: int = 512) -> np.ndarray:
    """[K] adaptive node map ``u(q) = 1 - 0.6 log(1 + q/0.02)/log(51) - 0.4 q^2`` (cross_sections.py:1025-1040)."""
    q = np.arange(num_knots) / (num_knots - 1)
    u = 1.0 - 0.6 * np.log1p(q / 0.02) / np.log(51.0) - 0.4 * q * q
    u[0], u[-1] = 1.0, 0.0
    return u


def table_nodes(num_nodes: int, eta_map=None) -> np.ndarray:
    """[N] node cosines, ascending (cross_sections.py:948-973): uniform, or ``eta_j = 1 - 2 q_j^2`` with q_j the
    inverse of the piecewise-linear map at ``u = j/(N-1)``."""
    n = int(num_nodes)
    if eta_map is None:
        return np.linspace(-1.0, 1.0, n)
    uk = np.asarray(eta_map, dtype=np.float64)
    K = uk.size
    qk = np.arange(K) / (K - 1)
    au, aq = uk[::-1], qk[::-1]
    u = np.arange(n) / (n - 1)
    j = np.clip(np.searchsorted(au, u, side="right"), 1, K - 1)
    q = aq[j - 1] + (aq[j] - aq[j - 1]) * (u - au[j - 1]) / (au[j] - au[j - 1])
    eta = 1.0 - 2.0 * q * q
    eta[0], eta[-1] = -1.0, 1.0
    return eta


def table_position(eta, num_nodes: int, eta_map=None):
    """(i0 int64, w1) of cosines eta in a table of ``num_nodes`` nodes (cross_sections.py:905-945)."""
    n = int(num_nodes)
    eta = np.clip(np.asarray(eta, dtype=np.float64), -1.0, 1.0)
    if eta_map is None:
        pos = (eta + 1.0) / 2.0 * (n - 1)
    else:
        knots = np.asarray(eta_map, dtype=np.float64)
        K = knots.size
        q = np.sqrt(np.maximum((1.0 - eta) * 0.5, 0.0))
        p = q * (K - 1)
        k0 = np.clip(np.floor(p), 0.0, K - 2.0)
        a = np.clip(p - k0, 0.0, 1.0)
        k0 = k0.astype(np.int64)
        u = knots[k0] + (knots[k0 + 1] - knots[k0]) * a
        pos = u * (n - 1)
    i0 = np.clip(np.floor(pos).astype(np.int64), 0, n - 2)
    w1 = np.clip(pos - i0, 0.0, 1.0)
    return i0, w1


def transfer_at(mu_s_eta, eta_map, rho, eta):
    """[Gi, Go, P] per-steradian transfer ``sum_m rho_m interp(mu_s_eta[m], eta)`` (cross_sections.py:976-995).

    mu_s_eta [M, Gi, Go, N], rho [M, P], eta [P]."""
    i0, w1 = table_position(eta, mu_s_eta.shape[-1], eta_map)
    v0 = mu_s_eta[..., i0]
    v1 = mu_s_eta[..., i0 + 1]
    v = v0 + (v1 - v0) * w1
    return np.einsum("mabp,mp->abp", v, rho)


def primary_directions(points, src):
    """[P, 3] unit vectors from the focal spot ``src`` to ``points`` [P, 3] (uncollided.py:74-96)."""
    d = np.asarray(points, dtype=np.float64) - np.asarray(src, dtype=np.float64)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def source_at_directions(mu_s_eta, eta_map, rho, phi_uc, omega_uc, omegas):
    """[K, G, P] first-collision source, downscatter only (uncollided.py:236-270).

    rho [M, P], phi_uc [G, P], omega_uc [P, 3], omegas [K, 3]."""
    G = phi_uc.shape[0]
    down = np.triu(np.ones((G, G)))
    out = np.empty((omegas.shape[0], G, phi_uc.shape[1]))
    for k, w in enumerate(np.asarray(omegas, dtype=np.float64)):
        T = transfer_at(mu_s_eta, eta_map, rho, omega_uc @ w)  # [Gi, Go, P]
        out[k] = np.einsum("abp,ab,ap->bp", T, down, phi_uc)
    return out


def exact_moments(mu_s_l, rho, phi_uc, l_max: int = 1):
    """[l_max+1, G, P] ``M_l[g] = sum_{g' <= g} sum_m rho_m mu_s_l^{g'->g} phi_UC,g'`` (uncollided.py:357-384)."""
    G = phi_uc.shape[0]
    down = np.triu(np.ones((G, G)))
    return np.einsum("mabl,ab,mp,ap->lbp", mu_s_l[..., : l_max + 1], down, rho, phi_uc)


def quadrature_sums(S, omega_uc, omegas, weights):
    """[3, G, P] ``I_j = sum_k w_k (omega_k . Omega_UC)^j S_k`` for j = 0, 1, 2 (uncollided.py:538-608)."""
    mu = np.asarray(omegas, dtype=np.float64) @ omega_uc.T  # [K, P]
    w = np.asarray(weights, dtype=np.float64)[:, None, None]
    return np.stack([(w * S).sum(0), (w * mu[:, None] * S).sum(0), (w * mu[:, None] ** 2 * S).sum(0)])


def correction_factors(I, E):
    """(scale, beta) [G, P] from the sums and the exact moments (uncollided.py:611-640): ``c = E1/E0``,
    ``beta = (c I0 - I1)/(I2 - c I1)`` clamped to [-1, 1] (0 where E0 <= 0 or the denominator is below
    1e-12 |I0|), ``scale = E0/(I0 + beta I1)`` (1 where that denominator is not positive)."""
    I0, I1, I2 = I
    E0, E1 = E
    pos = E0 > 0
    c = np.where(pos, E1 / np.where(pos, E0, 1.0), 0.0)
    den = I2 - c * I1
    ok = pos & (np.abs(den) > 1e-12 * np.abs(I0))
    beta = np.clip(np.where(ok, (c * I0 - I1) / np.where(ok, den, 1.0), 0.0), -1.0, 1.0)
    d0 = I0 + beta * I1
    scale = np.where(d0 > 0, E0 / np.where(d0 > 0, d0, 1.0), 1.0)
    return scale, beta


def apply_correction(S, scale, beta, omega_uc, omegas):
    """[K, G, P] ``S scale (1 + beta omega_k . Omega_UC)`` (uncollided.py:420-433)."""
    mu = np.asarray(omegas, dtype=np.float64) @ omega_uc.T
    return S * scale[None] * (1.0 + beta[None] * mu[:, None])


def restrict_ld(field, factor: int):
    """[..., 4, R/f, C/f, S/f] LD moments [0, x, y, z] of a field [..., R, C, S] on the refined grid
    (uncollided.py:311-352): sub-point local coordinates ``u_j = (2j + 1)/f - 1``, mean and
    ``mean(S u_a)/mean(u_a^2)``; x runs along cols, y along rows, z along slices."""
    f = int(factor)
    field = np.asarray(field, dtype=np.float64)
    lead = field.shape[:-3]
    R, C, S = (n // f for n in field.shape[-3:])
    x = field.reshape(*lead, R, f, C, f, S, f)
    nl = len(lead)
    ar, ac, as_ = nl + 1, nl + 3, nl + 5
    mean = x.mean(axis=(ar, ac, as_))
    if f == 1:
        z = np.zeros_like(mean)
        return np.stack([mean, z, z, z], axis=-4)
    u = (2.0 * np.arange(f) + 1.0) / f - 1.0
    scale = 1.0 / (u * u).mean()

    def slope(axis):
        shape = [1] * x.ndim
        shape[axis] = f
        return (x * u.reshape(shape)).mean(axis=(ar, ac, as_)) * scale

    return np.stack([mean, slope(ac), slope(ar), slope(as_)], axis=-4)
