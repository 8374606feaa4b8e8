"""Oracle linear-discontinuous (LD) cell equations (numpy float64). Test infrastructure only.

Follows ``citrus/scatter_solver/linear_discontinuous_cell.py`` (:14-107) and ``docs/source/method.rst:254-279``:
moments (0, x, y, z) per cell; ``c_a = |Omega_a| / Delta_a``, ``s_a = sign(Omega_a)`` (+1 for 0).
"""

from __future__ import annotations

import numpy as np


def cell_matrix(c: np.ndarray, s: np.ndarray, mu: float) -> np.ndarray:
    """Dense 4x4 LD cell matrix in moment order (0, x, y, z) (linear_discontinuous_cell.py:90-107)."""
    A = np.zeros((4, 4))
    csum = float(np.sum(c))
    A[0, 0] = csum + mu
    for a in range(3):
        A[0, 1 + a] = s[a] * c[a]
        A[1 + a, 0] = -3.0 * s[a] * c[a]
        A[1 + a, 1 + a] = 3.0 * c[a] + (csum - c[a]) + mu
    return A


def right_hand_sides(c, s, source, up):
    """R_0 and R_a from the source moments [..., 4] and the upstream neighbours' moments ``up[..., a, 4]``.

    ``up[..., a, :]`` holds the four moments of the neighbour along axis a (zero for vacuum)
    (linear_discontinuous_cell.py:14-56): ``R_0 = S_0 + sum_a c_a pb_a``, ``pb_a = up_a0 + s_a up_aa``,
    ``R_a = S_a - 3 s_a c_a pb_a + sum_{b != a} c_b up_{b,a}``.
    """
    pb = up[..., :, 0] + s * np.stack([up[..., a, 1 + a] for a in range(3)], axis=-1)  # [..., 3]
    r0 = source[..., 0] + np.sum(c * pb, axis=-1)
    ra = []
    for a in range(3):
        trans = sum(c[..., b] * up[..., b, 1 + a] for b in range(3) if b != a)
        ra.append(source[..., 1 + a] - 3.0 * s[..., a] * c[..., a] * pb[..., a] + trans)
    return r0, np.stack(ra, axis=-1)


def cell_solve(c, s, mu, r0, ra):
    """Closed-form solve (linear_discontinuous_cell.py:59-87): ``d_a = 3c_a + sum_{b!=a} c_b + mu``,
    ``psi_0 = (R_0 - sum s_a c_a R_a / d_a) / (sum c + mu + 3 sum c_a^2 / d_a)``, ``psi_a = (R_a + 3 s_a c_a psi_0)/d_a``.
    Returns [..., 4]."""
    csum = np.sum(c, axis=-1, keepdims=True)
    mu = np.asarray(mu)[..., None]
    d = 3.0 * c + (csum - c) + mu
    q = c / d
    den = csum[..., 0] + mu[..., 0] + 3.0 * np.sum(c * q, axis=-1)
    p0 = (r0 - np.sum(s * q * ra, axis=-1)) / den
    pa = (ra + 3.0 * s * c * p0[..., None]) / d
    return np.concatenate([p0[..., None], pa], axis=-1)
