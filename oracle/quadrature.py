"""Oracle direction sets and real spherical harmonics (numpy float64). Test infrastructure only.

Follows ``citrus/scatter_solver/angular_quadrature.py``: ``octant_of`` (:47-59), ``product_quadrature`` (:187-225),
the flat SH index ``n = l*l + k + l`` and the real orthonormal harmonics without Condon-Shortley phase (:475-521).
The level-symmetric S_N sets have no reference counterpart (the reference only has product sets); they follow the
published LQ_N construction (Lewis & Miller, Computational Methods of Neutron Transport, table 4-1).
"""

from __future__ import annotations

import math

import numpy as np
import scipy.special


def octant_of(omega: np.ndarray) -> np.ndarray:
    """Octant code ``(w_x<0)*4 + (w_y<0)*2 + (w_z<0)``, zero counting as non-negative (angular_quadrature.py:47)."""
    neg = (np.asarray(omega) < 0).astype(np.int64)
    return neg[..., 0] * 4 + neg[..., 1] * 2 + neg[..., 2]


def by_octant(omega: np.ndarray) -> np.ndarray:
    """Stable permutation listing octant 0 first, original order kept within an octant (angular_quadrature.py:121)."""
    return np.argsort(octant_of(omega), kind="stable")


def product_quadrature(num_polar: int, num_azimuthal: int) -> tuple[np.ndarray, np.ndarray]:
    """Gauss-Legendre in cos(theta) times equal azimuths (angular_quadrature.py:187-225); m = i*A + j."""
    mu, gw = np.polynomial.legendre.leggauss(num_polar)
    step = 2.0 * math.pi / num_azimuthal
    phi = (np.arange(num_azimuthal) + 0.5) * step
    st = np.sqrt(np.clip(1.0 - mu * mu, 0.0, None))
    omega = np.stack([st[:, None] * np.cos(phi)[None, :], st[:, None] * np.sin(phi)[None, :],
                      np.broadcast_to(mu[:, None], (num_polar, num_azimuthal))], axis=-1).reshape(-1, 3)
    weight = np.broadcast_to((gw * step)[:, None], (num_polar, num_azimuthal)).reshape(-1).copy()
    return omega, weight


# Published LQ_N first cosines and point weights (sum 1 per octant), Lewis & Miller table 4-1.
_LQN = {
    2: (0.5773503, [1.0]),
    4: (0.3500212, [0.3333333]),
    6: (0.2666355, [0.1761263, 0.1572071]),
    8: (0.2182179, [0.1209877, 0.0907407, 0.0925926]),
    12: (0.1672126, [0.0707626, 0.0558811, 0.0373377, 0.0502819, 0.0258513]),
}
# weight class of each sorted level triple (levels 1-based, i + j + k = N/2 + 2)
_LQN_CLASSES = {
    2: {(1, 1, 1): 0},
    4: {(2, 1, 1): 0},
    6: {(3, 1, 1): 0, (2, 2, 1): 1},
    8: {(4, 1, 1): 0, (3, 2, 1): 1, (2, 2, 2): 2},
    12: {(6, 1, 1): 0, (5, 2, 1): 1, (4, 3, 1): 2, (4, 2, 2): 3, (3, 3, 2): 4},
}


def level_symmetric_quadrature(order: int) -> tuple[np.ndarray, np.ndarray]:
    """Level-symmetric S_N set: N(N+2) directions, weights summing to 4 pi.

    Cosines ``mu_i^2 = mu_1^2 + (i-1) 2 (1 - 3 mu_1^2) / (N - 2)``; octant 0 lists the level triples (i, j, k) with
    i + j + k = N/2 + 2 in lexicographic order of (i, j) descending; the other octants are its reflections, in octant
    code order.
    """
    if order not in _LQN:
        raise ValueError(f"no level-symmetric set of order {order}")
    mu1, w = _LQN[order]
    half = order // 2
    if order == 2:
        mu = np.array([mu1])
    else:
        delta = 2.0 * (1.0 - 3.0 * mu1 * mu1) / (order - 2)
        mu = np.sqrt(mu1 * mu1 + delta * np.arange(half))
    w = np.asarray(w, dtype=np.float64)
    w = w / w[[_LQN_CLASSES[order][tuple(sorted(t, reverse=True))] for t in _triples(half)]].sum()
    points, weights = [], []
    for (i, j, k) in _triples(half):
        points.append((mu[i - 1], mu[j - 1], mu[k - 1]))
        weights.append(w[_LQN_CLASSES[order][tuple(sorted((i, j, k), reverse=True))]])
    base = np.array(points)
    base = base / np.linalg.norm(base, axis=1, keepdims=True)
    wb = np.array(weights) * (4.0 * math.pi / 8.0)
    omega, weight = [], []
    for o in range(8):
        sign = np.array([-1.0 if o & 4 else 1.0, -1.0 if o & 2 else 1.0, -1.0 if o & 1 else 1.0])
        omega.append(base * sign)
        weight.append(wb)
    return np.concatenate(omega), np.concatenate(weight)


def _triples(half: int) -> list[tuple[int, int, int]]:
    """Level triples of one octant of an S_N set with N = 2 half."""
    total = half + 2
    return [(i, j, total - i - j) for i in range(half, 0, -1) for j in range(half, 0, -1) if total - i - j >= 1]


def sh_index(L: int) -> list[tuple[int, int]]:
    """(l, k) of every flat index n = l*l + k + l (angular_quadrature.py:462-472)."""
    return [(l, k) for l in range(L + 1) for k in range(-l, l + 1)]


def real_spherical_harmonics(L: int, omega: np.ndarray) -> np.ndarray:
    """Real orthonormal SH without Condon-Shortley phase at unit directions, polar axis +z (angular_quadrature.py:475).

    Independent formulation: built from scipy's complex harmonics (which carry the phase (-1)^k, removed here).
    """
    omega = np.asarray(omega, dtype=np.float64)
    x, y, z = omega[..., 0], omega[..., 1], omega[..., 2]
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    phi = np.arctan2(y, x)
    cols = []
    for l, k in sh_index(L):
        m = abs(k)
        ylm = scipy.special.sph_harm_y(l, m, theta, phi)
        if k == 0:
            cols.append(ylm.real)
        elif k > 0:
            cols.append(math.sqrt(2.0) * (-1) ** m * ylm.real)
        else:
            cols.append(math.sqrt(2.0) * (-1) ** m * ylm.imag)
    return np.stack(cols, axis=-1)


def legendre(L: int, x: np.ndarray) -> np.ndarray:
    """P_0..P_L at x (scipy), [..., L+1]."""
    return np.stack([scipy.special.eval_legendre(l, x) for l in range(L + 1)], axis=-1)
