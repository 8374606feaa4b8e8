"""Oracle scattering-source moments and source iteration (numpy float64). Test infrastructure only.

Follows ``citrus/scatter_solver/angular_quadrature.py:600-677`` (``T[g, n] = sum_g' mu_s_l[g', g, l(n)] phi[g', n]``,
no (2l+1)/4pi factor, ``method.rst:191-245``) and ``citrus/scatter_solver/source_iteration.py:126-204`` (collision-order
loop; stop when ``||phi_n[:, 0, 0]|| / ||phi_total[:, 0, 0]|| < tol``).
"""

from __future__ import annotations

import math

import numpy as np

from .quadrature import sh_index
from .sweep import moments, sources_from_moments, sweep


def degree_of_moment(nsh: int) -> np.ndarray:
    """Degree l of every flat SH index (angular_quadrature.py:462)."""
    L = math.isqrt(nsh) - 1
    return np.array([l for l, _ in sh_index(L)])


def scatter_moments(mu_s_l: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """T[go, n, ...] = sum_gi mu_s_l[gi, go, l(n), *spatial] phi[gi, n, ...]; phi is [G, N, 4, *sp] or [G, N, *sp]."""
    Gi, Go = mu_s_l.shape[:2]
    nsh = phi.shape[1]
    deg = degree_of_moment(nsh)
    ld_axis = phi.ndim == mu_s_l.ndim  # [G, N, 4, R, C, S] vs transfer [Gi, Go, L+1, R, C, S]
    T = np.zeros((Go,) + phi.shape[1:], dtype=np.float64)
    for gi in range(Gi):
        for go in range(Go):
            rows = mu_s_l[gi, go][deg]  # [N, *sp]
            if ld_axis:
                rows = rows[:, None]
            T[go] += rows * phi[gi]
    return T


def source_iteration(shape, deltas, omega, weight, Y, mu_t, mu_s_l, fc_source, tol, max_iterations):
    """Collision-order source iteration from the first-collision source ``fc_source [M, G, 4, *sp]``.

    Returns (phi_total [G, N, 4, *sp], history list, num_iterations, converged).
    """
    phi_total = None
    history = []
    converged = False
    src = fc_source
    n = 0
    for n in range(1, max_iterations + 1):
        psi = sweep(shape, deltas, omega, mu_t, src)
        phi_n = moments(psi, weight, Y)
        phi_total = phi_n.copy() if phi_total is None else phi_total + phi_n
        norm_n = float(np.linalg.norm(phi_n[:, 0, 0]))
        norm_t = float(np.linalg.norm(phi_total[:, 0, 0]))
        rel = norm_n / norm_t if norm_t > 0 else 0.0
        history.append(rel)
        if rel < tol:
            converged = True
            break
        if n == max_iterations:
            break
        src = sources_from_moments(scatter_moments(mu_s_l, phi_n), Y)
    return phi_total, history, n, converged
