"""Independent closed-form checks that pin the oracle sweep (numpy float64). Test infrastructure only.

* ``slab_1d`` — textbook 1-D LD slab, a 2x2 solve per cell along the direction of travel
  (the idea of ``tests/oracles/linear_discontinuous_slab_1d.py:15-51``, restated).
* ``dense_global_solve`` — global upwind LD system of one direction and group assembled cell by cell from
  ``ld.cell_matrix`` plus the upwind couplings of ``method.rst:254-279`` and solved with ``numpy.linalg.solve``
  (the idea of ``tests/test_transport_sweep.py:182-233``).
* ``outflow_balance`` — particle balance: outflow through boundary faces plus absorption equals the source
  (``tests/test_transport_sweep.py:142-175``).
"""

from __future__ import annotations

import numpy as np

from .ld import cell_matrix


def slab_1d(mu, S0, S1, dx, mu_x):
    """1-D LD slab sweep; returns (psi0 [K], psi1 [K]) in the axis coordinate."""
    mu, S0, S1 = (np.asarray(v, dtype=np.float64) for v in (mu, S0, S1))
    K = mu.shape[0]
    s = 1.0 if mu_x > 0 else -1.0
    c = abs(mu_x) / dx
    p0 = np.zeros(K)
    p1 = np.zeros(K)
    inflow = 0.0
    for k in (range(K) if s > 0 else range(K - 1, -1, -1)):
        A = np.array([[c + mu[k], c], [-3.0 * c, 3.0 * c + mu[k]]])
        b = np.array([S0[k] + c * inflow, s * S1[k] - 3.0 * c * inflow])
        x0, x1 = np.linalg.solve(A, b)
        p0[k], p1[k] = x0, s * x1
        inflow = x0 + x1
    return p0, p1


def dense_global_solve(omega, deltas, mu, source):
    """Solve the global LD system of one direction. mu [R, C, S], source [4, R, C, S] -> psi [4, R, C, S]."""
    R, C, S = mu.shape
    N = R * C * S
    omega = np.asarray(omega, dtype=np.float64)
    s = np.where(omega < 0, -1.0, 1.0)
    c = np.abs(omega) / np.asarray(deltas, dtype=np.float64)
    A = np.zeros((4 * N, 4 * N))
    b = np.asarray(source, dtype=np.float64).reshape(4, N).T.reshape(-1).copy()

    def flat(i, j, k):
        return (i * C + j) * S + k

    for i in range(R):
        for j in range(C):
            for k in range(S):
                n = flat(i, j, k)
                A[4 * n:4 * n + 4, 4 * n:4 * n + 4] = cell_matrix(c, s, float(mu[i, j, k]))
                for a, (ii, jj, kk) in enumerate([(i, j - int(s[0]), k), (i - int(s[1]), j, k),
                                                   (i, j, k - int(s[2]))]):
                    if not (0 <= ii < R and 0 <= jj < C and 0 <= kk < S):
                        continue
                    m = flat(ii, jj, kk)
                    # inflow psi_0[m] + s_a psi_a[m] enters R_0 with +c_a and R_a with -3 s_a c_a
                    A[4 * n, 4 * m] -= c[a]
                    A[4 * n, 4 * m + 1 + a] -= c[a] * s[a]
                    A[4 * n + 1 + a, 4 * m] += 3.0 * s[a] * c[a]
                    A[4 * n + 1 + a, 4 * m + 1 + a] += 3.0 * c[a]
                    for bb in range(3):  # transverse: R_b += c_a psi_b[m]
                        if bb != a:
                            A[4 * n + 1 + bb, 4 * m + 1 + bb] -= c[a]
    x = np.linalg.solve(A, b)
    return x.reshape(N, 4).T.reshape(4, R, C, S)


def outflow_balance(psi_m, omega, deltas, mu, source_m):
    """(outflow + absorption, source) per group for one direction; psi_m, source_m [G, 4, R, C, S], mu [G, R, C, S]."""
    dx, dy, dz = deltas
    V = dx * dy * dz
    out = np.zeros(psi_m.shape[0])
    dims = {0: 3, 1: 2, 2: 4}  # axis -> dim of [G, 4, R, C, S]: x cols, y rows, z slices
    for a in range(3):
        s = 1.0 if omega[a] >= 0 else -1.0
        face = psi_m[:, 0] + s * psi_m[:, 1 + a]  # [G, R, C, S]
        idx = -1 if s > 0 else 0
        fb = np.take(face, idx, axis=dims[a] - 1)
        out += abs(omega[a]) * (V / deltas[a]) * fb.reshape(fb.shape[0], -1).sum(axis=1)
    lhs = out + V * (mu * psi_m[:, 0]).reshape(mu.shape[0], -1).sum(axis=1)
    rhs = V * source_m[:, 0].reshape(mu.shape[0], -1).sum(axis=1)
    return lhs, rhs
