"""Oracle of the interpolating density model (numpy float32, the reference's working dtype). Test infrastructure only.

Follows ``attenuation_to_densities`` (``citrus/utilities.py:710-753``): with x = max(X, 0) and the materials sorted by
nominal attenuation mu_s = rho_s^nom m_s, j = #{s: mu_s <= x}; inside the bracket (1 <= j <= M-1) the slots a = j-1
and b = j get (1 - f) rho_a^nom and f rho_b^nom with f = (x - mu_a) / (mu_b - mu_a); outside, slot a alone gets
x / m_a. Pinned in ``tests/test_oracle.py`` by what the model fixes: it reproduces the attenuation
(sum_m rho_m m_m = x), a nominal attenuation maps to that material at its nominal density, X <= 0 maps to nothing,
at most two adjacent materials are nonzero.
"""

from __future__ import annotations

import numpy as np


def attenuation_to_densities(X, order, mu, nominal, mass) -> np.ndarray:
    """X [...] float32 -> rho [M, ...] float32; ``order``, ``mu``, ``nominal``, ``mass`` are the [M] slot tables in
    bracketing order (float32, mu = nominal * mass)."""
    X = np.asarray(X, dtype=np.float32)
    mu = np.asarray(mu, dtype=np.float32)
    nominal = np.asarray(nominal, dtype=np.float32)
    mass = np.asarray(mass, dtype=np.float32)
    M = mu.shape[0]
    x = np.maximum(X, np.float32(0.0))
    j = (mu.reshape((M,) + (1,) * x.ndim) <= x[None]).sum(0)
    a = np.clip(j - 1, 0, M - 1)
    b = np.minimum(j, M - 1)
    inside = (j >= 1) & (j <= M - 1)
    with np.errstate(divide="ignore", invalid="ignore"):
        span = np.where(inside, mu[b] - mu[a], np.float32(1.0)).astype(np.float32)
        f = np.where(inside, (x - mu[a]) / span, np.float32(0.0)).astype(np.float32)
        rho_a = np.where(inside, (np.float32(1.0) - f) * nominal[a], x / mass[a]).astype(np.float32)
        rho_b = np.where(inside, f * nominal[b], np.float32(0.0)).astype(np.float32)
    out = np.zeros((M,) + x.shape, dtype=np.float32)
    order = np.asarray(order)
    for m in range(M):
        out[m] = np.where(order[a] == m, out[m] + rho_a, out[m])
        out[m] = np.where(order[b] == m, out[m] + rho_b, out[m])
    return out
