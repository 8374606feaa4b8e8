"""Oracle: direct transmission of a 1-D anti-scatter grid by brute force (test infrastructure only).

The direct part of the reference's grid model (citrus/scatter_solver/grid_model.py:270-311, the photons that cross
the grid without interacting), restated as the plain average it is: over entry offsets x across one strip pitch p
(midpoint rule) of exp(-(mu_s L_s(x) + mu_i (L - L_s(x)))), with the strip path L_s(x) found by sampling the
photon's track through the grid height (midpoint rule along the track) and testing each sample against the strips
[k p, k p + t). It shares no formula with the closed form it checks; its error is that of the two midpoint rules.
"""

from __future__ import annotations

import math

import numpy as np


def direct_transmission(pitch: float, thickness: float, height: float, mu_strip, mu_interspace, theta1, theta2,
                        n_offsets: int = 4000, n_track: int = 4000) -> np.ndarray:
    """[E, T1, T2] float64 transmission averaged over the entry offset (grid_model.py:270)."""
    ms = np.asarray(mu_strip, dtype=np.float64).reshape(-1)
    mi = np.asarray(mu_interspace, dtype=np.float64).reshape(-1)
    t1 = np.abs(np.asarray(theta1, dtype=np.float64).reshape(-1))
    t2 = np.abs(np.asarray(theta2, dtype=np.float64).reshape(-1))
    x = (np.arange(n_offsets) + 0.5) / n_offsets * pitch
    z = (np.arange(n_track) + 0.5) / n_track
    out = np.empty((ms.size, t1.size, t2.size))
    for a, th1 in enumerate(t1):
        shift = height * math.tan(th1)
        pos = x[:, None] + shift * z[None, :]  # [offsets, track] position across the strips
        share = (np.mod(pos, pitch) < thickness).mean(axis=1)  # [offsets] share of the track inside the strips
        for b, th2 in enumerate(t2):
            L = height * math.sqrt(1.0 + math.tan(th1) ** 2 + math.tan(th2) ** 2)
            for e in range(ms.size):
                out[e, a, b] = np.exp(-(ms[e] * share + mi[e] * (1.0 - share)) * L).mean()
    return out


def lamella_paths(pitch: float, thickness: float, height: float, theta1: float,
                  n_offsets: int = 20000, n_track: int = 4000) -> np.ndarray:
    """[n_offsets] lead path (cm) of a track tilted by theta1 from the focusing direction through the grid height,
    for entry offsets x across one pitch (midpoint rule), by sampling the track (midpoint rule) against the lamellae
    [k p, k p + t). The septum-shadow model of the reference's focused_1d_grid (detector.py:779-829) reduces the
    distribution to two numbers: the share of offsets whose path is zero (the open share) and the mean path over
    the others; both follow from these samples without its formulas."""
    x = (np.arange(n_offsets) + 0.5) / n_offsets * pitch
    z = (np.arange(n_track) + 0.5) / n_track
    pos = x[:, None] + height * math.tan(abs(theta1)) * z[None, :]
    share = (np.mod(pos, pitch) < thickness).mean(axis=1)
    return share * height / math.cos(theta1)
