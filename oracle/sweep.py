"""Oracle LD discrete-ordinates wavefront sweep (numpy; float64 by default). Test infrastructure only.

Follows ``citrus/scatter_solver/wavefront.py:75-125`` (diagonal planes per octant, upstream neighbours, vacuum
boundaries), ``citrus/scatter_solver/transport_sweep.py:271-305`` (per-plane gather, closed-form cell solve,
scatter) and ``:524-533`` (moments ``phi[g, n, k] = sum_m w_m Y_n(Omega_m) psi_m[g, k]``). Flat cell index
``(i*C + j)*S + k`` with i rows (y), j cols (x), k slices (z); axis a = 0, 1, 2 is x, y, z.
"""

from __future__ import annotations

import numpy as np

from . import ld
from .quadrature import octant_of


def octant_signs(o: int) -> tuple[int, int, int]:
    """(s_x, s_y, s_z) of an octant code (wavefront.py:16)."""
    return (-1 if o & 4 else 1, -1 if o & 2 else 1, -1 if o & 1 else 1)


def octant_plan(shape, o: int):
    """Cells of every diagonal plane of octant o, with upstream neighbours along x, y, z (-1 = vacuum).

    Returns a list over planes of (cells [P], upstream [P, 3]) (wavefront.py:92-111).
    """
    R, C, S = shape
    sx, sy, sz = octant_signs(o)
    i, j, k = np.meshgrid(np.arange(R), np.arange(C), np.arange(S), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    flat = (i * C + j) * S + k
    plane = (j if sx > 0 else C - 1 - j) + (i if sy > 0 else R - 1 - i) + (k if sz > 0 else S - 1 - k)
    nj, ni, nk = j - sx, i - sy, k - sz
    ux = np.where((nj >= 0) & (nj < C), (i * C + nj) * S + k, -1)
    uy = np.where((ni >= 0) & (ni < R), (ni * C + j) * S + k, -1)
    uz = np.where((nk >= 0) & (nk < S), (i * C + j) * S + nk, -1)
    up = np.stack([ux, uy, uz], axis=1)
    out = []
    for p in range(R + C + S - 2):
        sel = plane == p
        out.append((flat[sel], up[sel]))
    return out


def sweep(shape, deltas, omega, mu, source, dtype=np.float64):
    """LD sweep of every direction and group with vacuum boundaries.

    Args:
        shape: (R, C, S).
        deltas: (dx, dy, dz) in cm.
        omega: [M, 3] directions (any order).
        mu: [G, R, C, S] total attenuation in 1/cm.
        source: [M, G, 4, R, C, S] LD source moments per direction.
        dtype: arithmetic dtype (float64 for the oracle, float32 for the CPU timing baseline).

    Returns:
        psi [M, G, 4, R, C, S].
    """
    R, C, S = shape
    N = R * C * S
    omega = np.asarray(omega, dtype=np.float64)
    M = omega.shape[0]
    G = mu.shape[0]
    mu_flat = np.asarray(mu, dtype=dtype).reshape(G, N)
    src = np.asarray(source, dtype=dtype).reshape(M, G, 4, N)
    psi = np.zeros((M, G, 4, N), dtype=dtype)
    octs = octant_of(omega)
    c_all = (np.abs(omega) / np.asarray(deltas, dtype=np.float64)).astype(dtype)
    s_all = np.where(omega < 0, -1.0, 1.0).astype(dtype)
    for o in range(8):
        ms = np.nonzero(octs == o)[0]
        if ms.size == 0:
            continue
        c = c_all[ms][:, None, None, :]  # [Mo, 1, 1, 3]
        s = s_all[ms][:, None, None, :]
        src_o = src[ms]
        psi_o = np.zeros((ms.size, G, 4, N + 1), dtype=dtype)  # cell N is the zero vacuum cell
        for cells, up in octant_plan(shape, o):
            upn = np.where(up < 0, N, up)  # [P, 3]
            nb = psi_o[:, :, :, upn].transpose(0, 1, 3, 4, 2)  # [Mo, G, 4, P, 3] -> [Mo, G, P, 3(a), 4]
            sc = src_o[:, :, :, cells].transpose(0, 1, 3, 2)  # [Mo, G, P, 4]
            r0, ra = ld.right_hand_sides(c, s, sc, nb)
            sol = ld.cell_solve(c, s, mu_flat[:, cells][None], r0, ra)  # [Mo, G, P, 4]
            psi_o[:, :, :, cells] = sol.transpose(0, 1, 3, 2)
        psi[ms] = psi_o[..., :N]
    return psi.reshape(M, G, 4, R, C, S)


def moments(psi, weight, Y):
    """phi[g, n, k, ...] = sum_m w_m Y[m, n] psi[m, g, k, ...] (transport_sweep.py:524-533)."""
    return np.einsum("m,mn,mgk...->gnk...", np.asarray(weight, dtype=np.float64), Y, psi)


def sources_from_moments(T, Y):
    """S[m, g, k, ...] = sum_n Y[m, n] T[g, n, k, ...] (source_iteration.py:92-102)."""
    return np.einsum("mn,gnk...->mgk...", Y, T)
