"""Oracle group transmission and uncollided fluence (numpy float64). Test infrastructure only.

Follows ``citrus/scatter_solver/cross_sections.py:797-878`` (``T_g = sum_{e in g} w_e exp(-sum_m mu_m(e) P_m -
mu_f(e) P_f)``), ``citrus/scatter_solver/uncollided.py:160-197`` (``phi_UC = b / |r - r0|^2 T_g``) and
``citrus/scatter_solver/source.py:289-446`` / ``:484-510`` (panel point of the ray, collimated field with the
off-focal fraction past its edge, bilinear emission profile and filter thickness, constant beyond the outermost pixel
centres). Pinned by the vacuum ``f_g / r^2`` and polychromatic Beer's-law closed forms of the reference's tests
(tests/test_uncollided.py:69-137).
"""

from __future__ import annotations

import numpy as np

from .rays import panel_frame, source_position


def group_transmission(paths, fine_mu, fine_group, fine_weight, G, filter_path=None, filter_mu=None):
    """[G, N] from paths [M, N] in g/cm^2 and the fine tables ([M, E], [E], [E])."""
    tau = np.einsum("me,mn->en", fine_mu, paths)
    if filter_path is not None:
        tau = tau + np.outer(filter_mu, filter_path)
    t = fine_weight[:, None] * np.exp(-tau)
    out = np.zeros((G, paths.shape[1]))
    np.add.at(out, fine_group, t)
    return out


def panel_uv(d, sid, sdd, angle, curved):
    """(forward [N], u [N], v [N]) of the rays from the focal spot along d [N, 3]."""
    src = source_position(sid, angle)
    pc, uh, vh, nh = panel_frame(sid, sdd, angle)
    if not curved:
        t = ((pc - src) @ nh) / (d @ nh)
    else:
        a = d[:, 0] ** 2 + d[:, 1] ** 2
        with np.errstate(divide="ignore"):
            t = np.where(a > 0, sdd / np.sqrt(np.where(a > 0, a, 1.0)), np.inf)
    forward = np.isfinite(t) & (t > 0)
    hit = src + np.where(forward, t, 0.0)[:, None] * d
    if not curved:
        q = hit - pc
        return forward, q @ uh, q @ vh
    q = hit - src
    depth = -(q @ nh)
    return forward, sdd * np.arctan2(q @ uh, depth), sdd * (q @ vh) / depth


def bilinear(img, u, v, dch, drow, off_ch=0.0, off_row=0.0):
    rows, cols = img.shape

    def bracket(x, n):
        x = np.clip(x, 0.0, n - 1.0)
        i0 = np.clip(np.floor(x).astype(np.int64), 0, max(n - 2, 0))
        i1 = np.minimum(i0 + 1, n - 1)
        return i0, i1, np.clip(x - i0, 0.0, 1.0)

    c0, c1, wc = bracket((u + off_ch) / dch + (cols - 1) / 2.0, cols)
    r0, r1, wr = bracket((v + off_row) / drow + (rows - 1) / 2.0, rows)
    top = img[r0, c0] + (img[r0, c1] - img[r0, c0]) * wc
    bot = img[r1, c0] + (img[r1, c1] - img[r1, c0]) * wc
    return top + (bot - top) * wr


def field_weight(u, v, rows, cols, drow, dch, margin_ch=0.0, margin_row=0.0, off_dist=None, off_frac=None):
    hu, hv = cols * dch / 2, rows * drow / 2
    ou, ov = np.abs(u), np.abs(v)
    inside = (ou <= hu + margin_ch) & (ov <= hv + margin_row)
    if off_dist is None:
        return inside.astype(np.float64)
    pu = np.maximum(ou - (hu + margin_ch), 0.0)
    pv = np.maximum(ov - (hv + margin_row), 0.0)
    dist = np.sqrt(pu * pu + pv * pv)
    frac = np.interp(dist, off_dist, off_frac, left=off_frac[0], right=0.0)
    frac = np.where(dist > off_dist[-1], 0.0, frac)
    return np.where(inside, 1.0, frac)


def uncollided_fluence(paths, points, fine_mu, fine_group, fine_weight, G, sid, sdd, rows, cols, drow, dch, angle,
                       curved, profile=None, thickness=None, filter_density=0.0, filter_mu=None, margin_ch=0.0,
                       margin_row=0.0, off_dist=None, off_frac=None):
    """[G, N] uncollided fluence at points [N, 3] from the upstream paths [M, N]."""
    d = points - source_position(sid, angle)
    forward, u, v = panel_uv(d, sid, sdd, angle, curved)
    b = forward * field_weight(u, v, rows, cols, drow, dch, margin_ch, margin_row, off_dist, off_frac)
    if profile is not None:
        b = b * bilinear(profile, u, v, dch, drow)
    fp = None
    if thickness is not None:
        fp = bilinear(thickness, u, v, dch, drow) * filter_density
    T = group_transmission(paths, fine_mu, fine_group, fine_weight, G, fp, filter_mu)
    return T * (b / (d * d).sum(1))[None]
