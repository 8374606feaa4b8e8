"""Oracle collapsed-cone scatter image and primary image (numpy float64). Test infrastructure only.

Restates ``citrus/scatter_solver/scatter_image.py``:

- ``box_matrix`` (:952-1007): ``B[c, j] = |[j, j+1] n W_c| / |W_c|``, W_c the window of width w centred on pixel c
  clipped to [0, N];
- ``supervoxel_subpoint_emission`` (:829-875): ``e_j = S_j exp(-max(0, tau_cell - mu_cell w . (r_j - r_cell)))``;
- the finite-cone footprint (:1033-1138): widths ``hbar_k W_k / pitch``, hbar the emission-weighted mean panel height
  ``(r - panel_centre) . n_hat`` (the box-centre height where the emission vanishes);
- ``combined_scatter_image`` (:1141-1323): per direction the downstream optical depth from the cell centres, the
  multiple-collision source ``Y(w) T``, the attenuated total emission at the quadrature points, oblique line
  integrals through the pixel centres (ray_tracing.py:548-620), the blur, and the weighted sums fluence
  ``sum_k w_k I_k``, planar fluence ``sum_k w_k cos(theta0_k) I_k`` and signal ``sum_k s_kg I_k``;
- ``primary_from_path_lengths`` (:1419-1471) for a flat panel without filter or profile and ``scatter_fraction``
  (:1509).

Pinned in tests/test_oracle.py by: the single-voxel bookkeeping of the reference's test_scatter_image.py:131-167
(``sum_pixels A I = sum_k w_k E V / cos(theta_k)`` for a pixel-aligned cube in vacuum, planar ``sum_k w_k E V``);
convergence to a brute-force point-to-point volume integral as the cone set is refined (:75-127); the box matrices'
row sums, identity at width <= 1 and hand-computed overlaps; and the vacuum ``f_g / r^2`` and uniform-slab Beer's law
closed forms of the primary.
"""

from __future__ import annotations

import numpy as np

from . import rays


def box_matrix(N: int, w) -> np.ndarray:
    """[k, N, N] footprint filter matrices."""
    w = np.maximum(np.asarray(w, dtype=np.float64).reshape(-1), 1e-3)
    centers = np.arange(N) + 0.5
    half = 0.5 * w[:, None]
    lo = np.clip(centers[None] - half, 0.0, N)[..., None]
    hi = np.clip(centers[None] + half, 0.0, N)[..., None]
    j = np.arange(N)[None, None, :]
    overlap = np.maximum(np.minimum(hi, j + 1.0) - np.maximum(lo, j), 0.0)
    return overlap / (hi - lo)


def blur(image, wc, wr):
    """[k, G, Rd, Cd] images blurred along the channels then the rows."""
    k, G, Rd, Cd = image.shape
    bc = box_matrix(Cd, wc)
    br = box_matrix(Rd, wr)
    x = np.einsum("kgrj,kcj->kgrc", image, bc)
    return np.einsum("kri,kgic->kgrc", br, x)


def downstream_tau(mu, lo, delta, n, points, omega, sid, sdd, angle, curved):
    """[G, P] optical depth from points [P, 3] to the panel along omega (+inf if it does not reach the panel)."""
    _, _, _, nh = rays.panel_frame(sid, sdd, angle)
    G = mu.shape[0]
    if omega @ (-nh) <= 0:
        return np.full((G, points.shape[0]), np.inf)
    t = rays.panel_distance(points, omega, sid, sdd, angle, curved)
    t = np.broadcast_to(t, (points.shape[0],))
    return rays.line_integrals(mu, points, points + t[:, None] * omega, lo, delta, n)


def oblique_project(field, lo, delta, n, pixels, omega):
    """[G, Rd, Cd] line integrals of ``field [G, R, C, S]`` through the pixel centres along omega."""
    Rd, Cd = pixels.shape[:2]
    flat = pixels.reshape(-1, 3)
    hi = lo + delta * np.array([n[0], n[1], n[2]], dtype=np.float64)
    p0, p1 = rays.oblique_segments(flat, omega, lo, hi)
    return rays.line_integrals(field, p0, p1, lo, delta, n).reshape(field.shape[0], Rd, Cd)


def subcell_offsets(omega, deltas, f):
    """[f, f, f] offsets ``w . (r_j - r_cell)`` (axes rows, cols, slices)."""
    dx, dy, dz = deltas
    c = np.arange(f) - (f - 1) / 2.0
    return (omega[1] * dy * c[:, None, None] + omega[0] * dx * c[None, :, None] + omega[2] * dz * c[None, None, :])


def attenuated_emission(S_fc, S_mc, tau, mu_c, omega, deltas_q, f, coarse_shape):
    """[G, Rq, Cq, Sq] emission at the quadrature points of one direction.

    S_fc [G, Rq, Cq, Sq], S_mc and tau [G, Rc, Cc, Sc], mu_c [G, Rc, Cc, Sc]."""
    Rc, Cc, Sc = coarse_shape
    G = S_fc.shape[0]

    def up(x):
        return x.repeat(f, axis=1).repeat(f, axis=2).repeat(f, axis=3)

    if f == 1:
        t = tau
    else:
        off = subcell_offsets(omega, deltas_q, f)  # [f, f, f]
        offq = np.tile(off, (Rc, Cc, Sc))  # [Rq, Cq, Sq] (the pattern repeats per cell)
        t = np.maximum(up(tau) - up(mu_c) * offq[None], 0.0)
    return (S_fc + up(S_mc)) * np.exp(-(up(tau) if f == 1 else t))


def collapsed_cone_image(S_fc_fn, T, Y, mu_c, lo, deltas_q, shape_q, f, omegas, weights, cell_width, pixels,
                         pitch, heights, center_height, signal_weight, cos0, sid, sdd, angle, curved,
                         footprint=True, planar_weight=None, signal_pixel_weight=None):
    """(fluence, planar, signal_per_group) [G, Rd, Cd] of the total scatter of one view.

    S_fc_fn(k) -> [G, Rq, Cq, Sq] first-collision source at direction k; T [G, N_sh, Rc, Cc, Sc] cell-mean moments;
    Y [K, N_sh]; mu_c [G, Rc, Cc, Sc]; omegas [K, 3], weights [K], cell_width [K, 2]; pixels [Rd, Cd, 3]; pitch
    (channel, row) in cm; heights [Rq, Cq, Sq]; signal_weight [K, G] (flat) or [K, G, Cd] (curved, with
    planar_weight [K, Cd]); signal_pixel_weight(k) -> [G, Rd, Cd] an extra factor of the signal of direction k at
    each pixel (the transmission of an angle-dependent grid, scatter_image.py:373-390), applied to the blurred
    image."""
    G = T.shape[0]
    Rq, Cq, Sq = shape_q
    coarse = (Rq // f, Cq // f, Sq // f)
    dc = np.asarray(deltas_q, dtype=np.float64) * f
    centres = (lo + dc * (np.stack(np.meshgrid(np.arange(coarse[1]), np.arange(coarse[0]), np.arange(coarse[2]),
                                               indexing="ij"), -1).transpose(1, 0, 2, 3) + 0.5))
    # centres[r, c, s] = (x of col c, y of row r, z of slice s)
    pts = centres.reshape(-1, 3)
    n_c = (coarse[1], coarse[0], coarse[2])  # traversal order (cols, rows, slices) of rays.traverse
    n_q = (Cq, Rq, Sq)
    Rd, Cd = pixels.shape[:2]
    fl = np.zeros((G, Rd, Cd))
    pl = np.zeros((G, Rd, Cd))
    sg = np.zeros((G, Rd, Cd))
    for k, w in enumerate(np.asarray(omegas, dtype=np.float64)):
        tau = downstream_tau(mu_c, lo, dc, n_c, pts, w, sid, sdd, angle, curved).reshape(G, *coarse)
        S_mc = np.einsum("n,gnrcs->grcs", Y[k], T)
        e = attenuated_emission(S_fc_fn(k), S_mc, tau, mu_c, w, deltas_q, f, coarse)
        img = oblique_project(e, lo, np.asarray(deltas_q, dtype=np.float64), n_q, pixels, w)[None]  # [1, G, Rd, Cd]
        if footprint:
            tot = e.sum(0)
            s0 = tot.sum()
            hbar = (tot * heights).sum() / s0 if s0 > 0 else center_height
            wc = hbar * cell_width[k, 0] / pitch[0]
            wr = hbar * cell_width[k, 1] / pitch[1]
            img = blur(img, [wc], [wr])
        img = img[0]
        fl += weights[k] * img
        extra = 1.0 if signal_pixel_weight is None else signal_pixel_weight(k)
        if planar_weight is None:
            pl += weights[k] * cos0[k] * img
            sg += signal_weight[k][:, None, None] * img * extra
        else:
            pl += planar_weight[k][None, None, :] * img
            sg += signal_weight[k][:, None, :] * img * extra
    return fl, pl, sg


def primary_flat(path, fine_mu, fine_group, fine_weight, G, pixels, src, nh, energy, response_fn):
    """Primary of a flat panel without filter or profile: (fluence, planar, signal_per_group) [G, Rd, Cd] from the
    line integrals path [M', Rd, Cd]; ``response_fn(energy [G], theta [Rd, Cd]) -> [G, Rd, Cd]``."""
    from .uncollided import group_transmission

    Rd, Cd = path.shape[1:]
    T = group_transmission(path.reshape(path.shape[0], -1), fine_mu, fine_group, fine_weight, G).reshape(G, Rd, Cd)
    d = pixels - src
    r2 = (d * d).sum(-1)
    fluence = T / r2[None]
    cos0 = (d @ (-nh)) / np.sqrt(r2)
    planar = fluence * cos0[None]
    theta = np.arccos(np.clip(cos0, -1.0, 1.0))
    return fluence, planar, planar * response_fn(energy, theta)


def scatter_fraction(S, P):
    den = P + S
    return np.where(den > 0, S / np.where(den > 0, den, 1.0), 0.0)


def detector_response(energy_nodes, theta_nodes, table, energy, theta) -> np.ndarray:
    """[G, ...] detector response D(E, theta) of an [E, T] table (detector.py:226-256): linear in energy, then linear
    in |theta|, each clamped to its end nodes (a bilinear lookup is separable, so two numpy.interp passes give it)."""
    en = np.asarray(energy_nodes, dtype=np.float64)
    tn = np.asarray(theta_nodes, dtype=np.float64)
    tab = np.asarray(table, dtype=np.float64)
    e = np.asarray(energy, dtype=np.float64).reshape(-1)
    th = np.abs(np.asarray(theta, dtype=np.float64))
    at_energy = np.stack([np.interp(e, en, tab[:, t]) for t in range(tn.size)], axis=1)  # [G, T]
    return np.stack([np.interp(th, tn, at_energy[g]) for g in range(e.size)])


def channel_normals(sid, sdd, cols, dch, angle, off_ch=0.0) -> np.ndarray:
    """[cols, 3] panel normal at each channel of a curved panel: horizontal, from the pixel toward the vertical axis
    through the source (conventions.rst, "Curved panel"); its negative is the inward normal of the incidence angle."""
    pix = rays.pixel_centers(sid, sdd, 1, cols, 1.0, dch, angle, True, off_ch=off_ch)[0, :, 0, 0]  # [cols, 3]
    to_axis = rays.source_position(sid, angle) - pix
    to_axis[:, 2] = 0.0
    return to_axis / np.linalg.norm(to_axis, axis=1, keepdims=True)


def signal_weights(omega, weight, sid, sdd, angle, curved, energy, detector, cols=None, dch=None):
    """The emission-independent weights of the collapsed-cone image of one view (make_direction_context,
    scatter_image.py:542-690, without a grid): the directions facing the panel (``omega . -n_hat > 0``), their
    ``cos(theta0)`` to the panel normal at its centre, and

    - flat panel: signal weights [K', G] = ``w cos(theta0) D_g(theta0)``, theta0 the angle of the normalized
      direction;
    - curved panel: planar weights [K', Cd] = ``w clamp(omega . -n_c, 0, 1)`` with the channel's own normal n_c, and
      signal weights [K', G, Cd] = planar ``D_g(theta_c)``.

    ``detector`` is (energy nodes, theta nodes, table). Returns (keep [K], cos0 [K'], signal, planar or None)."""
    om = np.asarray(omega, dtype=np.float64)
    w = np.asarray(weight, dtype=np.float64)
    _, _, _, nh = rays.panel_frame(sid, sdd, angle)
    cos_all = om @ (-nh)
    keep = cos_all > 0
    om, w, cos0 = om[keep], w[keep], cos_all[keep]
    if not curved:  # theta0 from the normalized direction (detector_angle, detector.py:1054-1075)
        theta0 = np.arccos(np.clip(cos0 / np.linalg.norm(om, axis=1), -1.0, 1.0))
        d = detector_response(*detector, energy, theta0)  # [G, K']
        return keep, cos0, (w * cos0)[:, None] * d.T, None
    normals = channel_normals(sid, sdd, cols, dch, angle)
    cos_c = np.clip(om @ (-normals).T, 0.0, 1.0)  # [K', Cd]
    planar = w[:, None] * cos_c
    d = detector_response(*detector, energy, np.arccos(cos_c))  # [G, K', Cd]
    return keep, cos0, planar[:, None, :] * d.transpose(1, 0, 2), planar
