"""Oracle Siddon / Amanatides-Woo traversal in float64 (numpy). Test infrastructure only.

Follows the reference's fused traversal ``_trace`` (``citrus/scatter_solver/kernels/ray_tracing_kernels.py:216-327``)
and its notes (``docs/source/design_notes.rst:394-430``): slab clipping of the segment ``p0 -> p1`` to the grid box
(``a_lo = max(0, entry)``, ``a_hi = min(1, exit)``), entry voxel ``floor`` (``ceil - 1`` along a descending axis)
clamped to the grid, stepping to the nearest next plane with ties broken in the order x, y, z, plane parameters
recomputed from the voxel index at every step. Axes are (x, y, z) = (cols, rows, slices); flat voxel index
``(row * C + col) * S + slice``. Every operation is a single IEEE float64 operation (no fused multiply-add), so a GPU
traversal written with the same operations reproduces the voxel sequence bit for bit.

Pinned by the closed-form chords of a uniform box (slab method), brute-force midpoint sampling, and the edge cases of
the reference's tests (tests/test_ray_tracing.py:20-48, :307-331).
"""

from __future__ import annotations

import numpy as np

BIG = 1e30


def box(lo, delta, n):
    """(lo [3], hi [3], delta [3], n (nx, ny, nz)) in x, y, z order."""
    lo = np.asarray(lo, dtype=np.float64)
    delta = np.asarray(delta, dtype=np.float64)
    hi = lo + delta * np.asarray(n, dtype=np.float64)
    return lo, hi, delta, tuple(int(v) for v in n)


def traverse(p0, p1, lo, delta, n, max_steps=None):
    """Traverse rays; returns (flat index [N, K] (-1 padded), parameter length [N, K], |d| [N]).

    ``length * |d|`` is the chord length in the voxel; only entries with length > 0 are kept, in order.
    """
    lo, hi, delta, n = box(lo, delta, n)
    p0 = np.asarray(p0, dtype=np.float64).reshape(-1, 3)
    p1 = np.asarray(p1, dtype=np.float64).reshape(-1, 3)
    N = p0.shape[0]
    d = p1 - p0
    nz = d != 0.0
    inv = 1.0 / np.where(nz, d, 1.0)
    a1 = (lo - p0) * inv
    a2 = (hi - p0) * inv
    inside = (p0 >= lo) & (p0 <= hi)
    amin = np.where(nz, np.minimum(a1, a2), np.where(inside, -BIG, BIG))
    amax = np.where(nz, np.maximum(a1, a2), np.where(inside, BIG, -BIG))
    a_lo = np.maximum(np.maximum(np.maximum(amin[:, 0], amin[:, 1]), amin[:, 2]), 0.0)
    a_hi = np.minimum(np.minimum(np.minimum(amax[:, 0], amax[:, 1]), amax[:, 2]), 1.0)
    act = a_hi > a_lo
    a_lo = np.where(act, a_lo, 0.0)
    a_hi = np.where(act, a_hi, 0.0)
    nvec = np.asarray(n, dtype=np.float64)
    f = (p0 + a_lo[:, None] * d - lo) / delta
    idx = np.where(d < 0, np.ceil(f) - 1.0, np.floor(f))
    idx = np.minimum(np.maximum(idx, 0.0), nvec - 1.0).astype(np.int64)
    pos = d > 0.0
    step = np.where(pos, 1, np.where(d < 0.0, -1, 0))

    def plane(i, axis):
        k = (i + pos[:, axis].astype(np.int64)).astype(np.float64)
        return np.where(nz[:, axis], (lo[axis] + k * delta[axis] - p0[:, axis]) * inv[:, axis], BIG)

    an = [plane(idx[:, a], a) for a in range(3)]
    steps = max_steps if max_steps is not None else sum(n) + 3
    out_idx = np.full((N, steps), -1, dtype=np.int64)
    out_len = np.zeros((N, steps), dtype=np.float64)
    count = np.zeros(N, dtype=np.int64)
    a = a_lo.copy()
    R, C, S = n[1], n[0], n[2]
    rows = np.arange(N)
    for _ in range(steps):
        if not act.any():
            break
        nxt = np.minimum(np.minimum(an[0], an[1]), np.minimum(an[2], a_hi))
        seg = np.where(act, np.maximum(nxt - a, 0.0), 0.0)
        inb = act & (idx[:, 0] >= 0) & (idx[:, 0] < n[0]) & (idx[:, 1] >= 0) & (idx[:, 1] < n[1]) \
            & (idx[:, 2] >= 0) & (idx[:, 2] < n[2])
        keep = inb & (seg > 0.0)
        flat = (idx[:, 1] * C + idx[:, 0]) * S + idx[:, 2]
        sel = rows[keep]
        out_idx[sel, count[sel]] = flat[sel]
        out_len[sel, count[sel]] = seg[sel]
        count[sel] += 1
        sx = (an[0] <= an[1]) & (an[0] <= an[2])
        sy = (an[1] < an[0]) & (an[1] <= an[2])
        sz = (an[2] < an[0]) & (an[2] < an[1])
        for axis, s in ((0, sx), (1, sy), (2, sz)):
            idx[:, axis] += np.where(s, step[:, axis], 0)
            an[axis] = np.where(s, plane(idx[:, axis], axis), an[axis])
        act = act & (nxt < a_hi)
        a = nxt
    del R
    width = max(1, int(count.max()) if N else 1)
    return out_idx[:, :width], out_len[:, :width], np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2 + d[:, 2] ** 2)


def line_integrals(volume, p0, p1, lo, delta, n):
    """[G, N] integrals of ``volume [G, R, C, S]`` along the segments (float64)."""
    idx, seg, length = traverse(p0, p1, lo, delta, n)
    G = volume.shape[0]
    flat = np.asarray(volume, dtype=np.float64).reshape(G, -1)
    vals = np.where(idx[None] >= 0, flat[:, np.maximum(idx, 0)], 0.0)  # [G, N, K]
    return np.einsum("gnk,nk->gn", vals, seg) * length[None]


def box_chord(p0, p1, lo, hi):
    """Length of the segment p0 -> p1 inside the axis-aligned box [lo, hi] by the slab method (closed form)."""
    p0 = np.asarray(p0, dtype=np.float64)
    p1 = np.asarray(p1, dtype=np.float64)
    d = p1 - p0
    t0, t1 = 0.0, 1.0
    for a in range(3):
        if d[a] == 0:
            if p0[a] < lo[a] or p0[a] > hi[a]:
                return 0.0
            continue
        ta, tb = (lo[a] - p0[a]) / d[a], (hi[a] - p0[a]) / d[a]
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
    return max(0.0, t1 - t0) * float(np.linalg.norm(d))


# ---- cone-beam geometry (citrus/scatter_solver/geometry.py:531-705), float64 ----------------------------------------

def source_position(sid, angle, z=0.0):
    return np.array([np.sin(angle) * sid, np.cos(angle) * sid, z])


def panel_frame(sid, sdd, angle, z=0.0):
    """(panel centre, u_hat, v_hat, n_hat) of a view; n_hat points toward the source."""
    c, s = np.cos(angle), np.sin(angle)
    yv = sid - sdd
    return (np.array([s * yv, c * yv, z]), np.array([c, -s, 0.0]), np.array([0.0, 0.0, 1.0]),
            np.array([s, c, 0.0]))


def pixel_centers(sid, sdd, rows, cols, drow, dch, angle, curved, sub=1, off_row=0.0, off_ch=0.0):
    """[rows, cols, sub, sub, 3] (sub-)pixel centres: sub-pixel (sv, su) on the midpoint lattice of the pixel."""
    pc, uh, vh, nh = panel_frame(sid, sdd, angle)
    u = (np.arange(cols) - (cols - 1) / 2) * dch - off_ch
    v = (np.arange(rows) - (rows - 1) / 2) * drow - off_row
    o = (2 * np.arange(sub) + 1) / (2 * sub) - 0.5
    U = u[None, :, None, None] + (o * dch)[None, None, None, :]
    V = v[:, None, None, None] + (o * drow)[None, None, :, None]
    U, V = np.broadcast_arrays(U, V)
    if not curved:
        return pc + U[..., None] * uh + V[..., None] * vh
    g = U / sdd
    return (pc + (sdd * np.sin(g))[..., None] * uh + (2 * sdd * np.sin(0.5 * g) ** 2)[..., None] * nh
            + (V * np.cos(g))[..., None] * vh)


def panel_distance(points, omega, sid, sdd, angle, curved):
    """t with points + t omega on the panel (flat plane; larger cylinder root for a curved panel)."""
    pc, uh, vh, nh = panel_frame(sid, sdd, angle)
    if not curved:
        return ((pc - points) @ nh) / (omega @ nh)
    d = (points - source_position(sid, angle))[..., :2]
    w = omega[:2]
    a = w @ w
    b = d @ w
    c = (d * d).sum(-1) - sdd ** 2
    disc = b * b - a * c
    root = np.sqrt(np.maximum(disc, 0.0))
    den = np.where(b >= 0, b + root, a if a > 0 else 1.0)
    num = np.where(b >= 0, -c, root - b)
    ok = (a > 0) & (disc >= 0) & (den != 0)
    return np.where(ok, num / np.where(den != 0, den, 1.0), np.inf)


def oblique_segments(pixels, omega, lo, hi):
    """Segments through each pixel centre along omega, centred on the foot of the box centre (ray_tracing.py:592)."""
    center = lo + 0.5 * (hi - lo)
    half = 0.5 * np.sqrt(((hi - lo) ** 2).sum()) * (1 + 1e-6) + 1e-6
    w2 = omega @ omega
    cp = ((center - pixels) @ omega) / w2
    h = half / np.sqrt(w2)
    return pixels + (cp - h)[:, None] * omega, pixels + (cp + h)[:, None] * omega
