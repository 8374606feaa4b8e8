"""Oracle: conservative resampling between voxel grids by counting common sub-cells (test infrastructure only).

The reference's ``resample_conservative`` (citrus/scatter_solver/geometry.py:304-327) gives every target cell the
exact average of a voxel-wise constant field over its volume, zero outside the source box. When every edge of both
grids lies on one fine lattice, that average is a plain count: split space into the lattice's sub-cells, give each
the value of the source voxel holding its midpoint (zero outside), and average the sub-cells whose midpoints fall in
the target cell. No overlap formula is involved.
"""

from __future__ import annotations

import numpy as np


def _owner(edges: np.ndarray, points: np.ndarray) -> np.ndarray:
    """Index of the cell of ascending ``edges`` holding each point, -1 outside."""
    i = np.searchsorted(edges, points, side="right") - 1
    return np.where((points > edges[0]) & (points < edges[-1]), i, -1)


def resample_by_lattice(field: np.ndarray, edges_from, edges_to, step: float) -> np.ndarray:
    """[..., R', C', S'] averages on the target edges (rows y, cols x, slices z) of ``field [..., R, C, S]`` on the
    source edges, both sets of edges on multiples of ``step`` apart from a common origin."""
    f = np.asarray(field, dtype=np.float64)
    axes_from = [np.asarray(e, dtype=np.float64) for e in edges_from]
    axes_to = [np.asarray(e, dtype=np.float64) for e in edges_to]
    lo = min(min(a[0] for a in axes_from), min(a[0] for a in axes_to))
    out = f
    for axis in range(3):
        ef, et = axes_from[axis], axes_to[axis]
        hi = max(ef[-1], et[-1])
        n = int(round((hi - lo) / step))
        mid = lo + (np.arange(n) + 0.5) * step
        src, dst = _owner(ef, mid), _owner(et, mid)
        nt = et.size - 1
        weight = np.zeros((nt, ef.size - 1))
        counts = np.zeros(nt)
        for s, d in zip(src, dst):
            if d >= 0:
                counts[d] += 1
                if s >= 0:
                    weight[d, s] += 1
        weight /= counts[:, None]
        out = np.moveaxis(np.tensordot(weight, np.moveaxis(out, out.ndim - 3 + axis, 0), axes=(1, 0)), 0,
                          out.ndim - 3 + axis)
    return out
