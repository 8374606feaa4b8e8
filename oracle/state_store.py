"""Oracle of the compact per-view accumulator (numpy). Test infrastructure only.

Follows ``StateStore.put`` / ``StateStore.get`` (``citrus/citrus.py:686-737``): the linear-discontinuous cell mean of
every (group, moment) is stored as float16 of ``mean / scale`` with ``scale = max|mean| / 2^14`` over the cells (1
where the maximum is 0), computed in float32; ``get`` returns ``float16 * scale`` in float32 with zero slopes.
Pinned in ``tests/test_oracle.py`` by what the format fixes: the largest magnitude maps to exactly 2^14, zero and
all-zero moments round-trip exactly, values the float16 grid holds round-trip exactly, and every other value comes
back within half a float16 step (2^-11 relative to 2^14 scale units).
"""

from __future__ import annotations

import numpy as np

FLOAT16_TARGET = np.float32(2.0 ** 14)  # citrus.py:620


def pack(cell_mean: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Cell means [G, nsh, *cells] (float32) -> (q [G, nsh, *cells] float16, scale [G, nsh] float32)."""
    m = np.asarray(cell_mean, dtype=np.float32)
    G, nsh = m.shape[:2]
    flat = m.reshape(G, nsh, -1)
    maximum = np.abs(flat).max(axis=2)
    scale = np.where(maximum > 0, maximum / FLOAT16_TARGET, np.float32(1.0)).astype(np.float32)
    q = (flat / scale[:, :, None]).astype(np.float32).astype(np.float16)
    return q.reshape(m.shape), scale


def unpack(q: np.ndarray, scale: np.ndarray) -> np.ndarray:
    """(q [G, nsh, *cells] float16, scale [G, nsh]) -> cell means [G, nsh, *cells] float32."""
    q32 = np.asarray(q).astype(np.float32)
    s = np.asarray(scale, dtype=np.float32).reshape(scale.shape + (1,) * (q32.ndim - 2))
    return (q32 * s).astype(np.float32)
