"""Octant signs and the cell-level diagonal-plane plan of a grid (host-side, for inspection and tests).

Keeps ``OCTANT_SIGNS``, ``SweepPlan`` and ``build_sweep_plan`` of ``citrus/scatter_solver/wavefront.py:16-125``:
entries of all 8 octants sorted by (plane, octant) with a stable order, upstream neighbours along x, y, z or -1 for
vacuum. The HIP sweep itself runs a block-level schedule (``citrus_sweep_plan_create``) built natively from the same
plane definition.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch
from torch import Tensor

OCTANT_SIGNS = torch.tensor([[-1 if o & 4 else 1, -1 if o & 2 else 1, -1 if o & 1 else 1] for o in range(8)],
                            dtype=torch.int64)


@dataclass(frozen=True)
class SweepPlan:
    """Entries (octant, cell) of every diagonal plane with their upstream neighbours (wavefront.py:23-72)."""

    shape: tuple[int, int, int]
    num_planes: int
    plane_offsets: Tensor
    octant: Tensor
    cell: Tensor
    upstream: Tensor

    @property
    def num_cells(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]

    def to(self, device) -> "SweepPlan":
        return SweepPlan(self.shape, self.num_planes, self.plane_offsets.to(device), self.octant.to(device),
                         self.cell.to(device), self.upstream.to(device))

    def plane_entries(self, p: int) -> slice:
        return slice(int(self.plane_offsets[p]), int(self.plane_offsets[p + 1]))


def build_sweep_plan(shape: tuple[int, int, int], device=None) -> SweepPlan:
    """Plan of all 8 octants; plane of cell (i, j, k) for octant o is the sum of its coordinates mirrored into the
    octant's sweep direction (wavefront.py:75-125)."""
    R, C, S = (int(v) for v in shape)
    i, j, k = np.meshgrid(np.arange(R), np.arange(C), np.arange(S), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    flat = (i * C + j) * S + k
    octs, cells, ups, planes = [], [], [], []
    for o in range(8):
        sx, sy, sz = (int(v) for v in OCTANT_SIGNS[o])
        planes.append((j if sx > 0 else C - 1 - j) + (i if sy > 0 else R - 1 - i) + (k if sz > 0 else S - 1 - k))
        jn, in_, kn = j - sx, i - sy, k - sz
        ups.append(np.stack([np.where((jn >= 0) & (jn < C), (i * C + jn) * S + k, -1),
                             np.where((in_ >= 0) & (in_ < R), (in_ * C + j) * S + k, -1),
                             np.where((kn >= 0) & (kn < S), (i * C + j) * S + kn, -1)], axis=1))
        octs.append(np.full(flat.size, o))
        cells.append(flat)
    octant, cell, upstream, plane = (np.concatenate(v) for v in (octs, cells, ups, planes))
    order = np.argsort(plane * 8 + octant, kind="stable")
    num_planes = R + C + S - 2
    offsets = np.zeros(num_planes + 1, dtype=np.int64)
    offsets[1:] = np.cumsum(np.bincount(plane, minlength=num_planes))
    plan = SweepPlan((R, C, S), num_planes, torch.from_numpy(offsets), torch.from_numpy(octant[order]),
                     torch.from_numpy(cell[order]), torch.from_numpy(upstream[order]))
    return plan.to(device) if device is not None else plan
