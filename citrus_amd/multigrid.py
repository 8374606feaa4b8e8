"""Two-stage (coarse, fine) schedule of the CITRUS view steps, with the two stages' views in one sweep.

A view's first ``coarse_steps`` steps run on a coarse transport grid -- the same box in cells ``f`` times larger per
axis, so a sweep costs about ``f^3`` times less -- where the accumulator of the collided moments picks up its
low-order shape; then the accumulator is prolonged onto the fine grid and the view continues there. In a scan whose
views started at different times (a helical scan fed view by view), one outer step holds views of both stages: their
accumulator updates run in ONE persistent sweep over the two grids (``citrus.citrus_grids_step``).

There is no counterpart in the reference, which solves every view on one covering grid
(citrus/scatter_solver/scatter_solver.py:495-507); each stage's step is the reference's ``citrus_view_step``
(citrus/citrus.py:320-366) on that stage's grid.
"""

from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Sequence

import torch
from torch import Tensor

from .citrus import CitrusState, citrus_grids_step
from .scatter_solver.geometry import Grid
from .scatter_solver.scatter_solver import ScatterSolver


def refinement(coarse: Grid, fine: Grid) -> tuple[int, int, int]:
    """Integer refinement (f_rows, f_cols, f_slices) of ``coarse`` into ``fine`` over the same box; ValueError if
    the grids are not nested."""
    f = []
    (cx, cy, cz), (fx, fy, fz) = coarse.deltas, fine.deltas  # deltas are (x, y, z); shapes (rows y, cols x, slices z)
    for nc, nf, dc, df in zip(coarse.shape, fine.shape, (cy, cx, cz), (fy, fx, fz)):
        k = nf // nc
        if k < 1 or k * nc != nf or abs(dc - k * df) > 1e-9 * dc:
            raise ValueError(f"grids are not nested: {coarse.shape} at {coarse.deltas} vs {fine.shape} at "
                             f"{fine.deltas}")
        f.append(k)
    if any(abs(a - b) > 1e-9 for a, b in zip(coarse.lower_corner(), fine.lower_corner())):
        raise ValueError("grids do not cover the same box")
    return tuple(f)


def prolong_moments(phi: Tensor, coarse: Grid, fine: Grid) -> Tensor:
    """LD moments ``[..., 4, Rc, Cc, Sc]`` (order [0, x, y, z]; x along cols, y rows, z slices) of a coarse grid on a
    nested fine grid: each child cell takes the parent's linear expansion ``m0 + sum_a m_a u_a`` (u in [-1, 1]^3 over
    the parent) -- its mean is the expansion at the child centre ``u_a = (2 j_a + 1) / f_a - 1``, its slopes
    ``m_a / f_a``. Exact for fields linear in each coarse cell; the children's mean is the parent's mean."""
    fr, fc, fs = refinement(coarse, fine)
    Rc, Cc, Sc = coarse.shape
    if tuple(phi.shape[-4:]) != (4, Rc, Cc, Sc):
        raise ValueError(f"moments must end in [4, {Rc}, {Cc}, {Sc}], got {tuple(phi.shape)}")
    lead = phi.shape[:-4]
    p = phi.reshape(-1, 4, Rc, 1, Cc, 1, Sc, 1)

    def centres(f):
        return (2.0 * torch.arange(f, device=phi.device, dtype=phi.dtype) + 1.0) / f - 1.0

    uy = centres(fr).reshape(1, fr, 1, 1, 1, 1)
    ux = centres(fc).reshape(1, 1, 1, fc, 1, 1)
    uz = centres(fs).reshape(1, 1, 1, 1, 1, fs)
    m0, mx, my, mz = p[:, 0], p[:, 1], p[:, 2], p[:, 3]
    out = torch.empty(p.shape[0], 4, Rc, fr, Cc, fc, Sc, fs, device=phi.device, dtype=phi.dtype)
    out[:, 0] = m0 + mx * ux + my * uy + mz * uz
    out[:, 1] = (mx / fc).expand_as(out[:, 1])
    out[:, 2] = (my / fr).expand_as(out[:, 2])
    out[:, 3] = (mz / fs).expand_as(out[:, 3])
    return out.reshape(*lead, 4, Rc * fr, Cc * fc, Sc * fs)


@dataclass
class MultigridSchedule:
    """The two stages: ``coarse`` and ``fine`` solvers (the same library, sweep quadrature, Legendre order and
    geometry; the coarse transport cell a whole multiple of the fine one) and the steps a view spends on the coarse
    grid."""

    coarse: ScatterSolver
    fine: ScatterSolver
    coarse_steps: int = 2

    def __post_init__(self):
        if isinstance(self.coarse_steps, bool) or not isinstance(self.coarse_steps, int) or self.coarse_steps < 0:
            raise ValueError(f"coarse_steps must be an integer >= 0, got {self.coarse_steps!r}")
        if self.coarse.num_groups != self.fine.num_groups:
            raise ValueError("the two stages must have the same energy groups")

    def stage(self, state: CitrusState) -> str:
        """'coarse' for a view that has taken fewer than ``coarse_steps`` steps, else 'fine'."""
        return "coarse" if state.num_steps < self.coarse_steps else "fine"


def _slab(solver: ScatterSolver, grid_fine: Grid, angle: float) -> tuple[int, int]:
    """Fine-voxel slices [k0, k1) of the view's beam, in whole transport cells of ``solver``."""
    from .scatter_solver.cropping import beam_slice_range

    cell = solver.transport_grid(grid_fine)
    return beam_slice_range(solver.geometry, angle, grid_fine, max(1, round(cell.deltas[2] / grid_fine.deltas[2])))


def _at_switch(schedule: MultigridSchedule, state: CitrusState) -> bool:
    """Whether the state's accumulator still lives on the coarse grid although the view is in the fine stage."""
    return 0 < schedule.coarse_steps == state.num_steps and state.collided_moments is not None


def _prolonged(schedule: MultigridSchedule, state: CitrusState, rho_fine: Tensor, grid_fine: Grid, angle: float,
               crop: bool, fine_slab) -> CitrusState:
    """The state of a view entering the fine stage, its accumulator prolonged (and, cropped, cut to the fine slab)."""
    from .scatter_solver.cropping import crop_slices

    if crop:
        k0c, k1c = _slab(schedule.coarse, grid_fine, angle)
        box = crop_slices(grid_fine, rho_fine, k0c, k1c)[0]
    else:
        k0c, box = 0, grid_fine
    gc, gf = schedule.coarse.transport_grid(box), schedule.fine.transport_grid(box)
    phi = state.collided_moments
    if tuple(phi.shape[-3:]) != tuple(gc.shape):
        raise ValueError(f"the accumulator {tuple(phi.shape)} is not on the coarse stage's grid {gc.shape}")
    phi = prolong_moments(phi, gc, gf)
    if crop:
        per = round(gf.deltas[2] / grid_fine.deltas[2])  # fine voxels per fine transport cell along z
        k0f, k1f = fine_slab
        phi = phi[..., (k0f - k0c) // per:(k1f - k0c) // per].contiguous()
    return replace(state, collided_moments=phi, scatter_source=None, scatter_source_fields=None, _batch=None)


def citrus_multigrid_views_step(schedule: MultigridSchedule, states: Sequence[CitrusState], rho_fine: Tensor,
                                grid_fine: Grid, angles: Sequence[float], *, crop: bool = False,
                                primary: bool = True, contraction: str = "fp32", sweep_schedule: str = "auto"):
    """One CITRUS step of every view, each at its stage: views still in the coarse stage step on the coarse solver's
    transport grid, the others on the fine one; a view entering the fine stage has its accumulator prolonged
    (``prolong_moments``). With ``crop`` every view steps on its stage's grid cropped to the view's beam (helical
    scans, ``cropping.beam_slice_range``). Every view's accumulator update runs in one sweep over all these grids.
    Returns ([(scatter, primary)], [new state], [(stage, slab or None)]) per view; a state belongs to its stage's
    (cropped) grid."""
    from .scatter_solver.cropping import crop_slices

    if len(states) != len(angles):
        raise ValueError(f"need one angle per state, got {len(angles)} angles for {len(states)} states")
    stages = [schedule.stage(s) for s in states]
    tags, groups = [], {}
    for i, stage in enumerate(stages):
        solver = schedule.coarse if stage == "coarse" else schedule.fine
        slab = _slab(solver, grid_fine, angles[i]) if crop else None
        tags.append((stage, slab))
        groups.setdefault((stage, slab), []).append(i)
    batches, order = [], []
    for (stage, slab), idx in sorted(groups.items(), key=lambda kv: (kv[0][0], kv[0][1] or (0, 0))):
        solver = schedule.coarse if stage == "coarse" else schedule.fine
        if slab is None:
            fields = solver.material_fields(rho_fine, grid_fine)
        else:
            sub, rho_sub = crop_slices(grid_fine, rho_fine, *slab)
            fields = solver.material_fields(rho_sub, sub)
        sts = []
        for i in idx:
            st = states[i]
            if stage == "fine" and _at_switch(schedule, st):
                st = _prolonged(schedule, st, rho_fine, grid_fine, angles[i], crop, slab)
            sts.append(st)
        batches.append((solver, sts, fields, [angles[i] for i in idx]))
        order.append(idx)
    res, sts = citrus_grids_step(batches, primary=primary, contraction=contraction, schedule=sweep_schedule)
    results: list = [None] * len(states)
    new_states: list = [None] * len(states)
    for idx, r, s in zip(order, res, sts):
        for j, i in enumerate(idx):
            results[i], new_states[i] = r[j], s[j]
    return results, new_states, tags
