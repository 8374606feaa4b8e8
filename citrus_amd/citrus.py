"""The CITRUS scatter step of one view on the HIP kernels.

Keeps the interface of ``citrus/citrus.py:70-366``: ``CitrusState`` (the caller-owned accumulator Phi of the collided
moments of a view), ``uncollided_operator``, ``first_collision_operator``, ``accumulator_update_operator``
(``Phi_new = phi^(1) + K Phi`` in one sweep of the summed first-collision and scattering sources, plus optional extra
sweeps ``Phi_new <- Phi_new + K (Phi_new - Phi_old)``), ``project_scatter``, ``citrus_step`` and
``citrus_view_step`` (the uncollided fluence, the accumulator update, the signal-only scatter image and the primary).
With the densities held fixed, N steps from an empty state equal N source iterations. ``citrus_views_step`` runs the
steps of several views that share the densities with their accumulator updates batched in one sweep.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import torch
from torch import Tensor

from .scatter_solver.angular_quadrature import Quadrature
from .scatter_solver.geometry import Grid
from .scatter_solver.scatter_solver import MaterialFields, ScatterSolver
from .scatter_solver.source_iteration import (MixedOperand, scatter_moments_mixed, scatter_moments_views,
                                              summed_source_function)
from .scatter_solver.transport_sweep import sweep_views
from .scatter_solver.uncollided import DeferredFirstCollision, first_collision_deferred_samples


@dataclass
class CitrusState:
    """Caller-owned state of one view (citrus.py:70): angle, accumulator Phi [G, (L+1)^2, 4, Rc, Cc, Sc] (None is
    zero), cached cone set, step count and the relative change of Phi's scalar-flux cell means per step.
    ``citrus_views_step`` also keeps the scattering moments T = K Phi it formed for the image ([G, Rc, Cc, Sc,
    (L+1)^2, 4]) with the fields they belong to: the next step reuses them only for those same fields."""

    angle: Optional[float] = None
    collided_moments: Optional[Tensor] = None
    cone_quadrature: Optional[Quadrature] = None
    num_steps: int = 0
    history: list = field(default_factory=list)
    scatter_source: Optional[Tensor] = field(default=None, repr=False)
    scatter_source_fields: Optional[object] = field(default=None, repr=False)
    learned_residual: Optional[float] = None  # the judging sweep's relative change when the learned form started it
    _batch: Optional[tuple] = field(default=None, repr=False)


def _fields(solver: ScatterSolver, rho, grid: Grid, fields: Optional[MaterialFields]) -> MaterialFields:
    return solver.material_fields(rho, grid) if fields is None else fields


def _check_accumulator(solver: ScatterSolver, fields: MaterialFields, phi: Tensor) -> None:
    ctx = solver.sweep_context(fields.grid_coarse)
    want = (solver.num_groups, ctx.num_moments, 4) + tuple(fields.grid_coarse.shape)
    if tuple(phi.shape) != want:
        raise ValueError(f"collided moments must be {want}, got {tuple(phi.shape)}")


def _check_count(name: str, count) -> None:
    if isinstance(count, bool) or not isinstance(count, int) or count < 1:
        raise ValueError(f"{name} must be an integer >= 1, got {count!r}")


def uncollided_operator(solver: ScatterSolver, rho, grid: Grid, angle: float,
                        fields: Optional[MaterialFields] = None) -> Tensor:
    """[G, Rq, Cq, Sq] uncollided fluence G_1 at the quadrature points (citrus.py:108)."""
    return solver.uncollided_fluence(_fields(solver, rho, grid, fields), angle)


def first_collision_operator(solver: ScatterSolver, rho, grid: Grid, angle: float, phi_uncollided: Tensor,
                             fields: Optional[MaterialFields] = None) -> Tensor:
    """Once-collided moments phi^(1): one sweep of the first-collision source (citrus.py:126)."""
    fields = _fields(solver, rho, grid, fields)
    return solver.transport_sweep(fields, solver.first_collision_moment_source(fields, angle, phi_uncollided))


def accumulator_update_operator(solver: ScatterSolver, rho, grid: Grid, angle: float, phi_uncollided: Tensor,
                                collided_moments: Optional[Tensor], fields: Optional[MaterialFields] = None, *,
                                sweeps: int = 1) -> Tensor:
    """New accumulator ``phi^(1) + K Phi`` in one sweep of the summed sources (citrus.py:192-247); every further
    sweep adds K applied to the last change. The scattering moments are expanded inside the sweep kernel and the
    first-collision moments enter as the cell-major chunk source of the same sweep."""
    _check_count("sweeps", sweeps)
    fields = _fields(solver, rho, grid, fields)
    ctx = solver.sweep_context(fields.grid_coarse)
    fc = solver.first_collision_moment_source(fields, angle, phi_uncollided, ctx)
    if collided_moments is None:
        previous = None
        collided = solver.transport_sweep(fields, fc)
    else:
        _check_accumulator(solver, fields, collided_moments)
        previous = collided_moments.to(solver.device, solver.dtype)
        source = summed_source_function(solver.scatter_moment_source(fields, previous, ctx), fc)
        collided = solver.transport_sweep(fields, source)
        del source
    fc = None
    for _ in range(int(sweeps) - 1):
        change = collided if previous is None else collided - previous
        source = solver.scatter_moment_source(fields, change, ctx)
        previous = collided
        collided = solver.transport_sweep(fields, source).add_(previous)
        del source, change
    return collided


def project_scatter(solver: ScatterSolver, rho, grid: Grid, angle: float, phi_uncollided: Tensor,
                    collided_moments: Tensor, fields: Optional[MaterialFields] = None,
                    quadrature: Optional[Quadrature] = None, signal_only: bool = False) -> dict:
    """Total scatter image of one view in one collapsed-cone pass (citrus.py:263)."""
    fields = _fields(solver, rho, grid, fields)
    T = solver.multiple_collision_source_moments(fields, collided_moments.to(solver.device, solver.dtype))
    return solver.scatter_image(fields, angle, phi_uncollided, T, quadrature, signal_only=signal_only)


def citrus_view_step(solver: ScatterSolver, state: CitrusState, fields: MaterialFields, angle: float, *,
                     primary: bool = True, sweeps: int = 1):
    """One CITRUS scatter step of one view (citrus.py:320-366): (scatter signal [Rd, Cd], primary signal [Rd, Cd] or
    None, new state). Raises ``ValueError`` if ``angle`` differs from ``state.angle``."""
    if state.angle is not None and abs(float(state.angle) - float(angle)) > 1e-12:
        raise ValueError(f"state belongs to angle {state.angle}, got {angle}")
    grid = fields.grid_fine
    cone = state.cone_quadrature if state.cone_quadrature is not None else solver.cone_quadrature(grid, angle)
    phi_uc = uncollided_operator(solver, None, grid, angle, fields)
    new = accumulator_update_operator(solver, None, grid, angle, phi_uc, state.collided_moments, fields,
                                      sweeps=sweeps)
    scatter = project_scatter(solver, None, grid, angle, phi_uc, new, fields, cone, signal_only=True)["signal"]
    del phi_uc
    primary_signal = solver.primary(fields, angle)["signal"] if primary else None
    new_state = CitrusState(angle=float(angle), collided_moments=new, cone_quadrature=cone,
                            num_steps=state.num_steps + 1, history=state.history + [_change(new, state)])
    return scatter, primary_signal, new_state


def citrus_views_step(solver: ScatterSolver, states: Sequence[CitrusState], fields: MaterialFields,
                      angles: Sequence[float], *, primary: bool = True, contraction: str = "fp32",
                      direction_chunk=None, schedule: str = "auto", learned=None, learned_threshold: float = 0.01):
    """``citrus_view_step`` of V views that share the densities, batched for the GPU: per view the uncollided
    fluence and the corrected first-collision source (group-major, every sweep direction); the scattering moments of
    every accumulator; then the V accumulator updates ``Phi_new = G(S_FC + K Phi)`` in one sweep with the views
    batched in every launch (the reference loops over views one step at a time, citrus.py:1907-1970); then per view
    the signal-only image and the primary. Returns ([(scatter, primary or None)] per view, [new state] per view);
    each view's result equals its own ``citrus_view_step`` to float32 rounding. ``contraction="bf16"`` runs the
    sweep's scattering-source expansion in mixed precision (degrees l <= 1 in float32, l >= 2 as a bfloat16 matrix
    product, float32 accumulation): the scattering-moment kernel then writes that operand directly, and the next
    step's sweep reads it; ``direction_chunk`` and ``schedule`` are passed to ``sweep_views`` ("auto": the
    autotuner's choice).

    ``learned`` (a trained ``surrogate.UNet3D``) starts views from the learned form when every state is empty: the
    step's sweep from an empty accumulator is the first collision order phi^(1), and two more batched sweeps with one
    batched inference (``surrogate.learned_judged``) turn it into the judged estimate of the converged collided
    moments. A view whose judging sweep changes its scalar-flux cell means by more than ``learned_threshold`` keeps
    the plain first step; the residual goes to the new state's ``learned_residual``. Later steps continue the CITRUS
    iteration from whichever accumulator the view holds."""
    batch = _prepare_views(solver, states, fields, angles, contraction)
    collided = sweep_views(batch.ctx, fields.mu_t_coarse, batch.T, source=batch.source, contraction=contraction,
                           direction_chunk=direction_chunk, schedule=schedule)
    residuals = None
    if learned is not None and all(st.collided_moments is None for st in states):
        from .surrogate import learned_judged

        judged, residuals = learned_judged(solver, fields, batch.ctx, batch.source, collided, batch.T, learned)
        for v, r in enumerate(residuals):
            if r <= learned_threshold:
                collided[v].copy_(judged[v])
        del judged
    batch.source = None
    results, new_states = _finish_views(solver, batch, collided, primary, contraction)
    if residuals is not None:
        for st, r in zip(new_states, residuals):
            st.learned_residual = float(r)
    return results, new_states


@dataclass
class _ViewBatch:
    """Views of one step that share material fields, between the source preparation and the sweep."""

    states: Sequence[CitrusState]
    fields: MaterialFields
    angles: Sequence[float]
    ctx: object
    cones: list
    phi_uc: list
    source: Optional[Tensor]
    T: Tensor


def _prepare_views(solver: ScatterSolver, states: Sequence[CitrusState], fields: MaterialFields,
                   angles: Sequence[float], contraction: str = "fp32") -> _ViewBatch:
    """Uncollided fluence, group-major first-collision source and scattering moments of the accumulators."""
    V = len(states)
    if V < 1 or len(angles) != V:
        raise ValueError(f"need one angle per state, got {len(angles)} angles for {V} states")
    for state, angle in zip(states, angles):
        if state.angle is not None and abs(float(state.angle) - float(angle)) > 1e-12:
            raise ValueError(f"state belongs to angle {state.angle}, got {angle}")
    grid = fields.grid_fine
    ctx = solver.sweep_context(fields.grid_coarse)
    G, nsh, M = solver.num_groups, ctx.num_moments, ctx.num_directions
    R, C, S = fields.grid_coarse.shape
    device = solver.device
    cones = [st.cone_quadrature if st.cone_quadrature is not None else solver.cone_quadrature(grid, a)
             for st, a in zip(states, angles)]
    phi_uc = [uncollided_operator(solver, None, grid, a, fields) for a in angles]
    if fields.q == 1:
        # one point per cell: the first-collision cell means stay uncorrected; the sweep applies the sampling
        # correction as it reads them (bit for bit the corrected means)
        source = DeferredFirstCollision.empty(V, G, (R, C, S), M, device)
        omega = ctx.quadrature.omega.to(device, torch.float32)
        for v, a in enumerate(angles):
            first_collision_deferred_samples(solver.library, fields.rho_quadrature, phi_uc[v],
                                             fields.grid_quadrature, solver.geometry, a, omega,
                                             ctx.quadrature.weight, source, v, return_correction=False)
    else:
        source = torch.empty((V, G, R, C, S, M, 4), device=device, dtype=torch.float32)
        for v, a in enumerate(angles):
            solver.first_collision_group_major(fields, a, phi_uc[v], ctx, out=source[v])
    T = _previous_scatter_moments(states, fields)
    if isinstance(T, MixedOperand) and contraction != "bf16":
        T = None  # the float32 contraction needs the float32 moments: formed again below
    if T is None:
        T = torch.empty(V, G, R, C, S, nsh, 4, device=device, dtype=torch.float32)
        for v, st in enumerate(states):
            if st.collided_moments is None:
                T[v].zero_()
            elif isinstance(st.scatter_source, Tensor) and st.scatter_source_fields is fields:
                T[v].copy_(st.scatter_source)
            else:
                _check_accumulator(solver, fields, st.collided_moments)
                prev = st.collided_moments.to(device, torch.float32).permute(0, 3, 4, 5, 1, 2).contiguous()
                scatter_moments_views(fields.transfer_coarse, prev[None], out=T[v:v + 1])
    return _ViewBatch(states=states, fields=fields, angles=angles, ctx=ctx, cones=cones, phi_uc=phi_uc,
                      source=source, T=T)


def _finish_views(solver: ScatterSolver, batch: _ViewBatch, collided: Tensor, primary: bool,
                  contraction: str = "fp32"):
    """Scattering moments of the new accumulators (kept for the next step), images and new states."""
    fields = batch.fields
    # multiple-collision source of the images, kept for the next step's sweep while the fields stay the same (as
    # the mixed-precision operand with the bfloat16 contraction); its cell means also moment-major, the image's
    # expansion operand
    if contraction == "bf16" and batch.ctx.num_moments in (16, 36):
        T, means = scatter_moments_mixed(fields.transfer_coarse, collided, cell_means=True)
    else:
        T, means = scatter_moments_views(fields.transfer_coarse, collided, cell_means=True)
    results, news, norms = [], [], []
    for v, (st, a) in enumerate(zip(batch.states, batch.angles)):
        new = collided[v].permute(0, 4, 5, 1, 2, 3)
        image = solver.scatter_image(fields, a, batch.phi_uc[v], means[v].transpose(0, 1), batch.cones[v],
                                     signal_only=True, source_means=means[v])["signal"]
        primary_signal = solver.primary(fields, a)["signal"] if primary else None
        results.append((image, primary_signal))
        news.append(new)
        norms.append(_change_norms(new, st))
    del means
    # one host synchronization for the step's accumulator changes and the persistent sweep's status word, after
    # every view's work is queued; an aborted sweep raises here instead of handing on partial moments
    from .scatter_solver.kernels.sweep_kernels import decode_status, fold_status

    flat = decode_status(fold_status(torch.stack(norms)).tolist(), collided.device)
    changes = [_change_from_norms(flat[2 * v:2 * v + 2], st) for v, st in enumerate(batch.states)]
    new_states = [CitrusState(angle=float(a), collided_moments=new, cone_quadrature=batch.cones[v],
                              num_steps=st.num_steps + 1, history=st.history + [c],
                              scatter_source=T.view(v) if isinstance(T, MixedOperand) else T[v],
                              scatter_source_fields=fields, _batch=(T, v))
                  for v, (st, a, new, c) in enumerate(zip(batch.states, batch.angles, news, changes))]
    return results, new_states


def citrus_grids_step(batches, *, primary: bool = True, contraction: str = "fp32", schedule: str = "auto"):
    """The steps of view groups on different transport grids -- views cropped to their beams, coarse and fine
    multigrid stages -- with every group's accumulator updates in ONE sweep: one persistent launch (per 16 views)
    over all the grids when the schedule allows it (P3/P5, the persistent self-check passed), else one batched sweep
    per group. ``batches`` is a sequence of ``(solver, states, fields, angles)``; the solvers must share the sweep
    quadrature and groups (their transport cells may differ). Returns ([results], [new states]) per group, each as
    ``citrus_views_step`` returns them (bit for bit the same moments)."""
    from .scatter_solver.kernels.sweep_kernels import PERSISTENT_MOMENTS, persistent_usable
    from .scatter_solver.transport_sweep import sweep_grids

    prepared = [(solver, _prepare_views(solver, states, fields, angles, contraction))
                for solver, states, fields, angles in batches]
    if not prepared:
        return [], []
    nsh = prepared[0][1].ctx.num_moments
    device = prepared[0][0].device
    if schedule not in ("auto", "launches", "persistent"):
        raise ValueError(f"schedule must be 'auto', 'launches' or 'persistent', got {schedule!r}")
    one_launch = schedule == "persistent" or (schedule == "auto" and nsh in PERSISTENT_MOMENTS
                                              and persistent_usable(device)[0])
    if one_launch:
        collided = sweep_grids([(b.ctx, b.fields.mu_t_coarse, b.T, b.source) for _, b in prepared],
                               contraction=contraction)
    else:
        collided = [sweep_views(b.ctx, b.fields.mu_t_coarse, b.T, source=b.source, contraction=contraction,
                                schedule="launches") for _, b in prepared]
    out_results, out_states = [], []
    for (solver, b), phi in zip(prepared, collided):
        b.source = None
        r, s = _finish_views(solver, b, phi, primary, contraction)
        out_results.append(r)
        out_states.append(s)
    return out_results, out_states


def _previous_scatter_moments(states: Sequence[CitrusState], fields: MaterialFields) -> Optional[Tensor]:
    """The [V, ...] scattering moments of the previous batched step when every state still holds its slot of it, in
    order, for these same fields (no copy); else None."""
    batch = states[0]._batch
    if batch is None:
        return None
    views = batch[0].num_views if isinstance(batch[0], MixedOperand) else int(batch[0].shape[0])
    if views != len(states):
        return None
    for v, st in enumerate(states):
        same = st._batch is not None and st._batch[0] is batch[0] and st._batch[1] == v
        if not same or st.scatter_source_fields is not fields:
            return None
    return batch[0]


def citrus_cropped_views_step(solver: ScatterSolver, states: Sequence[CitrusState], rho_fine: Tensor,
                              grid_fine: Grid, angles: Sequence[float], *, primary: bool = True,
                              contraction: str = "fp32", schedule: str = "auto"):
    """The step of V views of a helical (or any) scan, each on the transport grid cropped to its beam
    (``cropping.beam_slice_range``, whole transport cells): views with the same slab share its material fields, and
    the accumulator updates of every slab run in one sweep (``citrus_grids_step``: one persistent launch over the
    uneven grids). Returns ([(scatter, primary or None)], [new state], [slab (k0, k1) of the fine grid]) per view; a
    view's state belongs to its slab."""
    from .scatter_solver.cropping import beam_slice_range, crop_slices

    if len(states) != len(angles):
        raise ValueError(f"need one angle per state, got {len(angles)} angles for {len(states)} states")
    cell = solver.transport_grid(grid_fine)
    factor = max(1, round(cell.deltas[2] / grid_fine.deltas[2]))
    slabs = [beam_slice_range(solver.geometry, a, grid_fine, factor) for a in angles]
    groups = []
    for slab in sorted(set(slabs)):
        idx = [i for i, s in enumerate(slabs) if s == slab]
        sub, rho_sub = crop_slices(grid_fine, rho_fine, *slab)
        fields = solver.material_fields(rho_sub, sub)
        groups.append((idx, (solver, [states[i] for i in idx], fields, [angles[i] for i in idx])))
    res, sts = citrus_grids_step([g for _, g in groups], primary=primary, contraction=contraction,
                                 schedule=schedule)
    results: list = [None] * len(angles)
    new_states: list = [None] * len(angles)
    for (idx, _), r, st in zip(groups, res, sts):
        for j, i in enumerate(idx):
            results[i], new_states[i] = r[j], st[j]
    return results, new_states, slabs


def _change_norms(new: Tensor, state: CitrusState) -> Tensor:
    """[2] device tensor (|new|, |new - previous|) of the scalar-flux cell means (no synchronization): float32 norms
    of the strided views, as citrus.py:355-361 forms them, with no float64 copy of either accumulator."""
    x = new[:, 0, 0]
    if state.collided_moments is None:
        return torch.stack([torch.linalg.vector_norm(x), torch.zeros((), device=x.device, dtype=x.dtype)])
    prev = state.collided_moments[:, 0, 0].to(x)
    return torch.stack([torch.linalg.vector_norm(x), torch.linalg.vector_norm(x - prev)])


def _change_from_norms(nd, state: CitrusState) -> float:
    """Relative change of the scalar-flux cell means over one step (1 from an empty state with nonzero fluence)."""
    n, d = nd
    if state.collided_moments is None:
        return 1.0 if n > 0 else 0.0
    return d / n if n > 0 else 0.0


def _change(new: Tensor, state: CitrusState) -> float:
    """Relative change of the scalar-flux cell means over one step (1 from an empty state with nonzero fluence)."""
    return _change_from_norms(_change_norms(new, state).tolist(), state)


def citrus_step(solver: ScatterSolver, state: CitrusState, rho, grid: Grid, angle: float,
                fields: Optional[MaterialFields] = None):
    """``citrus_view_step`` from the densities (citrus.py:293)."""
    return citrus_view_step(solver, state, _fields(solver, rho, grid, fields), angle)
