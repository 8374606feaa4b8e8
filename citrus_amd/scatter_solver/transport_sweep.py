"""Batched linear-discontinuous discrete-ordinates sweep on the transport grid (HIP).

Keeps the interface of ``citrus/scatter_solver/transport_sweep.py``: ``SweepContext`` (:73), ``make_sweep_context``
(:156), ``sweep`` (:457) with its source-function protocol (:38-43, :379-450): a source function maps [Dsel] reordered
direction indices to [Dsel, G, 4, R, C, S] moments, may write into an ``out=`` buffer, or (``cell_major = True``) writes
[Dsel, R, C, S, G, 4] into the sweep's buffer. It is called once per direction chunk with all 8 dc directions,
octant-major. The result ``phi [G, (L+1)^2, 4, R, C, S]`` is a view of group-major storage
``[G, R, C, S, (L+1)^2, 4]``, the layout the HIP kernels stream.

Backends: "auto" and "hip". The sweep runs only on the HIP kernels (float32 on a gfx950 device); there is no CPU
fallback, and an unusable kernel raises ``RuntimeError``. Sources that carry scattering-source moments
(``scatter_source_function``) are expanded inside the sweep kernel, so neither the per-direction source nor the
angular fluence is materialised.
"""

from __future__ import annotations

import inspect
from dataclasses import dataclass
from typing import Callable, Literal, Optional, Union

import torch
from torch import Tensor

from .. import _hip
from .angular_quadrature import Quadrature, SphericalHarmonicBasis, spherical_harmonic_basis
from .geometry import Grid
from .kernels import sweep_kernels
from .wavefront import OCTANT_SIGNS, SweepPlan, build_sweep_plan

__all__ = ["SourceFunction", "Backend", "SweepContext", "make_sweep_context", "sweep",
           "source_function_accepts_out", "source_is_cell_major"]

SourceFunction = Callable[[Tensor], Tensor]
Backend = Literal["auto", "hip"]
_BACKENDS = ("auto", "hip")


@dataclass(frozen=True, eq=False)
class HipTables:
    """Device tables of the HIP sweep: schedule, float32 SH tables and direction coefficients."""

    schedule: sweep_kernels.SweepSchedule
    Y: Tensor      # [M, nsh]
    Yw: Tensor     # [M, nsh]
    cdir: Tensor   # [M, 3]


@dataclass(frozen=True, eq=False)
class SweepContext:
    """Everything the sweep needs that depends only on the grid and the quadrature (transport_sweep.py:73-153).

    The quadrature is reordered by octant: reordered direction ``m = octant * D8 + d``.
    """

    shape: tuple[int, int, int]
    deltas: tuple[float, float, float]
    L: int
    plan: SweepPlan
    quadrature: Quadrature
    permutation: Tensor
    basis: SphericalHarmonicBasis
    directions_per_octant: int
    c: Tensor
    s: Tensor
    backend: str = "auto"
    hip: Optional[HipTables] = None

    @property
    def num_cells(self) -> int:
        return self.shape[0] * self.shape[1] * self.shape[2]

    @property
    def num_planes(self) -> int:
        return self.plan.num_planes

    @property
    def num_directions(self) -> int:
        return 8 * self.directions_per_octant

    @property
    def num_moments(self) -> int:
        return (self.L + 1) ** 2

    @property
    def device(self) -> torch.device:
        return self.c.device

    @property
    def dtype(self) -> torch.dtype:
        return self.c.dtype

    def direction_indices(self, start: int, stop: int, octants: Optional[Tensor] = None) -> Tensor:
        """Reordered indices ``octants[i] * D8 + start + j`` of the per-octant chunk [start, stop), octant-major."""
        d = torch.arange(start, stop, device=self.device, dtype=torch.int64)
        o = torch.arange(8, device=self.device, dtype=torch.int64) if octants is None else octants.to(self.device)
        return (o[:, None] * self.directions_per_octant + d[None, :]).reshape(-1)


def _check_backend(backend: str) -> None:
    if backend in ("torch", "triton"):
        raise ValueError(f"backend {backend!r} does not exist in this implementation: the sweep runs on the HIP "
                         "kernels only (backends 'auto' and 'hip')")
    if backend not in _BACKENDS:
        raise ValueError(f"backend must be one of {_BACKENDS}, got {backend!r}")


def make_sweep_context(grid: Grid, quadrature: Quadrature, L: int, device: Union[torch.device, str, None] = None,
                       dtype: torch.dtype = torch.float32, backend: Backend = "auto") -> SweepContext:
    """Build the sweep context of a grid and a quadrature, reordering the quadrature by octant.

    The quadrature must hold the same number of directions in every octant. On a GPU device with float32 tables the
    HIP schedule and device tables are built here (backend "hip" requires them and raises if they cannot be built).
    """
    _check_backend(backend)
    D8 = quadrature.directions_per_octant
    if D8 is None or D8 < 1:
        raise ValueError("the quadrature must hold the same (non-zero) number of directions in every octant")
    if L < 0:
        raise ValueError("L must be >= 0")
    if L > 7:
        raise ValueError("the HIP sweep supports L <= 7 ((L+1)^2 <= 64 moments)")
    device = torch.device(quadrature.device if device is None else device)
    if backend == "hip" and device.type != "cuda":
        raise ValueError("backend='hip' requires a GPU device")
    permutation, _ = quadrature.by_octant()
    reordered64 = quadrature.reordered(permutation).to(dtype=torch.float64)
    expected = torch.arange(8).repeat_interleave(D8)
    if not torch.equal(reordered64.octant.cpu(), expected):
        raise RuntimeError("octant reordering failed")
    basis64 = spherical_harmonic_basis(L, reordered64.to(device="cpu"))
    deltas = tuple(float(v) for v in grid.deltas)
    omega = reordered64.omega.cpu().reshape(8, D8, 3)
    c = omega.abs() / torch.tensor(deltas, dtype=torch.float64)
    s = torch.where(omega < 0, -torch.ones_like(omega), torch.ones_like(omega))
    if not torch.equal(s.to(torch.int64), OCTANT_SIGNS[:, None, :].expand(8, D8, 3)):
        raise RuntimeError("direction signs disagree with the octant signs")
    hip = None
    if device.type == "cuda" and dtype == torch.float32:
        try:
            schedule = sweep_kernels.SweepSchedule(tuple(grid.shape), device)
            hip = HipTables(schedule=schedule,
                            Y=basis64.Y.to(device=device, dtype=torch.float32).contiguous(),
                            Yw=basis64.Y_weighted.to(device=device, dtype=torch.float32).contiguous(),
                            cdir=c.reshape(-1, 3).to(device=device, dtype=torch.float32).contiguous())
        except RuntimeError:
            if backend == "hip":
                raise
    elif backend == "hip":
        raise ValueError("backend='hip' needs float32 tables on a GPU device")
    basis = SphericalHarmonicBasis(L=L, Y=basis64.Y.to(device=device, dtype=dtype),
                                   Y_weighted=basis64.Y_weighted.to(device=device, dtype=dtype),
                                   degree_of_moment=basis64.degree_of_moment.to(device))
    return SweepContext(shape=tuple(grid.shape), deltas=deltas, L=L, plan=build_sweep_plan(tuple(grid.shape)),
                        quadrature=reordered64.to(device=device, dtype=dtype), permutation=permutation.to(device),
                        basis=basis, directions_per_octant=D8, c=c.to(device=device, dtype=dtype).contiguous(),
                        s=s.to(device=device, dtype=dtype).contiguous(), backend=backend, hip=hip)


def source_function_accepts_out(source_function: SourceFunction) -> bool:
    """Whether a source function takes an ``out`` keyword (transport_sweep.py:379-393)."""
    try:
        params = inspect.signature(source_function).parameters
    except (TypeError, ValueError):
        return False
    p = params.get("out")
    return p is not None and p.kind in (inspect.Parameter.POSITIONAL_OR_KEYWORD, inspect.Parameter.KEYWORD_ONLY)


def source_is_cell_major(source_function: SourceFunction) -> bool:
    """Whether a source function writes cell-major [Dsel, R, C, S, G, 4] moments (its ``cell_major`` attribute)."""
    return bool(getattr(source_function, "cell_major", False))


def _resolve_chunk(context: SweepContext, direction_chunk: Optional[int]) -> int:
    D8 = context.directions_per_octant
    if direction_chunk is None:
        chunk = D8
    else:
        chunk = int(direction_chunk)
        if chunk < 1:
            raise ValueError("direction_chunk must be >= 1")
    return min(chunk, D8, sweep_kernels.MAX_CHUNK)


def _fill_source(context: SweepContext, source_function: SourceFunction, d0: int, d1: int, G: int,
                 out: Tensor) -> None:
    """Evaluate the source of chunk [d0, d1) of every octant into the cell-major buffer ``out`` [8 dc, N, G, 4]
    (transport_sweep.py:408-450)."""
    chunk = d1 - d0
    N = context.num_cells
    index = context.direction_indices(d0, d1)
    if source_is_cell_major(source_function):
        view = out.view((8 * chunk,) + tuple(context.shape) + (G, 4))
        got = source_function(index, out=view)
        if got.data_ptr() != view.data_ptr() or tuple(got.shape) != tuple(view.shape):
            raise ValueError("a cell-major source function must write its [Dsel, R, C, S, G, 4] moments into `out`")
        return
    expected = (8 * chunk, G, 4) + tuple(context.shape)
    if source_function_accepts_out(source_function):
        got = source_function(index, out=torch.empty(expected, device=out.device, dtype=out.dtype))
    else:
        got = source_function(index)
    if tuple(got.shape) != expected:
        raise ValueError(f"source_function returned shape {tuple(got.shape)}, expected {expected}")
    out.view(8 * chunk, N, 4 * G).copy_(got.reshape(8 * chunk, 4 * G, N).transpose(1, 2))


def _resolve_hip(context: SweepContext, backend: Optional[str], mu_t: Tensor) -> None:
    chosen = context.backend if backend is None else backend
    _check_backend(chosen)
    if mu_t.device.type != "cuda" or mu_t.dtype != torch.float32:
        raise ValueError(f"the HIP sweep needs float32 inputs on a GPU device, got {mu_t.dtype} on {mu_t.device}")
    if context.hip is None:
        ok, why = sweep_kernels.hip_available()
        raise RuntimeError(f"the sweep context has no HIP tables (float32 GPU context needed): {why}")
    ok, why = sweep_kernels.kernels_usable(mu_t.device)
    if not ok:
        raise RuntimeError(f"the HIP sweep kernel is unusable: {why}")


def sweep(sweep_context: SweepContext, mu_t: Tensor, source_function: SourceFunction,
          direction_chunk: Optional[int] = None, return_psi: bool = False, backend: Optional[Backend] = None
          ) -> Union[Tensor, tuple[Tensor, Tensor]]:
    """LD discrete-ordinates sweep of all directions and groups with vacuum boundaries (transport_sweep.py:457).

    Args:
        sweep_context: context of the grid and the quadrature.
        mu_t: [G, R, C, S] float32 total attenuation in 1/cm on the GPU.
        source_function: source function (see the module docstring); one carrying ``hip_moments`` (from
            ``scatter_source_function``) is expanded inside the kernel.
        direction_chunk: directions per octant swept together. Defaults to all (at most 64).
        return_psi: also return the angular fluence [M, G, 4, R, C, S] in reordered direction order.
        backend: "auto" or "hip". Defaults to the context's backend.

    Returns:
        phi [G, (L+1)^2, 4, R, C, S] (a view of [G, R, C, S, (L+1)^2, 4] storage), and psi with ``return_psi``.
    """
    if mu_t.dim() != 4 or tuple(mu_t.shape[1:]) != tuple(sweep_context.shape):
        raise ValueError(f"mu_t must be [G, {sweep_context.shape}], got {tuple(mu_t.shape)}")
    _resolve_hip(sweep_context, backend, mu_t)
    return _sweep_unchecked(sweep_context, mu_t, source_function, direction_chunk, return_psi)


def _sweep_unchecked(context: SweepContext, mu_t: Tensor, source_function: SourceFunction,
                     direction_chunk: Optional[int] = None, return_psi: bool = False):
    G = int(mu_t.shape[0])
    R, C, S = context.shape
    N = context.num_cells
    D8 = context.directions_per_octant
    nsh = context.num_moments
    hip = context.hip
    device = mu_t.device
    mu = mu_t.reshape(G, N).contiguous()
    Dc = _resolve_chunk(context, direction_chunk)
    T = getattr(source_function, "hip_moments", None)
    extra = getattr(source_function, "extra", None) if T is not None else None
    if T is not None:
        if T.shape != (G, R, C, S, nsh, 4) or not T.is_contiguous() or T.device != device:
            raise ValueError(f"moment source must carry contiguous [G, R, C, S, {nsh}, 4] moments on {device}")
    phi_store = torch.empty(G, R, C, S, nsh, 4, device=device, dtype=torch.float32)
    psi_out = torch.empty(context.num_directions, G, 4, N, device=device, dtype=torch.float32) if return_psi else None
    # an explicit source that can hand over all directions at once, group-major [G, R, C, S, 8 D8, 4] (or its cell
    # means [G, R, C, S, 8 D8] when the slopes vanish), is read by the kernel in place (no per-chunk buffer)
    explicit = source_function if T is None else extra
    group_major = None
    group_layout = 1
    if explicit is not None and not return_psi and hasattr(explicit, "hip_group_major"):
        group_major = explicit.hip_group_major()
        want = (G, R, C, S, context.num_directions, 4)
        if tuple(group_major.shape) == want[:-1]:
            group_layout = 2
        elif tuple(group_major.shape) != want:
            raise ValueError(f"hip_group_major must return a contiguous {want} or {want[:-1]} tensor on {device}")
        if not group_major.is_contiguous() or group_major.device != device:
            raise ValueError(f"hip_group_major must return a contiguous {want} or {want[:-1]} tensor on {device}")
    for d0 in range(0, D8, Dc):
        d1 = min(d0 + Dc, D8)
        dc = d1 - d0
        if group_major is not None:
            hip.schedule.run(mu=mu, phi=phi_store, Y=hip.Y, Yw=hip.Yw, cdir=hip.cdir, G=G, nsh=nsh, D8=D8, d0=d0,
                             dc=dc, T=T, source=group_major, accumulate=d0 > 0, source_layout=group_layout)
            continue
        buf = None
        if T is None or extra is not None or return_psi:
            buf = torch.empty(8 * dc, N, G, 4, device=device, dtype=torch.float32)
            if T is None:
                _fill_source(context, source_function, d0, d1, G, buf)
            elif extra is not None:
                _fill_source(context, extra, d0, d1, G, buf)
            else:
                buf.zero_()
        hip.schedule.run(mu=mu, phi=phi_store, Y=hip.Y, Yw=hip.Yw, cdir=hip.cdir, G=G, nsh=nsh, D8=D8, d0=d0, dc=dc,
                         T=T, source=buf, accumulate=d0 > 0, write_psi=return_psi)
        if psi_out is not None:
            psi_out[context.direction_indices(d0, d1)] = buf.permute(0, 2, 3, 1)
    phi = phi_store.permute(0, 4, 5, 1, 2, 3)
    if psi_out is not None:
        return phi, psi_out.view(context.num_directions, G, 4, R, C, S)
    return phi


def _check_moments(moments, G: int, spatial: tuple, nsh: int) -> tuple:
    """(V, mixed) of float32 moments [V, G, R, C, S, nsh, 4] or a mixed-precision operand; raises ValueError on a
    shape, dtype or layout mismatch."""
    from .source_iteration import MixedOperand

    if isinstance(moments, MixedOperand):
        V = moments.num_views
        moments.check(V, G, spatial, nsh)
        return V, True
    if moments.dim() != 7 or tuple(moments.shape[1:]) != (G,) + tuple(spatial) + (nsh, 4) \
            or not moments.is_contiguous() or moments.dtype != torch.float32:
        raise ValueError(f"moments must be contiguous float32 [V, {G}, *{tuple(spatial)}, {nsh}, 4]")
    return int(moments.shape[0]), False


def _mixed(moments):
    """The mixed-precision operand of float32 moments (converted once), or the operand itself."""
    from .source_iteration import MixedOperand, bf16_operand

    return moments if isinstance(moments, MixedOperand) else bf16_operand(moments)


def sweep_views(sweep_context: SweepContext, mu_t: Tensor, moments: Tensor, out: Optional[Tensor] = None,
                direction_chunk: Optional[int] = None, source: Optional[Tensor] = None,
                contraction: str = "fp32", schedule: str = "auto", waves_per_cu: int = 0) -> Tensor:
    """Sweep V views that share the material fields, each from its own scattering-source moments, in one launch
    sequence: the batched form of one ``sweep`` per view of an outer iteration (citrus.py:1920-1935).

    Args:
        sweep_context: context of the grid and the quadrature (float32, GPU).
        mu_t: [G, R, C, S] float32 total attenuation shared by the views.
        moments: [V, G, R, C, S, nsh, 4] float32 contiguous scattering-source moments (group-major layout), or their
            mixed-precision operand (``source_iteration.MixedOperand``, which implies ``contraction="bf16"``).
        out: optional [V, G, R, C, S, nsh, 4] float32 output buffer.
        direction_chunk: directions per octant per pass. Defaults to all; "auto" takes the fastest chunk (and, with
            ``schedule="auto"``, schedule) measured by the autotuner for this configuration.
        source: optional explicit source added to the moments' expansion, every direction of every view:
            [V, G, R, C, S, M, 4] LD moments or [V, G, R, C, S, M] cell means (zero slopes), M the reordered
            directions (the summed source of the CITRUS step, citrus.py:232-236), or the first-collision cell means
            before their sampling correction (``uncollided.DeferredFirstCollision``): the persistent schedule applies
            the correction as it reads them, the launch sequence after ``apply_()``; the moments are the same.
        contraction: "fp32", or "bf16" for the mixed-precision source expansion (P3 and P5): the moments of degree
            l <= 1 in float32 and those of degree l >= 2 as a bfloat16 matrix product, all accumulated in float32.
        schedule: "launches" (one launch per block diagonal), "persistent" (one launch of resident waves per
            direction chunk and at most 16 views, P3 and P5) or "auto" (persistent where it is built and passed its
            self-check on this device, else launches); all give the same moments bit for bit.
        waves_per_cu: the persistent schedule's resident waves per CU (0: the kernel's occupancy); the moments do
            not depend on it.

    Returns:
        [V, G, R, C, S, nsh, 4] fluence moments.
    """
    if mu_t.dim() != 4 or tuple(mu_t.shape[1:]) != tuple(sweep_context.shape):
        raise ValueError(f"mu_t must be [G, {sweep_context.shape}], got {tuple(mu_t.shape)}")
    _resolve_hip(sweep_context, None, mu_t)
    G = int(mu_t.shape[0])
    R, C, S = sweep_context.shape
    nsh = sweep_context.num_moments
    V, mixed = _check_moments(moments, G, (R, C, S), nsh)
    if mixed:
        contraction = "bf16"
    layout = 0
    deferred = source if _is_deferred(source) else None
    if deferred is not None:
        _check_deferred(deferred, V, G, (R, C, S), sweep_context.num_directions)
        source = None  # block order: the persistent sweep reads it as it is, the launch sequence after apply_()
    if source is not None:
        M = sweep_context.num_directions
        base = (V, G, R, C, S, M)
        if tuple(source.shape) == base + (4,):
            layout = 1
        elif tuple(source.shape) == base:
            layout = 2
        else:
            raise ValueError(f"source must be {base + (4,)} or {base}, got {tuple(source.shape)}")
        if not source.is_contiguous() or source.device != mu_t.device or source.dtype != torch.float32:
            raise ValueError("source must be a contiguous float32 tensor on the moments' device")
    if contraction not in ("fp32", "bf16"):
        raise ValueError(f"contraction must be 'fp32' or 'bf16', got {contraction!r}")
    if schedule not in ("auto", "launches", "persistent"):
        raise ValueError(f"schedule must be 'auto', 'launches' or 'persistent', got {schedule!r}")
    shape = (V, G, R, C, S, nsh, 4)
    if out is None:
        out = torch.empty(shape, device=mu_t.device, dtype=torch.float32)
    elif tuple(out.shape) != shape or not out.is_contiguous() or out.dtype != torch.float32:
        raise ValueError(f"out must be a contiguous float32 {list(shape)} tensor")
    if schedule == "auto":
        from .kernels.sweep_kernels import PERSISTENT_MOMENTS, persistent_usable

        if direction_chunk == "auto":  # both measured once per configuration (autotune.SweepTuner)
            from .autotune import default_tuner

            schedule, direction_chunk = default_tuner().best_config(sweep_context, mu_t, V, contraction)
            if schedule == "persistent" and waves_per_cu == 0:
                waves_per_cu = default_tuner().best_waves(sweep_context, mu_t, V, direction_chunk, contraction)
        elif nsh in PERSISTENT_MOMENTS and persistent_usable(mu_t.device)[0]:
            schedule = "persistent"
        else:
            schedule = "launches"
    if schedule == "persistent":
        return _sweep_grids_persistent([(sweep_context, mu_t, moments, deferred if deferred is not None else source,
                                         out)], direction_chunk, contraction, waves_per_cu)[0]
    if deferred is not None:
        source, layout = deferred.apply_(), 2
    D8 = sweep_context.directions_per_octant
    if direction_chunk == "auto":  # measured once per configuration (autotune.SweepTuner)
        from .autotune import default_tuner

        direction_chunk = default_tuner().best_chunk(sweep_context, mu_t, V, contraction)
    Dc = _resolve_chunk(sweep_context, direction_chunk)
    hip = sweep_context.hip
    mu = mu_t.reshape(G, -1).contiguous()
    T, T16, tlo = moments, None, None
    if contraction == "bf16":
        op = _mixed(moments)
        T, T16, tlo = None, op.high, op.low
    for d0 in range(0, D8, Dc):
        dc = min(d0 + Dc, D8) - d0
        hip.schedule.run(mu=mu, phi=out, Y=hip.Y, Yw=hip.Yw, cdir=hip.cdir, G=G, nsh=nsh, D8=D8, d0=d0, dc=dc, V=V,
                         T=T, accumulate=d0 > 0, source=source, source_layout=layout, T16=T16, T_low=tlo)
    return out


def sweep_grids(items, direction_chunk: Optional[int] = None, contraction: str = "fp32",
                waves_per_cu: int = 0) -> list:
    """Sweep the views of several grids (views cropped to their beams, coarse and fine multigrid stages) together in
    persistent launches of at most 16 views each: every view's block tasks in one launch per direction chunk.

    Args:
        items: sequence of ``(sweep_context, mu_t, moments, source)`` with the ``sweep_views`` shapes for that
            context's grid (``source`` may be None; all items then have none, else all the same layout). The
            contexts share the quadrature and the SH order; their grids may differ in shape and cell size.
        direction_chunk: directions per octant per launch (default all).
        contraction: "fp32" or "bf16" as in ``sweep_views``.
        waves_per_cu: resident waves per CU as in ``sweep_views``.

    Returns:
        One [V_i, G, R_i, C_i, S_i, nsh, 4] moment tensor per item, equal bit for bit to ``sweep_views`` of the item.
    """
    if not items:
        return []
    prepared = []
    for ctx, mu_t, moments, source in items:
        if mu_t.dim() != 4 or tuple(mu_t.shape[1:]) != tuple(ctx.shape):
            raise ValueError(f"mu_t must be [G, {ctx.shape}], got {tuple(mu_t.shape)}")
        _resolve_hip(ctx, None, mu_t)
        V, _ = _check_moments(moments, int(mu_t.shape[0]), tuple(ctx.shape), ctx.num_moments)
        out = torch.empty((V, int(mu_t.shape[0])) + tuple(ctx.shape) + (ctx.num_moments, 4), device=mu_t.device,
                          dtype=torch.float32)
        prepared.append((ctx, mu_t, moments, source, out))
    return _sweep_grids_persistent(prepared, direction_chunk, contraction, waves_per_cu)


def _is_deferred(source) -> bool:
    from .uncollided import DeferredFirstCollision

    return isinstance(source, DeferredFirstCollision)


def _check_deferred(src, V: int, G: int, spatial: tuple, M: int) -> None:
    from .uncollided import fc_blocks

    if tuple(src.shape) != tuple(spatial):
        raise ValueError(f"deferred first-collision source of grid {tuple(src.shape)}, expected {tuple(spatial)}")
    if M % 8:
        raise ValueError(f"the sweep's directions come in 8 octants, got {M}")
    nb, dp, pad = fc_blocks(spatial), -(-(M // 8) // 4) * 4, _hip.FC_GROUP_PAD
    want = {"samples": (V, G, nb * 8 * 64 * dp + pad), "scale": (V, G, nb * 64 + pad),
            "beta": (V, G, nb * 64 + pad), "u": (V, nb * 64, 4), "omega": (M, 4)}
    for name, shape in want.items():
        t = getattr(src, name)
        if tuple(t.shape) != shape or not t.is_contiguous() or t.dtype != torch.float32:
            raise ValueError(f"deferred first-collision {name} must be contiguous float32 {list(shape)}, got "
                             f"{tuple(t.shape)} {t.dtype}")


def _sweep_grids_persistent(items, direction_chunk, contraction: str, waves_per_cu: int = 0) -> list:
    """Persistent sweeps of (context, mu_t, moments, source, out) items, 16 views per launch. A
    ``DeferredFirstCollision`` source not yet applied runs as source layout 3 (the sweep applies the correction)."""
    from .kernels.sweep_kernels import PERSISTENT_MAX_VIEWS, PERSISTENT_MOMENTS, persistent_schedule

    if contraction not in ("fp32", "bf16"):
        raise ValueError(f"contraction must be 'fp32' or 'bf16', got {contraction!r}")
    ctx0 = items[0][0]
    D8, nsh, M = ctx0.directions_per_octant, ctx0.num_moments, ctx0.num_directions
    G = int(items[0][1].shape[0])
    if nsh not in PERSISTENT_MOMENTS:
        raise ValueError(f"the persistent sweep is built for {PERSISTENT_MOMENTS} moments (P3, P5), got {nsh}")
    layouts = set()
    omega = None  # the directions of the deferred first-collision sources (layout 3)
    views = []  # (shape, spec) per view
    for ctx, mu_t, moments, source, out in items:
        if ctx.directions_per_octant != D8 or ctx.num_moments != nsh or ctx.hip.Y.shape != ctx0.hip.Y.shape \
                or not torch.equal(ctx.hip.Y, ctx0.hip.Y):
            raise ValueError("the grids of one persistent sweep must share the quadrature and the SH order")
        if int(mu_t.shape[0]) != G:
            raise ValueError("the grids of one persistent sweep must have the same number of groups")
        R, C, S = ctx.shape
        V, mixed = _check_moments(moments, G, (R, C, S), nsh)
        if mixed and contraction != "bf16":
            raise ValueError("a mixed-precision operand needs contraction='bf16'")
        layout = -1
        fc = None
        if _is_deferred(source):
            _check_deferred(source, V, G, (R, C, S), M)
            if source.applied:
                source = source.corrected
            else:
                fc, source = source, None
        if source is not None:
            base = (V, G, R, C, S, M)
            layout = 1 if tuple(source.shape) == base + (4,) else 2 if tuple(source.shape) == base else 0
            if layout == 0:
                raise ValueError(f"source must be {base + (4,)} or {base}, got {tuple(source.shape)}")
            if not source.is_contiguous() or source.dtype != torch.float32:
                raise ValueError("source must be a contiguous float32 tensor")
        if fc is not None:  # block order, checked above
            layout = 3
            source = fc.samples
            if omega is None:
                omega = fc.omega
            elif not torch.equal(omega, fc.omega):
                raise ValueError("the deferred first-collision sources of one sweep must share the directions")
        layouts.add(layout)
        mu = mu_t.reshape(G, -1).contiguous()
        op = _mixed(moments) if contraction == "bf16" else None
        for v in range(V):
            spec = dict(mu=mu, phi=out[v], cdir=ctx.hip.cdir, source=None if source is None else source[v])
            if fc is not None:
                spec.update(fc_scale=fc.scale[v], fc_beta=fc.beta[v], fc_u=fc.u[v])
            if op is None:
                spec["T"] = moments[v]
            else:
                spec["T16"], spec["T_low"] = op.high[v], op.low[v]
            views.append((tuple(ctx.shape), spec))
    if len(layouts) != 1:
        raise ValueError("every item of one persistent sweep needs a source of the same layout, or none")
    layout = layouts.pop()
    Dc = _resolve_chunk(ctx0, None if direction_chunk == "auto" else direction_chunk)
    device = items[0][1].device
    for b0 in range(0, len(views), PERSISTENT_MAX_VIEWS):
        batch = views[b0:b0 + PERSISTENT_MAX_VIEWS]
        sched = persistent_schedule([s for s, _ in batch], G, device)
        for d0 in range(0, D8, Dc):
            dc = min(d0 + Dc, D8) - d0
            sched.run(views=[spec for _, spec in batch], Y=ctx0.hip.Y, Yw=ctx0.hip.Yw, nsh=nsh, D8=D8, d0=d0,
                      dc=dc, accumulate=d0 > 0, source_layout=layout, fc_omega=omega if layout == 3 else None,
                      waves_per_cu=waves_per_cu)
    return [it[4] for it in items]


class SweepGraph:
    """``sweep_views`` captured once in a HIP graph on fixed buffers and replayed: the launches of a batched
    view-sweep (184-odd block-plane launches, or the persistent schedule's flag clear and one launch per chunk, and
    the bfloat16 operand's conversion) become one graph launch. The caller writes each step's
    moments (and explicit source) into ``moments`` (``source``) and calls ``replay()``; the result is in ``out``.
    The sweep records no memset and its kernels are deterministic, so every replay equals the eager sweep of the
    same inputs bit for bit.
    """

    def __init__(self, sweep_context: SweepContext, mu_t: Tensor, moments: Tensor, source: Optional[Tensor] = None,
                 contraction: str = "fp32", direction_chunk: Optional[int] = None, schedule: str = "auto"):
        if _is_deferred(source):
            raise ValueError("a graph replays fixed buffers: pass the corrected source tensor, not a deferred one")
        self.context = sweep_context
        self.mu_t = mu_t
        self.moments = moments
        self.source = source
        self.out = torch.empty_like(moments)
        kwargs = dict(out=self.out, direction_chunk=direction_chunk, source=source, contraction=contraction,
                      schedule=schedule)
        device = moments.device
        side = torch.cuda.Stream(device=device)  # default priority
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):  # warm-up: usability check, face buffers, operand buffers
            for _ in range(2):
                sweep_views(sweep_context, mu_t, moments, **kwargs)
        torch.cuda.current_stream(device).wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            sweep_views(sweep_context, mu_t, moments, **kwargs)

    def replay(self) -> Tensor:
        """One batched view-sweep of the current contents of ``moments`` (and ``source``) into ``out``."""
        self.graph.replay()
        return self.out
