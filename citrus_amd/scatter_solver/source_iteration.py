"""Source iteration over collision order and the scattering-source moments (HIP).

Keeps the interface of ``citrus/scatter_solver/source_iteration.py``: ``SourceIterationResult`` (:28),
``scatter_moments`` (:49), ``scatter_source_function`` (:65), ``multiple_collision_source_moments`` (:109),
``source_iteration`` (:138). Scattering-source moments live in the group-major layout ``[G, R, C, S, (L+1)^2, 4]``
of the HIP kernels; the source functions built here carry them (``hip_moments``) so the sweep kernel expands
``sum_n Y_n(Omega) T_n`` itself. The convergence norms are taken in float64 on the device.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Callable, Optional

import torch
from torch import Tensor

from .. import _hip
from .transport_sweep import SourceFunction, SweepContext, sweep

__all__ = ["SourceIterationResult", "MaterialTransfer", "MixedOperand", "source_iteration", "scatter_moments",
           "scatter_source_function", "summed_source_function", "multiple_collision_source_moments", "bf16_operand",
           "scatter_moments_mixed"]


@dataclass
class SourceIterationResult:
    """Summed moments, number of sweeps, relative-change history, per-sweep norms, convergence flag (:28-46)."""

    phi_total: Tensor
    num_iterations: int
    history: list[float] = field(default_factory=list)
    phi_norms: list[float] = field(default_factory=list)
    converged: bool = False


@dataclass(frozen=True, eq=False)
class MaterialTransfer:
    """Legendre transfer moments in factored form, ``mu_s_l(cell) = sum_m rho[m, cell] sigma[m, gi, go, l]``.

    The dense ``[G_in, G_out, L+1, R, C, S]`` volume of ``transfer_moments_volume`` (cross_sections.py:886) is never
    formed: the HIP kernel combines the densities with the per-material table on the fly.

    Attributes:
        rho: [M, R, C, S] float32 densities on the GPU in g/cm^3.
        sigma: [M, G, G, L+1] float32 mass transfer moments in cm^2/g.
    """

    rho: Tensor
    sigma: Tensor
    downscatter_only: bool = field(init=False, default=False)

    def __post_init__(self):
        # one look at the small table when the transfer is built, so no later call synchronizes the host
        s = self.sigma.detach().cpu()
        lower = torch.tril(torch.ones(s.shape[1], s.shape[2], dtype=torch.bool), diagonal=-1)  # g_in > g_out
        object.__setattr__(self, "downscatter_only", bool((s[:, lower] == 0).all()))


def _group_major(phi: Tensor) -> Tensor:
    """Contiguous [G, R, C, S, nsh, 4] storage of moments given as [G, nsh, 4, R, C, S] (free for sweep outputs)."""
    return phi.permute(0, 3, 4, 5, 1, 2).contiguous()


def scatter_moments(mu_s_l_volume, phi: Tensor) -> Tensor:
    """Scattering-source moments ``T[g, n] = sum_g' mu_s_l[g', g, l(n)] phi[g', n]`` (source_iteration.py:49).

    Args:
        mu_s_l_volume: dense ``[G_in, G_out, L+1, R, C, S]`` float32 transfer volume, or a ``MaterialTransfer``.
        phi: [G, (L+1)^2, 4, R, C, S] float32 moments on the GPU.

    Returns:
        T [G, (L+1)^2, 4, R, C, S], a view of group-major storage.
    """
    G, nsh = int(phi.shape[0]), int(phi.shape[1])
    spatial = tuple(phi.shape[3:])
    N = spatial[0] * spatial[1] * spatial[2]
    src = _group_major(phi)
    _hip.require_gpu_tensor(src, "phi")
    out = torch.empty_like(src)
    lib = _hip.library()
    stream = _hip.stream_of(phi.device)
    with torch.cuda.device(phi.device):
        if isinstance(mu_s_l_volume, MaterialTransfer):
            rho = mu_s_l_volume.rho.contiguous()
            sigma = mu_s_l_volume.sigma.contiguous()
            _hip.require_gpu_tensor(rho, "rho")
            _hip.require_gpu_tensor(sigma, "sigma")
            if tuple(rho.shape[1:]) != spatial or sigma.shape[1] != G or sigma.shape[2] != G:
                raise ValueError("MaterialTransfer shapes do not match phi")
            _hip.check(lib.citrus_scatter_moments_factored(
                rho.data_ptr(), int(rho.shape[0]), sigma.data_ptr(), int(sigma.shape[3]) - 1,
                int(mu_s_l_volume.downscatter_only), src.data_ptr(), out.data_ptr(), G, nsh, N, 1, None, stream),
                "citrus_scatter_moments_factored")
        else:
            transfer = mu_s_l_volume.contiguous()
            _hip.require_gpu_tensor(transfer, "mu_s_l_volume")
            if transfer.shape[0] != G or transfer.shape[1] != G or tuple(transfer.shape[3:]) != spatial:
                raise ValueError(f"transfer volume {tuple(transfer.shape)} does not match phi {tuple(phi.shape)}")
            _hip.check(lib.citrus_scatter_moments_dense(transfer.data_ptr(), int(transfer.shape[2]) - 1,
                                                        src.data_ptr(), out.data_ptr(), G, nsh, N, 1, stream),
                       "citrus_scatter_moments_dense")
    return out.permute(0, 4, 5, 1, 2, 3)


class MomentSource:
    """Source function ``S(Omega_m) = sum_n Y_n(Omega_m) T_n`` (+ an optional extra cell-major source).

    The sweep recognises ``hip_moments`` and expands the moments inside its kernel. Called directly, it follows the
    cell-major source protocol of ``scatter_source_function`` (source_iteration.py:94-103).
    """

    cell_major = True

    def __init__(self, context: SweepContext, moments: Tensor, extra: Optional[SourceFunction] = None):
        self.context = context
        self.hip_moments = moments  # [G, R, C, S, nsh, 4]
        self.extra = extra

    def __call__(self, direction_index: Tensor, out: Optional[Tensor] = None, accumulate: bool = False) -> Tensor:
        G, R, C, S, nsh, _ = self.hip_moments.shape
        Y = self.context.basis.Y.to(device=self.hip_moments.device, dtype=self.hip_moments.dtype)
        Y_sel = Y.index_select(0, direction_index.to(Y.device))
        matrix = self.hip_moments.permute(4, 1, 2, 3, 0, 5).reshape(nsh, -1)  # [nsh, (R, C, S, G, 4)]
        shape = (Y_sel.shape[0], R, C, S, G, 4)
        if out is None:
            result = (Y_sel @ matrix).view(shape)
            if self.extra is not None:
                result = result + self.extra(direction_index, out=torch.empty_like(result))
            return result
        flat = out.view(Y_sel.shape[0], -1)
        if self.extra is not None:
            self.extra(direction_index, out=out)
            flat.addmm_(Y_sel, matrix)
        elif accumulate:
            flat.addmm_(Y_sel, matrix)
        else:
            torch.mm(Y_sel, matrix, out=flat)
        return out


def scatter_source_function(sweep_context: SweepContext, T: Tensor) -> MomentSource:
    """Source function of the scattering-source moments T [G, (L+1)^2, 4, R, C, S] (source_iteration.py:65)."""
    if T.shape[1] != sweep_context.num_moments:
        raise ValueError(f"T has {T.shape[1]} spherical-harmonic moments, context has {sweep_context.num_moments}")
    if T.dim() != 6 or T.shape[2] != 4 or tuple(T.shape[3:]) != tuple(sweep_context.shape):
        raise ValueError(f"T must be [G, {sweep_context.num_moments}, 4, {sweep_context.shape}], got {tuple(T.shape)}")
    return MomentSource(sweep_context, _group_major(T))


def summed_source_function(moment_source: MomentSource, extra: SourceFunction) -> MomentSource:
    """``extra + moment_source`` in one sweep: the summed source of the CITRUS step (citrus.py:232-236)."""
    if not source_is_cell_major_fn(extra):
        raise ValueError("the extra source must be cell-major")
    return MomentSource(moment_source.context, moment_source.hip_moments, extra=extra)


def source_is_cell_major_fn(fn) -> bool:
    return bool(getattr(fn, "cell_major", False))


def scatter_moments_views(transfer: MaterialTransfer, phi_views: Tensor, out: Optional[Tensor] = None,
                          cell_means: bool = False):
    """Batched ``scatter_moments`` of V views sharing one factored transfer: [V, G, R, C, S, nsh, 4] in and out.
    With ``cell_means`` also returns the cell means of the result moment-major, [V, nsh, G, R, C, S] (the operand of
    the image's multiple-collision expansion, written by the same kernel)."""
    V, G, R, C, S, nsh, four = (int(v) for v in phi_views.shape)
    _hip.require_gpu_tensor(phi_views, "phi_views")
    rho, sigma = transfer.rho.contiguous(), transfer.sigma.contiguous()
    if out is None:
        out = torch.empty_like(phi_views)
    means = torch.empty(V, nsh, G, R, C, S, device=phi_views.device, dtype=torch.float32) if cell_means else None
    lib = _hip.library()
    with torch.cuda.device(phi_views.device):
        _hip.check(lib.citrus_scatter_moments_factored(
            rho.data_ptr(), int(rho.shape[0]), sigma.data_ptr(), int(sigma.shape[3]) - 1,
            int(transfer.downscatter_only), phi_views.data_ptr(), out.data_ptr(), G, nsh, R * C * S, V,
            _hip.ptr(means), _hip.stream_of(phi_views.device)), "citrus_scatter_moments_factored")
    return (out, means) if cell_means else out


def scatter_moments_stored(transfer: MaterialTransfer, q: Tensor, scale: Tensor,
                           out: Optional[Tensor] = None) -> Tensor:
    """``scatter_moments_views`` of stored accumulators (the state store's format: q [V, G, nsh, R, C, S] float16
    and scale [V, G, nsh], cell means ``q scale``, zero LD slopes), equal bit for bit to the contraction of their
    unpacked form (``state_store.unpack_accumulators``) without it being formed: [V, G, R, C, S, nsh, 4]."""
    if q.dtype != torch.float16 or q.dim() != 6 or not q.is_contiguous():
        raise ValueError("q must be a contiguous float16 [V, G, nsh, R, C, S] tensor")
    _hip.require_gpu_tensor(q, "q", torch.float16)
    _hip.require_gpu_tensor(scale, "scale")
    V, G, nsh, R, C, S = (int(v) for v in q.shape)
    if tuple(scale.shape) != (V, G, nsh) or not scale.is_contiguous() or scale.dtype != torch.float32:
        raise ValueError(f"scale must be a contiguous float32 [{V}, {G}, {nsh}] tensor")
    rho, sigma = transfer.rho.contiguous(), transfer.sigma.contiguous()
    shape = (V, G, R, C, S, nsh, 4)
    if out is None:
        out = torch.empty(shape, device=q.device, dtype=torch.float32)
    elif tuple(out.shape) != shape or not out.is_contiguous() or out.dtype != torch.float32:
        raise ValueError(f"out must be a contiguous float32 {list(shape)} tensor")
    with torch.cuda.device(q.device):
        _hip.check(_hip.library().citrus_scatter_moments_stored(
            rho.data_ptr(), int(rho.shape[0]), sigma.data_ptr(), int(sigma.shape[3]) - 1,
            int(transfer.downscatter_only), q.data_ptr(), scale.data_ptr(), out.data_ptr(), G, nsh, R * C * S, V,
            _hip.stream_of(q.device)), "citrus_scatter_moments_stored")
    return out


@dataclass(frozen=True, eq=False)
class MixedOperand:
    """The mixed-precision operand of the sweep's source expansion for V views: the scattering-source moments of
    degree l <= 1 in float32, ``low [V, G, R, C, S, 4 components, 4 moments]``, and those of degree l >= 2 in
    bfloat16, ``high [V, G, R, C, S, 4 components, KH]`` (KH = nsh - 4 rounded up to 16, zero padded). The sweep
    expands it as ``low Y + high Y`` (a float32 and a bfloat16 matrix product, both accumulated in float32)."""

    low: Tensor
    high: Tensor
    nsh: int

    @property
    def num_views(self) -> int:
        return int(self.low.shape[0])

    @property
    def spatial(self) -> tuple:
        return tuple(int(v) for v in self.low.shape[2:5])

    @property
    def device(self) -> torch.device:
        return self.low.device

    def view(self, v: int) -> "MixedOperand":
        """The operand of view v alone (views of the same storage)."""
        return MixedOperand(self.low[v:v + 1], self.high[v:v + 1], self.nsh)

    def check(self, V: int, G: int, spatial: tuple, nsh: int) -> None:
        kh = 16 * ((nsh - 4 + 15) // 16)
        if self.nsh != nsh or tuple(self.low.shape) != (V, G) + tuple(spatial) + (4, 4) \
                or tuple(self.high.shape) != (V, G) + tuple(spatial) + (4, kh):
            raise ValueError(f"mixed operand must be low [{V}, {G}, *{spatial}, 4, 4] and high [{V}, {G}, *{spatial}, "
                             f"4, {kh}] for {nsh} moments")
        if self.low.dtype != torch.float32 or self.high.dtype != torch.bfloat16 or not self.low.is_contiguous() \
                or not self.high.is_contiguous() or self.low.device.type != "cuda" or self.high.device != self.low.device:
            raise ValueError("mixed operand: contiguous float32 low and bfloat16 high on one GPU")


def empty_mixed_operand(V: int, G: int, spatial: tuple, nsh: int, device) -> MixedOperand:
    if nsh not in (16, 36):
        raise ValueError("the mixed-precision operand is built for P3 and P5 (16 and 36 moments)")
    kh = 16 * ((nsh - 4 + 15) // 16)
    return MixedOperand(torch.empty((V, G) + tuple(spatial) + (4, 4), device=device, dtype=torch.float32),
                        torch.empty((V, G) + tuple(spatial) + (4, kh), device=device, dtype=torch.bfloat16), nsh)


def bf16_operand(moments: Tensor) -> MixedOperand:
    """The mixed-precision operand (``citrus_bf16_operand``) of float32 scattering-source moments
    [V, G, R, C, S, nsh, 4]: degrees l <= 1 copied, degrees l >= 2 rounded to bfloat16 (nearest, ties to even)."""
    V, G, R, C, S, nsh, four = (int(v) for v in moments.shape)
    _hip.require_gpu_tensor(moments, "moments")
    if not moments.is_contiguous() or four != 4:
        raise ValueError("moments must be contiguous [V, G, R, C, S, nsh, 4]")
    op = empty_mixed_operand(V, G, (R, C, S), nsh, moments.device)
    with torch.cuda.device(moments.device):
        _hip.check(_hip.library().citrus_bf16_operand(moments.data_ptr(), G, nsh, R * C * S, V, op.low.data_ptr(),
                                                      op.high.data_ptr(), _hip.stream_of(moments.device)),
                   "citrus_bf16_operand")
    return op


def scatter_moments_mixed(transfer: MaterialTransfer, phi_views: Tensor, out: Optional[MixedOperand] = None,
                          cell_means: bool = False):
    """``scatter_moments_views`` whose result is written directly as the mixed-precision operand of the next sweep
    (``citrus_scatter_moments_mixed``): returns the operand, and with ``cell_means`` also the float32 cell means
    [V, nsh, G, R, C, S] of the image's expansion."""
    V, G, R, C, S, nsh, four = (int(v) for v in phi_views.shape)
    _hip.require_gpu_tensor(phi_views, "phi_views")
    rho, sigma = transfer.rho.contiguous(), transfer.sigma.contiguous()
    out = empty_mixed_operand(V, G, (R, C, S), nsh, phi_views.device) if out is None else out
    out.check(V, G, (R, C, S), nsh)
    means = torch.empty(V, nsh, G, R, C, S, device=phi_views.device, dtype=torch.float32) if cell_means else None
    with torch.cuda.device(phi_views.device):
        _hip.check(_hip.library().citrus_scatter_moments_mixed(
            rho.data_ptr(), int(rho.shape[0]), sigma.data_ptr(), int(sigma.shape[3]) - 1,
            int(transfer.downscatter_only), phi_views.data_ptr(), out.low.data_ptr(), out.high.data_ptr(), G, nsh,
            R * C * S, V, _hip.ptr(means), _hip.stream_of(phi_views.device)), "citrus_scatter_moments_mixed")
    return (out, means) if cell_means else out


def multiple_collision_source_moments(mu_s_l_volume_coarse, phi_total: Tensor) -> Tensor:
    """Scattering operator applied to the total collided fluence (source_iteration.py:109)."""
    return scatter_moments(mu_s_l_volume_coarse, phi_total)


def _scalar_norms(*phis: Tensor) -> list[float]:
    """float64 norms of the scalar-flux cell means ``phi[:, 0, 0]``, one host synchronization (:126-135) that also
    reads the persistent sweep's status word (an aborted sweep raises ``RuntimeError``)."""
    from .kernels.sweep_kernels import decode_status, fold_status

    norms = torch.stack([torch.linalg.vector_norm(p[:, 0, 0].to(torch.float64)) for p in phis])
    return decode_status(fold_status(norms).tolist(), norms.device)


def source_iteration(sweep_context: SweepContext, mu_t_coarse: Tensor, mu_s_l_volume_coarse,
                     first_collision_source_function: SourceFunction, tolerance: float, max_iterations: int,
                     direction_chunk: Optional[int] = None,
                     callback: Optional[Callable[[int, Tensor, float], None]] = None) -> SourceIterationResult:
    """Source iteration over collision order from the first-collision source (source_iteration.py:138-204)."""
    if max_iterations < 1:
        raise ValueError("max_iterations must be >= 1")
    phi_total: Optional[Tensor] = None
    history: list[float] = []
    norms: list[float] = []
    converged = False
    source = first_collision_source_function
    n = 0
    for n in range(1, max_iterations + 1):
        phi_n = sweep(sweep_context, mu_t_coarse, source, direction_chunk=direction_chunk)
        source = None
        phi_total = phi_n.clone() if phi_total is None else phi_total.add_(phi_n)
        norm_n, norm_total = _scalar_norms(phi_n, phi_total)
        rel = norm_n / norm_total if norm_total > 0 else 0.0
        history.append(rel)
        norms.append(norm_n)
        if callback is not None:
            callback(n, phi_n, rel)
        if rel < tolerance:
            converged = True
            break
        if n == max_iterations:
            break
        source = scatter_source_function(sweep_context, scatter_moments(mu_s_l_volume_coarse, phi_n))
        del phi_n
    assert phi_total is not None
    return SourceIterationResult(phi_total=phi_total, num_iterations=n, history=history, phi_norms=norms,
                                 converged=converged)
