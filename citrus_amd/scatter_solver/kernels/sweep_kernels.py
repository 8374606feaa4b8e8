"""Host wrappers of the HIP LD sweep (``citrus_sweep_*`` in include/citrus_hip.h).

Counterpart of ``citrus/scatter_solver/kernels/sweep_kernels.py``: ``kernels_usable`` (:85-108, cached per device,
never raises), the self-check (:111-143) and the launch wrapper ``sweep_planes`` (:265-290). Here the self-check
verifies exact particle balance (outflow + absorption = source, a property of the LD equations) of a small random
problem swept on the device; there is no fallback path, so an unusable kernel makes ``sweep`` raise.
"""

from __future__ import annotations

import ctypes
import os
import weakref
from typing import Optional

import numpy as np
import torch
from torch import Tensor

from ... import _hip

SELF_CHECK_REL_TOL = 1e-4
BLOCK_EDGE = 4
MAX_CHUNK = 32
_SELF_CHECK: dict[str, tuple[bool, str]] = {}


class _HeldBuffers:
    """Scratch buffers a captured HIP graph still points to. A schedule's scratch grows on demand; a graph captured
    while an older buffer was current replays into that buffer, so once any capture has used the schedule its
    replaced buffers are kept here until the schedule itself goes, instead of returning to torch's allocator."""

    def __init__(self) -> None:
        self.captured = False
        self.buffers: list = []

    def note_use(self) -> None:
        if torch.cuda.is_available() and torch.cuda.is_current_stream_capturing():
            self.captured = True

    def retire(self, old: Optional[Tensor]) -> None:
        if old is not None and (self.captured or (torch.cuda.is_available()
                                                   and torch.cuda.is_current_stream_capturing())):
            self.buffers.append(old)


class SweepSchedule:
    """Owner of a native ``CitrusSweepPlan`` (block-diagonal task tables on the device) and its face scratch."""

    def __init__(self, shape: tuple[int, int, int], device: torch.device, block: int = BLOCK_EDGE):
        lib = _hip.library()
        handle = ctypes.c_void_p()
        with torch.cuda.device(device):
            _hip.check(lib.citrus_sweep_plan_create(int(shape[0]), int(shape[1]), int(shape[2]), int(block),
                                                    ctypes.byref(handle)), "citrus_sweep_plan_create")
        self.handle = handle
        self.shape = tuple(int(v) for v in shape)
        self.device = device
        self.block = block
        self._faces: Optional[Tensor] = None
        self._held = _HeldBuffers()
        self._finalizer = weakref.finalize(self, lib.citrus_sweep_plan_destroy, handle)

    def faces(self, G: int, dc: int, V: int) -> Tensor:
        """Face scratch for (G, dc, V), grown on demand and reused across sweeps. A buffer a captured graph has
        recorded is kept alive when a larger one replaces it (``_HeldBuffers``)."""
        n = int(_hip.library().citrus_sweep_face_floats(self.handle, G, dc, V))
        if n < 0:
            raise ValueError("invalid face-buffer request")
        if self._faces is None or self._faces.numel() < n:
            self._held.retire(self._faces)
            self._faces = torch.empty(n, device=self.device, dtype=torch.float32)
        self._held.note_use()
        return self._faces

    def run(self, *, mu: Tensor, phi: Tensor, Y: Tensor, Yw: Tensor, cdir: Tensor, G: int, nsh: int, D8: int,
            d0: int, dc: int, V: int = 1, T: Optional[Tensor] = None, source: Optional[Tensor] = None,
            accumulate: bool = False, write_psi: bool = False, source_layout: int = 0,
            T16: Optional[Tensor] = None, T_low: Optional[Tensor] = None) -> None:
        """Enqueue one chunk sweep on torch's current stream (no host synchronization). ``T16``/``T_low`` (the two
        parts of the mixed-precision operand, ``MixedOperand.high`` / ``.low``) replace ``T`` for the
        mixed-precision source contraction."""
        N = self.shape[0] * self.shape[1] * self.shape[2]
        for name, t in (("mu", mu), ("phi", phi), ("Y", Y), ("Yw", Yw), ("cdir", cdir)):
            _hip.require_gpu_tensor(t, name)
        if mu.numel() != G * N or phi.numel() != V * G * N * nsh * 4:
            raise ValueError("mu or phi size does not match the grid")
        if T is not None:
            _hip.require_gpu_tensor(T, "T")
            if T.numel() != V * G * N * nsh * 4:
                raise ValueError("T size does not match [V, G, N, nsh, 4]")
        if T16 is not None:
            kh = 16 * ((nsh - 4 + 15) // 16)
            if T is not None or T_low is None:
                raise ValueError("T16 replaces T and needs T_low")
            if T16.dtype != torch.bfloat16 or T16.device.type != "cuda" or not T16.is_contiguous() \
                    or T16.numel() != V * G * N * 4 * kh:
                raise ValueError(f"T16 must be contiguous bfloat16 [V, G, N, 4, {kh}] on the GPU")
            _hip.require_gpu_tensor(T_low, "T_low")
            if T_low.numel() != V * G * N * 16:
                raise ValueError("T_low must hold [V, G, N, 4, 4] floats")
        if source is not None:
            _hip.require_gpu_tensor(source, "source")
            if source_layout == 0 and source.numel() != V * 8 * dc * N * G * 4:
                raise ValueError("source size does not match [V, 8 dc, N, G, 4]")
            if source_layout == 1 and source.numel() != V * G * N * 8 * D8 * 4:
                raise ValueError("group-major source size does not match [V, G, N, 8 D8, 4]")
            if source_layout == 2 and source.numel() != V * G * N * 8 * D8:
                raise ValueError("group-major cell-mean source size does not match [V, G, N, 8 D8]")
            if source_layout not in (0, 1, 2):
                raise ValueError("source_layout must be 0, 1 or 2")
            if source_layout != 0 and write_psi:
                raise ValueError("write_psi needs the cell-major source layout")
        if Y.shape != (8 * D8, nsh) or Yw.shape != (8 * D8, nsh) or cdir.shape != (8 * D8, 3):
            raise ValueError("direction tables do not match the quadrature")
        faces = self.faces(G, dc, V)
        args = _hip.SweepArgs(G=G, nsh=nsh, D8=D8, d0=d0, dc=dc, V=V, accumulate=int(accumulate),
                              write_psi=int(write_psi), mu_t=mu.data_ptr(), T=_hip.ptr(T), source=_hip.ptr(source),
                              phi=phi.data_ptr(), Y=Y.data_ptr(), Yw=Yw.data_ptr(), cdir=cdir.data_ptr(),
                              faces=faces.data_ptr(), source_layout=int(source_layout), T_bf16=_hip.ptr(T16),
                              T_low=_hip.ptr(T_low))
        lib = _hip.library()
        with torch.cuda.device(self.device):
            _hip.check(lib.citrus_sweep_run(self.handle, ctypes.byref(args), _hip.stream_of(self.device)),
                       "citrus_sweep_run")


PERSISTENT_MAX_VIEWS = 16
PERSISTENT_MOMENTS = (16, 36)  # moment counts (P3, P5) the persistent kernel is built for
PERSISTENT_STATUS = {0: "ok", 1: "a dependency wait passed its spin limit", 2: "a wave ran on an XCD id outside 0..7"}


class PersistentSchedule:
    """Owner of a native ``CitrusPersistentPlan`` (per-XCD task queues of a list of view grids) and its workspace.

    One ``run`` sweeps one direction chunk of every view in a single launch of resident waves; the views' grids may
    differ in shape and cell size. The moments equal those of ``SweepSchedule.run`` on each view's grid bit for bit.
    """

    def __init__(self, shapes, G: int, device: torch.device):
        shapes = [tuple(int(v) for v in s) for s in shapes]
        if not 1 <= len(shapes) <= PERSISTENT_MAX_VIEWS:
            raise ValueError(f"a persistent sweep takes 1..{PERSISTENT_MAX_VIEWS} views, got {len(shapes)}")
        lib = _hip.library()
        flat = (ctypes.c_int * (3 * len(shapes)))(*[v for s in shapes for v in s])
        handle = ctypes.c_void_p()
        with torch.cuda.device(device):
            _hip.check(lib.citrus_persistent_plan_create(len(shapes), flat, int(G), ctypes.byref(handle)),
                       "citrus_persistent_plan_create")
        self.handle = handle
        self.shapes = shapes
        self.G = int(G)
        self.device = device
        self._workspace: Optional[Tensor] = None
        self._held = _HeldBuffers()
        self._finalizer = weakref.finalize(self, lib.citrus_persistent_plan_destroy, handle)

    @property
    def num_tasks(self) -> int:
        """Block tasks of one chunk sweep."""
        return int(_hip.library().citrus_persistent_num_tasks(self.handle))

    def workspace(self, dc: int) -> Tensor:
        """256-byte aligned workspace (flags, tickets, faces) for chunks of ``dc`` directions, grown on demand."""
        n = int(_hip.library().citrus_persistent_workspace_bytes(self.handle, int(dc)))
        if n < 0:
            raise ValueError("invalid persistent workspace request")
        if self._workspace is None or self._workspace.numel() < n:
            self._held.retire(getattr(self, "_raw", None))
            raw = torch.empty(n + 256, device=self.device, dtype=torch.uint8)
            off = (-raw.data_ptr()) % 256
            self._raw = raw
            self._workspace = raw[off:off + n]
        self._held.note_use()
        return self._workspace

    def run(self, *, views, Y: Tensor, Yw: Tensor, nsh: int, D8: int, d0: int, dc: int, accumulate: bool = False,
            source_layout: int = -1, spin_limit: int = 0, fc_omega: Optional[Tensor] = None,
            waves_per_cu: int = 0) -> None:
        """Enqueue one chunk sweep of every view on torch's current stream. ``views`` holds one dict per view of the
        plan with the device tensors ``mu`` [G, N], ``phi`` [G, N, nsh, 4], ``cdir`` [8 D8, 3], and either ``T``
        [G, N, nsh, 4] or the mixed operand ``T16`` [G, N, 4, KH] with ``T_low`` [G, N, 4, 4], and optionally
        ``source`` (group-major, the same ``source_layout`` 1, 2 or 3 for every view; -1 without). Layout 3 holds the
        cell means before the first-collision sampling correction, which the sweep applies, in the sweep's block
        order (``uncollided.to_block_order``): ``source`` [G, blocks * 8 * 64 * dp + pad] (dp = D8 rounded up to 4,
        pad = ``FC_GROUP_PAD``), and ``fc_scale`` and ``fc_beta`` [G, blocks * 64 + pad] and ``fc_u`` [blocks * 64, 4]
        in each view, ``fc_omega`` [8 D8, 4] the directions. A sweep that aborts ORs its status code into the device's status word
        (``persistent_status_sink``), which the callers' next host read reports (``decode_status``); ``spin_limit``
        > 0 replaces the default bound of each dependency wait; ``waves_per_cu`` > 0 caps the resident waves per CU
        (0: the kernel's occupancy)."""
        if len(views) != len(self.shapes):
            raise ValueError(f"the plan has {len(self.shapes)} views, got {len(views)}")
        if nsh not in PERSISTENT_MOMENTS:
            raise ValueError(f"the persistent sweep is built for {PERSISTENT_MOMENTS} moments, got {nsh}")
        if source_layout not in (-1, 1, 2, 3):
            raise ValueError("source_layout must be -1, 1, 2 or 3")
        if (source_layout == 3) != (fc_omega is not None):
            raise ValueError("fc_omega goes with source_layout 3, and only with it")
        if fc_omega is not None:
            _hip.require_gpu_tensor(fc_omega, "fc_omega")
            if tuple(fc_omega.shape) != (8 * D8, 4) or not fc_omega.is_contiguous():
                raise ValueError(f"fc_omega must be a contiguous float32 [{8 * D8}, 4] tensor")
        G, M, kh = self.G, 8 * D8, 16 * ((nsh - 4 + 15) // 16)
        arr = (_hip.SweepView * len(views))()
        for v, (spec, shape) in enumerate(zip(views, self.shapes)):
            N = shape[0] * shape[1] * shape[2]
            for name in ("mu", "phi", "cdir"):
                _hip.require_gpu_tensor(spec[name], f"views[{v}].{name}")
            if spec["mu"].numel() != G * N or spec["phi"].numel() != G * N * nsh * 4:
                raise ValueError(f"views[{v}]: mu or phi size does not match the grid {shape}")
            if tuple(spec["cdir"].shape) != (M, 3):
                raise ValueError(f"views[{v}]: cdir must be [{M}, 3]")
            T, T16, tlo, src = spec.get("T"), spec.get("T16"), spec.get("T_low"), spec.get("source")
            if T16 is not None:
                if T is not None or tlo is None or T16.dtype != torch.bfloat16 or not T16.is_contiguous() \
                        or T16.numel() != G * N * 4 * kh or tlo.numel() != G * N * 16:
                    raise ValueError(f"views[{v}]: T16 must be contiguous bfloat16 [G, N, 4, {kh}] with T_low "
                                     "[G, N, 4, 4]")
                _hip.require_gpu_tensor(tlo, f"views[{v}].T_low")
            else:
                if T is None:
                    raise ValueError(f"views[{v}]: need T or T16")
                _hip.require_gpu_tensor(T, f"views[{v}].T")
                if T.numel() != G * N * nsh * 4:
                    raise ValueError(f"views[{v}]: T size does not match [G, N, nsh, 4]")
            if (src is None) != (source_layout == -1):
                raise ValueError(f"views[{v}]: source does not match source_layout {source_layout}")
            nb = int(np.prod([-(-int(n) // 4) for n in shape]))  # blocks of the block order (layout 3)
            if src is not None:
                _hip.require_gpu_tensor(src, f"views[{v}].source")
                want = G * (nb * 8 * 64 * (-(-D8 // 4) * 4) + _hip.FC_GROUP_PAD) if source_layout == 3 else \
                    G * N * M * (4 if source_layout == 1 else 1)
                if src.numel() != want or not src.is_contiguous():
                    raise ValueError(f"views[{v}]: source size does not match layout {source_layout}")
            fc = [spec.get(k) for k in ("fc_scale", "fc_beta", "fc_u")]
            if source_layout == 3:
                cells = G * (nb * 64 + _hip.FC_GROUP_PAD)
                for name, t, n in zip(("fc_scale", "fc_beta", "fc_u"), fc, (cells, cells, 4 * nb * 64)):
                    if t is None:
                        raise ValueError(f"views[{v}]: source_layout 3 needs {name}")
                    _hip.require_gpu_tensor(t, f"views[{v}].{name}")
                    if t.numel() != n or not t.is_contiguous():
                        raise ValueError(f"views[{v}]: {name} must be a contiguous float32 tensor of {n} values")
            else:
                fc = [None, None, None]
            arr[v] = _hip.SweepView(mu_t=spec["mu"].data_ptr(), T=_hip.ptr(T), T_bf16=_hip.ptr(T16),
                                    T_low=_hip.ptr(tlo), source=_hip.ptr(src), phi=spec["phi"].data_ptr(),
                                    cdir=spec["cdir"].data_ptr(), fc_scale=_hip.ptr(fc[0]), fc_beta=_hip.ptr(fc[1]),
                                    fc_u=_hip.ptr(fc[2]))
        if Y.shape != (M, nsh) or Yw.shape != (M, nsh):
            raise ValueError("direction tables do not match the quadrature")
        ws = self.workspace(dc)
        if spin_limit < 0:
            raise ValueError("spin_limit must be >= 0")
        if waves_per_cu < 0:
            raise ValueError("waves_per_cu must be >= 0")
        args = _hip.PersistentArgs(G=G, nsh=nsh, D8=D8, d0=d0, dc=dc, accumulate=int(accumulate),
                                   source_layout=int(source_layout), Y=Y.data_ptr(), Yw=Yw.data_ptr(), views=arr,
                                   status=persistent_status_sink(self.device).data_ptr(), spin_limit=int(spin_limit),
                                   fc_omega=_hip.ptr(fc_omega), waves_per_cu=int(waves_per_cu))
        lib = _hip.library()
        with torch.cuda.device(self.device):
            _hip.check(lib.citrus_persistent_sweep_run(self.handle, ctypes.byref(args), ws.data_ptr(), ws.numel(),
                                                       _hip.stream_of(self.device)), "citrus_persistent_sweep_run")

    def check(self) -> tuple[int, int]:
        """(status, unfinished tasks) of the last sweep; synchronizes torch's current stream. (0, 0) when it ran
        through."""
        if self._workspace is None:
            raise RuntimeError("no sweep has run on this schedule")
        status, unfinished = ctypes.c_int(), ctypes.c_int64()
        with torch.cuda.device(self.device):
            _hip.check(_hip.library().citrus_persistent_check(self.handle, self._workspace.data_ptr(),
                                                              ctypes.byref(status), ctypes.byref(unfinished),
                                                              _hip.stream_of(self.device)), "citrus_persistent_check")
        return int(status.value), int(unfinished.value)


_PERSISTENT: dict = {}
_PERSISTENT_CHECK: dict[str, tuple[bool, str]] = {}
_STATUS_SINKS: dict[str, Tensor] = {}


def persistent_status_sink(device) -> Tensor:
    """[1] int32 device word per device into which every persistent sweep that aborts ORs its status code. The
    library never clears it: a step's single host read (``fold_status`` / ``decode_status``) sees the aborts of every
    chunk and sweep since the last read."""
    device = torch.device(device)
    key = str(device)
    sink = _STATUS_SINKS.get(key)
    if sink is None:
        sink = _STATUS_SINKS[key] = torch.zeros(1, device=device, dtype=torch.int32)
    return sink


def fold_status(values: Tensor) -> Tensor:
    """``values`` flattened to float64 with the device's persistent-sweep status word appended (0 when no persistent
    sweep ran there), for one host read of both (no synchronization here)."""
    flat = values.reshape(-1).to(torch.float64)
    sink = _STATUS_SINKS.get(str(values.device))
    tail = sink.to(torch.float64) if sink is not None else torch.zeros(1, device=values.device, dtype=torch.float64)
    return torch.cat([flat, tail])


def decode_status(host_values, device) -> list:
    """The payload of a ``fold_status`` read; raises ``RuntimeError`` naming the abort when the status is nonzero
    (after clearing the device word, so the next step starts clean). A persistent sweep that aborted left partial
    moments, so nothing computed from it may be used."""
    values = list(host_values)
    status = int(values[-1]) if values else 0
    if status:
        sink = _STATUS_SINKS.get(str(torch.device(device)))
        if sink is not None:
            sink.zero_()
        reasons = [why for code, why in PERSISTENT_STATUS.items() if code and status & code] or [f"code {status}"]
        raise RuntimeError("persistent sweep aborted (" + "; ".join(reasons) + "): its moments are incomplete")
    return values[:-1]


def check_persistent_status(device) -> None:
    """Host read of the device's status word (synchronizes); raises ``RuntimeError`` after an abort."""
    sink = _STATUS_SINKS.get(str(torch.device(device)))
    if sink is not None:
        decode_status([float(sink.item())], device)


def persistent_usable(device) -> tuple[bool, str]:
    """Whether the persistent sweep runs through on ``device``: a small two-grid problem swept persistently must
    complete every task (status 0) and equal the launch-sequence sweep bit for bit. Cached per device; never
    raises."""
    device = torch.device(device)
    key = str(device)
    if key not in _PERSISTENT_CHECK:
        ok, why = kernels_usable(device)
        if ok:
            try:
                ok, why = _persistent_self_check(device)
            except Exception as e:  # noqa: BLE001
                ok, why = False, f"persistent self-check failed ({type(e).__name__}: {str(e)[:300]})"
        _PERSISTENT_CHECK[key] = (ok, why)
    return _PERSISTENT_CHECK[key]


def _persistent_self_check(device: torch.device) -> tuple[bool, str]:
    from ..angular_quadrature import level_symmetric_quadrature
    from ..geometry import Grid
    from ..transport_sweep import make_sweep_context, sweep_grids, sweep_views

    q = level_symmetric_quadrature(4)
    gen = torch.Generator().manual_seed(13)
    items = []
    for shape in ((5, 6, 7), (8, 4, 4)):
        ctx = make_sweep_context(Grid(shape, 0.5), q, 3, device=device)
        mu = (torch.rand(2, *shape, generator=gen) + 0.3).to(device)
        T = torch.randn(2, 2, *shape, 16, 4, generator=gen).to(device)
        items.append((ctx, mu, T, None))
    outs = sweep_grids(items)
    shapes = [tuple(c.shape) for c, _, T, _ in items for _ in range(int(T.shape[0]))]
    status, unfinished = persistent_schedule(shapes, 2, device).check()
    if status or unfinished:
        return False, f"persistent sweep did not run through: {PERSISTENT_STATUS.get(status, status)}, " \
                      f"{unfinished} tasks unfinished"
    for (ctx, mu, T, _), got in zip(items, outs):
        if not torch.equal(got, sweep_views(ctx, mu, T, schedule="launches")):
            return False, "persistent sweep differs from the launch-sequence sweep"
    return True, "persistent sweep runs through and equals the launch-sequence sweep"


def persistent_schedule(shapes, G: int, device: torch.device) -> PersistentSchedule:
    """The cached ``PersistentSchedule`` of a list of view grid shapes."""
    key = (tuple(tuple(int(v) for v in s) for s in shapes), int(G), str(device))
    sched = _PERSISTENT.get(key)
    if sched is None:
        sched = _PERSISTENT[key] = PersistentSchedule(shapes, G, device)
    return sched


def hip_available() -> tuple[bool, str]:
    """Whether the HIP library loads (honouring the ``CITRUS_DISABLE_HIP`` kill switch)."""
    if os.environ.get(_hip.DISABLE_ENV_VAR, "") not in ("", "0"):
        return False, f"disabled by {_hip.DISABLE_ENV_VAR}"
    try:
        _hip.library()
    except RuntimeError as e:
        return False, str(e)
    if not torch.cuda.is_available():
        return False, "no GPU device"
    return True, "HIP library loaded"


def kernels_usable(device) -> tuple[bool, str]:
    """Whether the HIP sweep is usable on ``device``: library loaded, gfx950 device and the self-check passed.

    The verdict is cached per device; the function never raises.
    """
    ok, why = hip_available()
    if not ok:
        return ok, why
    device = torch.device(device)
    if device.type != "cuda":
        return False, "not a GPU device"
    key = str(device)
    if key not in _SELF_CHECK:
        try:
            arch = torch.cuda.get_device_properties(device).gcnArchName
            if not arch.startswith("gfx950"):
                _SELF_CHECK[key] = (False, f"device is {arch}, the library is built for gfx950")
            else:
                _SELF_CHECK[key] = _self_check(device)
        except Exception as e:  # noqa: BLE001
            _SELF_CHECK[key] = (False, f"self-check failed ({type(e).__name__}: {str(e)[:300]})")
    return _SELF_CHECK[key]


def _self_check(device: torch.device) -> tuple[bool, str]:
    """Particle balance of every direction and group of a small random problem, evaluated in float64."""
    from ..angular_quadrature import product_quadrature
    from ..geometry import Grid
    from ..transport_sweep import _sweep_unchecked, make_sweep_context

    grid = Grid((3, 4, 5), 0.7, 1.3, 0.9)
    context = make_sweep_context(grid, product_quadrature(4, 8), 1, device=device, dtype=torch.float32)
    G = 3
    gen = torch.Generator().manual_seed(11)
    mu = (torch.rand(G, *grid.shape, generator=gen) * 2.0 + 0.1)
    src = torch.randn(context.num_directions, G, 4, *grid.shape, generator=gen)
    src[:, :, 0] = src[:, :, 0].abs() + 0.1
    _, psi = _sweep_unchecked(context, mu.to(device), lambda idx: src.to(device).index_select(0, idx),
                              direction_chunk=3, return_psi=True)
    psi = psi.double().cpu()
    mu64, src64 = mu.double(), src.double()
    dx, dy, dz = grid.deltas
    V = dx * dy * dz
    omega = context.quadrature.omega.double().cpu()
    worst = 0.0
    for m in range(context.num_directions):
        out = torch.zeros(G, dtype=torch.float64)
        for a, dim in ((0, 2), (1, 1), (2, 3)):  # axis -> dim of [G, R, C, S]
            s = 1.0 if omega[m, a] >= 0 else -1.0
            face = (psi[m, :, 0] + s * psi[m, :, 1 + a]).select(dim, -1 if s > 0 else 0)
            out += abs(float(omega[m, a])) * (V / grid.deltas[a]) * face.reshape(G, -1).sum(1)
        lhs = out + V * (mu64 * psi[m, :, 0]).reshape(G, -1).sum(1)
        rhs = V * src64[m, :, 0].reshape(G, -1).sum(1)
        worst = max(worst, float(((lhs - rhs).abs() / rhs.abs().clamp_min(1e-30)).max()))
    if not worst <= SELF_CHECK_REL_TOL:
        return False, f"self-check: particle balance of the HIP sweep off by {worst:.2e} (rel)"
    return True, f"HIP sweep passes the particle-balance self-check (max rel dev {worst:.1e})"
