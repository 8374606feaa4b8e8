"""Compact per-view accumulators and the scan-level CITRUS step.

``StateStore`` keeps the interface of the reference's ``StateStore`` (``citrus/citrus.py:623-757``): one accumulator
per view slot, stored as its linear-discontinuous cell mean in float16 with one float32 scale per (group, moment)
(``scale = max|mean| / 2^14``), returned by ``get`` as float32 with zero slopes. Packing and unpacking run in HIP
kernels (``citrus_state_pack`` / ``citrus_state_unpack``) straight from the sweep's group-major moments (no copy of
the sweep's output), and ``put_views`` packs a batch of group-major moments in one launch. On an MI355X (288 GB of
HBM) the default
``location="auto"`` keeps every slot resident on the GPU up to ``gpu_budget_mib``; ``"cpu"`` keeps them in pinned host
memory (packed on the GPU, copied down asynchronously).

``citrus_scan_step`` is the per-outer-iteration scatter stage over all views of a scan
(``scatter_stage``, ``citrus/citrus.py:1907-1970``, restricted to the age-ordered recursion): views in batches, each
batch's accumulators from the store, one batched ``citrus_views_step``, the new accumulators back into the store; it
returns the scatter and primary signals and their ratio. ``scatter_corrected_counts`` is the correction the outer loop
applies to the measured counts, ``Y / (1 + SPR)`` (``citrus/citrus.py:1506``).
"""

from __future__ import annotations

from dataclasses import replace
from typing import Optional, Sequence

import torch
from torch import Tensor

from . import _hip

FLOAT16_TARGET = 2.0 ** 14  # the per-(group, moment) maximum maps here (citrus.py:620)


def _group_major(acc: Tensor) -> Tensor:
    """[G, N, 4, R, C, S] accumulator -> contiguous float32 [G, R, C, S, N, 4] (no copy for the sweep's own views)."""
    t = acc.permute(0, 3, 4, 5, 1, 2)
    if t.dtype != torch.float32 or not t.is_contiguous():
        t = t.to(torch.float32).contiguous()
    return t


def pack_accumulators(moments: Tensor) -> tuple[Tensor, Tensor]:
    """Group-major moments [V, G, R, C, S, nsh, 4] (float32, GPU, contiguous) -> (q [V, G, nsh, R, C, S] float16,
    scale [V, G, nsh] float32): the compact form of every view's cell means."""
    _hip.require_gpu_tensor(moments, "moments")
    if moments.dim() != 7 or int(moments.shape[-1]) != 4:
        raise ValueError(f"moments must be [V, G, R, C, S, nsh, 4], got {tuple(moments.shape)}")
    V, G, R, C, S, nsh, _ = (int(n) for n in moments.shape)
    N = R * C * S
    lib = _hip.library()
    ws = torch.empty(int(lib.citrus_state_pack_workspace_floats(V, G, nsh, N)), device=moments.device,
                     dtype=torch.float32)
    q = torch.empty(V, G, nsh, R, C, S, device=moments.device, dtype=torch.float16)
    scale = torch.empty(V, G, nsh, device=moments.device, dtype=torch.float32)
    with torch.cuda.device(moments.device):
        _hip.check(lib.citrus_state_pack(moments.data_ptr(), V, G, nsh, N, scale.data_ptr(), q.data_ptr(),
                                         ws.data_ptr(), ws.numel(), _hip.stream_of(moments.device)),
                   "citrus_state_pack")
    return q, scale


def unpack_accumulators(q: Tensor, scale: Tensor, out: Optional[Tensor] = None) -> Tensor:
    """(q [V, G, nsh, R, C, S] float16, scale [V, G, nsh]) -> group-major moments [V, G, R, C, S, nsh, 4] float32
    with the cell means ``q * scale`` and zero LD slopes."""
    if q.dtype != torch.float16 or q.device.type != "cuda" or not q.is_contiguous() or q.dim() != 6:
        raise ValueError("q must be a contiguous float16 [V, G, nsh, R, C, S] tensor on the GPU")
    _hip.require_gpu_tensor(scale, "scale")
    V, G, nsh, R, C, S = (int(n) for n in q.shape)
    if tuple(scale.shape) != (V, G, nsh):
        raise ValueError(f"scale must be [{V}, {G}, {nsh}], got {tuple(scale.shape)}")
    if out is None:
        out = torch.empty(V, G, R, C, S, nsh, 4, device=q.device, dtype=torch.float32)
    elif tuple(out.shape) != (V, G, R, C, S, nsh, 4) or not out.is_contiguous() or out.dtype != torch.float32:
        raise ValueError("out must be a contiguous float32 [V, G, R, C, S, nsh, 4] tensor")
    lib = _hip.library()
    with torch.cuda.device(q.device):
        _hip.check(lib.citrus_state_unpack(q.data_ptr(), scale.data_ptr(), V, G, nsh, R * C * S, out.data_ptr(),
                                           _hip.stream_of(q.device)), "citrus_state_unpack")
    return out


class StateStore:
    """Compact storage of one CITRUS accumulator per view: the LD cell mean in scaled float16 (citrus.py:623).

    Args:
        num_slots: number of stored accumulators (views).
        full_shape: accumulator shape [G, N, 4, Rc, Cc, Sc].
        device: GPU device of the accumulators passed to ``put`` and returned by ``get``.
        location: ``"gpu"``, ``"cpu"`` (pinned host memory) or ``"auto"`` (the GPU when the total fits
            ``gpu_budget_mib``).
        gpu_budget_mib: largest total in MiB kept on the GPU by ``"auto"``.
        pinned_memory_limit_gib: largest total in GiB stored in pinned host memory.

    Attributes:
        location (str): ``"gpu"`` or ``"cpu"``.
        pinned (bool): whether the host buffers are pinned.
        stored_shape (tuple): [G, N, 1, Rc, Cc, Sc].
        transfer_seconds (float): seconds spent in put and get (measured with events on the device's stream; reading
            it waits for them).
    """

    def __init__(self, num_slots: int, full_shape: Sequence[int], device, location: str = "auto",
                 gpu_budget_mib: float = 64 * 1024.0, pinned_memory_limit_gib: float = 8.0) -> None:
        self.num_slots = int(num_slots)
        self.full_shape = tuple(int(n) for n in full_shape)
        if len(self.full_shape) != 6:
            raise ValueError(f"accumulator shape must be [G, N, 4, Rc, Cc, Sc], got {self.full_shape}")
        if self.full_shape[2] != 4:
            raise ValueError("the accumulator's third axis holds the 4 LD moments")
        self.stored_shape = self.full_shape[:2] + (1,) + self.full_shape[3:]
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise ValueError("the state store packs on the GPU: device must be a GPU")
        if location == "auto":
            location = "gpu" if self.total_bytes <= gpu_budget_mib * 2 ** 20 else "cpu"
        if location not in ("gpu", "cpu"):
            raise ValueError(f"location must be 'auto', 'gpu' or 'cpu', got {location!r}")
        self.location = location
        self.pinned = location == "cpu" and self.total_bytes <= pinned_memory_limit_gib * 2 ** 30
        self._data: list[Optional[Tensor]] = [None] * self.num_slots
        self._scale: list[Optional[Tensor]] = [None] * self.num_slots
        self._events: list = []
        self._seconds = 0.0

    @property
    def bytes_per_view(self) -> int:
        """Stored bytes of one accumulator, its scales included."""
        G, N = self.full_shape[:2]
        cells = self.full_shape[3] * self.full_shape[4] * self.full_shape[5]
        return G * N * cells * 2 + G * N * 4

    @property
    def total_bytes(self) -> int:
        return self.bytes_per_view * self.num_slots

    @property
    def transfer_seconds(self) -> float:
        for a, b in self._events:
            b.synchronize()
            self._seconds += a.elapsed_time(b) * 1e-3
        self._events = []
        return self._seconds

    def _slot(self, slot: int) -> int:
        s = int(slot)
        if not 0 <= s < self.num_slots:
            raise IndexError(f"slot {slot} outside [0, {self.num_slots})")
        return s

    def _timed(self):
        stream = torch.cuda.current_stream(self.device)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record(stream)
        return stream, a, b

    def put(self, slot: int, accumulator: Tensor) -> None:
        """Store an accumulator [G, N, 4, Rc, Cc, Sc] in a slot as its cell mean in scaled float16."""
        self.put_views([slot], [accumulator])

    def get(self, slot: int) -> Optional[Tensor]:
        """The decoded accumulator of a slot ([G, N, 4, Rc, Cc, Sc] float32 on ``device``, zero slopes), or None
        before its first ``put``."""
        return self.get_views([slot])[0]

    def put_views(self, slots: Sequence[int], accumulators) -> None:
        """Store several accumulators: a sequence of [G, N, 4, Rc, Cc, Sc] tensors (the sweep's own permuted views
        are packed in place, one launch each), or group-major moments [V, G, Rc, Cc, Sc, N, 4] (one launch)."""
        slots = [self._slot(s) for s in slots]
        G, N = self.full_shape[:2]
        if isinstance(accumulators, Tensor) and accumulators.dim() == 7:
            want = (G,) + self.full_shape[3:] + (N, 4)
            if tuple(accumulators.shape[1:]) != want:
                raise ValueError(f"moments must be [V, *{want}], got {tuple(accumulators.shape)}")
            batches = [accumulators.to(self.device, torch.float32).contiguous()]
        else:
            batches = []
            for acc in accumulators:
                if tuple(acc.shape) != self.full_shape:
                    raise ValueError(f"accumulator must be {self.full_shape}, got {tuple(acc.shape)}")
                batches.append(_group_major(acc.to(self.device))[None])
        if sum(int(m.shape[0]) for m in batches) != len(slots):
            raise ValueError(f"{len(slots)} slots for {sum(int(m.shape[0]) for m in batches)} accumulators")
        stream, a, b = self._timed()
        j = 0
        for moments in batches:
            q, scale = pack_accumulators(moments)
            for v in range(int(moments.shape[0])):
                s = slots[j]
                j += 1
                qs = q[v].reshape(self.stored_shape)
                sc = scale[v].reshape(G, N, 1, 1, 1, 1)
                if self.location == "gpu":
                    self._data[s], self._scale[s] = qs, sc
                    continue
                if self._data[s] is None:
                    self._data[s] = torch.empty(self.stored_shape, dtype=torch.float16, pin_memory=self.pinned)
                    self._scale[s] = torch.empty(G, N, 1, 1, 1, 1, dtype=torch.float32, pin_memory=self.pinned)
                self._data[s].copy_(qs, non_blocking=self.pinned)
                self._scale[s].copy_(sc, non_blocking=self.pinned)
        b.record(stream)
        if self.location == "cpu":  # the host buffers hold the data before anything reuses them
            b.synchronize()
        self._events.append((a, b))

    def get_views(self, slots: Sequence[int]) -> list:
        """Decoded accumulators of several slots ([G, N, 4, Rc, Cc, Sc] float32 views of group-major storage, zero
        slopes; None for a slot never put), one unpacking launch per slot."""
        slots = [self._slot(s) for s in slots]
        G, N = self.full_shape[:2]
        cells = self.full_shape[3:]
        out = []
        stream, a, b = self._timed()
        for s in slots:
            if self._data[s] is None:
                out.append(None)
                continue
            q = self._data[s].to(self.device, non_blocking=self.pinned).reshape(1, G, N, *cells)
            scale = self._scale[s].to(self.device, non_blocking=self.pinned).reshape(1, G, N)
            out.append(unpack_accumulators(q, scale)[0].permute(0, 4, 5, 1, 2, 3))
        b.record(stream)
        self._events.append((a, b))
        return out

    def stored_views(self, slots: Sequence[int]) -> list:
        """The stored form of several slots on the device, without decoding: (q [1, G, N, Rc, Cc, Sc] float16, scale
        [1, G, N] float32) per slot, or None for a slot never put."""
        slots = [self._slot(s) for s in slots]
        G, N = self.full_shape[:2]
        cells = self.full_shape[3:]
        out = []
        for s in slots:
            if self._data[s] is None:
                out.append(None)
                continue
            q = self._data[s].to(self.device, non_blocking=self.pinned).reshape(1, G, N, *cells)
            scale = self._scale[s].to(self.device, non_blocking=self.pinned).reshape(1, G, N)
            out.append((q, scale))
        return out

    def close(self) -> None:
        """Release the stored accumulators."""
        self._data = [None] * self.num_slots
        self._scale = [None] * self.num_slots

    def describe(self) -> dict:
        """``location``, ``pinned``, ``stored_shape``, ``full_shape``, ``bytes_per_view``, ``num_views``,
        ``total_gib`` (citrus.py:746-757)."""
        return {"location": self.location, "pinned": self.pinned, "stored_shape": list(self.stored_shape),
                "full_shape": list(self.full_shape), "bytes_per_view": self.bytes_per_view,
                "num_views": self.num_slots, "total_gib": self.total_bytes / 2 ** 30}


def citrus_scan_step(solver, store: StateStore, states: Sequence, fields, angles: Sequence[float], *,
                     views_per_batch: int = 4, primary: bool = True, contraction: str = "fp32",
                     schedule: str = "auto"):
    """One scatter stage over all views of a scan (the age-ordered recursion of ``scatter_stage``,
    citrus.py:1907-1935): views in batches of ``views_per_batch``, each view's accumulator from ``store``, one batched
    ``citrus_views_step`` per batch, the new accumulators back into ``store``. ``states`` are the views'
    ``CitrusState`` without their accumulators (the store holds those).

    Returns:
        (scatter [V, Rd, Cd], primary [V, Rd, Cd] or None, ratio scatter / primary or None, new states).
    """
    from .citrus import citrus_views_step

    V = len(angles)
    if len(states) != V or store.num_slots < V:
        raise ValueError(f"need one state and one store slot per view: {len(states)} states, {store.num_slots} "
                         f"slots, {V} views")
    if isinstance(views_per_batch, bool) or int(views_per_batch) < 1:
        raise ValueError("views_per_batch must be >= 1")
    B = int(views_per_batch)
    scatter, prim, new_states = [None] * V, [None] * V, list(states)
    for b0 in range(0, V, B):
        idx = list(range(b0, min(V, b0 + B)))
        stored = store.stored_views(idx)
        if all(x is not None for x in stored) and contraction == "fp32":
            batch = _states_from_store(store, [states[i] for i in idx], stored, fields)
        else:
            accs = store.get_views(idx)
            batch = [replace(states[i], collided_moments=accs[j], scatter_source=None, scatter_source_fields=None,
                             _batch=None) for j, i in enumerate(idx)]
        results, stepped = citrus_views_step(solver, batch, fields, [angles[i] for i in idx], primary=primary,
                                             contraction=contraction, schedule=schedule)
        store.put_views(idx, [s.collided_moments for s in stepped])
        for j, i in enumerate(idx):
            scatter[i], prim[i] = results[j]
            new_states[i] = replace(stepped[j], collided_moments=None, scatter_source=None,
                                    scatter_source_fields=None, _batch=None)
    scatter = torch.stack(scatter)
    if not primary:
        return scatter, None, None, new_states
    primary_t = torch.stack(prim)
    tiny = torch.finfo(torch.float32).tiny
    return scatter, primary_t, scatter / primary_t.clamp_min(tiny), new_states


def _states_from_store(store: StateStore, states: Sequence, stored: list, fields) -> list:
    """The views' states for one batched step straight from their stored accumulators: the step's scattering-source
    moments formed from the stored cell means by one kernel per view (``scatter_moments_stored``: bit for bit those
    of the unpacked accumulators, which are never formed), shared by the batch as its previous step's moments, and
    as the accumulator only the decoded scalar-flux cell means [G, 1, 1, Rc, Cc, Sc] that the step's change norms
    read."""
    from .scatter_solver.source_iteration import scatter_moments_stored

    G, N = store.full_shape[:2]
    cells = store.full_shape[3:]
    stream, a, b = store._timed()
    T = torch.empty((len(states), G) + tuple(cells) + (N, 4), device=store.device, dtype=torch.float32)
    out = []
    for j, (st, (q, scale)) in enumerate(zip(states, stored)):
        scatter_moments_stored(fields.transfer_coarse, q, scale, out=T[j:j + 1])
        means = (q[0, :, 0].float() * scale[0, :, 0].reshape(G, 1, 1, 1))[:, None, None]  # as citrus_state_unpack
        out.append(replace(st, collided_moments=means, scatter_source=T[j], scatter_source_fields=fields,
                           _batch=(T, j)))
    b.record(stream)
    store._events.append((a, b))
    return out


def scatter_corrected_counts(counts: Tensor, spr: Optional[Tensor]) -> Tensor:
    """Measured counts with the estimated scatter removed, ``Y / (1 + SPR)`` (citrus.py:1506); ``spr`` None leaves
    them as they are."""
    return counts if spr is None else counts / (1.0 + spr)
