"""Autotuning of the batched sweep's launch shape.

The matrix-core sweep takes a direction chunk (directions per octant per pass, at most 32): all D8 directions at
once, or fewer, larger-occupancy passes repeated per chunk. It runs under one of two schedules with identical
results: a sequence of block-diagonal launches, or one persistent launch per chunk (P3 and P5). The best pair depends
on the grid, the groups, the views per launch and the contraction, so it is measured: each candidate sweeps synthetic
moments of the real shape (one warm-up, then the fastest of ``repeats`` timed runs with HIP events on torch's stream),
and the fastest is kept per configuration for the life of the process. For the persistent schedule the resident waves
per CU are measured as well (``best_waves``: the kernel's occupancy or a cap below it; the moments do not depend on
them). The tile of the cell planes (4^3 blocks) and one group per wave are fixed by the kernel.
"""

from __future__ import annotations

from typing import Callable, Optional, Sequence

import torch
from torch import Tensor


def chunk_candidates(D8: int) -> list[int]:
    """Direction chunks worth timing: all directions of an octant (when <= 32), about half and a third of them, and 8."""
    out = []
    for c in (D8, (D8 + 1) // 2, (D8 + 2) // 3, 8):
        c = min(c, 32)
        if 1 <= c <= D8 and c not in out:
            out.append(c)
    return out


PERSISTENT_MARGIN = 1.03  # the persistent schedule is chosen when within 3% of the fastest configuration
WAVE_CANDIDATES = (0, 6, 4)  # resident waves per CU of the persistent schedule: 0 is the kernel's occupancy


def pick_config(times: dict) -> tuple[str, int]:
    """The (schedule, chunk) to use from measured milliseconds per configuration: the fastest, except that the
    fastest persistent configuration wins when within ``PERSISTENT_MARGIN`` of it. The persistent schedule also reads
    deferred first-collision samples (the launch sequence corrects them in a pass of its own first) and needs no
    per-chunk launches, so a tie on the sweep alone goes to it."""
    if not times:
        raise ValueError("no configuration was measured")
    choice = min(times, key=times.get)
    persistent = [k for k in times if k[0] == "persistent"]
    if persistent:
        fastest = min(persistent, key=times.get)
        if times[fastest] <= PERSISTENT_MARGIN * times[choice]:
            choice = fastest
    return choice


class SweepTuner:
    """Fastest direction chunk of the launch-sequence ``sweep_views`` (``best_chunk``) and fastest (schedule, chunk)
    over both schedules (``best_config``) per (grid shape, moments, directions, groups, views, contraction)."""

    def __init__(self, repeats: int = 2):
        self.repeats = int(repeats)
        self.cache: dict = {}
        self.timings: dict = {}

    def key(self, context, G: int, V: int, contraction: str) -> tuple:
        return (tuple(context.shape), context.num_moments, context.directions_per_octant, int(G), int(V), contraction)

    def best_chunk(self, context, mu_t: Tensor, V: int, contraction: str = "fp32",
                   candidates: Optional[Sequence[int]] = None,
                   run: Optional[Callable[[int], None]] = None) -> int:
        """The fastest chunk for this configuration (measured on first use, then cached)."""
        G = int(mu_t.shape[0])
        key = self.key(context, G, V, contraction)
        if key in self.cache:
            return self.cache[key]
        from .transport_sweep import sweep_views

        cands = list(candidates) if candidates is not None else chunk_candidates(context.directions_per_octant)
        if run is None:
            R, C, S = context.shape
            T = torch.zeros(V, G, R, C, S, context.num_moments, 4, device=mu_t.device, dtype=torch.float32)
            T[..., 0, 0] = 1.0
            out = torch.empty_like(T)

            def run(chunk: int) -> None:
                sweep_views(context, mu_t, T, out=out, direction_chunk=chunk, contraction=contraction,
                            schedule="launches")

        stream = torch.cuda.current_stream(mu_t.device)
        times = {}
        for c in cands:
            run(c)  # warm-up
            best = float("inf")
            for _ in range(self.repeats):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                run(c)
                b.record(stream)
                b.synchronize()
                best = min(best, a.elapsed_time(b))
            times[c] = best
        choice = min(times, key=times.get)
        self.cache[key] = choice
        self.timings[key] = times
        return choice

    def best_config(self, context, mu_t: Tensor, V: int, contraction: str = "fp32",
                    run: Optional[Callable[[str, int], None]] = None) -> tuple[str, int]:
        """The fastest (schedule, chunk) for this configuration, over both schedules (measured once, then cached)."""
        G = int(mu_t.shape[0])
        key = ("config",) + self.key(context, G, V, contraction)
        if key in self.cache:
            return self.cache[key]
        from .kernels.sweep_kernels import PERSISTENT_MAX_VIEWS, PERSISTENT_MOMENTS, persistent_usable
        from .transport_sweep import sweep_views

        chunks = chunk_candidates(context.directions_per_octant)
        cands = [("launches", c) for c in chunks]
        if context.num_moments in PERSISTENT_MOMENTS and V <= PERSISTENT_MAX_VIEWS and \
                persistent_usable(mu_t.device)[0]:
            cands += [("persistent", c) for c in chunks]
        if run is None:
            R, C, S = context.shape
            T = torch.zeros(V, G, R, C, S, context.num_moments, 4, device=mu_t.device, dtype=torch.float32)
            T[..., 0, 0] = 1.0
            out = torch.empty_like(T)

            def run(schedule: str, chunk: int) -> None:
                sweep_views(context, mu_t, T, out=out, direction_chunk=chunk, contraction=contraction,
                            schedule=schedule)

        stream = torch.cuda.current_stream(mu_t.device)
        times = {}
        for sched, c in cands:
            run(sched, c)  # warm-up
            best = float("inf")
            for _ in range(self.repeats):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                run(sched, c)
                b.record(stream)
                b.synchronize()
                best = min(best, a.elapsed_time(b))
            times[(sched, c)] = best
        choice = pick_config(times)
        self.cache[key] = choice
        self.timings[key] = times
        return choice

    def best_waves(self, context, mu_t: Tensor, V: int, chunk: int, contraction: str = "fp32",
                   candidates: Sequence[int] = WAVE_CANDIDATES,
                   run: Optional[Callable[[int], None]] = None) -> int:
        """The fastest resident-wave cap of the persistent schedule at this chunk (measured once, then cached); ties
        go to the occupancy (0)."""
        G = int(mu_t.shape[0])
        key = ("waves", int(chunk)) + self.key(context, G, V, contraction)
        if key in self.cache:
            return self.cache[key]
        from .transport_sweep import sweep_views

        if run is None:
            R, C, S = context.shape
            T = torch.zeros(V, G, R, C, S, context.num_moments, 4, device=mu_t.device, dtype=torch.float32)
            T[..., 0, 0] = 1.0
            out = torch.empty_like(T)

            def run(waves: int) -> None:
                sweep_views(context, mu_t, T, out=out, direction_chunk=chunk, contraction=contraction,
                            schedule="persistent", waves_per_cu=waves)

        stream = torch.cuda.current_stream(mu_t.device)
        times = {}
        for w in candidates:
            run(w)  # warm-up
            best = float("inf")
            for _ in range(self.repeats):
                a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                a.record(stream)
                run(w)
                b.record(stream)
                b.synchronize()
                best = min(best, a.elapsed_time(b))
            times[int(w)] = best
        fastest = min(times, key=times.get)
        choice = 0 if 0 in times and times[0] <= PERSISTENT_MARGIN * times[fastest] else fastest
        self.cache[key] = choice
        self.timings[key] = times
        return choice


_TUNER = SweepTuner()


def default_tuner() -> SweepTuner:
    """The process-wide tuner ``sweep_views(direction_chunk="auto")`` consults."""
    return _TUNER
