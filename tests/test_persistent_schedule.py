"""The persistent sweep's schedule, checked on the CPU through ``citrus_persistent_schedule_table`` (the host build of
the queues and of the dependency rule the kernel waits on):

- every (view, group, octant, block) task appears once, and each (view, group) lives in one XCD queue;
- every dependency precedes its task in the same queue, so the smallest unfinished ticket can always run (no wave can
  wait on a later ticket: deadlock freedom for any number of resident waves);
- under random interleavings of task starts (faces and moments read) and finishes (faces and moments written) that
  the dependencies allow, every task reads the faces of its own upstream task, no two visits of a block overlap, and
  each block's octants add into its moments in the launch sequence's order (ld_sweep.hip ``citrus_sweep_plan_create``:
  pass p sweeps octants p and 7-p, the one whose block plane comes first visits first, octant p on the middle plane).
"""

from __future__ import annotations

import ctypes
import random

import numpy as np
import pytest

from citrus_amd import _hip

SHAPES = [
    [(5, 6, 7)],                              # 2x2x2 blocks, odd K
    [(9, 13, 6)],                             # 3x4x2 blocks, even K: split middle plane
    [(8, 8, 8), (4, 12, 6), (7, 3, 10)],      # three grids of different shapes in one plan
]


def schedule_table(shapes, G):
    if not _hip.LIBRARY_PATH.exists():
        pytest.skip("library not built")
    lib = _hip.library()
    flat = (ctypes.c_int * (3 * len(shapes)))(*[v for s in shapes for v in s])
    n = int(lib.citrus_persistent_schedule_table(len(shapes), flat, G, None, None, None))
    assert n > 0
    tasks = np.zeros((n, 4), np.int32)
    qoff = np.zeros(9, np.int32)
    deps = np.zeros((n, 7), np.int32)
    assert lib.citrus_persistent_schedule_table(len(shapes), flat, G, tasks.ctypes.data, qoff.ctypes.data,
                                                deps.ctypes.data) == n
    return tasks, qoff, deps


def blocks_of(shape):
    return tuple((s + 3) // 4 for s in shape)


def decode(row):
    x = int(row[0])
    return x & 0xFF, (x >> 8) & 0xFF, (x >> 16) & 7, int(row[1]), int(row[2]), int(row[3])


def plane(o, b, nb):
    bi, bj, bk = b
    nbr, nbc, nbs = nb
    a = nbc - 1 - bj if o & 4 else bj
    c = nbr - 1 - bi if o & 2 else bi
    e = nbs - 1 - bk if o & 1 else bk
    return a + c + e


def launch_visit_order(b, nb):
    """Octants in the order the launch sequence visits block b."""
    K = sum(n - 1 for n in nb)
    order = []
    for p in range(4):
        oa, ob = p, 7 - p
        order += [oa, ob] if 2 * plane(oa, b, nb) <= K else [ob, oa]
    return order


def signs(o):
    return (-1 if o & 4 else 1, -1 if o & 2 else 1, -1 if o & 1 else 1)  # x (cols), y (rows), z (slices)


def axis_neighbours(b, o, nb, sign):
    """Neighbour along x, y, z in direction sign * (octant's sweep direction), or None outside."""
    sx, sy, sz = signs(o)
    bi, bj, bk = b
    out = []
    for di, dj, dk in ((0, sign * sx, 0), (sign * sy, 0, 0), (0, 0, sign * sz)):
        n = (bi + di, bj + dj, bk + dk)
        out.append(n if all(0 <= n[a] < nb[a] for a in range(3)) else None)
    return out


@pytest.mark.parametrize("shapes", SHAPES)
@pytest.mark.parametrize("G", [1, 3])
def test_tasks_once_and_queues_topological(shapes, G):
    tasks, qoff, deps = schedule_table(shapes, G)
    n = len(tasks)
    expect = sum(8 * G * np.prod(blocks_of(s)) for s in shapes)
    assert n == expect and qoff[0] == 0 and qoff[8] == n and np.all(np.diff(qoff) >= 0)
    seen = set()
    queue_of_domain = {}
    for q in range(8):
        for i in range(qoff[q], qoff[q + 1]):
            v, g, o, bi, bj, bk = decode(tasks[i])
            key = (v, g, o, bi, bj, bk)
            assert key not in seen
            seen.add(key)
            assert queue_of_domain.setdefault((v, g), q) == q
            for d in deps[i]:
                if d >= 0:
                    assert qoff[q] <= d < i, (i, d)  # earlier ticket of the same queue
                    assert decode(tasks[d])[:2] == (v, g)
    assert len(queue_of_domain) == len(shapes) * G
    if len(shapes) * G >= 8:  # every XCD gets work
        assert np.all(np.diff(qoff) > 0)


def simulate(shapes, G, seed):
    """Random interleaving of task starts and finishes allowed by the dependencies, checking every read."""
    tasks, qoff, deps = schedule_table(shapes, G)
    n = len(tasks)
    info = [decode(t) for t in tasks]
    rng = random.Random(seed)
    done = np.zeros(n, bool)
    started = np.zeros(n, bool)
    in_flight = set()
    busy_blocks = set()
    face = {}      # (v, g, slot, block, axis) -> writer (octant, block)
    visits = {}    # (v, g, block) -> octants in order of their finishes
    nbs = [blocks_of(s) for s in shapes]

    def ready(i):
        return not started[i] and all(d < 0 or done[d] for d in deps[i])

    pending = set(range(n))
    while pending or in_flight:
        starts = [i for i in pending if ready(i)]
        events = [("s", i) for i in starts] + [("f", i) for i in in_flight]
        assert events, "no task can start or finish: the dependencies deadlock"
        kind, i = rng.choice(events)
        v, g, o, bi, bj, bk = info[i]
        b, nb = (bi, bj, bk), nbs[v]
        slot = 0 if o < 4 else 1
        if kind == "s":
            assert (v, g, b) not in busy_blocks, "two visits of one block in flight"
            busy_blocks.add((v, g, b))
            for ax, ub in enumerate(axis_neighbours(b, o, nb, -1)):
                if ub is not None:
                    assert face.get((v, g, slot, ub, ax)) == (o, ub), f"task {i} reads a stale or foreign face"
            started[i] = True
            pending.discard(i)
            in_flight.add(i)
        else:
            for ax, db in enumerate(axis_neighbours(b, o, nb, +1)):
                if db is not None:
                    face[(v, g, slot, b, ax)] = (o, b)
            visits.setdefault((v, g, b), []).append(o)
            busy_blocks.discard((v, g, b))
            in_flight.discard(i)
            done[i] = True
    for (v, g, b), order in visits.items():
        assert order == launch_visit_order(b, nbs[v]), (v, g, b, order)
    assert len(visits) == sum(G * np.prod(blocks_of(s)) for s in shapes)


@pytest.mark.parametrize("shapes", SHAPES)
def test_random_interleavings_have_no_hazard(shapes):
    for seed in range(3):
        simulate(shapes, 2, seed)


def test_ticket_order_with_few_waves_never_stalls():
    """Waves that take tickets in queue order and block on their task's dependencies always make progress."""
    shapes, G = [(9, 13, 6), (8, 8, 8)], 2
    tasks, qoff, deps = schedule_table(shapes, G)
    rng = random.Random(4)
    for waves in (1, 2, 5):
        done = np.zeros(len(tasks), bool)
        for q in range(8):
            nxt = qoff[q]
            holding = []
            while nxt < qoff[q + 1] or holding:
                while len(holding) < waves and nxt < qoff[q + 1]:
                    holding.append(nxt)
                    nxt += 1
                runnable = [t for t in holding if all(d < 0 or done[d] for d in deps[t])]
                assert runnable, f"queue {q} stalls with {waves} waves"
                t = rng.choice(runnable)
                done[t] = True
                holding.remove(t)
        assert done.all()


def test_schedule_table_rejects_bad_arguments():
    if not _hip.LIBRARY_PATH.exists():
        pytest.skip("library not built")
    lib = _hip.library()
    flat = (ctypes.c_int * 3)(4, 0, 4)
    assert lib.citrus_persistent_schedule_table(1, flat, 2, None, None, None) == -1
    flat = (ctypes.c_int * 3)(4, 4, 4)
    assert lib.citrus_persistent_schedule_table(0, flat, 2, None, None, None) == -1
    assert lib.citrus_persistent_schedule_table(1, flat, 0, None, None, None) == -1
    assert lib.citrus_persistent_schedule_table(17, flat, 2, None, None, None) == -1
