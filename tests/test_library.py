"""The C-ABI library loads and exports every entry point declared in include/*.h; host-side API checks (CPU)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest
import numpy as np
import torch

from citrus_amd import _hip
from citrus_amd.scatter_solver.angular_quadrature import level_symmetric_quadrature, product_quadrature
from citrus_amd.scatter_solver.geometry import Grid
from citrus_amd.scatter_solver.transport_sweep import make_sweep_context, sweep

ROOT = Path(__file__).resolve().parents[1]


def declared_functions() -> list[str]:
    names = []
    for header in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names += re.findall(r"\b(citrus_\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    if not _hip.LIBRARY_PATH.exists():
        pytest.skip("library not built")
    lib = ctypes.CDLL(str(_hip.LIBRARY_PATH))
    declared = declared_functions()
    assert declared, "no declarations found"
    missing = [n for n in declared if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_hip.exported_symbols()) == set(declared), "the ctypes bindings and the header disagree"
    assert _hip.library().citrus_abi_version() == _hip.ABI_VERSION


def test_host_side_errors_without_device():
    if not _hip.LIBRARY_PATH.exists():
        pytest.skip("library not built")
    lib = _hip.library()
    plan = ctypes.c_void_p()
    assert lib.citrus_sweep_plan_create(0, 4, 4, 4, ctypes.byref(plan)) == 1
    assert b"positive" in lib.citrus_last_error()
    assert lib.citrus_sweep_plan_create(4, 4, 4, 5, ctypes.byref(plan)) == 3
    assert lib.citrus_sweep_run(None, None, None) == 1


def test_cpu_context_and_sweep_rejects_cpu():
    grid = Grid((3, 4, 5), 0.5)
    ctx = make_sweep_context(grid, level_symmetric_quadrature(4), 2, device="cpu")
    assert ctx.num_directions == 24 and ctx.directions_per_octant == 3 and ctx.hip is None
    assert ctx.num_planes == 3 + 4 + 5 - 2
    with pytest.raises(ValueError):
        sweep(ctx, torch.ones(2, 3, 4, 5), lambda idx: torch.zeros(idx.numel(), 2, 4, 3, 4, 5))
    with pytest.raises(ValueError):
        make_sweep_context(grid, product_quadrature(4, 8), 1, device="cpu", backend="hip")
    with pytest.raises(ValueError):
        make_sweep_context(grid, product_quadrature(4, 8), 1, device="cpu", backend="triton")


def test_wavefront_plan_covers_each_entry_once():
    from citrus_amd.scatter_solver.wavefront import build_sweep_plan

    plan = build_sweep_plan((3, 4, 2))
    N = plan.num_cells
    pairs = set(zip(plan.octant.tolist(), plan.cell.tolist()))
    assert len(pairs) == 8 * N and plan.num_planes == 7
    # upstream neighbours lie on the previous plane of the same octant
    plane_of = {}
    for p in range(plan.num_planes):
        sl = plan.plane_entries(p)
        for o, c in zip(plan.octant[sl].tolist(), plan.cell[sl].tolist()):
            plane_of[(o, c)] = p
    for e in range(8 * N):
        o, c = int(plan.octant[e]), int(plan.cell[e])
        for u in plan.upstream[e].tolist():
            if u >= 0:
                assert plane_of[(o, u)] == plane_of[(o, c)] - 1


def test_product_sh_matches_oracle():
    import numpy as np

    from citrus_amd.scatter_solver.angular_quadrature import real_spherical_harmonics
    from oracle import quadrature as oq

    for q in (level_symmetric_quadrature(12), product_quadrature(8, 16)):
        Y = real_spherical_harmonics(7, q.omega).numpy()
        assert np.abs(Y - oq.real_spherical_harmonics(7, q.omega.numpy())).max() < 1e-12
    om, w = oq.level_symmetric_quadrature(12)
    q = level_symmetric_quadrature(12)
    assert np.abs(q.omega.numpy() - om).max() < 1e-15 and np.abs(q.weight.numpy() - w).max() < 1e-15


def test_status_fold_and_decode_on_the_cpu():
    """The plumbing of the persistent sweep's status word without a GPU: folded reads carry the word as their last
    value, a zero word returns the payload, a nonzero word raises RuntimeError naming the abort and clears the word."""
    from citrus_amd.scatter_solver.kernels import sweep_kernels as sk

    values = torch.tensor([[1.0, 2.0], [3.0, 4.0]], dtype=torch.float64)
    assert sk.fold_status(values).tolist() == [1.0, 2.0, 3.0, 4.0, 0.0]  # no sink on this device: status 0
    sink = torch.tensor([1], dtype=torch.int32)
    sk._STATUS_SINKS["cpu"] = sink
    try:
        folded = sk.fold_status(values).tolist()
        assert folded[-1] == 1.0
        with pytest.raises(RuntimeError, match="spin limit"):
            sk.decode_status(folded, "cpu")
        assert int(sink.item()) == 0
        assert sk.decode_status(sk.fold_status(values).tolist(), "cpu") == [1.0, 2.0, 3.0, 4.0]
        sink.fill_(3)
        with pytest.raises(RuntimeError, match="XCD"):
            sk.check_persistent_status("cpu")
    finally:
        sk._STATUS_SINKS.pop("cpu", None)


def _struct_fields(header_text: str, name: str) -> list:
    body = re.search(r"typedef struct " + name + r" \{(.*?)\} " + name + ";", header_text, flags=re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if decl:
            fields += [re.findall(r"(\w+)\s*$", part.strip())[0] for part in decl.split(",")]
    return fields


def test_bindings_and_documented_stub_match_the_header_structs():
    """The ctypes mirrors of the argument structs, and the reference-side stub documented in INTEGRATION.md, list the
    header's fields in the header's order (a short struct would make the library read past its end)."""
    header = (ROOT / "include" / "citrus_hip.h").read_text()
    assert [f for f, _ in _hip.SweepArgs._fields_] == _struct_fields(header, "CitrusSweepArgs")
    assert [f for f, _ in _hip.SweepView._fields_] == _struct_fields(header, "CitrusSweepView")
    assert [f for f, _ in _hip.PersistentArgs._fields_] == _struct_fields(header, "CitrusPersistentArgs")
    doc = (ROOT / "INTEGRATION.md").read_text()
    block = re.search(r"```python\n(.*?)```", doc, flags=re.S).group(1)
    stub = block[block.index("class CitrusSweepArgs"):block.index("lib.citrus_sweep_plan_create")]
    namespace = {"ctypes": ctypes}
    exec(stub, namespace)  # the documented class definition only
    assert [f for f, _ in namespace["CitrusSweepArgs"]._fields_] == _struct_fields(header, "CitrusSweepArgs")


def test_deferred_first_collision_shapes_are_checked():
    """A deferred first-collision source whose buffers disagree with the sweep's shape is refused before any launch."""
    from citrus_amd.scatter_solver.transport_sweep import _check_deferred
    from citrus_amd.scatter_solver.uncollided import DeferredFirstCollision

    d = DeferredFirstCollision.empty(2, 3, (2, 3, 4), 8, "cpu")
    pad = _hip.FC_GROUP_PAD
    assert tuple(d.samples.shape) == (2, 3, 8 * 64 * 4 + pad) and tuple(d.u.shape) == (2, 64, 4)
    assert tuple(d.block_samples().shape) == (2, 3, 1, 8, 64, 4)
    _check_deferred(d, 2, 3, (2, 3, 4), 8)
    for args in ((1, 3, (2, 3, 4), 8), (2, 4, (2, 3, 4), 8), (2, 3, (2, 3, 5), 8), (2, 3, (2, 3, 4), 16),
                 (2, 3, (2, 3, 4), 12)):
        with pytest.raises(ValueError):
            _check_deferred(d, *args)
    d.u = d.u[..., :3]
    with pytest.raises(ValueError, match="u"):
        _check_deferred(d, 2, 3, (2, 3, 4), 8)
    e = DeferredFirstCollision.empty(1, 2, (5, 8, 9), 8 * 21, "cpu")  # S12: 21 directions per octant, padded to 24
    assert tuple(e.block_samples().shape) == (1, 2, 2 * 2 * 3, 8, 64, 24)
    assert tuple(e.scale.shape) == (1, 2, 12 * 64 + pad)


def test_block_order_round_trip_and_cell_placement():
    """The sweep's block order: blocks of 4^3 cells row-major over (row, column, slice) blocks, a block's cells
    row-major, partial blocks padded with zeros; from_block_order undoes it."""
    from citrus_amd.scatter_solver.uncollided import fc_blocks, from_block_order, to_block_order

    gen = torch.Generator().manual_seed(4)
    for shape in [(8, 8, 8), (6, 5, 7), (4, 9, 2), (1, 1, 1)]:
        R, C, S = shape
        x = torch.randn((3,) + shape + (2,), generator=gen)
        b = to_block_order(x, shape)
        assert tuple(b.shape) == (3, fc_blocks(shape), 64, 2)
        assert torch.equal(from_block_order(b, shape, 1), x)
        nbc, nbs = -(-C // 4), -(-S // 4)
        for r, c, s in [(0, 0, 0), (R - 1, C - 1, S - 1), (R // 2, C - 1, 0)]:
            blk = ((r // 4) * nbc + c // 4) * nbs + s // 4
            assert torch.equal(b[:, blk, ((r % 4) * 4 + c % 4) * 4 + s % 4], x[:, r, c, s])
        assert int((b != 0).sum()) == x.numel()  # every value once, the padding cells zero


def test_octant_task_slots_follow_each_octants_sweep():
    """The cells of octant o's run of block-order samples come in o's sweep order: each run is a permutation of the
    block's 64 cells; the octant's upstream corner is first and the opposite corner last; a cell comes after its
    upstream neighbour along every axis (the direction of travel given by the octant's signs); the planes of equal
    depth are contiguous, 1, 3, 6, 10, 12, 12, 10, 6, 3, 1 cells long."""
    from citrus_amd.scatter_solver.uncollided import octant_task_slots

    slots = octant_task_slots()
    assert slots.shape == (8, 64)
    sizes = [1, 3, 6, 10, 12, 12, 10, 6, 3, 1]
    starts = np.cumsum([0] + sizes)
    for o in range(8):
        assert sorted(slots[o].tolist()) == list(range(64))
        step = {"c": -1 if o & 4 else 1, "r": -1 if o & 2 else 1, "s": -1 if o & 1 else 1}
        corner = lambda lo: tuple(0 if (step[a] > 0) == lo else 3 for a in "rcs")  # noqa: E731
        lc = lambda r, c, s: (r * 4 + c) * 4 + s  # noqa: E731
        assert slots[o, lc(*corner(True))] == 0 and slots[o, lc(*corner(False))] == 63
        for r in range(4):
            for c in range(4):
                for s in range(4):
                    here = slots[o, lc(r, c, s)]
                    for axis, (ur, uc, us) in (("r", (r - step["r"], c, s)), ("c", (r, c - step["c"], s)),
                                               ("s", (r, c, s - step["s"]))):
                        if 0 <= ur < 4 and 0 <= uc < 4 and 0 <= us < 4:
                            assert slots[o, lc(ur, uc, us)] < here, (o, axis)
                    dr = r if step["r"] > 0 else 3 - r
                    dc = c if step["c"] > 0 else 3 - c
                    ds = s if step["s"] > 0 else 3 - s
                    depth = dr + dc + ds
                    assert starts[depth] <= here < starts[depth + 1]
