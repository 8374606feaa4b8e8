"""Parity of the HIP LD sweep with the FP64 oracle (GPU).

Mirrors the reference's sweep tests (tests/test_transport_sweep.py): 1-D slab (:91), particle balance (:155), dense
global assembly (:237), on-the-fly moments (:340), chunk invariance (:355), heterogeneous random problem (:400),
source called once per chunk (:461), cell-major sources (:496); plus the fused moment source and the configs'
shapes. The float32 bar is ``F32_REL = 1e-5`` of the field maximum."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from citrus_amd.scatter_solver.angular_quadrature import level_symmetric_quadrature, product_quadrature
from citrus_amd.scatter_solver.geometry import Grid
from citrus_amd.scatter_solver.source_iteration import scatter_source_function
from citrus_amd.scatter_solver.transport_sweep import make_sweep_context, source_is_cell_major, sweep
from oracle import checks, quadrature as oq, sweep as osw
from tests.conftest import gpu_device
from tests.helpers import (F32_REL, direction_source_function, octant_set, oracle_sweep, random_problem,
                           rel_err)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1, -1])
def test_slab_oracle(axis, sign):
    dev = gpu_device()
    rng = np.random.default_rng(10 + axis)
    K, G = 9, 2
    deltas = (0.7, 1.3, 0.9)
    shape = [3, 3, 3]
    vol_dim = {0: 1, 1: 0, 2: 2}[axis]
    shape[vol_dim] = K
    grid = Grid(tuple(shape), deltas[0], deltas[1] / deltas[0], deltas[2] / deltas[0])
    d = [0.0, 0.0, 0.0]
    d[axis] = float(sign)
    o = int(oq.octant_of(np.array(d)))
    q = octant_set(replace={o: d})
    ctx = make_sweep_context(grid, q, 1, device=dev)
    mu1 = rng.random((G, K)) * 1.5 + 0.05
    S0, S1 = rng.random((G, K)), rng.standard_normal((G, K)) * 0.3

    def along(v):
        view = [G, 1, 1, 1]
        view[1 + vol_dim] = K
        return torch.from_numpy(np.broadcast_to(v.reshape(view), (G, *shape)).copy())

    src = torch.zeros(G, 4, *shape, dtype=torch.float64)
    src[:, 0] = along(S0)
    src[:, 1 + axis] = along(S1)
    src32 = src.to(dev, torch.float32)
    _, psi = sweep(ctx, along(mu1).to(dev, torch.float32), lambda idx: src32[None].expand(idx.numel(), *src32.shape),
                   return_psi=True)
    psi = psi.cpu().double()
    for g in range(G):
        p0, p1 = checks.slab_1d(mu1[g], S0[g], S1[g], deltas[axis], float(sign))
        r0 = along(np.broadcast_to(p0, (G, K)).copy())[g]
        r1 = along(np.broadcast_to(p1, (G, K)).copy())[g]
        tol = F32_REL * max(1.0, float(r0.abs().max()))
        assert float((psi[o, g, 0] - r0).abs().max()) < tol
        assert float((psi[o, g, 1 + axis] - r1).abs().max()) < tol


def test_particle_balance_and_dense_assembly():
    dev = gpu_device()
    shape = (3, 4, 3)
    grid = Grid(shape, 0.9, 1.2 / 0.9, 0.75 / 0.9)
    q = octant_set(base=(0.55, 0.38, 0.74))
    ctx = make_sweep_context(grid, q, 1, device=dev)
    mu, src = random_problem(shape, 2, 8, dev, seed=3)
    _, psi = sweep(ctx, mu, direction_source_function(src), return_psi=True)
    psi, mu64, src64 = psi.cpu().double().numpy(), mu.cpu().double().numpy(), src.cpu().double().numpy()
    om, _ = (q.omega[ctx.permutation.cpu()].numpy(), None)
    for m in range(8):
        lhs, rhs = checks.outflow_balance(psi[m], om[m], grid.deltas, mu64, src64[m])
        assert np.allclose(lhs, rhs, rtol=1e-5, atol=0)
        for g in range(2):
            ref = checks.dense_global_solve(om[m], grid.deltas, mu64[g], src64[m, g])
            assert rel_err(psi[m, g], ref) < F32_REL


@pytest.mark.parametrize("chunk", [None, 5])
def test_random_heterogeneous_matches_oracle(chunk):
    """12x10x14 (partial blocks), G=6, 128 directions, L=3, thin-to-thick cells, sparse signed sources."""
    dev = gpu_device()
    shape, G = (12, 10, 14), 6
    gen = torch.Generator().manual_seed(21)
    mu = torch.exp(torch.empty(G, *shape, dtype=torch.float64).uniform_(-3.0, 1.5, generator=gen))
    grid = Grid(shape, 0.8, 1.15 / 0.8, 0.9 / 0.8)
    q = product_quadrature(8, 16)
    ctx = make_sweep_context(grid, q, 3, device=dev)
    src = torch.randn(ctx.num_directions, G, 4, *shape, generator=gen, dtype=torch.float64)
    src[:, :, 0] = src[:, :, 0].abs() * (torch.rand(G, *shape, generator=gen, dtype=torch.float64) < 0.3)
    mu32, src32 = mu.to(dev, torch.float32), src.to(dev, torch.float32)
    phi, psi = sweep(ctx, mu32, direction_source_function(src32), direction_chunk=chunk, return_psi=True)
    psi_ref, phi_ref = oracle_sweep(ctx, q, mu32, src32)
    assert rel_err(psi, psi_ref) < F32_REL and rel_err(phi, phi_ref) < F32_REL
    assert float(psi[:, :, 0].min()) < 0.0  # negative cell means are exercised


def test_moments_match_ordinates_and_chunks():
    dev = gpu_device()
    grid = Grid((3, 4, 5), 0.6, 1.1 / 0.6, 0.9 / 0.6)
    q = product_quadrature(4, 8)
    ctx = make_sweep_context(grid, q, 3, device=dev)
    mu, src = random_problem(grid.shape, 2, ctx.num_directions, dev)
    phi, psi = sweep(ctx, mu, direction_source_function(src), return_psi=True)
    assert phi.shape == (2, 16, 4, 3, 4, 5)
    Yw = ctx.basis.Y_weighted.to(dev)
    ref = torch.einsum("mn,mgk...->gnk...", Yw.double(), psi.double())
    assert rel_err(phi, ref) < F32_REL
    for chunk in (1, 3):
        assert rel_err(sweep(ctx, mu, direction_source_function(src), direction_chunk=chunk), phi) < F32_REL


@pytest.mark.parametrize("chunk", [None, 3])
def test_source_called_once_per_chunk_and_cell_major(chunk):
    dev = gpu_device()
    grid = Grid((4, 5, 3), 0.7, 1.1 / 0.7, 0.9 / 0.7)
    q = product_quadrature(4, 8)
    ctx = make_sweep_context(grid, q, 2, device=dev)
    mu, src = random_problem(grid.shape, 3, ctx.num_directions, dev)
    calls = []

    def plain(index):
        calls.append(index.clone())
        return src.index_select(0, index)

    cm_src = src.permute(0, 3, 4, 5, 1, 2).contiguous()

    def cell_major(index, out=None):
        assert out is not None and out.shape == (index.numel(),) + tuple(cm_src.shape[1:])
        return out.copy_(cm_src.index_select(0, index))

    cell_major.cell_major = True
    assert source_is_cell_major(cell_major)
    ref = sweep(ctx, mu, direction_source_function(src), direction_chunk=chunk)
    assert torch.equal(sweep(ctx, mu, plain, direction_chunk=chunk), ref)
    assert torch.equal(sweep(ctx, mu, cell_major, direction_chunk=chunk), ref)
    dc = 4 if chunk is None else 3
    starts = list(range(0, 4, dc))
    assert len(calls) == len(starts)
    for d0, index in zip(starts, calls):
        assert torch.equal(index, ctx.direction_indices(d0, min(d0 + dc, 4)))


@pytest.mark.parametrize("kernel", ["mfma", "vector"])
@pytest.mark.parametrize("shape,G,order,L,chunk", [((5, 7, 6), 3, 8, 3, None), ((9, 4, 6), 2, 12, 5, 7),
                                                   ((8, 12, 7), 2, 12, 5, None), ((4, 5, 9), 2, 4, 1, 2)])
def test_fused_moment_source_matches_oracle(shape, G, order, L, chunk, kernel, monkeypatch):
    """Source expanded from SH moments inside the kernel vs the oracle's explicit expansion, on the matrix-core
    kernel (ragged blocks, 10 and 21 directions per octant, chunks of 7 and 2) and on the vector kernel."""
    monkeypatch.setenv("CITRUS_SWEEP_KERNEL", kernel)
    dev = gpu_device()
    grid = Grid(shape, 0.5)
    q = level_symmetric_quadrature(order)
    ctx = make_sweep_context(grid, q, L, device=dev)
    gen = torch.Generator().manual_seed(7)
    mu = (torch.rand(G, *shape, generator=gen) + 0.3).to(dev)
    T = torch.randn(G, ctx.num_moments, 4, *shape, generator=gen).to(dev)
    phi = sweep(ctx, mu, scatter_source_function(ctx, T), direction_chunk=chunk)
    om, w = q.omega[ctx.permutation.cpu()].numpy(), q.weight[ctx.permutation.cpu()].numpy()
    Y = oq.real_spherical_harmonics(L, om)
    src = osw.sources_from_moments(T.double().cpu().numpy(), Y)
    psi = osw.sweep(shape, grid.deltas, om, mu.double().cpu().numpy(), src)
    assert rel_err(phi, osw.moments(psi, w, Y)) < F32_REL


def test_config0_shape_fused_sweep():
    """32^3 cells, 8 groups, S8 (80 directions), P3 (16 moments): one fused sweep vs the oracle."""
    dev = gpu_device()
    shape, G, L = (32, 32, 32), 8, 3
    grid = Grid(shape, 0.4)
    q = level_symmetric_quadrature(8)
    ctx = make_sweep_context(grid, q, L, device=dev)
    gen = torch.Generator().manual_seed(1)
    mu = (torch.rand(G, *shape, generator=gen) * 0.45 + 0.15).to(dev)
    T = (torch.randn(G, ctx.num_moments, 4, *shape, generator=gen) * 0.1).to(dev)
    T[:, 0, 0] = T[:, 0, 0].abs() + 0.05
    phi = sweep(ctx, mu, scatter_source_function(ctx, T))
    om, w = q.omega[ctx.permutation.cpu()].numpy(), q.weight[ctx.permutation.cpu()].numpy()
    Y = oq.real_spherical_harmonics(L, om)
    psi = osw.sweep(shape, grid.deltas, om, mu.double().cpu().numpy(),
                    osw.sources_from_moments(T.double().cpu().numpy(), Y))
    err = rel_err(phi, osw.moments(psi, w, Y))
    print(f"config0 fused sweep rel err {err:.2e}")
    assert err < F32_REL


def test_backend_errors():
    dev = gpu_device()
    grid = Grid((3, 3, 4), 0.5)
    q = product_quadrature(4, 8)
    ctx = make_sweep_context(grid, q, 1, device=dev)
    mu, src = random_problem(grid.shape, 2, 32, dev)
    with pytest.raises(ValueError):
        sweep(ctx, mu, direction_source_function(src), backend="torch")
    with pytest.raises(ValueError):
        sweep(ctx, mu.double(), direction_source_function(src.double()))
    with pytest.raises(ValueError):
        make_sweep_context(grid, q, 1, device=dev, backend="bogus")


@pytest.mark.parametrize("kernel", ["mfma", "vector"])
@pytest.mark.parametrize("chunk", [None, 3])
@pytest.mark.parametrize("means", [False, True])
def test_group_major_explicit_source_with_moments(kernel, chunk, means, monkeypatch):
    """Moment source plus an explicit source handed over group-major for every direction (the CITRUS step's summed
    source) equals the cell-major chunk path and the oracle, on both kernels and in direction chunks; with zero
    slopes the source may be handed over as its cell means alone."""
    from citrus_amd.scatter_solver.source_iteration import summed_source_function

    monkeypatch.setenv("CITRUS_SWEEP_KERNEL", kernel)
    dev = gpu_device()
    shape, G, L = (6, 5, 7), 2, 3
    grid = Grid(shape, 0.5)
    q = level_symmetric_quadrature(8)
    ctx = make_sweep_context(grid, q, L, device=dev)
    gen = torch.Generator().manual_seed(17)
    mu = (torch.rand(G, *shape, generator=gen) + 0.3).to(dev)
    T = (torch.randn(G, ctx.num_moments, 4, *shape, generator=gen) * 0.2).to(dev)
    extra_cm = torch.randn(ctx.num_directions, *shape, G, 4, generator=gen).to(dev)  # [M, R, C, S, G, 4] reordered
    if means:
        extra_cm[..., 1:] = 0.0

    def plain(index, out=None):
        return out.copy_(extra_cm.index_select(0, index))

    plain.cell_major = True

    def grouped(index, out=None):
        return out.copy_(extra_cm.index_select(0, index))

    grouped.cell_major = True
    grouped.hip_group_major = (lambda: extra_cm[..., 0].permute(4, 1, 2, 3, 0).contiguous()) if means else \
        (lambda: extra_cm.permute(4, 1, 2, 3, 0, 5).contiguous())
    phi_gm = sweep(ctx, mu, summed_source_function(scatter_source_function(ctx, T), grouped), direction_chunk=chunk)
    phi_cm = sweep(ctx, mu, summed_source_function(scatter_source_function(ctx, T), plain), direction_chunk=chunk)
    assert rel_err(phi_gm, phi_cm) < F32_REL
    om, w = q.omega[ctx.permutation.cpu()].numpy(), q.weight[ctx.permutation.cpu()].numpy()
    Y = oq.real_spherical_harmonics(L, om)
    src = osw.sources_from_moments(T.double().cpu().numpy(), Y) + \
        extra_cm.double().cpu().permute(0, 4, 5, 1, 2, 3).numpy()
    psi = osw.sweep(shape, grid.deltas, om, mu.double().cpu().numpy(), src)
    assert rel_err(phi_gm, osw.moments(psi, w, Y)) < F32_REL


@pytest.mark.parametrize("nsh", [16, 36])
def test_mixed_operand_is_exact_low_and_rounded_high(nsh):
    """The mixed-precision operand keeps the degrees l <= 1 (moments 0..3) in float32 bit for bit and holds the
    degrees l >= 2 as bfloat16 round-to-nearest-even (torch's conversion), component-major and zero padded; the
    scattering-moment kernel that writes the operand directly gives the same bits as converting its float32 result."""
    from citrus_amd.scatter_solver.source_iteration import (MaterialTransfer, bf16_operand, scatter_moments_mixed,
                                                            scatter_moments_views)

    dev = gpu_device()
    gen = torch.Generator().manual_seed(21)
    V, G, shape = 2, 3, (3, 4, 5)
    T = torch.randn(V, G, *shape, nsh, 4, generator=gen)
    T *= torch.tensor([1e-3, 5.0, 3e4]).reshape(1, G, 1, 1, 1, 1, 1)  # groups of very different magnitudes
    op = bf16_operand(T.to(dev))
    kh = 16 * ((nsh - 4 + 15) // 16)
    assert tuple(op.high.shape) == (V, G, *shape, 4, kh) and op.high.dtype == torch.bfloat16
    Tt = T.transpose(-1, -2).contiguous()  # [..., 4 comp, nsh]
    assert torch.equal(op.low.cpu(), Tt[..., :4])
    want = torch.zeros(V, G, *shape, 4, kh, dtype=torch.bfloat16)
    want[..., :nsh - 4] = Tt[..., 4:].to(torch.bfloat16)
    assert torch.equal(op.high.cpu().view(torch.int16), want.view(torch.int16))
    rho = torch.rand(2, *shape, generator=gen).to(dev)
    L = int(round(nsh ** 0.5)) - 1
    sigma = torch.rand(2, G, G, L + 1, generator=gen).tril().transpose(1, 2).contiguous().to(dev)  # gi <= go only
    transfer = MaterialTransfer(rho, sigma)
    phi = torch.randn(V, G, *shape, nsh, 4, generator=gen).to(dev)
    T32, means32 = scatter_moments_views(transfer, phi, cell_means=True)
    direct, means = scatter_moments_mixed(transfer, phi, cell_means=True)
    converted = bf16_operand(T32)
    assert torch.equal(direct.low, converted.low) and torch.equal(means, means32)
    assert torch.equal(direct.high.view(torch.int16), converted.high.view(torch.int16))


@pytest.mark.parametrize("L", [3, 5])
def test_bf16_contraction_sweep(L):
    """Mixed-precision source expansion (degrees l <= 1 in float32, l >= 2 as a bfloat16 matrix product, float32
    accumulation): within the bfloat16 rounding of the higher degrees of the float32 sweep and of the FP64 oracle;
    the same sweep from the converted operand or from float32 moments; unsupported orders refused."""
    from citrus_amd.scatter_solver.transport_sweep import sweep_views

    dev = gpu_device()
    shape, G = (6, 5, 7), 3
    grid = Grid(shape, 0.5)
    q = level_symmetric_quadrature(8)
    ctx = make_sweep_context(grid, q, L, device=dev)
    gen = torch.Generator().manual_seed(23)
    mu = (torch.rand(G, *shape, generator=gen) + 0.3).to(dev)
    T = torch.randn(2, G, *shape, ctx.num_moments, 4, generator=gen) * 0.2
    T[..., 0, 0] = T[..., 0, 0].abs() + 0.5
    T *= torch.tensor([1.0, 1e-3, 40.0]).reshape(1, G, 1, 1, 1, 1, 1)
    Td = T.to(dev).contiguous()
    f32 = sweep_views(ctx, mu, Td)
    b16 = sweep_views(ctx, mu, Td, contraction="bf16")
    from citrus_amd.scatter_solver.source_iteration import bf16_operand

    assert torch.equal(sweep_views(ctx, mu, bf16_operand(Td)), b16)  # the operand implies the mixed contraction
    errs = [rel_err(b16[:, g], f32[:, g]) for g in range(G)]
    print(f"mixed contraction L={L}: per-group error vs float32 {errs}", flush=True)
    for g in range(G):  # each group against its own magnitude
        assert errs[g] < 2e-3
    om = q.omega[ctx.permutation.cpu()].numpy()
    w = q.weight[ctx.permutation.cpu()].numpy()
    Y = oq.real_spherical_harmonics(L, om)
    for v in range(2):
        Tv = T[v].double().permute(0, 4, 5, 1, 2, 3).numpy()
        psi = osw.sweep(shape, grid.deltas, om, mu.double().cpu().numpy(), osw.sources_from_moments(Tv, Y))
        ref = osw.moments(psi, w, Y)
        got = b16[v].permute(0, 4, 5, 1, 2, 3)
        for g in range(G):
            assert rel_err(got[g], ref[g]) < 4e-3
    ctx2 = make_sweep_context(grid, q, 2, device=dev)
    with pytest.raises(ValueError, match="P3 and P5"):
        sweep_views(ctx2, mu, torch.zeros(1, G, *shape, 9, 4, device=dev), contraction="bf16")


@pytest.mark.parametrize("contraction,with_source", [("fp32", False), ("fp32", True), ("bf16", True)])
def test_sweep_graph_replays_equal_eager(contraction, with_source):
    """A captured batched sweep, replayed with new inputs and eager work between the replays, equals the eager sweep
    of the same inputs bit for bit on every replay."""
    from citrus_amd.scatter_solver.transport_sweep import SweepGraph, sweep_views

    dev = gpu_device()
    shape, G, V = (6, 5, 7), 2, 2
    ctx = make_sweep_context(Grid(shape, 0.5), level_symmetric_quadrature(6), 3, device=dev)
    gen = torch.Generator().manual_seed(29)
    mu = (torch.rand(G, *shape, generator=gen) + 0.3).to(dev)
    T = torch.zeros(V, G, *shape, ctx.num_moments, 4, device=dev)
    src = torch.zeros(V, G, *shape, ctx.num_directions, device=dev) if with_source else None
    graph = SweepGraph(ctx, mu, T, source=src, contraction=contraction)
    for step in range(3):
        T.copy_(torch.randn(T.shape, generator=gen))
        if src is not None:
            src.copy_(torch.rand(src.shape, generator=gen))
        noise = torch.randn(256, 256, device=dev) @ torch.randn(256, 256, device=dev)  # eager work between replays
        torch.cuda.synchronize()
        got = graph.replay().clone()
        want = sweep_views(ctx, mu, T, source=src, contraction=contraction)
        torch.cuda.synchronize()
        assert torch.equal(got, want), step
        del noise


@pytest.mark.parametrize("schedule", ["persistent", "launches"])
def test_sweep_graph_survives_scratch_growth(schedule):
    """A graph captured on a schedule whose scratch later grows (an eager sweep of the same grid with more directions
    per chunk, or more views) keeps replaying into the buffer it recorded: the replay still equals the eager sweep of
    its inputs bit for bit, and tensors allocated after the growth are not written by it."""
    from citrus_amd.scatter_solver.transport_sweep import SweepGraph, sweep_views

    dev = gpu_device()
    shape, G, V = (7, 6, 5), 2, 2
    ctx = make_sweep_context(Grid(shape, 0.5), level_symmetric_quadrature(6), 3, device=dev)
    gen = torch.Generator().manual_seed(37)
    mu = (torch.rand(G, *shape, generator=gen) + 0.3).to(dev)
    T = torch.randn(V, G, *shape, ctx.num_moments, 4, generator=gen).to(dev)
    graph = SweepGraph(ctx, mu, T, direction_chunk=2, schedule=schedule)
    want = sweep_views(ctx, mu, T, direction_chunk=2, schedule=schedule)
    # eager sweeps that need larger scratch on the same schedule objects: all directions per chunk, more views
    big = torch.randn(2 * V, G, *shape, ctx.num_moments, 4, generator=gen).to(dev)
    sweep_views(ctx, mu, T, schedule=schedule)
    sweep_views(ctx, mu, big, schedule="launches")
    torch.cuda.synchronize()
    junk = [torch.full((1 << 20,), 7.0, device=dev) for _ in range(8)]  # reuse of any freed block shows here
    for _ in range(2):
        got = graph.replay().clone()
        torch.cuda.synchronize()
        assert torch.equal(got, want)
    assert all(bool((j == 7.0).all()) for j in junk)


def test_sweep_autotuner_picks_and_caches_a_chunk():
    from citrus_amd.scatter_solver.autotune import SweepTuner, chunk_candidates, default_tuner
    from citrus_amd.scatter_solver.transport_sweep import sweep_views

    dev = gpu_device()
    shape, G, V = (8, 8, 8), 2, 2
    ctx = make_sweep_context(Grid(shape, 0.5), level_symmetric_quadrature(8), 2, device=dev)
    gen = torch.Generator().manual_seed(31)
    mu = (torch.rand(G, *shape, generator=gen) + 0.3).to(dev)
    tuner = SweepTuner(repeats=1)
    calls = []
    chunk = tuner.best_chunk(ctx, mu, V)
    assert chunk in chunk_candidates(ctx.directions_per_octant)
    assert tuner.best_chunk(ctx, mu, V, run=lambda c: calls.append(c)) == chunk and not calls  # cached
    T = torch.randn(V, G, *shape, ctx.num_moments, 4, generator=gen).to(dev)
    auto = sweep_views(ctx, mu, T, direction_chunk="auto")
    schedule, chosen = default_tuner().cache[("config",) + default_tuner().key(ctx, G, V, "fp32")]
    assert schedule == "launches"  # P2: the persistent kernel is built for P3 and P5 only
    assert torch.equal(auto, sweep_views(ctx, mu, T, direction_chunk=chosen, schedule=schedule))
    assert rel_err(auto, sweep_views(ctx, mu, T)) < F32_REL


def test_persistent_sweep_with_fewer_resident_waves():
    """The persistent schedule with its resident waves per CU capped (down to one) gives the moments of the full
    occupancy bit for bit; the tuner measures the caps and keeps one of them; a negative cap is refused."""
    from citrus_amd.scatter_solver.autotune import WAVE_CANDIDATES, SweepTuner
    from citrus_amd.scatter_solver.transport_sweep import sweep_views

    dev = gpu_device()
    shape, G, V = (9, 8, 6), 3, 2
    ctx = make_sweep_context(Grid(shape, 0.5), level_symmetric_quadrature(6), 3, device=dev)
    gen = torch.Generator().manual_seed(37)
    mu = (torch.rand(G, *shape, generator=gen) + 0.3).to(dev)
    T = torch.randn(V, G, *shape, ctx.num_moments, 4, generator=gen).to(dev)
    full = sweep_views(ctx, mu, T, schedule="persistent")
    for waves in (1, 3, 64):
        assert torch.equal(sweep_views(ctx, mu, T, schedule="persistent", waves_per_cu=waves), full)
    with pytest.raises(ValueError):
        sweep_views(ctx, mu, T, schedule="persistent", waves_per_cu=-1)
    tuner = SweepTuner(repeats=1)
    chunk = ctx.directions_per_octant
    waves = tuner.best_waves(ctx, mu, V, chunk)
    key = ("waves", chunk) + tuner.key(ctx, G, V, "fp32")
    assert waves in WAVE_CANDIDATES and set(tuner.timings[key]) == set(WAVE_CANDIDATES)
    calls = []
    assert tuner.best_waves(ctx, mu, V, chunk, run=lambda w: calls.append(w)) == waves and not calls  # cached


def test_bf16_operand_past_two_to_the_32_work_items():
    """The operand conversion of 8 views x 16 groups x 64^3 cells x P5 spans 4.8e9 (cell, component, moment) slots,
    more than one dispatch's 2^32 work-items: the kernel loops with a stride, so the last view, which lies past the
    first 2^32 slots, is converted exactly (low: copied; high: bfloat16, round to nearest even)."""
    from citrus_amd.scatter_solver.source_iteration import bf16_operand

    dev = gpu_device()
    V, G, n, nsh = 8, 16, 64, 36
    T = torch.zeros(V, G, n, n, n, nsh, 4, device=dev)
    gen = torch.Generator(device=dev).manual_seed(5)
    T[V - 1].normal_(generator=gen)
    op = bf16_operand(T)
    last = T[V - 1]  # [G, R, C, S, nsh, 4]
    assert torch.equal(op.low[V - 1], last[..., :4, :].transpose(-1, -2).contiguous())
    assert torch.equal(op.high[V - 1, ..., :nsh - 4], last[..., 4:, :].transpose(-1, -2).to(torch.bfloat16))
    assert not bool(op.low[:V - 1].any()) and not bool(op.high[:V - 1].float().any())
