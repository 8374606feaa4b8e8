"""First-collision source, sampling correction and LD moments on the HIP kernels vs the FP64 oracle (GPU).

Reference tests mirrored: hand evaluation of the source (tests/test_uncollided.py:181-213), the restriction
(:217-244) and the conservation of the corrected source (tests/test_scatter_solver.py:282-323). FP32 bar
``F32_REL = 1e-5`` of the field maximum."""

from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
import torch

from citrus_amd import _hip
from citrus_amd.scatter_solver.angular_quadrature import level_symmetric_quadrature
from citrus_amd.scatter_solver.cross_sections import MultigroupLibrary, seeded_library, synthetic_library
from citrus_amd.scatter_solver.geometry import ConeGeometry, Grid
from citrus_amd.scatter_solver.uncollided import (
    first_collision_correction_factors, first_collision_exact_moments, first_collision_moments_at_directions,
    first_collision_quadrature_sums, first_collision_sampling_correction, first_collision_source_at_directions,
    primary_directions_to_voxels, restrict_to_linear_discontinuous_moments, first_collision_corrected_samples)
from oracle import first_collision as ofc
from tests.conftest import gpu_device
from tests.helpers import F32_REL, rel_err

pytestmark = pytest.mark.gpu

GEOMETRY = ConeGeometry(60.0, 100.0, 32, 32, 0.5, 0.5)
ANGLE = 0.7


def _six_material_library(G=5):
    """Six materials built from the synthetic pair (scaled copies), for the per-point material slots."""
    base = synthetic_library(G, 3, num_eta=41)
    scales = torch.tensor([1.0, 1.0, 0.8, 1.3, 0.6, 1.1], dtype=torch.float64)
    pick = torch.tensor([0, 1, 0, 1, 0, 1])

    def rows(t):
        return t[pick] * scales.reshape(-1, *([1] * (t.dim() - 1)))

    return MultigroupLibrary(base.group_edges, rows(base.mu_t), rows(base.mu_s_l), base.eta_grid, rows(base.mu_s_eta),
                             tuple(f"m{m}" for m in range(6)), base.spectrum_fraction, base.fine_energy,
                             rows(base.fine_mu_t), base.fine_weight, base.fine_group)


def _problem(library, shape, seed, sparse=False):
    gen = torch.Generator().manual_seed(seed)
    M, G = library.num_materials, library.num_groups
    rho = torch.rand(M, *shape, generator=gen, dtype=torch.float64) + 0.1
    if sparse:  # at most two nonzero materials per point
        keep = torch.zeros(M, *shape, dtype=torch.bool)
        a = torch.randint(0, M, shape, generator=gen)
        b = torch.randint(0, M, shape, generator=gen)
        keep.scatter_(0, a[None], True)
        keep.scatter_(0, b[None], True)
        rho = rho * keep
    phi = torch.rand(G, *shape, generator=gen, dtype=torch.float64) + 0.05
    return rho, phi


def _oracle_inputs(library, grid, rho, phi):
    pts = grid.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    ou = ofc.primary_directions(pts, np.array(GEOMETRY.source_position(ANGLE)))
    emap = None if library.eta_map is None else library.eta_map.numpy()
    # the kernels see float32 inputs: the oracle runs on the same rounded values in float64
    r = rho.float().double().reshape(rho.shape[0], -1).numpy()
    p = phi.float().double().reshape(phi.shape[0], -1).numpy()
    return ou, emap, r, p


@pytest.mark.parametrize("kind", ["seeded", "adaptive", "six_materials"])
def test_fc_source_matches_oracle(kind):
    dev = gpu_device()
    library = {"seeded": lambda: seeded_library(6, 3, 120.0, 20.0, seed=11, num_eta=65),
               "adaptive": lambda: synthetic_library(4, 3, num_eta=33, angular_table_grid="adaptive"),
               "six_materials": _six_material_library}[kind]()
    grid = Grid((6, 7, 5), 0.6)
    rho, phi = _problem(library, grid.shape, 1, sparse=kind == "six_materials")
    q = level_symmetric_quadrature(8)
    S = first_collision_source_at_directions(library, rho.float().to(dev), phi.float().to(dev), grid, GEOMETRY, ANGLE,
                                             q.omega)
    ou, emap, r, p = _oracle_inputs(library, grid, rho, phi)
    ref = ofc.source_at_directions(library.mu_s_eta.numpy(), emap, r, p, ou, q.omega.numpy())
    assert S.shape == (q.omega.shape[0], library.num_groups) + grid.shape
    assert rel_err(S.reshape(ref.shape), ref) < F32_REL


@pytest.mark.parametrize("shape", [(8, 6, 7), (2, 8, 32)])
def test_fc_sums_and_factors_match_oracle(shape):
    dev = gpu_device()
    library = seeded_library(8, 3, 120.0, 20.0, seed=12, num_eta=129)
    grid = Grid(shape, 0.5)
    rho, phi = _problem(library, grid.shape, 2)
    q = level_symmetric_quadrature(12)
    rho_d, phi_d = rho.float().to(dev), phi.float().to(dev)
    I = first_collision_quadrature_sums(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega, q.weight)
    E = first_collision_exact_moments(library, rho_d.double(), phi_d.double(), l_max=1)
    ou, emap, r, p = _oracle_inputs(library, grid, rho, phi)
    S = ofc.source_at_directions(library.mu_s_eta.numpy(), emap, r, p, ou, q.omega.numpy())
    I_ref = ofc.quadrature_sums(S, ou, q.omega.numpy(), q.weight.numpy())
    E_ref = ofc.exact_moments(library.mu_s_l.numpy(), r, p, 1)
    for j in range(3):
        assert rel_err(I[j].reshape(I_ref[j].shape), I_ref[j]) < F32_REL
    assert rel_err(E.reshape(E_ref.shape), E_ref) < 1e-12
    corr = first_collision_sampling_correction(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega, q.weight)
    s_ref, b_ref = ofc.correction_factors(I_ref, E_ref)
    assert rel_err(corr.scale.reshape(s_ref.shape), s_ref) < F32_REL
    assert float(np.abs(corr.beta.reshape(b_ref.shape).cpu().double().numpy() - b_ref).max()) < 1e-4
    assert float(corr.beta.abs().max()) <= 1.0
    # factors from the sums: the host float64 formula and the oracle agree on the same sums
    s2, b2 = first_collision_correction_factors(torch.from_numpy(I_ref), torch.from_numpy(E_ref))
    assert np.allclose(s2.numpy(), s_ref, rtol=1e-14) and np.allclose(b2.numpy(), b_ref, atol=1e-14)


@pytest.mark.parametrize("factor,shape", [(1, (6, 4, 8)), (2, (6, 4, 8)), (1, (2, 8, 16))])
def test_fc_moments_match_oracle_in_every_layout(factor, shape):
    dev = gpu_device()
    library = seeded_library(5, 3, 120.0, 20.0, seed=13, num_eta=97)
    grid = Grid(shape, 0.4)
    rho, phi = _problem(library, grid.shape, 3)
    q = level_symmetric_quadrature(6)
    rho_d, phi_d = rho.float().to(dev), phi.float().to(dev)
    corr = first_collision_sampling_correction(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega, q.weight)
    ou, emap, r, p = _oracle_inputs(library, grid, rho, phi)
    S = ofc.source_at_directions(library.mu_s_eta.numpy(), emap, r, p, ou, q.omega.numpy())
    Sc = ofc.apply_correction(S, corr.scale.double().cpu().reshape(5, -1).numpy(),
                              corr.beta.double().cpu().reshape(5, -1).numpy(), ou, q.omega.numpy())
    K, G = S.shape[:2]
    ref = ofc.restrict_ld(Sc.reshape(K, G, *grid.shape), factor)  # [K, G, 4, Rc, Cc, Sc]
    mm = first_collision_moments_at_directions(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega, corr, factor)
    assert rel_err(mm, ref) < F32_REL
    cm = first_collision_moments_at_directions(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega, corr, factor,
                                               cell_major=True)
    assert torch.equal(cm, mm.permute(0, 3, 4, 5, 1, 2).contiguous())
    gm = first_collision_moments_at_directions(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega, corr, factor,
                                               layout="group_major")
    assert torch.equal(gm, mm.permute(1, 3, 4, 5, 0, 2).contiguous())
    if factor == 1:  # one point per cell: zero slopes, and the cell-mean layout holds the means alone
        assert not bool(mm[:, :, 1:].any())
        gmm = first_collision_moments_at_directions(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega, corr,
                                                    factor, layout="group_major_means")
        assert torch.equal(gmm, gm[..., 0].contiguous())
        # samples stored by the sums kernel and corrected in place: the same values from one source evaluation
        fused, corr2 = first_collision_corrected_samples(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega,
                                                         q.weight)
        assert torch.equal(corr2.scale, corr.scale) and torch.equal(corr2.beta, corr.beta)
        assert torch.equal(fused, gmm)
    else:
        with pytest.raises(ValueError):
            first_collision_moments_at_directions(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega, corr, factor,
                                                  layout="group_major_means")
    # the fused kernel equals correction + restriction applied to the sampled source on the device
    S_dev = first_collision_source_at_directions(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega)
    two_step = restrict_to_linear_discontinuous_moments(corr.apply(S_dev, q.omega), grid, factor)
    assert rel_err(mm, two_step) < F32_REL


def test_fc_correction_conserves_the_exact_moments():
    """Corrected, restricted source integrates (quadrature) to the restricted exact l = 0 moment; the l = 1 moment
    about Omega_UC is matched where beta is not clamped (test_scatter_solver.py:282-323)."""
    dev = gpu_device()
    library = seeded_library(6, 5, 120.0, 20.0, seed=14, num_eta=257)
    grid = Grid((8, 8, 8), 0.5)
    rho, phi = _problem(library, grid.shape, 4)
    q = level_symmetric_quadrature(8)
    rho_d, phi_d = rho.float().to(dev), phi.float().to(dev)
    corr = first_collision_sampling_correction(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega, q.weight)
    E = first_collision_exact_moments(library, rho_d.double(), phi_d.double(), l_max=1)
    w = q.weight.to(dev, torch.float64)
    mm = first_collision_moments_at_directions(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega, corr, 2)
    got = torch.einsum("k,kg...->g...", w, mm.double())
    exact = restrict_to_linear_discontinuous_moments(E[0], grid, 2)
    assert float((got - exact).abs().max()) < 1e-5 * float(exact[:, 0].abs().max())
    S = first_collision_source_at_directions(library, rho_d, phi_d, grid, GEOMETRY, ANGLE, q.omega).double()
    Sc = corr.apply(S, q.omega.to(dev, torch.float64))
    assert float(Sc.min()) >= 0.0
    mu = torch.einsum("rcsx,kx->krcs", primary_directions_to_voxels(grid, GEOMETRY, ANGLE, dev, torch.float64),
                      q.omega.to(dev, torch.float64))
    J = torch.einsum("k,krcs,kgrcs->grcs", w, mu, Sc)
    free = corr.beta.abs() < 1
    assert float(free.double().mean()) > 0.5
    assert float((J - E[1])[free].abs().max()) < 1e-5 * float(E[1].abs().max())


@pytest.mark.parametrize("sn,L,contraction,shape", [(6, 3, "fp32", (8, 8, 8)), (12, 5, "fp32", (8, 8, 8)),
                                                     (12, 5, "bf16", (8, 8, 8)), (8, 3, "bf16", (8, 8, 8)),
                                                     (12, 5, "fp32", (6, 5, 7)), (6, 3, "bf16", (9, 4, 6))])
def test_deferred_correction_equals_the_corrected_samples_and_sweeps_alike(sn, L, contraction, shape):
    """The sampling correction applied away from the sums kernel: by ``citrus_fc_correct_means`` or by the
    persistent sweep as it reads the stored samples (source layout 3, block order). Both give the corrected samples
    of ``first_collision_corrected_samples`` (and their sweep) bit for bit; the launch sequence applies it first. The
    block-order samples equal the grid-order ones of ``citrus_fc_sums`` bit for bit (their padding zero). S12 puts 21
    directions in a chunk (two direction tiles); bf16 runs the mixed-precision expansion; 6x5x7 and 9x4x6 leave
    partial blocks."""
    from citrus_amd.scatter_solver.source_iteration import bf16_operand
    from citrus_amd.scatter_solver.transport_sweep import make_sweep_context, sweep_views
    from citrus_amd.scatter_solver.uncollided import (DeferredFirstCollision, _fc_args, _fc_sums_raw, fc_blocks,
                                                      first_collision_deferred_samples, from_block_order,
                                                      to_block_order)

    dev = gpu_device()
    library = seeded_library(6, 3, 120.0, 20.0, seed=15, num_eta=97)
    grid = Grid(shape, 0.5)
    ctx = make_sweep_context(grid, level_symmetric_quadrature(sn), L, device=dev)
    om, w = ctx.quadrature.omega.to(dev, torch.float32), ctx.quadrature.weight
    V, G, M = 2, library.num_groups, ctx.num_directions
    deferred = DeferredFirstCollision.empty(V, G, grid.shape, M, dev)
    corrected = []
    for v in range(V):
        rho, phi = _problem(library, grid.shape, 20 + v)
        if v == 1:
            rho[:, 0] = 0.0  # a slab of void cells: E0 = 0, so scale 1 and beta 0
        rho_d, phi_d = rho.float().to(dev), phi.float().to(dev)
        angle = ANGLE + 0.9 * v
        corr = first_collision_deferred_samples(library, rho_d, phi_d, grid, GEOMETRY, angle, om, w, deferred, v)
        raw = torch.empty((G,) + tuple(shape) + (M,), device=dev)
        _fc_sums_raw(library, rho_d, phi_d, grid, GEOMETRY, angle, om, w, samples=raw)
        D8 = M // 8
        blocked = deferred.cell_samples()[v].permute(0, 1, 3, 2, 4)  # [G, blocks, 64, 8, dp] (cells past the
        # grid unused)
        unblocked = from_block_order(blocked, shape, 1)  # [G, R, C, S, 8, dp]
        assert torch.equal(unblocked[..., :D8].reshape(raw.shape), raw)
        if unblocked.shape[-1] > D8:
            assert float(unblocked[..., D8:].abs().max()) == 0.0  # the padding of the runs
        ref, ref_corr = first_collision_corrected_samples(library, rho_d, phi_d, grid, GEOMETRY, angle, om, w)
        # the factors the sums kernel forms in float64 equal the host formula's, and land in block order with the
        # cells past the grid zero
        assert torch.equal(corr.scale, ref_corr.scale) and torch.equal(corr.beta, ref_corr.beta)
        cells = fc_blocks(shape) * 64
        assert torch.equal(deferred.scale[v, :, :cells], to_block_order(ref_corr.scale, shape).reshape(G, -1))
        assert torch.equal(deferred.beta[v, :, :cells], to_block_order(ref_corr.beta, shape).reshape(G, -1))
        if v == 1:
            assert bool((ref_corr.scale[:, 0] == 1.0).all()) and bool((ref_corr.beta[:, 0] == 0.0).all())
        assert rel_err(corr.omega_uncollided, ref_corr.omega_uncollided) < 1e-6
        u_want = torch.zeros(tuple(shape) + (4,), device=dev)
        with torch.cuda.device(dev):
            a, keep, _ = _fc_args(library, rho_d, phi_d, grid, GEOMETRY, angle, om)
            _hip.check(_hip.library().citrus_fc_primary_directions(ctypes.byref(a), u_want.data_ptr(),
                                                                   _hip.stream_of(dev)), "primary directions")
        assert torch.equal(deferred.u[v], to_block_order(u_want, shape).reshape(-1, 4))
        corrected.append(ref)
    corrected = torch.stack(corrected).contiguous()
    gen = torch.Generator(device=dev).manual_seed(3)
    mu = torch.rand(G, *grid.shape, device=dev, generator=gen) * 0.5 + 0.2
    T = torch.randn(V, G, *grid.shape, ctx.num_moments, 4, device=dev, generator=gen) * 0.01
    if contraction == "bf16":
        T = bf16_operand(T)
    print(f"deferred correction: S{sn} P{L} {contraction}", flush=True)
    want = sweep_views(ctx, mu, T, source=corrected, schedule="persistent", contraction=contraction)
    got = sweep_views(ctx, mu, T, source=deferred, schedule="persistent", contraction=contraction)
    assert bool(torch.isfinite(want).all()) and bool(torch.isfinite(got).all())
    assert not deferred.applied and torch.equal(got, want)
    got = sweep_views(ctx, mu, T, source=deferred, schedule="launches", contraction=contraction)
    assert deferred.applied and torch.equal(deferred.apply_(), corrected) and torch.equal(got, want)


def test_blocked_sums_refuse_misaligned_outputs():
    """citrus_fc_sums_blocked stores the samples and the primary directions four floats at a time, so a samples or u
    pointer off a 16-byte boundary is refused with an error before any launch."""
    from citrus_amd.scatter_solver.uncollided import DeferredFirstCollision, _fc_args, _fc_sums_raw

    dev = gpu_device()
    library = seeded_library(4, 2, 120.0, 20.0, seed=15, num_eta=65)
    grid = Grid((4, 4, 4), 0.5)
    q = level_symmetric_quadrature(4)
    om, w = q.omega.to(dev, torch.float32), q.weight
    rho, phi = _problem(library, grid.shape, 3)
    d = DeferredFirstCollision.empty(1, library.num_groups, grid.shape, int(om.shape[0]), dev)
    a, keep, mu_s_l = _fc_args(library, rho.float().to(dev), phi.float().to(dev), grid, GEOMETRY, ANGLE, om)
    wt = w.to(dev, torch.float32).contiguous()
    out = torch.empty(5, library.num_groups, *grid.shape, device=dev)
    lib = _hip.library()
    stream = _hip.stream_of(dev)
    D8 = int(om.shape[0]) // 8
    good = [d.samples[0].data_ptr(), d.scale[0].data_ptr(), d.beta[0].data_ptr(), d.u[0].data_ptr()]
    assert lib.citrus_fc_sums_blocked(ctypes.byref(a), wt.data_ptr(), mu_s_l.data_ptr(), int(mu_s_l.shape[-1]),
                                      out.data_ptr(), good[0], D8, good[1], good[2], good[3], stream) == 0
    for bad in ([good[0] + 4] + good[1:], good[:3] + [good[3] + 8]):
        rc = lib.citrus_fc_sums_blocked(ctypes.byref(a), wt.data_ptr(), mu_s_l.data_ptr(), int(mu_s_l.shape[-1]),
                                        out.data_ptr(), bad[0], D8, bad[1], bad[2], bad[3], stream)
        assert rc != 0 and "16-byte" in lib.citrus_last_error().decode()
    torch.cuda.synchronize()
    del keep
