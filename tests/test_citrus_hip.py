"""The CITRUS view step on the HIP kernels (GPU).

Reference tests mirrored: N CITRUS steps from an empty state equal N source-iteration sweeps
(tests/test_citrus_step.py:43); plus the accumulator of two steps against the FP64 oracle's restatement of the whole
step (uncollided fluence, corrected first-collision moments, sweeps of the summed sources). FP32 bar ``F32_REL``."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from citrus_amd.citrus import (CitrusState, citrus_view_step, citrus_views_step, first_collision_operator,
                               uncollided_operator)
from citrus_amd.scatter_solver.config import SolverConfig
from citrus_amd.scatter_solver.cross_sections import seeded_library
from citrus_amd.scatter_solver.geometry import ConeGeometry, Grid
from citrus_amd.scatter_solver.scatter_solver import ScatterSolver
from citrus_amd.utilities import ellipsoid_phantom
from oracle import first_collision as ofc
from oracle import quadrature as oq
from oracle import rays as orays
from oracle import scattering as osc
from oracle import sweep as osw
from oracle import uncollided as ounc
from tests.conftest import gpu_device
from tests.helpers import F32_REL, rel_err

pytestmark = pytest.mark.gpu


def _solver(dev, spacing=0.5):
    library = seeded_library(4, 3, 120.0, 20.0, seed=31, num_eta=129)
    config = SolverConfig(legendre_order=3, sweep_quadrature="level_symmetric", sn_order=4, transport_cell_cm=1.0,
                          quadrature_point_spacing_cm=spacing, cone_num_polar=3, cone_num_azimuthal=6,
                          max_source_iterations=3, source_iteration_tolerance=1e-12)
    geometry = ConeGeometry(30.0, 50.0, 16, 16, 0.5, 0.5)
    solver = ScatterSolver(library, config, geometry, device=dev)
    grid = Grid((8, 8, 8), 0.5)
    rho = ellipsoid_phantom(grid, seed=3).to(dev)
    return solver, grid, rho


def test_citrus_steps_equal_source_iteration():
    dev = gpu_device()
    solver, grid, rho = _solver(dev)
    fields = solver.material_fields(rho, grid)
    assert fields.q == 2 and fields.grid_coarse.shape == (4, 4, 4)
    angle = 0.4
    state = CitrusState()
    signals = []
    for _ in range(3):
        scatter, primary, state = citrus_view_step(solver, state, fields, angle)
        signals.append(scatter)
    assert state.num_steps == 3 and len(state.history) == 3 and state.history[0] == 1.0
    assert state.history[1] > state.history[2] > 0.0
    phi_uc = uncollided_operator(solver, None, grid, angle, fields)
    result = solver.run_source_iteration(fields, solver.first_collision_moment_source(fields, angle, phi_uc))
    assert result.num_iterations == 3
    assert rel_err(state.collided_moments, result.phi_total) < F32_REL
    assert bool((primary > 0).all()) and bool((signals[-1] > 0).all())
    assert float(signals[-1].sum()) > float(signals[0].sum())  # more collision orders, more scatter
    with pytest.raises(ValueError):
        citrus_view_step(solver, state, fields, angle + 0.1)


def test_two_view_steps_match_oracle():
    dev = gpu_device()
    solver, grid, rho = _solver(dev)
    fields = solver.material_fields(rho, grid)
    angle = 0.9
    _, _, s1 = citrus_view_step(solver, CitrusState(), fields, angle, primary=False)
    _, _, s2 = citrus_view_step(solver, s1, fields, angle, primary=False)
    lib, geo = solver.library, solver.geometry
    G = lib.num_groups
    gq, gc = fields.grid_quadrature, fields.grid_coarse
    rho_q = fields.rho_quadrature.double().cpu().numpy().reshape(2, -1)
    rho_c = fields.rho_coarse.double().cpu().numpy()
    # uncollided fluence at the quadrature points, traced on the quadrature grid
    pts = gq.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    src = orays.source_position(geo.source_iso_dist, angle)
    paths = orays.line_integrals(rho_q.reshape(2, *gq.shape), np.broadcast_to(src, pts.shape), pts,
                                 np.array(gq.lower_corner()), np.array(gq.deltas),
                                 (gq.num_cols, gq.num_rows, gq.num_slices))
    phi_uc = ounc.uncollided_fluence(paths, pts, lib.fine_mu_t.numpy(), lib.fine_group.numpy(),
                                     lib.fine_weight.numpy(), G, geo.source_iso_dist, geo.source_detector_dist,
                                     geo.num_det_rows, geo.num_det_channels, geo.delta_det_row,
                                     geo.delta_det_channel, angle, False)
    # corrected, restricted first-collision source at the sweep directions
    om, w = oq.level_symmetric_quadrature(4)
    ou = ofc.primary_directions(pts, src)
    S = ofc.source_at_directions(lib.mu_s_eta.numpy(), None, rho_q, phi_uc, ou, om)
    I = ofc.quadrature_sums(S, ou, om, w)
    E = ofc.exact_moments(lib.mu_s_l.numpy(), rho_q, phi_uc, 1)
    scale, beta = ofc.correction_factors(I, E)
    fc = ofc.restrict_ld(ofc.apply_correction(S, scale, beta, ou, om).reshape(om.shape[0], G, *gq.shape), 2)
    mu_c = np.einsum("mg,mijk->gijk", lib.mu_t.numpy(), rho_c)
    Y = oq.real_spherical_harmonics(3, om)
    psi1 = osw.sweep(gc.shape, gc.deltas, om, mu_c, fc)
    phi1 = osw.moments(psi1, w, Y)
    assert rel_err(s1.collided_moments, phi1) < F32_REL
    mu_s = np.einsum("mabl,mijk->ablijk", lib.mu_s_l.numpy()[..., :4], rho_c)
    T = osc.scatter_moments(mu_s, phi1)
    psi2 = osw.sweep(gc.shape, gc.deltas, om, mu_c, fc + osw.sources_from_moments(T, Y))
    phi2 = osw.moments(psi2, w, Y)
    assert rel_err(s2.collided_moments, phi2) < F32_REL


def test_first_collision_operator_is_the_first_step():
    dev = gpu_device()
    solver, grid, rho = _solver(dev)
    fields = solver.material_fields(rho, grid)
    phi_uc = uncollided_operator(solver, None, grid, 0.2, fields)
    phi1 = first_collision_operator(solver, None, grid, 0.2, phi_uc, fields)
    _, _, s1 = citrus_view_step(solver, CitrusState(), fields, 0.2, primary=False)
    assert rel_err(s1.collided_moments, phi1) < 1e-6


@pytest.mark.parametrize("spacing,q", [(0.5, 2), (1.0, 1)])
def test_views_step_equals_view_steps(spacing, q):
    """The batched step of V views (one sweep launch sequence for all accumulator updates) equals each view's own
    citrus_view_step, over two steps, with one (cell means of the source) and two quadrature points per cell."""
    dev = gpu_device()
    solver, grid, rho = _solver(dev, spacing)
    fields = solver.material_fields(rho, grid)
    assert fields.q == q
    angles = [0.3, 1.9, 4.0]
    batched = [CitrusState() for _ in angles]
    single = [CitrusState() for _ in angles]
    for _ in range(2):
        results, batched = citrus_views_step(solver, batched, fields, angles)
        for v, a in enumerate(angles):
            scatter, primary, single[v] = citrus_view_step(solver, single[v], fields, a)
            assert rel_err(batched[v].collided_moments, single[v].collided_moments) < F32_REL
            assert rel_err(results[v][0], scatter) < F32_REL
            assert torch.equal(results[v][1], primary)
            assert abs(batched[v].history[-1] - single[v].history[-1]) < 1e-5
    assert [s.num_steps for s in batched] == [2, 2, 2]
    # new densities: the scattering moments kept by the states belong to the old fields and are formed again
    fields2 = solver.material_fields(rho * 1.1, grid)
    results, batched = citrus_views_step(solver, batched, fields2, angles)
    for v, a in enumerate(angles):
        scatter, _, single[v] = citrus_view_step(solver, single[v], fields2, a)
        assert rel_err(batched[v].collided_moments, single[v].collided_moments) < F32_REL
        assert rel_err(results[v][0], scatter) < F32_REL
    with pytest.raises(ValueError):
        citrus_views_step(solver, batched, fields, angles[:2])
    with pytest.raises(ValueError):
        citrus_views_step(solver, batched, fields, [a + 0.1 for a in angles])


def test_views_step_bf16_contraction_error():
    """The view step with the mixed-precision source contraction: scatter image within 1e-3 and accumulator within the
    bfloat16 rounding of the float32 step (two steps, so the contraction acts on a nonzero accumulator)."""
    dev = gpu_device()
    solver, grid, rho = _solver(dev, 1.0)
    fields = solver.material_fields(rho, grid)
    angles = [0.4, 2.0]
    s32 = [CitrusState() for _ in angles]
    s16 = [CitrusState() for _ in angles]
    for _ in range(2):
        r32, s32 = citrus_views_step(solver, s32, fields, angles)
        r16, s16 = citrus_views_step(solver, s16, fields, angles, contraction="bf16")
    for v in range(len(angles)):
        assert rel_err(r16[v][0], r32[v][0]) < 1e-3
        assert rel_err(s16[v].collided_moments, s32[v].collided_moments) < 4e-3


def test_bf16_view_step_meets_the_image_bar_at_convergence():
    """The mixed-precision contraction over 20 view steps of the CPU-runnable configuration's shapes (64^3 phantom at
    0.2 cm, 32^3 transport cells, 8 groups over 120-20 keV, P3, S8, flat 128 x 128 panel of 1.6 mm pixels, SID 100 cm,
    SDD 150 cm, two views): at every step from the fifth on, and at the converged state, the scatter image is within
    1e-3 of the float32 steps from the same empty states (which the oracle tests pin to the FP64 restatement at 1e-5)
    and the scalar-flux cell means within 2e-3."""
    dev = gpu_device()
    library = seeded_library(8, 3, 120.0, 20.0, seed=17, num_eta=257)
    config = SolverConfig(num_groups=8, legendre_order=3, sweep_quadrature="level_symmetric", sn_order=8,
                          transport_cell_cm=0.4, quadrature_point_spacing_cm=0.5)
    geometry = ConeGeometry(100.0, 150.0, 128, 128, 0.16, 0.16)
    solver = ScatterSolver(library, config, geometry, device=dev)
    fine = Grid((64, 64, 64), 0.2)
    fields = solver.material_fields(ellipsoid_phantom(fine, seed=1234).to(dev), fine)
    assert fields.grid_coarse.shape == (32, 32, 32) and fields.q == 1
    angles = [0.0, 1.3]
    s32 = [CitrusState() for _ in angles]
    s16 = [CitrusState() for _ in angles]
    worst = []
    for step in range(20):
        r32, s32 = citrus_views_step(solver, s32, fields, angles)
        r16, s16 = citrus_views_step(solver, s16, fields, angles, contraction="bf16")
        image = max(rel_err(r16[v][0], r32[v][0]) for v in range(len(angles)))
        flux = max(rel_err(s16[v].collided_moments[:, 0, 0], s32[v].collided_moments[:, 0, 0])
                   for v in range(len(angles)))
        worst.append((image, flux))
        print(f"bf16 step {step}: image error {image:.2e}, scalar-flux error {flux:.2e}, change "
              f"{s32[0].history[-1]:.1e}", flush=True)
    assert s32[0].history[-1] < 1e-4  # converged
    for image, flux in worst[4:]:
        assert image < 1e-3 and flux < 2e-3


def test_cropped_helical_step_equals_the_covering_grid_without_the_rest():
    """Each helical view on its beam slab equals the covering-grid step of the same densities with everything outside
    the slab removed (scatter image and the slab's collided moments)."""
    from citrus_amd.citrus import citrus_cropped_views_step
    from citrus_amd.scatter_solver.cropping import crop_slices

    dev = gpu_device()
    library = seeded_library(4, 3, 120.0, 20.0, seed=31, num_eta=129)
    config = SolverConfig(legendre_order=3, sweep_quadrature="level_symmetric", sn_order=4, transport_cell_cm=1.0,
                          quadrature_point_spacing_cm=1.0, cone_num_polar=3, cone_num_azimuthal=6)
    angles = [0.0, 2.5, 5.0]
    geometry = ConeGeometry(30.0, 50.0, 8, 16, 0.5, 0.5).with_helical_z_shifts(angles, [-3.0, 0.0, 3.0])
    solver = ScatterSolver(library, config, geometry, device=dev)
    grid = Grid((16, 16, 32), 0.5)
    rho = ellipsoid_phantom(grid, seed=8).to(dev)
    # one cone set per view for both (a slab's own covering cone set is narrower: another discretization)
    cones = [solver.cone_quadrature(grid, a) for a in angles]
    results, states, slabs = citrus_cropped_views_step(solver, [CitrusState(cone_quadrature=c) for c in cones], rho,
                                                       grid, angles)
    assert len(set(slabs)) == 3 and all(k1 - k0 < 32 for k0, k1 in slabs)
    for v, (a, (k0, k1)) in enumerate(zip(angles, slabs)):
        kept = torch.zeros_like(rho)
        kept[..., k0:k1] = rho[..., k0:k1]
        fields = solver.material_fields(kept, grid)
        (full,), (fstate,) = citrus_views_step(solver, [CitrusState(cone_quadrature=cones[v])], fields, [a])
        assert rel_err(results[v][0], full[0]) < F32_REL
        sub = fstate.collided_moments[..., k0 // 2:k1 // 2]
        assert rel_err(states[v].collided_moments, sub) < F32_REL  # a photon leaving the convex slab never returns


def test_bf16_view_step_at_p5_sixteen_groups_and_s12():
    """The mixed-precision view step at the benchmark's angular and energy shape (16 groups, P5, S12: 21 directions per
    chunk, two direction tiles) on a small grid (16^3 transport cells), two views, four steps: every state stays finite
    and the scatter image is within 1e-3 of the float32 steps."""
    dev = gpu_device()
    library = seeded_library(16, 5, 140.0, 20.0, seed=99, num_eta=257)
    config = SolverConfig(num_groups=16, legendre_order=5, sweep_quadrature="level_symmetric", sn_order=12,
                          transport_cell_cm=1.28, quadrature_point_spacing_cm=2.0, cone_num_polar=4,
                          cone_num_azimuthal=8)
    geometry = ConeGeometry(100.0, 150.0, 64, 64, 0.48, 0.48, use_curved_detector=True)
    solver = ScatterSolver(library, config, geometry, device=dev)
    fine = Grid((64, 64, 64), 0.32)
    fields = solver.material_fields(ellipsoid_phantom(fine, seed=1234).to(dev), fine)
    assert fields.grid_coarse.shape == (16, 16, 16) and fields.q == 1
    angles = [0.0, 1.3]
    s32 = [CitrusState() for _ in angles]
    s16 = [CitrusState() for _ in angles]
    for step in range(4):
        r32, s32 = citrus_views_step(solver, s32, fields, angles)
        r16, s16 = citrus_views_step(solver, s16, fields, angles, contraction="bf16")
        for v in range(len(angles)):
            assert bool(torch.isfinite(s32[v].collided_moments).all()), (step, v, "fp32")
            assert bool(torch.isfinite(s16[v].collided_moments).all()), (step, v, "bf16")
        image = max(rel_err(r16[v][0], r32[v][0]) for v in range(len(angles)))
        print(f"P5 bf16 step {step}: image error {image:.2e}", flush=True)
        assert image < 1e-3
