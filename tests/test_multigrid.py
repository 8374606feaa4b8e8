"""The two-stage (coarse, fine) schedule: prolongation of LD moments (CPU: exact for piecewise-linear fields,
conservative, refuses non-nested grids) and the mixed-stage step on the GPU (both stages' accumulator updates in one
persistent sweep equal each stage's own batched step)."""

from __future__ import annotations

import pytest
import torch

from citrus_amd.multigrid import MultigridSchedule, citrus_multigrid_views_step, prolong_moments, refinement
from citrus_amd.scatter_solver.geometry import Grid


def _linear_moments(grid: Grid, coef) -> torch.Tensor:
    """LD moments [4, R, C, S] of f = a + b x + c y + d z: the mean is f at the cell centre, the slope along an axis
    the change of f over half the cell (u in [-1, 1] spans the cell)."""
    a, b, c, d = coef
    R, C, S = grid.shape
    dx, dy, dz = grid.deltas
    x0, y0, z0 = grid.lower_corner()
    x = x0 + (torch.arange(C, dtype=torch.float64) + 0.5) * dx
    y = y0 + (torch.arange(R, dtype=torch.float64) + 0.5) * dy
    z = z0 + (torch.arange(S, dtype=torch.float64) + 0.5) * dz
    mean = a + b * x[None, :, None] + c * y[:, None, None] + d * z[None, None, :]
    out = torch.empty(4, R, C, S, dtype=torch.float64)
    out[0] = mean
    out[1] = b * dx / 2
    out[2] = c * dy / 2
    out[3] = d * dz / 2
    return out


def test_prolongation_is_exact_for_linear_fields_and_conservative():
    coarse = Grid((3, 4, 2), 0.8, 1.0, 1.5)
    fine = Grid((6, 12, 4), 0.8 / 3, 1.5, 2.25)  # rows x2, cols x3, slices x2 (anisotropic cells)
    assert refinement(coarse, fine) == (2, 3, 2)
    coef = (0.7, -0.3, 0.45, 0.2)
    got = prolong_moments(_linear_moments(coarse, coef)[None, None], coarse, fine)[0, 0]
    assert torch.allclose(got, _linear_moments(fine, coef), rtol=0, atol=1e-12)
    # arbitrary moments: the children's mean is the parent's mean
    gen = torch.Generator().manual_seed(5)
    phi = torch.randn(2, 3, 4, *coarse.shape, generator=gen, dtype=torch.float64)
    child = prolong_moments(phi, coarse, fine)
    parent_mean = child[:, :, 0].reshape(2, 3, 3, 2, 4, 3, 2, 2).mean(dim=(3, 5, 7))
    assert torch.allclose(parent_mean, phi[:, :, 0], rtol=0, atol=1e-12)
    assert torch.allclose(child[:, :, 1, ::2, ::3, ::2], phi[:, :, 1] / 3, rtol=0, atol=1e-15)


def test_prolongation_refuses_grids_that_are_not_nested():
    with pytest.raises(ValueError):
        refinement(Grid((4, 4, 4), 1.0), Grid((6, 6, 6), 2.0 / 3))  # 1.5 fine cells per coarse cell
    with pytest.raises(ValueError):
        prolong_moments(torch.zeros(4, 3, 3, 3), Grid((4, 4, 4), 1.0), Grid((8, 8, 8), 0.5))


@pytest.mark.gpu
def test_mixed_stage_step_in_one_sweep_equals_each_stage():
    from citrus_amd.citrus import CitrusState, citrus_views_step
    from citrus_amd.scatter_solver.config import SolverConfig
    from citrus_amd.scatter_solver.cross_sections import seeded_library
    from citrus_amd.scatter_solver.geometry import ConeGeometry
    from citrus_amd.scatter_solver.scatter_solver import ScatterSolver
    from citrus_amd.utilities import ellipsoid_phantom
    from tests.conftest import gpu_device
    from tests.helpers import F32_REL, rel_err

    dev = gpu_device()
    library = seeded_library(4, 3, 120.0, 20.0, seed=31, num_eta=129)
    geometry = ConeGeometry(30.0, 50.0, 16, 16, 0.5, 0.5)

    def solver(cell):
        config = SolverConfig(legendre_order=3, sweep_quadrature="level_symmetric", sn_order=4,
                              transport_cell_cm=cell, quadrature_point_spacing_cm=1.0, cone_num_polar=3,
                              cone_num_azimuthal=6)
        return ScatterSolver(library, config, geometry, device=dev)

    grid = Grid((8, 8, 8), 0.5)
    rho = ellipsoid_phantom(grid, seed=3).to(dev)
    sched = MultigridSchedule(coarse=solver(1.0), fine=solver(0.5), coarse_steps=1)
    angles = [0.3, 1.9, 4.0, 5.1]
    cones = [sched.fine.cone_quadrature(grid, a) for a in angles]
    # views 2 and 3 take their coarse step first; then all four step: 0, 1 coarse, 2, 3 fine (prolonged)
    states = [CitrusState(cone_quadrature=c) for c in cones]
    _, early, tags = citrus_multigrid_views_step(sched, states[2:], rho, grid, angles[2:])
    assert [s for s, _ in tags] == ["coarse", "coarse"] and tuple(early[0].collided_moments.shape[-3:]) == (4, 4, 4)
    mixed = states[:2] + early
    results, new, tags = citrus_multigrid_views_step(sched, mixed, rho, grid, angles)
    assert [s for s, _ in tags] == ["coarse", "coarse", "fine", "fine"]
    assert [tuple(s.collided_moments.shape[-3:]) for s in new] == [(4, 4, 4)] * 2 + [(8, 8, 8)] * 2
    # each stage on its own
    fc = sched.coarse.material_fields(rho, grid)
    rc, sc = citrus_views_step(sched.coarse, states[:2], fc, angles[:2])
    ff = sched.fine.material_fields(rho, grid)
    grid_c, grid_f = fc.grid_coarse, ff.grid_coarse
    pro = [CitrusState(angle=s.angle, collided_moments=prolong_moments(s.collided_moments, grid_c, grid_f),
                       cone_quadrature=s.cone_quadrature, num_steps=s.num_steps, history=s.history) for s in early]
    rf, sf = citrus_views_step(sched.fine, pro, ff, angles[2:])
    for got, want in zip(new, sc + sf):
        assert torch.equal(got.collided_moments, want.collided_moments)
    for got, want in zip(results, rc + rf):
        assert rel_err(got[0], want[0]) < F32_REL
        assert torch.equal(got[1], want[1])


@pytest.mark.gpu
def test_cropped_multigrid_helical_views_in_one_launch():
    """Helical views on their beam slabs at mixed stages: the one-launch sweep over the four uneven grids equals the
    same step with one launch sequence per grid, bit for bit; states live on their stage's cropped grid."""
    from citrus_amd.citrus import CitrusState
    from citrus_amd.scatter_solver.config import SolverConfig
    from citrus_amd.scatter_solver.cross_sections import seeded_library
    from citrus_amd.scatter_solver.geometry import ConeGeometry
    from citrus_amd.scatter_solver.scatter_solver import ScatterSolver
    from citrus_amd.utilities import ellipsoid_phantom
    from tests.conftest import gpu_device

    dev = gpu_device()
    library = seeded_library(4, 3, 120.0, 20.0, seed=31, num_eta=129)
    angles = [0.0, 1.2, 2.5, 5.0]
    geometry = ConeGeometry(30.0, 50.0, 8, 16, 0.5, 0.5).with_helical_z_shifts(angles, [-3.0, -1.5, 0.0, 3.0])

    def solver(cell):
        config = SolverConfig(legendre_order=3, sweep_quadrature="level_symmetric", sn_order=4,
                              transport_cell_cm=cell, quadrature_point_spacing_cm=2.0, cone_num_polar=3,
                              cone_num_azimuthal=6)
        return ScatterSolver(library, config, geometry, device=dev)

    grid = Grid((16, 16, 32), 0.5)
    rho = ellipsoid_phantom(grid, seed=8).to(dev)
    sched = MultigridSchedule(coarse=solver(2.0), fine=solver(1.0), coarse_steps=1)
    cones = [sched.fine.cone_quadrature(grid, a) for a in angles]
    first = [CitrusState(cone_quadrature=c) for c in cones[:2]]
    _, first, _ = citrus_multigrid_views_step(sched, first, rho, grid, angles[:2], crop=True)
    states = first + [CitrusState(cone_quadrature=c) for c in cones[2:]]  # two fine-stage, two coarse-stage views
    runs = []
    for sweep_schedule in ("auto", "launches"):
        res, new, tags = citrus_multigrid_views_step(sched, states, rho, grid, angles, crop=True,
                                                     sweep_schedule=sweep_schedule)
        runs.append((res, new, tags))
    (res, new, tags), (res2, new2, _) = runs
    assert [s for s, _ in tags] == ["fine", "fine", "coarse", "coarse"]
    for (stage, (k0, k1)), st in zip(tags, new):
        per = 2 if stage == "fine" else 4  # fine voxels (0.5 cm) per transport cell along z
        assert st.collided_moments.shape[-1] == (k1 - k0) // per and (k1 - k0) < 32
    for a, b in zip(new, new2):
        assert torch.equal(a.collided_moments, b.collided_moments)
    for a, b in zip(res, res2):
        assert torch.equal(a[0], b[0]) and bool(torch.isfinite(a[0]).all())
