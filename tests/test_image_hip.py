"""Collapsed-cone scatter image, primary image and SPR on the HIP kernels vs the FP64 oracle (GPU).

Reference tests mirrored: single-voxel bookkeeping (tests/test_scatter_image.py:131-167), the combined image of the
first-collision and multiple-collision sources (:729-787), the footprint (tests/test_footprint.py) and the primary
closed forms. FP32 bar ``F32_REL = 1e-5`` of the field maximum."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from citrus_amd import _hip
from citrus_amd.scatter_solver.angular_quadrature import Quadrature, hemisphere_quadrature, real_spherical_harmonics
from citrus_amd.scatter_solver.cross_sections import seeded_library, synthetic_library
from citrus_amd.scatter_solver.detector import energy_weighted_response, ideal_counting_response, no_grid
from citrus_amd.scatter_solver.geometry import ConeGeometry, Grid
from citrus_amd.scatter_solver.ray_tracing import HipProjectors
from citrus_amd.scatter_solver.scatter_image import (
    box_matrix, combined_scatter_image, make_direction_context, panel_direction_set, panel_height_field,
    primary_image, scatter_fraction, scatter_image_from_source)
from citrus_amd.scatter_solver.source import SourceModel
from oracle import first_collision as ofc
from oracle import image as oim
from oracle import rays as orays
from tests.conftest import gpu_device
from tests.helpers import F32_REL, rel_err

pytestmark = pytest.mark.gpu


def test_box_matrix_kernel_matches_oracle():
    dev = gpu_device()
    w = torch.tensor([0.2, 1.0, 1.7, 3.3, 9.0, 40.0])
    got = box_matrix(23, w.to(dev))
    assert rel_err(got, oim.box_matrix(23, w.numpy())) < 1e-6


@pytest.mark.parametrize("angle", [0.0, math.pi / 2])
def test_single_voxel_bookkeeping_on_gpu(angle):
    dev = gpu_device()
    grid = Grid((3, 3, 3), 1.0)
    geometry = ConeGeometry(20.0, 30.0, 40, 40, 1.0, 1.0)
    E0 = torch.tensor([2.0, 0.5], device=dev)
    field = torch.zeros(2, *grid.shape, device=dev)
    field[:, 1, 1, 1] = E0
    n_hat = torch.tensor(geometry.panel_axes(angle)[2], dtype=torch.float64)
    q = hemisphere_quadrature(-n_hat, 6, 12, math.radians(55.0), dtype=torch.float32)
    lib = synthetic_library(2, 1)
    out = scatter_image_from_source(lambda d: field.expand(d.shape[0], -1, -1, -1, -1), grid, geometry, angle, q,
                                    torch.zeros_like(field), HipProjectors(), energy_weighted_response(), no_grid(),
                                    lib, chunk=16, footprint=False)
    cos = q.omega.double() @ -n_hat
    expected = E0.double().cpu() * (q.weight.double() / cos).sum()
    assert torch.allclose(out["fluence"].sum((1, 2)).double().cpu() * geometry.pixel_area, expected, rtol=1e-5)
    planar = E0.double().cpu() * q.weight.double().sum()
    assert torch.allclose(out["planar_fluence"].sum((1, 2)).double().cpu() * geometry.pixel_area, planar, rtol=1e-5)
    sig = float(out["signal"].sum()) * geometry.pixel_area
    assert sig == pytest.approx(float((planar * lib.group_energy).sum()), rel=1e-5)


def _image_problem(curved, q_factor, seed=5, shape=(4, 5, 3)):
    gen = torch.Generator().manual_seed(seed)
    library = seeded_library(3, 2, 120.0, 20.0, seed=21, num_eta=65)
    coarse = Grid(shape, 0.8)
    gq = Grid(tuple(n * q_factor for n in coarse.shape), 0.8 / q_factor)
    rho = torch.rand(2, *gq.shape, generator=gen, dtype=torch.float64) * 1.2
    phi = torch.rand(3, *gq.shape, generator=gen, dtype=torch.float64) * 0.01 + 0.001
    T = torch.randn(3, 9, *coarse.shape, generator=gen, dtype=torch.float64) * 0.002
    T[:, 0] = T[:, 0].abs() + 0.01
    mu_c = torch.rand(3, *coarse.shape, generator=gen, dtype=torch.float64) * 0.3 + 0.1
    geometry = ConeGeometry(40.0, 60.0, 12, 14, 0.6, 0.6, use_curved_detector=curved)
    return library, coarse, gq, rho, phi, T, mu_c, geometry


@pytest.mark.parametrize("curved,q_factor,shape", [(False, 1, (4, 5, 3)), (False, 2, (4, 5, 3)), (True, 1, (4, 5, 3)),
                                                   (True, 1, (2, 8, 16))])
def test_combined_scatter_image_matches_oracle(curved, q_factor, shape):
    dev = gpu_device()
    library, coarse, gq, rho, phi, T, mu_c, geometry = _image_problem(curved, q_factor, shape=shape)
    angle = 0.35
    quad = panel_direction_set(geometry, angle, (3, 6), max_polar_angle=gq)
    det, grid_response = energy_weighted_response(), no_grid()
    ctx = make_direction_context(quad, geometry, angle, det, grid_response, library, dev, torch.float32, chunk=5)
    f32 = lambda t: t.float().to(dev)
    out = combined_scatter_image(library, f32(rho), f32(phi), gq, q_factor, f32(T), 2, geometry, angle, quad,
                                 f32(mu_c), HipProjectors(), det, grid_response, context=ctx, grid_coarse=coarse)
    sig_only = combined_scatter_image(library, f32(rho), f32(phi), gq, q_factor, f32(T), 2, geometry, angle, quad,
                                      f32(mu_c), HipProjectors(), det, grid_response, context=ctx,
                                      grid_coarse=coarse, signal_only=True)
    # oracle on the same (float32-rounded) inputs; the direction tables (the panel-facing directions, cos(theta0),
    # signal and planar weights) from the oracle's own restatement of the direction context, from the cone set alone
    d64 = lambda t: t.float().double().cpu().numpy()
    table = (det.energy_kev.numpy(), det.theta_rad.numpy(), det.response.numpy())
    keep, cos0, signal, planar = oim.signal_weights(d64(quad.omega), d64(quad.weight), geometry.source_iso_dist,
                                                   geometry.source_detector_dist, angle, curved,
                                                   library.group_energy.numpy(), table,
                                                   cols=geometry.num_det_channels, dch=geometry.delta_det_channel)
    om = d64(quad.omega)[keep]
    assert np.array_equal(om, ctx.omega.double().cpu().numpy())
    pts = gq.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    ou = ofc.primary_directions(pts, np.array(geometry.source_position(angle)))
    S = ofc.source_at_directions(library.mu_s_eta.numpy(), None, d64(rho).reshape(2, -1), d64(phi).reshape(3, -1),
                                 ou, om).reshape(om.shape[0], 3, *gq.shape)
    Y = real_spherical_harmonics(2, torch.from_numpy(om)).numpy()
    pixels = geometry.pixel_centers(angle, dtype=torch.float64).numpy()
    heights = panel_height_field(gq, geometry, angle, dtype=torch.float64).numpy()
    _, _, n_hat = geometry.panel_axes(angle)
    pc = geometry.panel_center(angle)
    lo_, ext = gq.lower_corner(), gq.extent
    hs = [sum((lo_[a] + c[a] * ext[a] - pc[a]) * n_hat[a] for a in range(3))
          for c in [(i, j, k) for i in (0, 1) for j in (0, 1) for k in (0, 1)]]
    fl, pl, sg = oim.collapsed_cone_image(
        lambda k: S[k], d64(T), Y, d64(mu_c), np.array(gq.lower_corner()), gq.deltas, gq.shape, q_factor, om,
        d64(quad.weight)[keep], d64(quad.cell_width)[keep], pixels,
        (geometry.delta_det_channel, geometry.delta_det_row), heights, 0.5 * (min(hs) + max(hs)),
        signal, cos0, geometry.source_iso_dist, geometry.source_detector_dist, angle, curved, footprint=True,
        planar_weight=planar)
    assert rel_err(out["fluence"], fl) < F32_REL
    assert rel_err(out["planar_fluence"], pl) < F32_REL
    assert rel_err(out["signal_per_group"], sg) < F32_REL
    assert rel_err(sig_only["signal"], sg.sum(0)) < F32_REL
    assert float(out["fluence"].min()) > 0.0
    # optical depths from the two density channels along shared crossing sequences: same image as tracing mu_t
    rho_c = torch.rand(2, *coarse.shape, generator=torch.Generator().manual_seed(9), dtype=torch.float64) + 0.2
    mu_mix = torch.einsum("mg,mrcs->grcs", library.mu_t.double(), rho_c)
    args = (library, f32(rho), f32(phi), gq, q_factor, f32(T), 2, geometry, angle, quad, f32(mu_mix), HipProjectors(),
            det, grid_response)
    per_group = combined_scatter_image(*args, context=ctx, grid_coarse=coarse, signal_only=True)
    mixed = combined_scatter_image(*args, context=ctx, grid_coarse=coarse, signal_only=True,
                                   density_coarse=f32(rho_c))
    assert rel_err(mixed["signal"], per_group["signal"]) < F32_REL


def test_primary_image_and_spr():
    dev = gpu_device()
    lib = synthetic_library(3, 1, num_fine_per_group=3, fine_spread=0.3)
    grid = Grid((6, 8, 5), 0.5)
    geometry = ConeGeometry(40.0, 70.0, 10, 12, 0.8, 0.8)
    angle = 0.3
    rho = torch.zeros(2, *grid.shape, dtype=torch.float64)
    rho[0] = 1.0
    rho[1, 2:4, 3:6, 1:4] = 1.5
    out = primary_image(lib, rho.float().to(dev), grid, geometry, angle, SourceModel(), HipProjectors(),
                        ideal_counting_response(), no_grid())
    pixels = orays.pixel_centers(40.0, 70.0, 10, 12, 0.8, 0.8, angle, False)[:, :, 0, 0]
    src = orays.source_position(40.0, angle)
    lo = np.array(grid.lower_corner())
    n = (grid.num_cols, grid.num_rows, grid.num_slices)
    path = orays.line_integrals(rho.numpy(), np.broadcast_to(src, (120, 3)), pixels.reshape(-1, 3), lo,
                                np.array(grid.deltas), n).reshape(2, 10, 12)
    _, _, _, nh = orays.panel_frame(40.0, 70.0, angle)
    counting = lambda e, th: np.ones((len(e),) + th.shape)
    fl, pl, sg = oim.primary_flat(path, lib.fine_mu_t.numpy(), lib.fine_group.numpy(), lib.fine_weight.numpy(), 3,
                                  pixels, src, nh, lib.group_energy.numpy(), counting)
    assert rel_err(out["fluence"], fl) < F32_REL
    assert rel_err(out["planar_fluence"], pl) < F32_REL
    assert rel_err(out["signal"], sg.sum(0)) < F32_REL
    spr = scatter_fraction(0.3 * out["signal"], out["signal"])
    assert torch.allclose(spr, torch.full_like(spr, 0.3 / 1.3), rtol=1e-6)


def test_primary_image_through_a_focused_grid():
    """The primary behind an angle-dependent grid (focused 1-D septa, direct transmission, pinned on the CPU) is the
    primary without a grid times the grid's transmission at each pixel's incidence angles, per group, to float32
    rounding."""
    from citrus_amd.scatter_solver.detector import analytic_1d_grid

    dev = gpu_device()
    lib = synthetic_library(3, 1, num_fine_per_group=3, fine_spread=0.3)
    grid = Grid((6, 8, 5), 0.5)
    angle = 0.3
    rho = torch.zeros(2, *grid.shape, dtype=torch.float32)
    rho[0] = 1.0
    rho[1, 2:4, 3:6, 1:4] = 1.5
    septa = analytic_1d_grid(8.0, 0.0036, 0.15, lambda e: 3000.0 / e, lambda e: 0.0 * e + 0.05,
                            focal_distance_cm=60.0)  # focused off the source: every pixel has its own angles
    for curved in (False, True):
        geometry = ConeGeometry(40.0, 70.0, 10, 12, 0.8, 0.8, use_curved_detector=curved)
        args = (lib, rho.to(dev), grid, geometry, angle, SourceModel(), HipProjectors(), ideal_counting_response())
        plain = primary_image(*args, no_grid())["signal_per_group"]
        gridded = primary_image(*args, septa)["signal_per_group"]
        pixels = geometry.pixel_centers(angle, None, torch.float64)
        omega = torch.nn.functional.normalize(pixels - torch.tensor(geometry.source_position(angle),
                                                                    dtype=torch.float64), dim=-1)
        trans = septa.transmission_at(lib.group_energy.double(), omega, pixels, geometry, angle)  # [G, rows, cols]
        assert float(trans.max() - trans.min()) > 0.02  # the pixels see the septa at different angles
        assert rel_err(gridded, plain.double().cpu() * trans) < F32_REL


@pytest.mark.parametrize("curved,crossed", [(False, False), (True, False), (False, True)])
def test_scatter_image_through_an_angle_dependent_grid(curved, crossed):
    """The collapsed-cone image behind septa whose transmission depends on the photon's angles (focused off the
    source, so every (direction, pixel) has its own grid angles; crossed: a second set along the rows): every
    direction's blurred image weighted at each pixel by the signal weight times the grid's transmission there, against
    the oracle's image with that factor (the transmission from transmission_at in float64, pinned on the CPU), in
    the full image and the signal-only form; without the grid the same problem gives a different signal."""
    from dataclasses import replace as dc_replace

    from citrus_amd.scatter_solver.detector import analytic_1d_grid

    dev = gpu_device()
    library, coarse, gq, rho, phi, T, mu_c, geometry = _image_problem(curved, 1)
    angle = 0.35
    quad = panel_direction_set(geometry, angle, (3, 6), max_polar_angle=gq)
    det = energy_weighted_response()
    septa = analytic_1d_grid(8.0, 0.0036, 0.15, lambda e: 3000.0 / e, lambda e: 0.0 * e + 0.05,
                            focal_distance_cm=45.0)
    if crossed:
        septa = dc_replace(septa, crossed=dc_replace(septa, septa_along="rows", name="rows"))
    assert septa.depends_on_angles
    ctx = make_direction_context(quad, geometry, angle, det, septa, library, dev, torch.float32, chunk=5)
    assert ctx.pixel_grid is septa and ctx.signal_group_weight is None
    f32 = lambda t: t.float().to(dev)
    args = (library, f32(rho), f32(phi), gq, 1, f32(T), 2, geometry, angle, quad, f32(mu_c), HipProjectors(), det)
    out = combined_scatter_image(*args, septa, context=ctx, grid_coarse=coarse)
    sig_only = combined_scatter_image(*args, septa, context=ctx, grid_coarse=coarse, signal_only=True)
    plain = combined_scatter_image(*args, no_grid(), grid_coarse=coarse, signal_only=True)
    d64 = lambda t: t.float().double().cpu().numpy()
    table = (det.energy_kev.numpy(), det.theta_rad.numpy(), det.response.numpy())
    keep, cos0, signal, planar = oim.signal_weights(d64(quad.omega), d64(quad.weight), geometry.source_iso_dist,
                                                   geometry.source_detector_dist, angle, curved,
                                                   library.group_energy.numpy(), table,
                                                   cols=geometry.num_det_channels, dch=geometry.delta_det_channel)
    om = d64(quad.omega)[keep]
    pts = gq.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    ou = ofc.primary_directions(pts, np.array(geometry.source_position(angle)))
    S = ofc.source_at_directions(library.mu_s_eta.numpy(), None, d64(rho).reshape(2, -1), d64(phi).reshape(3, -1),
                                 ou, om).reshape(om.shape[0], 3, *gq.shape)
    Y = real_spherical_harmonics(2, torch.from_numpy(om)).numpy()
    pixels = geometry.pixel_centers(angle, dtype=torch.float64)
    trans = septa.transmission_at(library.group_energy.double(), torch.from_numpy(om), pixels, geometry,
                                  angle).numpy()  # [G, K, Rd, Cd]
    assert float(trans.max() - trans.min()) > 0.05
    heights = panel_height_field(gq, geometry, angle, dtype=torch.float64).numpy()
    _, _, n_hat = geometry.panel_axes(angle)
    pc = geometry.panel_center(angle)
    lo_, ext = gq.lower_corner(), gq.extent
    hs = [sum((lo_[a] + c[a] * ext[a] - pc[a]) * n_hat[a] for a in range(3))
          for c in [(i, j, k) for i in (0, 1) for j in (0, 1) for k in (0, 1)]]
    fl, pl, sg = oim.collapsed_cone_image(
        lambda k: S[k], d64(T), Y, d64(mu_c), np.array(gq.lower_corner()), gq.deltas, gq.shape, 1, om,
        d64(quad.weight)[keep], d64(quad.cell_width)[keep], pixels.numpy(),
        (geometry.delta_det_channel, geometry.delta_det_row), heights, 0.5 * (min(hs) + max(hs)),
        signal, cos0, geometry.source_iso_dist, geometry.source_detector_dist, angle, curved, footprint=True,
        planar_weight=planar, signal_pixel_weight=lambda k: trans[:, k])
    assert rel_err(out["fluence"], fl) < F32_REL
    assert rel_err(out["planar_fluence"], pl) < F32_REL
    assert rel_err(out["signal_per_group"], sg) < F32_REL
    assert rel_err(sig_only["signal"], sg.sum(0)) < F32_REL
    assert rel_err(plain["signal"], sg.sum(0)) > 0.02


def test_sh_expand_matches_float64_product():
    """The image's multiple-collision expansion on the HIP kernel equals the float64 product to float32 rounding
    (sums of nsh products, four per matrix-core step) for P5 and P3, whole and partial direction tiles, column counts
    with and without the 16-byte stores, and refuses mismatched shapes."""
    from citrus_amd.scatter_solver.scatter_image import sh_expand

    dev = gpu_device()
    gen = torch.Generator().manual_seed(44)
    for K, nsh, ncol in ((64, 36, 3 * 1000 + 17), (64, 36, 4096 + 256), (37, 16, 999), (5, 9, 64), (1, 1, 3)):
        Y = torch.randn(K, nsh, generator=gen, dtype=torch.float64)
        B = torch.randn(nsh, ncol, generator=gen, dtype=torch.float64)
        got = sh_expand(Y.to(dev, torch.float32), B.to(dev, torch.float32)).double().cpu()
        want = Y.float().double() @ B.float().double()
        scale = (Y.abs().float().double() @ B.abs().float().double())
        assert float(((got - want).abs() / scale).max()) < 1e-6, (K, nsh, ncol)
    Y = torch.randn(64, 36, device=dev)
    with pytest.raises(ValueError):
        sh_expand(Y, torch.randn(16, 100, device=dev))
    with pytest.raises(ValueError):
        sh_expand(Y, torch.randn(100, 36, device=dev).t())


def test_multiple_emission_matches_float64():
    """The collapsed multiple-collision emission (expansion, attenuation, group sums in one pass) equals the float64
    evaluation of sum_g w exp(-tau_g) Y.means_g and sum_g exp(-tau_g) Y.means_g: per-group optical depths and one or
    two traced channels with a mix table (+inf channel 0 = no path), direction counts across a 64-direction
    workgroup and cell counts without the 16-byte accesses."""
    from citrus_amd.scatter_solver.scatter_image import multiple_emission

    dev = gpu_device()
    gen = torch.Generator().manual_seed(45)
    for K, nsh, G, N, channels in ((64, 36, 16, 4096, 2), (70, 16, 5, 1001, 0), (3, 9, 3, 64, 1), (17, 36, 32, 130, 2)):
        Y = torch.randn(K, nsh, generator=gen, dtype=torch.float64)
        means = torch.randn(nsh, G, N, generator=gen, dtype=torch.float64)
        w = torch.rand(K, G, generator=gen, dtype=torch.float64)
        if channels:
            tau = 3.0 * torch.rand(K, channels, N, generator=gen, dtype=torch.float64)
            tau[:, 0, ::7] = math.inf
            mix = torch.rand(G, channels, generator=gen, dtype=torch.float64)
            mix[0, -1] = 0.0  # 0 x inf in the mix
            tg = torch.einsum("gc,kcn->kgn", mix.float().double(), tau.float().double())
            tg[:, :, ::7] = math.inf
        else:
            tau = 3.0 * torch.rand(K, G, N, generator=gen, dtype=torch.float64)
            mix = None
            tg = tau.float().double()
        s = torch.einsum("kl,lgn->kgn", Y.float().double(), means.float().double())
        e = torch.exp(-tg) * s
        want = torch.stack([torch.einsum("kg,kgn->kn", w.float().double(), e), e.sum(1)], 1)
        s_abs = torch.einsum("kl,lgn->kgn", Y.float().double().abs(), means.float().double().abs())
        scale = torch.stack([torch.einsum("kg,kgn->kn", w.float().double(), torch.exp(-tg) * s_abs),
                             (torch.exp(-tg) * s_abs).sum(1)], 1) + 1e-30
        f32 = lambda t: t.to(dev, torch.float32).contiguous()
        got = multiple_emission(f32(Y), f32(means), f32(tau), f32(w), None if mix is None else f32(mix))
        err = float(((got.double().cpu() - want).abs() / scale).max())
        assert err < 2e-6, (K, nsh, G, N, channels, err)
    with pytest.raises(ValueError):
        multiple_emission(f32(Y), f32(means), f32(tau[:, :1]), f32(w), f32(mix))


def test_footprint_filters_match_the_torch_footprint():
    """citrus_footprint_filters (mean heights from the emission kernel's per-block sums, widths, both box filters)
    equals the torch form mean_height_from_sums -> widths -> box_matrix to float32 rounding, including a direction
    whose emission vanishes (the box-centre height)."""
    from citrus_amd.scatter_solver.scatter_image import _Footprint, box_matrix

    dev = gpu_device()
    library, coarse, gq, rho, phi, T, mu_c, geometry = _image_problem(True, 1)
    angle = 0.35
    quad = panel_direction_set(geometry, angle, (3, 6), max_polar_angle=gq)
    ctx = make_direction_context(quad, geometry, angle, energy_weighted_response(), no_grid(), library, dev,
                                 torch.float32, chunk=5)
    fp = _Footprint(gq, geometry, angle, ctx)
    gen = torch.Generator().manual_seed(46)
    k = ctx.chunk_bounds[1][1] - ctx.chunk_bounds[1][0]
    partial = torch.rand(k, 37, 2, generator=gen)
    partial[:, :, 1] *= 10.0
    partial[0] = 0.0  # no emission: the box-centre height
    partial = partial.to(dev)
    bct, br = fp.filters(1, partial)
    w = fp.widths(1, fp.mean_height_from_sums(partial.sum(1)))
    Rd, Cd = ctx.image_shape
    want_c = box_matrix(Cd, w[:, 0]).transpose(1, 2)
    want_r = box_matrix(Rd, w[:, 1])
    assert float((bct - want_c).abs().max()) < 1e-5
    assert float((br - want_r).abs().max()) < 1e-5
    image = torch.rand(k, 1, Rd, Cd, generator=gen).to(dev)
    from citrus_amd.scatter_solver.scatter_image import box_blur_rows_cols
    blurred = box_blur_rows_cols(image, w[:, 0], w[:, 1])
    assert rel_err(fp.blur_with(image, bct, br), blurred) < F32_REL
    # channel weights folded into the channel filters; the sum over the chunk's images accumulated in place
    pw = torch.rand(k, Cd, generator=gen).to(dev)
    acc = torch.ones(Rd, Cd, device=dev)
    fp.accumulate_blurred(acc, image, *fp.filters(1, partial, pw))
    assert rel_err(acc, 1.0 + torch.einsum("kgrc,kc->rc", blurred, pw)) < F32_REL
    with pytest.raises(ValueError):
        fp.filters(1, partial, pw[:, 1:].contiguous())


@pytest.mark.parametrize("height_scale", [1.0, 0.01, 40.0])
def test_banded_footprint_blur_equals_the_dense_filters(height_scale):
    """citrus_footprint_blur (the box filters summed window by window, the directions summed per pixel) equals the
    dense filter matrices' two batched products and direction sum (accumulate_blurred), with and without channel
    weights; the height scale makes windows of a few pixels, below one pixel (clamped) and wider than the panel
    (clipped and renormalised); a direction without emission takes the box-centre height."""
    from citrus_amd.scatter_solver.scatter_image import _Footprint

    dev = gpu_device()
    library, coarse, gq, rho, phi, T, mu_c, geometry = _image_problem(True, 1)
    angle = 0.35
    quad = panel_direction_set(geometry, angle, (3, 6), max_polar_angle=gq)
    ctx = make_direction_context(quad, geometry, angle, energy_weighted_response(), no_grid(), library, dev,
                                 torch.float32, chunk=5)
    fp = _Footprint(gq, geometry, angle, ctx)
    gen = torch.Generator().manual_seed(47)
    k = ctx.chunk_bounds[1][1] - ctx.chunk_bounds[1][0]
    partial = torch.rand(k, 37, 2, generator=gen)
    partial[:, :, 1] *= 10.0 * height_scale
    partial[0] = 0.0
    partial = partial.to(dev)
    Rd, Cd = ctx.image_shape
    image = torch.rand(k, 1, Rd, Cd, generator=gen).to(dev)
    pw = torch.rand(k, Cd, generator=gen).to(dev)
    for weight in (None, pw):
        want = torch.full((Rd, Cd), 0.5, device=dev)
        fp.accumulate_blurred(want, image, *fp.filters(1, partial, weight))
        got = torch.full((Rd, Cd), 0.5, device=dev)
        fp.blur_accumulate(got, image, 1, partial, weight)
        assert rel_err(got, want) < F32_REL, (height_scale, weight is None)
    with pytest.raises(ValueError):
        fp.blur_accumulate(got, image, 1, partial, pw[:, 1:].contiguous())


def test_multiple_emission_at_the_view_step_size():
    """At the 64^3 view step's size (K = 64 directions, P5, 16 groups, 262,144 cells, 2 traced channels) the
    collapsed multiple-collision emission equals the expansion on sh_expand followed by the attenuation and group sums
    in float64, relative to the sums of magnitudes (bar 2e-6): the 32-bit moment offsets and the 64-cell workgroups
    hold at full size."""
    from citrus_amd.scatter_solver.scatter_image import multiple_emission, sh_expand

    dev = gpu_device()
    gen = torch.Generator(device=dev).manual_seed(47)
    K, nsh, G, N = 64, 36, 16, 64 ** 3
    Y = torch.randn(K, nsh, device=dev, generator=gen)
    means = torch.randn(nsh, G, N, device=dev, generator=gen)
    tau = 3.0 * torch.rand(K, 2, N, device=dev, generator=gen)
    mix = torch.rand(G, 2, device=dev, generator=gen)
    w = torch.rand(K, G, device=dev, generator=gen)
    got = multiple_emission(Y, means, tau, w, mix).double()
    S = sh_expand(Y, means.view(nsh, G * N)).view(K, G, N)
    S_abs = sh_expand(Y.abs(), means.abs().view(nsh, G * N)).view(K, G, N)
    del means
    e_t = torch.exp(-torch.einsum("gc,kcn->kgn", mix.double(), tau.double()))
    for c, weight in ((0, w.double()), (1, torch.ones_like(w, dtype=torch.float64))):
        want = torch.einsum("kg,kgn->kn", weight, e_t * S.double())
        scale = torch.einsum("kg,kgn->kn", weight, e_t * S_abs.double()) + 1e-30
        assert float(((got[:, c] - want).abs() / scale).max()) < 2e-6, c


def _cone_directions(grid, geometry, angle, K, half_angle_deg, gen):
    """K unit vectors within ``half_angle_deg`` of the mean primary direction of the grid's voxel centres: a narrow
    window of eta nodes, as the panel-facing directions of one image chunk give."""
    pts = grid.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    axis = ofc.primary_directions(pts, np.array(geometry.source_position(angle))).mean(0)
    axis /= np.linalg.norm(axis)
    a = np.cross(axis, [0.0, 0.0, 1.0])
    a /= np.linalg.norm(a)
    b = np.cross(axis, a)
    th = np.radians(half_angle_deg) * np.sqrt(torch.rand(K, generator=gen, dtype=torch.float64).numpy())
    ph = 2 * np.pi * torch.rand(K, generator=gen, dtype=torch.float64).numpy()
    om = (np.cos(th)[:, None] * axis + np.sin(th)[:, None] * (np.cos(ph)[:, None] * a + np.sin(ph)[:, None] * b))
    return torch.from_numpy(om).float()


def _node_windows(library, grid, geometry, angle, om, K):
    """Largest eta-node window (lower nodes spanned + 2) over the emission kernel's point blocks and direction tiles,
    from the FP64 restatement of the table position (uniform nodes)."""
    pts = grid.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    u = ofc.primary_directions(pts, np.array(geometry.source_position(angle)))
    n = int(library.mu_s_eta.shape[-1])
    eta = np.clip(u @ om.double().numpy().T, -1.0, 1.0)
    i0 = np.clip(np.floor((eta + 1.0) / 2.0 * (n - 1)), 0, n - 2)
    blk, widest = _hip.EMISSION_BLOCK, 0
    for b in range(0, len(pts), blk):
        for t in range(0, K, 64):
            w = i0[b:b + blk, t:t + 64]
            widest = max(widest, int(w.max() - w.min()) + 2)
    return widest


@pytest.mark.parametrize("G", [5, 16, 32])
@pytest.mark.parametrize("spread", ["cone", "wide", "sphere"])
def test_emission_source_equals_the_first_collision_source(G, spread):
    """The emission kernel's first-collision source, read off with zero depths and no multiple-collision part, equals
    citrus_fc_source's bit for bit: narrow direction cones take the node-contracted path (each point's source at the
    nodes of the window formed once in LDS and interpolated per direction), directions over the sphere the per-pair
    evaluation; 70 directions make a whole and a partial tile of 64, the 990 points a partial last block."""
    from citrus_amd.scatter_solver.scatter_image import _attenuated_emission
    from citrus_amd.scatter_solver.uncollided import first_collision_source_at_directions

    dev = gpu_device()
    library = seeded_library(G, 3, 120.0, 20.0, seed=29, num_eta=129)
    grid = Grid((9, 10, 11), 0.5)
    geometry = ConeGeometry(40.0, 60.0, 12, 14, 0.6, 0.6)
    angle = 0.7
    gen = torch.Generator().manual_seed(11)
    P, K = grid.num_voxels, 70
    rho = (torch.rand(2, *grid.shape, generator=gen, dtype=torch.float64) * 1.2).float()
    rho[0, :2] = 0.0  # points without the first material
    phi = (torch.rand(G, *grid.shape, generator=gen, dtype=torch.float64) * 0.01 + 0.001).float()
    if spread in ("cone", "wide"):  # wide: windows of 24-27 nodes (12 at 32 groups), near the LDS limit of 30 (15)
        om = _cone_directions(grid, geometry, angle, K, 6.0 if spread == "cone" else (28.0 if G > 16 else 45.0), gen)
        widest = _node_windows(library, grid, geometry, angle, om, K)
        assert widest <= (13 if G > 16 else 27) and (spread == "cone" or widest >= 10)
    else:
        om = torch.randn(K, 3, generator=gen, dtype=torch.float64)
        om = (om / om.norm(dim=1, keepdim=True)).float()
        assert _node_windows(library, grid, geometry, angle, om, K) > 40
    d = lambda t: t.to(dev)
    want = first_collision_source_at_directions(library, d(rho), d(phi), grid, geometry, angle, d(om))
    out, partial = _attenuated_emission(library, d(rho), d(phi), grid, geometry, angle, d(om),
                                        d(torch.zeros(K, G, P)), d(torch.zeros(G, P)), d(torch.zeros(K, G, P)), 1,
                                        None, None)
    assert torch.equal(out, want)
    sums = want.reshape(K, G, P).sum(1).double().cpu()
    nb = -(-P // _hip.EMISSION_BLOCK)
    ref = torch.stack([sums[:, b * _hip.EMISSION_BLOCK:(b + 1) * _hip.EMISSION_BLOCK].sum(1) for b in range(nb)], 1)
    assert rel_err(partial[..., 0], ref) < F32_REL
    assert float(partial[..., 1].abs().max()) == 0.0  # no heights: h = 0


@pytest.mark.parametrize("spread", ["sphere", "cone"])
@pytest.mark.parametrize("G", [5, 8, 16])
def test_collapsed_emission_with_split_groups_matches_oracle(G, spread):
    """The view step's collapsed emission against the FP64 restatement: sum_g w_kg S_g(p, w_k) exp(-sum_c mix_gc
    tau_kc) plus the collapsed multiple-collision part, unreachable cells (tau = inf) contributing nothing, and the
    per-block sums of the footprint's emission-weighted heights. 8 and 16 groups fill the kernel's group width (the
    view step's straight-line form), 5 do not (the general form); directions over the sphere and in a cone (the
    node-contracted path)."""
    from citrus_amd.scatter_solver.scatter_image import _attenuated_emission

    dev = gpu_device()
    library = seeded_library(G, 3, 120.0, 20.0, seed=23, num_eta=129)
    grid = Grid((6, 8, 9), 0.5)
    geometry = ConeGeometry(40.0, 60.0, 12, 14, 0.6, 0.6)
    angle = 0.4
    gen = torch.Generator().manual_seed(8)
    P, K = grid.num_voxels, 7
    rho = (torch.rand(2, *grid.shape, generator=gen, dtype=torch.float64) * 1.2).float()
    phi = (torch.rand(G, *grid.shape, generator=gen, dtype=torch.float64) * 0.01 + 0.001).float()
    if spread == "cone":  # the node-contracted path
        om = _cone_directions(grid, geometry, angle, K, 8.0, gen)
    else:
        om = torch.randn(K, 3, generator=gen, dtype=torch.float64)
        om = (om / om.norm(dim=1, keepdim=True)).float()
    tau = torch.rand(K, 2, P, generator=gen) * 2.0
    tau[0, :, :5] = float("inf")
    mix = torch.rand(G, 2, generator=gen)
    multiple = torch.randn(K, 2, P, generator=gen) * 1e-3
    w = torch.rand(K, G, generator=gen)
    heights = torch.randn(P, generator=gen)
    d = lambda t: t.to(dev)
    out, partial = _attenuated_emission(library, d(rho), d(phi), grid, geometry, angle, d(om), d(tau),
                                        d(torch.zeros(G, P)), d(multiple), 1, d(heights), d(w), tau_mix=d(mix),
                                        collapsed=True)
    pts = grid.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    ou = ofc.primary_directions(pts, np.array(geometry.source_position(angle)))
    S = ofc.source_at_directions(library.mu_s_eta.numpy(), None, rho.double().reshape(2, -1).numpy(),
                                 phi.double().reshape(G, -1).numpy(), ou, om.double().numpy())  # [K, G, P]
    t = np.einsum("gc,kcp->kgp", mix.double().numpy(), tau.double().numpy())
    e = S * np.exp(-t)
    want = np.einsum("kg,kgp->kp", w.double().numpy(), e) + multiple[:, 0].double().numpy()
    esum = e.sum(1) + multiple[:, 1].double().numpy()
    assert rel_err(out.reshape(K, P), want) < F32_REL
    assert bool(torch.isfinite(out).all()) and float(np.abs(e[0, :, :5]).max()) == 0.0
    nb = int(partial.shape[1])
    sums = np.zeros((K, nb, 2))
    blk = _hip.EMISSION_BLOCK
    assert nb == -(-P // blk)
    for b in range(nb):
        sl = slice(blk * b, min(P, blk * (b + 1)))
        sums[:, b, 0] = esum[:, sl].sum(1)
        sums[:, b, 1] = (esum[:, sl] * heights[sl].double().numpy()).sum(1)
    assert rel_err(partial, sums) < F32_REL
