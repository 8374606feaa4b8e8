"""HIP ray traversal vs the FP64 oracle (GPU): bit-exact voxel sequences, line integrals, and the five projector
operations (reference tests: tests/test_ray_tracing.py:20-48 chords and midpoints, :126-259 closed forms, :307-331
edge cases, :335-439 fused vs torch vs f64)."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from citrus_amd.scatter_solver.geometry import ConeGeometry, Grid
from citrus_amd.scatter_solver.ray_tracing import HipProjectors, line_integrals, trace_voxel_sequences
from oracle import rays as orays
from tests.conftest import gpu_device
from tests.helpers import F32_REL, rel_err

pytestmark = pytest.mark.gpu


def _grid_box(grid):
    lo = np.array(grid.lower_corner())
    delta = np.array(grid.deltas)
    n = (grid.num_cols, grid.num_rows, grid.num_slices)
    return lo, delta, n


def _edge_segments(grid, rng, count=3000):
    """Random segments plus the reference's edge cases: axis-aligned, on planes, through corners, zero length."""
    lo, delta, n = _grid_box(grid)
    hi = lo + delta * np.array(n)
    ext = hi - lo
    p0 = lo - 0.3 * ext + rng.random((count, 3)) * 1.6 * ext
    p1 = lo - 0.3 * ext + rng.random((count, 3)) * 1.6 * ext
    extra0, extra1 = [], []
    for a in range(3):  # axis-aligned through cell centres and along planes
        for t in (0.5, 1.0, 2.0, 3.5):
            q0 = lo + 0.37 * ext
            q0[(a + 1) % 3] = lo[(a + 1) % 3] + t * delta[(a + 1) % 3]
            q1 = q0.copy()
            q0[a], q1[a] = lo[a] - 1.0, hi[a] + 1.0
            extra0 += [q0, q1]
            extra1 += [q1, q0]
    extra0 += [lo - 1.0, hi + 1.0, lo + 0.5 * ext, lo]  # diagonal through corners, zero length, corner point
    extra1 += [hi + 1.0, lo - 1.0, lo + 0.5 * ext, lo]
    # rays from a far source through lattice points (upstream-like, many near-ties)
    src = np.array([0.37 * ext[0], 40.0, 0.0])
    pts = lo + delta * (rng.integers(0, 4, (200, 3)) + rng.choice([0.0, 0.5], (200, 3)))
    extra0 += [src] * 200
    extra1 += list(pts)
    return np.concatenate([p0, np.array(extra0)]), np.concatenate([p1, np.array(extra1)])


@pytest.mark.parametrize("shape,pitch", [((5, 6, 4), (0.5, 0.4, 0.8)), ((16, 16, 16), (0.32, 0.32, 0.32))])
def test_voxel_sequences_bit_exact(shape, pitch):
    dev = gpu_device()
    grid = Grid(shape, pitch[0], pitch[1] / pitch[0], pitch[2] / pitch[0])
    rng = np.random.default_rng(0)
    p0, p1 = _edge_segments(grid, rng)
    lo, delta, n = _grid_box(grid)
    ref_idx, ref_len, _ = orays.traverse(p0, p1, lo, delta, n)
    idx, seg, count = trace_voxel_sequences(torch.from_numpy(p0).to(dev), torch.from_numpy(p1).to(dev), grid)
    idx, seg, count = idx.cpu().numpy(), seg.cpu().numpy(), count.cpu().numpy()
    ref_count = (ref_idx >= 0).sum(1)
    assert np.array_equal(count, ref_count)
    K = ref_idx.shape[1]
    assert np.array_equal(np.where(np.arange(idx.shape[1])[None] < count[:, None], idx, -1)[:, :K], ref_idx)
    ok = ref_idx >= 0
    assert np.allclose(seg[:, :K][ok], ref_len[ok], rtol=1e-6, atol=0)


def test_line_integrals_uniform_box_and_random_field():
    dev = gpu_device()
    grid = Grid((5, 6, 4), 0.5, 0.8, 1.6)
    rng = np.random.default_rng(1)
    p0, p1 = _edge_segments(grid, rng, 2000)
    lo, delta, n = _grid_box(grid)
    hi = lo + delta * np.array(n)
    ones = torch.ones(1, *grid.shape, device=dev)
    chords = line_integrals(ones, torch.from_numpy(p0), torch.from_numpy(p1), grid).cpu().double().numpy()[0]
    exact = np.array([orays.box_chord(a, b, lo, hi) for a, b in zip(p0, p1)])
    assert np.abs(chords - exact).max() <= 1e-5 * max(1.0, exact.max())
    field = rng.random((3, *grid.shape)).astype(np.float32)
    got = line_integrals(torch.from_numpy(field).to(dev), torch.from_numpy(p0), torch.from_numpy(p1), grid)
    ref = orays.line_integrals(field, p0, p1, lo, delta, n)
    assert rel_err(got, ref) < F32_REL


def _geometry(curved):
    return ConeGeometry(100.0, 150.0, 24, 32, 0.16 * 4, 0.16 * 4, use_curved_detector=curved)


@pytest.mark.parametrize("curved", [False, True])
def test_projector_operations_match_oracle(curved):
    dev = gpu_device()
    grid = Grid((12, 14, 10), 0.4, 1.1, 0.9)
    geo = _geometry(curved)
    angle = 0.7
    rng = np.random.default_rng(2)
    mu = rng.random((2, *grid.shape)).astype(np.float32) * 0.3
    mu_t = torch.from_numpy(mu).to(dev)
    lo, delta, n = _grid_box(grid)
    hi = lo + delta * np.array(n)
    proj = HipProjectors()
    src = orays.source_position(geo.source_iso_dist, angle)
    centers = grid.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    # upstream to voxel centres
    up = proj.upstream_cumulative(mu_t, grid, geo, angle)
    ref = orays.line_integrals(mu, np.broadcast_to(src, centers.shape), centers, lo, delta, n)
    assert rel_err(up.reshape(2, -1), ref) < F32_REL
    # upstream to arbitrary points
    pts = lo + rng.random((500, 3)) * (hi - lo) * 1.2 - 0.1 * (hi - lo)
    got = proj.upstream_to_points(mu_t, grid, geo, angle, torch.from_numpy(pts))
    assert rel_err(got, orays.line_integrals(mu, np.broadcast_to(src, pts.shape), pts, lo, delta, n)) < F32_REL
    # primary, one and 2x2 sub-rays
    for sub in (1, 2):
        pix = orays.pixel_centers(geo.source_iso_dist, geo.source_detector_dist, geo.num_det_rows,
                                  geo.num_det_channels, geo.delta_det_row, geo.delta_det_channel, angle, curved, sub)
        ends = pix.reshape(-1, 3)
        tau = orays.line_integrals(mu, np.broadcast_to(src, ends.shape), ends, lo, delta, n)
        tau = tau.reshape(2, geo.num_det_rows, geo.num_det_channels, sub * sub)
        ref = -np.log(np.exp(-tau).mean(-1))
        got = proj.primary_project(mu_t, grid, geo, angle, samples_per_pixel_axis=sub)
        assert rel_err(got, ref) < F32_REL
    # downstream along cone-like directions, including ones that do not head toward the panel
    _, _, _, nh = orays.panel_frame(geo.source_iso_dist, geo.source_detector_dist, angle)
    w = rng.standard_normal((6, 3))
    w[:4] -= 2.5 * nh
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    down = proj.downstream_cumulative(mu_t, grid, geo, angle, torch.from_numpy(w)).cpu().double().numpy()
    diag = math.sqrt(((hi - lo) ** 2).sum()) * (1 + 1e-9) + 1e-9
    for k in range(6):
        if not (w[k] @ (-nh) > 0):
            assert np.isinf(down[k]).all()
            continue
        t = np.clip(orays.panel_distance(centers, w[k], geo.source_iso_dist, geo.source_detector_dist, angle, curved),
                    0.0, diag)
        ref = orays.line_integrals(mu, centers, centers + t[:, None] * w[k], lo, delta, n)
        assert rel_err(down[k].reshape(2, -1), ref) < F32_REL
    # oblique through pixel centres, shared field and per-direction fields
    pix = orays.pixel_centers(geo.source_iso_dist, geo.source_detector_dist, geo.num_det_rows,
                              geo.num_det_channels, geo.delta_det_row, geo.delta_det_channel, angle, curved)[:, :, 0, 0]
    pix = pix.reshape(-1, 3)
    got = proj.oblique_project(mu_t, grid, geo, angle, torch.from_numpy(w[:3]))
    fields = rng.random((3, 2, *grid.shape)).astype(np.float32)
    got_k = proj.oblique_project(torch.from_numpy(fields).to(dev), grid, geo, angle, torch.from_numpy(w[:3]))
    for k in range(3):
        a, b = orays.oblique_segments(pix, w[k], lo, hi)
        ref = orays.line_integrals(mu, a, b, lo, delta, n).reshape(2, geo.num_det_rows, geo.num_det_channels)
        assert rel_err(got[k], ref) < F32_REL
        ref_k = orays.line_integrals(fields[k], a, b, lo, delta, n).reshape(2, geo.num_det_rows,
                                                                               geo.num_det_channels)
        assert rel_err(got_k[k], ref_k) < F32_REL


@pytest.mark.parametrize("curved,angle", [(False, 0.7), (True, 0.7), (False, 0.0), (True, 1.5 * math.pi)])
def test_downstream_shared_sequence_matches_oracle(curved, angle):
    """One crossing sequence per direction, gathered per voxel and cut at each voxel's panel distance, against the
    FP64 per-ray traversal; with a [G, M] mix it gives the group optical depths from the density channels."""
    dev = gpu_device()
    grid = Grid((12, 14, 10), 0.4, 1.1, 0.9)
    geo = _geometry(curved)
    rng = np.random.default_rng(5)
    rho = rng.random((2, *grid.shape)).astype(np.float32)
    mix = (rng.random((3, 2)) * 0.4).astype(np.float32)
    lo, delta, n = _grid_box(grid)
    hi = lo + delta * np.array(n)
    centers = grid.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    _, _, _, nh = orays.panel_frame(geo.source_iso_dist, geo.source_detector_dist, angle)
    w = rng.standard_normal((7, 3))
    w[:5] -= 2.5 * nh
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    w[5] = -nh  # along the panel normal: axis-aligned at angles that are multiples of pi / 2
    proj = HipProjectors()
    rho_t = torch.from_numpy(rho).to(dev)
    got = proj.downstream_cumulative_mixed(rho_t, None, grid, geo, angle, torch.from_numpy(w)).cpu().double().numpy()
    got_mix = proj.downstream_cumulative_mixed(rho_t, torch.from_numpy(mix), grid, geo, angle,
                                               torch.from_numpy(w)).cpu().double().numpy()
    diag = math.sqrt(((hi - lo) ** 2).sum()) * (1 + 1e-9) + 1e-9
    for k in range(7):
        if not (w[k] @ (-nh) > 0):
            assert np.isinf(got[k]).all() and np.isinf(got_mix[k]).all()
            continue
        t = np.clip(orays.panel_distance(centers, w[k], geo.source_iso_dist, geo.source_detector_dist, angle, curved),
                    0.0, diag)
        ref = orays.line_integrals(rho, centers, centers + t[:, None] * w[k], lo, delta, n)
        assert rel_err(got[k].reshape(2, -1), ref) < F32_REL
        assert rel_err(got_mix[k].reshape(3, -1), mix.astype(np.float64) @ ref) < F32_REL
    # the group fields traced channel by channel agree with the mixed densities
    mu = torch.einsum("gm,mrcs->grcs", torch.from_numpy(mix), torch.from_numpy(rho)).contiguous().to(dev)
    per_group = proj.downstream_cumulative(mu, grid, geo, angle, torch.from_numpy(w[:5])).cpu().double().numpy()
    assert rel_err(got_mix[:5], per_group) < F32_REL
    assert proj.downstream_cumulative_mixed(rho_t, None, grid, geo, angle, torch.zeros(0, 3)).shape == (0, 2, 12, 14,
                                                                                                       10)
    # 1 to 4 channels take their samples eight at a time, more than 4 one at a time: the same sums bit for bit
    extra = torch.from_numpy(rng.random((3, *grid.shape)).astype(np.float32)).to(dev)
    five = proj.downstream_cumulative_mixed(torch.cat([rho_t, extra]), None, grid, geo, angle, torch.from_numpy(w))
    assert torch.equal(five[:, :2].cpu(), torch.from_numpy(got).float())
    one = proj.downstream_cumulative_mixed(rho_t[1:].contiguous(), None, grid, geo, angle, torch.from_numpy(w))
    assert torch.equal(one[:, 0], five[:, 1])


def test_downstream_shared_sequence_limits():
    dev = gpu_device()
    geo = _geometry(False)
    proj = HipProjectors()
    big = Grid((1, 600, 1), 0.1)
    w = torch.tensor([[0.0, 0.0, -1.0]], dtype=torch.float64)
    with pytest.raises(ValueError, match="511"):
        proj.downstream_cumulative_mixed(torch.ones(1, 1, 600, 1, device=dev), None, big, geo, 0.0, w)
    grid = Grid((4, 4, 4), 0.5)
    with pytest.raises(ValueError):
        proj.downstream_cumulative_mixed(torch.ones(2, 4, 4, 4, device=dev), torch.ones(3, 4), grid, geo, 0.0, w)
    seq = proj.crossing_sequences(grid, w.to(dev))
    with pytest.raises(ValueError, match="sequences"):  # records of another grid
        proj.downstream_cumulative_mixed(torch.ones(2, 4, 4, 5, device=dev), None, Grid((4, 4, 5), 0.5), geo, 0.0,
                                         w, seq)
    rec = seq.cpu().view(torch.int32)[0]
    assert int(rec[0]) == 4  # straight down the slices of a 4-voxel column: 4 samples of 0.5 cm
    assert torch.equal(rec[4:8], torch.tensor([0, -1, -2, -3], dtype=torch.int32))
    nbytes = int(seq.shape[1])
    lib = __import__("citrus_amd._hip", fromlist=["library"]).library()
    buf = torch.zeros(1, nbytes, dtype=torch.uint8, device=dev)
    vol, box, wd = torch.ones(600, device=dev), np.array([0, 0, 0, 0.1, 0.1, 0.1]), w.to(dev)
    rc = lib.citrus_ray_downstream_shared(vol.data_ptr(), 600, 1, 1, 1, box.ctypes.data, None, None, wd.data_ptr(),
                                          1, buf.data_ptr(), None, 1, buf.data_ptr(), None)
    assert rc != 0 and "511" in lib.citrus_last_error().decode()


def test_upstream_closed_form_uniform_box():
    """Uniform attenuation: upstream depth = mu * chord of the source-to-centre segment inside the box."""
    dev = gpu_device()
    grid = Grid((8, 8, 8), 0.5)
    geo = _geometry(False)
    mu = torch.full((1, *grid.shape), 0.2, device=dev)
    up = HipProjectors().upstream_cumulative(mu, grid, geo, 0.3).cpu().double().numpy().reshape(-1)
    lo, delta, n = _grid_box(grid)
    hi = lo + delta * np.array(n)
    src = orays.source_position(100.0, 0.3)
    centers = grid.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    exact = np.array([0.2 * orays.box_chord(src, c, lo, hi) for c in centers])
    assert np.abs(up - exact).max() < 1e-5 * exact.max()


def test_held_channel_last_copy_is_scoped_to_its_block():
    """``HipProjectors.holding`` reuses one channel-last copy of a field for the traversals inside its block (the same
    depths as without it) and drops it on leaving: a field changed in place afterwards is traced as it is then."""
    dev = gpu_device()
    grid = Grid((6, 7, 5), 0.4)
    geo = _geometry(False)
    rng = np.random.default_rng(8)
    rho = torch.from_numpy(rng.random((2, *grid.shape)).astype(np.float32)).to(dev)
    _, _, _, nh = orays.panel_frame(geo.source_iso_dist, geo.source_detector_dist, 0.4)
    w = rng.standard_normal((3, 3)) - 2.5 * nh
    w = torch.from_numpy(w / np.linalg.norm(w, axis=1, keepdims=True))
    proj = HipProjectors()
    plain = proj.downstream_cumulative_mixed(rho, None, grid, geo, 0.4, w)
    with proj.holding(rho, grid):
        held = proj.downstream_cumulative_mixed(rho, None, grid, geo, 0.4, w)
        other = proj.downstream_cumulative_mixed(rho * 3.0, None, grid, geo, 0.4, w)  # another tensor: not held
    assert torch.equal(held, plain)
    assert rel_err(other, 3.0 * plain.double()) < F32_REL
    rho.mul_(2.0)
    assert rel_err(proj.downstream_cumulative_mixed(rho, None, grid, geo, 0.4, w), 2.0 * plain.double()) < F32_REL
