"""Anti-scatter grid: the direct transmission of a 1-D grid against the brute-force oracle and its closed special
cases, the trilinear table lookup, and the grid angles against geometric invariants (CPU)."""

from __future__ import annotations

import math
from dataclasses import replace

import numpy as np
import pytest
import torch

from citrus_amd.scatter_solver.detector import (GridResponse, analytic_1d_grid, focused_1d_grid, grid_angles,
                                                grid_direct_transmission, lead_mass_attenuation, no_grid,
                                                septum_shadow_transmission)
from citrus_amd.scatter_solver.geometry import ConeGeometry
from oracle import grid as ogrid

PITCH, THICK, HEIGHT = 0.0336, 0.0036, 0.3


def test_direct_transmission_matches_brute_force():
    """Closed form = the oracle's average over entry offsets of the sampled strip path (two midpoint rules, 4000
    points each: bar 1e-3), at strip-crossing shifts below, near and above one pitch."""
    mu_s = np.array([60.0, 8.0, 0.5])
    mu_i = np.array([0.05, 0.0, 0.2])
    th1 = np.array([0.0, 0.004, 0.05, 0.112, 0.3])  # h tan(theta1) = 0, 0.04 p, 0.45 p, 1.0 p, 2.8 p
    th2 = np.array([0.0, 0.4])
    got = grid_direct_transmission(PITCH, THICK, HEIGHT, torch.tensor(mu_s), torch.tensor(mu_i), torch.tensor(th1),
                                   torch.tensor(th2)).numpy()
    want = ogrid.direct_transmission(PITCH, THICK, HEIGHT, mu_s, mu_i, th1, th2)
    assert np.abs(got - want).max() < 1e-3
    # symmetric in the sign of the angles
    neg = grid_direct_transmission(PITCH, THICK, HEIGHT, torch.tensor(mu_s), torch.tensor(mu_i), torch.tensor(-th1),
                                   torch.tensor(-th2)).numpy()
    assert np.array_equal(neg, got)


def test_direct_transmission_special_cases():
    """Normal incidence: ((p - t) e^{-mu_i h} + t e^{-mu_s h}) / p; no strips: e^{-mu_i L}; opaque strips and a
    shift D < p - t: only the entry offsets whose whole track stays in the interspace, (p - t - D) / p, get through."""
    f64 = dict(dtype=torch.float64)
    mu_s, mu_i = torch.tensor([30.0], **f64), torch.tensor([0.1], **f64)
    t0 = grid_direct_transmission(PITCH, THICK, HEIGHT, mu_s, mu_i, torch.tensor([0.0]), torch.tensor([0.0]))
    want = ((PITCH - THICK) * math.exp(-0.1 * HEIGHT) + THICK * math.exp(-30.0 * HEIGHT)) / PITCH
    assert abs(float(t0) - want) < 1e-14
    th1, th2 = torch.tensor([0.02, 0.2], **f64), torch.tensor([0.1], **f64)
    open_ = grid_direct_transmission(PITCH, 0.0, HEIGHT, mu_s, mu_i, th1, th2)
    L = HEIGHT * torch.sqrt(1 + torch.tan(th1) ** 2 + torch.tan(th2) ** 2)
    assert torch.allclose(open_[0, :, 0], torch.exp(-0.1 * L), rtol=1e-13, atol=0)
    opaque = grid_direct_transmission(PITCH, THICK, HEIGHT, torch.tensor([1e8], **f64), torch.tensor([0.0], **f64),
                                      torch.tensor([0.01], **f64), torch.tensor([0.0], **f64))
    shift = HEIGHT * math.tan(0.01)  # < p - t: the photons entering [t, p - D) never meet a strip
    assert abs(float(opaque) - (PITCH - THICK - shift) / PITCH) < 1e-6
    with pytest.raises(ValueError):
        grid_direct_transmission(PITCH, 2 * PITCH, HEIGHT, mu_s, mu_i, th1, th2)


def test_table_lookup_is_trilinear_and_symmetric():
    """GridResponse.at reproduces a table linear in (E, |theta1|, |theta2|) between the nodes, clamps at the ends,
    uses |theta|, and a crossed grid multiplies the two transmissions."""
    e = torch.tensor([10.0, 50.0, 150.0], dtype=torch.float64)
    t1 = torch.tensor([0.0, 0.1, 0.4], dtype=torch.float64)
    t2 = torch.tensor([0.0, 0.5], dtype=torch.float64)
    lin = lambda E, a, b: 0.2 + 0.001 * E + 0.5 * a + 0.3 * b
    table = lin(e[:, None, None], t1[None, :, None], t2[None, None, :])
    g = GridResponse(e, t1, t2, table)
    assert g.depends_on_angles
    E = torch.tensor([20.0, 120.0], dtype=torch.float64)
    a = torch.tensor([0.05, -0.3, 0.7], dtype=torch.float64)  # the last beyond the nodes: clamped to 0.4
    b = torch.tensor([0.25, -0.1, 0.9], dtype=torch.float64)
    got = g.at(E, a, b)
    want = lin(E[:, None], a.abs().clamp(max=0.4)[None], b.abs().clamp(max=0.5)[None])
    assert torch.allclose(got, want, rtol=0, atol=1e-14)
    crossed = GridResponse(e, t1, t2, torch.full_like(table, 0.5), septa_along="rows")
    both = replace(g, crossed=crossed)
    geo = ConeGeometry(30.0, 50.0, 4, 6, 0.5, 0.5)
    pix = geo.pixel_centers(0.3, None, torch.float64)
    om = torch.nn.functional.normalize(pix - torch.tensor(geo.source_position(0.3), dtype=torch.float64), dim=-1)
    assert torch.allclose(both.transmission_at(E, om, pix, geo, 0.3), 0.5 * g.transmission_at(E, om, pix, geo, 0.3),
                          rtol=1e-15, atol=0)
    with pytest.raises(ValueError):
        GridResponse(e, t1, t2, table[:, :2])
    with pytest.raises(ValueError):
        replace(g, crossed=replace(crossed, septa_along="channels"))
    with pytest.raises(NotImplementedError):
        g.at_energy(E, "cpu", torch.float64)


@pytest.mark.parametrize("curved", [False, True])
def test_grid_angles_of_focal_spot_photons(curved):
    """A grid focused at the source-detector distance sees every photon from the focal spot at theta1 = 0. theta2
    is the angle along the septa: on a flat panel it grows across the channels (septa along the channels); on the
    curved panel, whose surface already faces the source across the channels, it is 0 there, and with septa along
    the rows it is the rows' vertical angle. An unfocused grid sees -n at (0, 0)."""
    geo = ConeGeometry(60.0, 100.0, 5, 7, 0.8, 0.8, use_curved_detector=curved)
    ang = 0.7
    pix = geo.pixel_centers(ang, None, torch.float64)
    src = torch.tensor(geo.source_position(ang), dtype=torch.float64)
    om = torch.nn.functional.normalize(pix - src, dim=-1)
    th1, th2 = grid_angles(om, pix, geo, ang)
    assert float(th1.abs().max()) < 1e-12
    th1r, th2r = grid_angles(om, pix, geo, ang, septa_along="rows")
    if curved:
        assert float(th2.abs().max()) < 1e-12
        assert float(th1r.abs().max()) < 1e-12
        assert float(th2r[2].abs().max()) < 1e-12 and float(th2r[0].abs().min()) > 0.005  # off the central row
    else:
        assert float(th2[:, 0].abs().min()) > 0.01 and float(th2[:, 3].abs().max()) < 1e-12
    u, v, n = (torch.tensor(x, dtype=torch.float64) for x in geo.panel_axes(ang))
    flat = replace(geo, use_curved_detector=False)
    a1, a2 = grid_angles(-n[None].expand(3, 3), flat.pixel_centers(ang, None, torch.float64)[0, :3], flat, ang,
                         focal_distance=math.inf)
    assert float(a1.abs().max()) < 1e-15 and float(a2.abs().max()) < 1e-15
    # [K, 3] directions with [rows, channels, 3] pixels give [K, rows, channels]
    k1, _ = grid_angles(om[0, :2], pix, geo, ang)
    assert tuple(k1.shape) == (2, 5, 7)


def test_analytic_grid_table_and_no_grid():
    """analytic_1d_grid tabulates the direct transmission on its nodes (pitch t + h / ratio); no_grid stays
    angle-independent and transmits 1."""
    g = analytic_1d_grid(10.0, THICK, HEIGHT, lambda e: 2000.0 / e, lambda e: 0.0 * e,
                        energy_kev=np.array([30.0, 60.0]), theta1_rad=np.array([0.0, 0.05]),
                        theta2_rad=np.array([0.0]))
    pitch = THICK + HEIGHT / 10.0
    f64 = dict(dtype=torch.float64)
    want = grid_direct_transmission(pitch, THICK, HEIGHT, torch.tensor([2000 / 30.0, 2000 / 60.0], **f64),
                                    torch.zeros(2, **f64), torch.tensor([0.0, 0.05], **f64), torch.tensor([0.0], **f64))
    assert torch.equal(g.transmission, want)
    assert g.depends_on_angles and not no_grid().depends_on_angles
    geo = ConeGeometry(30.0, 50.0, 4, 6, 0.5, 0.5)
    pix = geo.pixel_centers(0.0, None, torch.float64)
    t = no_grid().transmission_at(torch.tensor([40.0, 80.0]), torch.zeros(4, 6, 3, dtype=torch.float64), pix, geo, 0.0)
    assert tuple(t.shape) == (2, 4, 6) and bool((t == 1).all())


@pytest.mark.parametrize("theta_deg", [0.0, 1.0, 3.0, 8.0, 25.0])
def test_septum_shadow_model_against_track_sampling(theta_deg):
    """The reference-signature focused_1d_grid (the lamella-shadow model, detector.py:779): its open share and mean
    lead path equal the brute-force sampled tracks' share of zero paths and mean non-zero path (oracle.grid), and the
    tabulated transmission is open + shadow exp(-mu rho L) of those two numbers at the lead table's attenuation."""
    p, t, h = 1.0 / 60.0, 0.0036, 0.15
    th = math.radians(theta_deg)
    paths = ogrid.lamella_paths(p, t, h, th)
    hit = paths > 1e-12
    open_share = 1.0 - hit.mean()
    mean_path = paths[hit].mean()
    mu = np.array([0.5, 30.0, 300.0])  # 1/cm
    got = septum_shadow_transmission(p, t, h, mu, np.array([th]))[:, 0]
    want = open_share + (1 - open_share) * np.exp(-mu * mean_path)
    assert np.allclose(got, want, rtol=2e-3, atol=2e-4)
    g = focused_1d_grid(energies=np.array([40.0, 90.0]), theta1=np.array([0.0, th]) if th > 0 else None)
    assert g.focal_distance_cm == 150.0 and "10:1" in g.name and not torch.equal(g.transmission[0], g.transmission[1])
    i = int(np.argmin(np.abs(g.theta1_rad.numpy() - th)))
    mu_pb = lead_mass_attenuation(np.array([40.0, 90.0])) * 11.35
    assert np.allclose(g.transmission[:, i, 0].numpy(),
                       septum_shadow_transmission(p, t, h, mu_pb, g.theta1_rad.numpy()[i:i + 1])[:, 0], rtol=1e-12)
    assert torch.equal(g.transmission[..., 0], g.transmission[..., 1])  # no theta2 dependence
    with pytest.raises(ValueError):
        focused_1d_grid(lamella_thickness_cm=0.02)
