"""Host-side setup of the view step (CPU): configuration, cone direction sets, detector tables, density coarsening."""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from citrus_amd.scatter_solver.angular_quadrature import hemisphere_quadrature
from citrus_amd.scatter_solver.config import SolverConfig
from citrus_amd.scatter_solver.detector import (absorber_response, detector_angle, energy_weighted_response,
                                                ideal_counting_response, no_grid)
from citrus_amd.scatter_solver.geometry import ConeGeometry, Grid
from citrus_amd.scatter_solver.scatter_image import covering_polar_angle, panel_direction_set
from citrus_amd.scatter_solver.scatter_solver import coarsen_density


def test_config_validation_and_quadrature_points():
    c = SolverConfig(transport_cell_cm=0.4, quadrature_point_spacing_cm=0.5)
    assert c.quadrature_points(0.2) == 1  # configs[0]: q = min(max(1, round(0.8)), max(1, round(2))) = 1
    assert SolverConfig(transport_cell_cm=1.0, quadrature_point_spacing_cm=0.5).quadrature_points(0.5) == 2
    for bad in (dict(legendre_order=8), dict(sweep_quadrature="gauss"), dict(uncollided_trace_grid="medium"),
                dict(max_source_iterations=0), dict(direction_chunk=0)):
        with pytest.raises(ValueError):
            SolverConfig(**bad).validate()


def test_reference_solver_config_constructs_unchanged():
    """The configuration of the reference's quadrature-grid tests (test_quadrature_grid.py:38-40) constructs and
    validates as written; the reference's validation rules reject what they reject there (config.py:177-222)."""
    from dataclasses import replace

    c = SolverConfig(num_groups=4, legendre_order=3, num_polar=4, num_azimuthal=8, cone_num_polar=8,
                     cone_num_azimuthal=16, transport_cell_cm=4.0, quadrature_point_spacing_cm=2.0,
                     source_iteration_tolerance=1e-4, max_source_iterations=40, direction_chunk=8)
    assert c.validate() is c and c.quadrature_points(2.0) == 2 and c.materials[3] == "water"
    assert replace(c, uncollided_trace_grid="fine").validate().uncollided_trace_grid == "fine"
    assert SolverConfig(group_edges_kev=[120, 80, 60, 40, 30, 25, 20]).group_edges_kev == (120, 80, 60, 40, 30, 25, 20)
    assert SolverConfig(grouping_medium={"water": 1.0, "air": 0.0}).validate().grouping_medium == (
        ("air", 0.0), ("water", 1.0))
    assert c.to_dict()["num_groups"] == 4
    for bad in (dict(num_polar=3), dict(num_azimuthal=6), dict(num_groups=0), dict(group_edges_kev=(120.0, 20.0)),
                dict(group_edges_method="linear"), dict(grouping_object_size_cm=0.0),
                dict(grouping_medium=(("water", -1.0),)), dict(grouping_medium="lead"), dict(num_eta_table=1),
                dict(angular_table_grid="log"), dict(materials=("water",)), dict(dtype="float16"),
                dict(group_averaged_responses=1)):
        with pytest.raises(ValueError):
            replace(c, **bad).validate()


def test_cone_set_weights_and_axis():
    axis = torch.tensor([0.3, -0.9, 0.2], dtype=torch.float64)
    q = hemisphere_quadrature(axis, 5, 10, 0.7, tangent_axes=([1.0, 0.0, 0.0], [0.0, 0.0, 1.0]))
    a = axis / axis.norm()
    assert abs(float(q.weight.sum()) - 2 * math.pi * (1 - math.cos(0.7))) < 1e-12
    cosines = q.omega @ a
    assert float(cosines.min()) > math.cos(0.7) and torch.allclose(q.omega.norm(dim=-1), torch.ones(50, dtype=torch.float64))
    assert q.cell_width.shape == (50, 2) and bool((q.cell_width > 0).all())
    # a cell's gnomonic width is about its angular size: wider cells far from the axis
    assert float(q.cell_width[-1].max()) > float(q.cell_width[0].max())


def test_covering_polar_angle_contains_the_box():
    grid = Grid((6, 8, 5), 0.7)
    geometry = ConeGeometry(40.0, 60.0, 20, 24, 0.8, 0.8)
    angle = 0.6
    cover = covering_polar_angle(grid, geometry, angle)
    g = torch.Generator().manual_seed(0)
    lo, ext = torch.tensor(grid.lower_corner(), dtype=torch.float64), torch.tensor(grid.extent, dtype=torch.float64)
    r = lo + torch.rand(300, 3, generator=g, dtype=torch.float64) * ext
    uh, vh, nh = (torch.tensor(a, dtype=torch.float64) for a in geometry.panel_axes(angle))
    pc = torch.tensor(geometry.panel_center(angle), dtype=torch.float64)
    hu, hv = geometry.panel_half_extent()
    uv = (2 * torch.rand(300, 2, generator=g, dtype=torch.float64) - 1) * torch.tensor([hu, hv], dtype=torch.float64)
    p = pc + uv[:, :1] * uh + uv[:, 1:] * vh
    d = p[:, None] - r[None]
    theta = torch.arccos((d @ -nh) / torch.linalg.norm(d, dim=-1))
    assert float(theta.max()) < cover
    q = panel_direction_set(geometry, angle, (4, 8), max_polar_angle=grid, dtype=torch.float64)
    assert float((q.omega @ -nh).min()) >= math.cos(cover) - 1e-12


def test_detector_tables():
    e = torch.tensor([30.0, 60.0])
    th = torch.tensor([[0.0, 0.5], [1.0, 1.4]])
    assert torch.equal(ideal_counting_response().at(e, th), torch.ones(2, 2, 2))
    assert torch.allclose(energy_weighted_response().at(e, th), e[:, None, None].expand(2, 2, 2))
    mu = lambda en: 5.0 * (30.0 / en) ** 2.5
    d = absorber_response(0.05, 4.5, mu, energies=np.array([20.0, 30.0, 60.0, 100.0]),
                          thetas=np.radians(np.linspace(0, 80, 81)))
    got = d.at(torch.tensor([30.0], dtype=torch.float64), torch.tensor([math.radians(40.0)], dtype=torch.float64))
    exact = 30.0 * (1 - math.exp(-mu(30.0) * 4.5 * 0.05 / math.cos(math.radians(40.0))))
    assert abs(float(got) - exact) < 1e-12  # on the nodes the table is exact
    assert torch.equal(no_grid().at_energy(e, "cpu", torch.float64), torch.ones(2, dtype=torch.float64))
    geo = ConeGeometry(40.0, 60.0, 4, 4, 1.0, 1.0)
    nh = torch.tensor(geo.panel_axes(0.3)[2], dtype=torch.float64)
    assert float(detector_angle(-nh[None], geo, 0.3)) < 1e-7


def test_coarsen_density_block_means():
    fine = Grid((4, 6, 2), 0.5)
    coarse = Grid((2, 3, 1), 1.0)
    rho = torch.arange(2 * 48, dtype=torch.float64).reshape(2, 4, 6, 2)
    c = coarsen_density(rho, fine, coarse)
    assert c.shape == (2, 2, 3, 1) and float(c[1, 1, 2, 0]) == float(rho[1, 2:4, 4:6, 0:2].mean())
    assert torch.allclose(c.sum() * 8, rho.sum())
    with pytest.raises(ValueError):
        coarsen_density(rho, fine, Grid((3, 3, 1), 1.0))


def test_sweep_tuner_candidates():
    from citrus_amd.scatter_solver.autotune import chunk_candidates

    assert chunk_candidates(21) == [21, 11, 7, 8]
    assert chunk_candidates(10) == [10, 5, 4, 8]
    assert chunk_candidates(3) == [3, 2, 1]
    assert chunk_candidates(40) == [32, 20, 14, 8]


def test_helical_scan_and_view_angles():
    from citrus_amd.utilities import helical_scan, view_angles

    a = view_angles(8, 0.0, 4 * math.pi)
    assert np.allclose(a, np.arange(8) * 4 * math.pi / 8)
    geo = helical_scan(ConeGeometry(40.0, 60.0, 8, 8, 1.0, 1.0), a, pitch_cm=2.0, center_cm=0.5)
    mid = 0.5 * (a[0] + a[-1])
    for ang in a:
        assert abs(geo.z_shift(ang) - (0.5 + 2.0 * (ang - mid) / (2 * math.pi))) < 1e-12
        assert geo.source_position(ang)[2] == geo.z_shift(ang) == geo.panel_center(ang)[2]
    with pytest.raises(ValueError):
        view_angles(0)


def test_beam_slab_contains_the_beam_and_crops_align():
    from citrus_amd.scatter_solver.cropping import beam_slice_range, crop_slices

    grid = Grid((12, 14, 40), 0.5)
    geo = ConeGeometry(30.0, 50.0, 10, 16, 0.6, 0.6, use_curved_detector=True).with_helical_z_shifts(
        [0.0, 10.0], [-4.0, 4.0])
    g = torch.Generator().manual_seed(4)
    for angle in (0.0, 2.3, 7.9):
        k0, k1 = beam_slice_range(geo, angle, grid, multiple=2)
        assert k0 % 2 == 0 and (k1 % 2 == 0 or k1 == grid.num_slices) and 0 <= k0 < k1 <= grid.num_slices
        zl = grid.lower_corner()[2] + k0 * 0.5
        zh = grid.lower_corner()[2] + k1 * 0.5
        # rays from the source to panel points: inside the volume's xy box they stay in [zl, zh]
        u, v = geo.pixel_uv(dtype=torch.float64)
        uu = u[torch.randint(0, u.numel(), (400,), generator=g)]
        vv = v[torch.randint(0, v.numel(), (400,), generator=g)]
        pts = geo.panel_points(uu, vv, angle)
        src = torch.tensor(geo.source_position(angle), dtype=torch.float64)
        t = torch.linspace(0, 1, 2001, dtype=torch.float64)[:, None, None]
        ray = src + t * (pts - src)  # [2001, 400, 3]
        inside = (ray[..., 0].abs() <= 3.5) & (ray[..., 1].abs() <= 3.0)
        zs = ray[..., 2][inside]
        assert float(zs.min()) >= zl - 1e-9 and float(zs.max()) <= zh + 1e-9
        rho = torch.arange(2 * 12 * 14 * 40, dtype=torch.float64).reshape(2, 12, 14, 40)
        sub, rs = crop_slices(grid, rho, k0, k1)
        zc = sub.axis_coordinates(dtype=torch.float64)[2]
        assert torch.allclose(zc, grid.axis_coordinates(dtype=torch.float64)[2][k0:k1])
        assert torch.equal(rs, rho[..., k0:k1])


def test_library_file_round_trip_and_reference_format(tmp_path):
    """``MultigroupLibrary.save``/``load`` use the reference's field dict (cross_sections.py:624-650): a round trip
    keeps every field, a dict written with the reference's field names loads, a file without ``eta_map`` loads as a
    uniform table and a file missing another field raises KeyError."""
    import torch

    from citrus_amd.scatter_solver.cross_sections import MultigroupLibrary, seeded_library, synthetic_library

    for lib in (seeded_library(4, 3, 120.0, 20.0, seed=5, num_eta=33), synthetic_library(3, 2, num_eta=17,
                                                                                           angular_table_grid="adaptive")):
        path = tmp_path / "lib.pt"
        lib.save(str(path))
        back = MultigroupLibrary.load(str(path))
        for k in MultigroupLibrary.FILE_FIELDS:
            a, b = getattr(lib, k), getattr(back, k)
            if isinstance(a, torch.Tensor):
                assert torch.equal(a, b), k
            else:
                assert a == b, k
    fields = torch.load(str(path), weights_only=True)
    assert set(fields) == set(MultigroupLibrary.FILE_FIELDS)
    fields.pop("eta_map")
    torch.save(fields, str(tmp_path / "no_map.pt"))
    assert MultigroupLibrary.load(str(tmp_path / "no_map.pt")).eta_map is None
    fields.pop("nominal_density")
    torch.save(fields, str(tmp_path / "old.pt"))
    import pytest

    with pytest.raises(KeyError):
        MultigroupLibrary.load(str(tmp_path / "old.pt"))


def test_density_model_host_tables():
    """Effective energy, log-log mass attenuation (tabulated energies exact, interpolation between), material order."""
    import math

    import torch

    from citrus_amd.density import effective_energy, mass_attenuation_at, material_order, nominal_attenuation
    from citrus_amd.scatter_solver.cross_sections import seeded_library

    lib = seeded_library(4, 3, 120.0, 20.0, seed=5, num_eta=33)
    w, E = lib.fine_weight.double(), lib.fine_energy.double()
    assert abs(effective_energy(lib) - float((w * E).sum() / w.sum())) < 1e-12
    order = torch.argsort(E)
    e0, e1 = float(E[order[3]]), float(E[order[4]])
    assert torch.equal(mass_attenuation_at(lib, e0), lib.fine_mu_t[:, order[3]].double())
    mid = math.sqrt(e0 * e1)  # halfway in log E: the geometric mean of the coefficients
    want = torch.sqrt(lib.fine_mu_t[:, order[3]].double() * lib.fine_mu_t[:, order[4]].double())
    assert torch.allclose(mass_attenuation_at(lib, mid), want, rtol=1e-12)
    assert material_order(lib) == torch.argsort(nominal_attenuation(lib)).tolist() == [0, 1]
    import pytest

    with pytest.raises(ValueError):
        mass_attenuation_at(lib, 1000.0)


def test_shipped_surrogate_weights_match_the_layer_shapes(tmp_path):
    """The shipped 64^3 weights load on the CPU with every array of the shape and dtype the convolution kernels
    index; a truncated array, a float64 array or a missing entry raises ValueError before anything reaches a GPU."""
    from pathlib import Path

    from citrus_amd.surrogate import LAYERS, UNet3D, layer_shapes

    path = Path(__file__).resolve().parents[1] / "citrus_amd" / "weights" / "surrogate_64cube_s12_p5_g16.npz"
    (cin, cout, width), arrays = UNet3D.read_arrays(str(path))
    shapes = layer_shapes(cin, cout, width)
    assert (cin, cout, width) == (2 + 4 * 16, 4 * 16, 16)  # two densities + P1 moments of 16 groups -> P1 of 16
    assert set(arrays) == {f"{n}.{k}" for n in LAYERS for k in "wb"}
    for name, (co, ci, k) in shapes.items():
        assert arrays[f"{name}.w"].shape == (co, ci, k ** 3) and arrays[f"{name}.b"].shape == (co,)
    assert sum(a.size for a in arrays.values()) > 100_000
    base = {"shape": np.array([cin, cout, width], dtype=np.int64), **arrays}
    for key, value in (("e0a.w", arrays["e0a.w"][:, :, :26]), ("head.b", arrays["head.b"].astype(np.float64))):
        np.savez(tmp_path / "bad.npz", **{**base, key: value})
        with pytest.raises(ValueError):
            UNet3D.read_arrays(str(tmp_path / "bad.npz"))
    np.savez(tmp_path / "bad.npz", **{k: v for k, v in base.items() if k != "b1.w"})
    with pytest.raises(ValueError, match="missing"):
        UNet3D.read_arrays(str(tmp_path / "bad.npz"))


def test_csi_and_lead_tables_and_csi_response():
    """The embedded CsI and lead tables (detector.py:58-93) interpolate linearly in (log E, log mu): exact at the
    nodes, the geometric mean at the geometric-mean energy of two nodes, the end segments' power laws beyond the
    table; csi_response is E (1 - exp(-mu rho t / cos theta)) on its nodes (detector.py:258-291)."""
    from citrus_amd.scatter_solver.detector import (CSI_TABLE, LEAD_TABLE, RHO_CSI, csi_mass_attenuation,
                                                    csi_response, lead_mass_attenuation)

    for table, f in ((CSI_TABLE, csi_mass_attenuation), (LEAD_TABLE, lead_mass_attenuation)):
        assert np.allclose(f(table[:, 0]), table[:, 1], rtol=1e-13)
        e_mid = np.sqrt(table[:-1, 0] * table[1:, 0])
        assert np.allclose(f(e_mid), np.sqrt(table[:-1, 1] * table[1:, 1]), rtol=1e-12)
        lo = table[0, 1] * (0.5 ** (math.log(table[1, 1] / table[0, 1]) / math.log(table[1, 0] / table[0, 0])))
        assert math.isclose(float(f([table[0, 0] * 0.5])[0]), lo, rel_tol=1e-12)
        hi = table[-1, 1] * (2.0 ** (math.log(table[-1, 1] / table[-2, 1]) / math.log(table[-1, 0] / table[-2, 0])))
        assert math.isclose(float(f([table[-1, 0] * 2.0])[0]), hi, rel_tol=1e-12)
    assert float(csi_mass_attenuation([33.17 + 1e-3])[0]) == 21.6  # above the iodine K edge
    with pytest.raises(ValueError):
        csi_mass_attenuation([0.0])
    r = csi_response()
    assert tuple(r.response.shape) == (281, 181) and not r.angle_independent
    e, th = r.energy_kev.numpy(), r.theta_rad.numpy()
    want = e[:, None] * (1 - np.exp(-csi_mass_attenuation(e)[:, None] * RHO_CSI * 0.06 / np.cos(th)[None, :]))
    assert np.allclose(r.response.numpy(), want, rtol=1e-12)
    g = torch.tensor([30.0, 35.98 + 1e-3, 70.0], dtype=torch.float64)
    d0 = r.at(g, torch.zeros(1, dtype=torch.float64))[:, 0]
    assert float(d0[1]) > float(d0[0])  # the Cs K edge raises the absorbed share


def test_dataset_round_trip_and_the_reference_layout(tmp_path):
    """save_dataset / load_dataset (utilities.py:1480-1527): arrays and meta round-trip; a file in the reference's
    layout (np.savez with the meta as the JSON string meta_json) reads; format < 3 and a missing meta are refused."""
    import json

    from citrus_amd.utilities import load_dataset, save_dataset

    data = {"sinogram": np.arange(24, dtype=np.float32).reshape(2, 3, 4), "angles": np.array([0.0, 0.5]),
            "meta": {"description": "test", "data_format": 3, "num_views": 2}}
    files = save_dataset(data, tmp_path / "sub" / "scan.npz")
    assert files == [str(tmp_path / "sub" / "scan.npz"), str(tmp_path / "sub" / "scan.json")]
    back = load_dataset(tmp_path / "sub" / "scan")
    assert np.array_equal(back["sinogram"], data["sinogram"]) and back["sinogram"].dtype == np.float32
    assert back["meta"]["num_views"] == 2 and json.loads((tmp_path / "sub" / "scan.json").read_text())["num_views"] == 2
    # written the way the reference writes it
    meta = {"data_format": 3, "spr": 0.2}
    np.savez(tmp_path / "ref.npz", meta_json=np.asarray(json.dumps(meta)), counts=np.ones(5))
    ref = load_dataset(tmp_path / "ref.npz")
    assert ref["meta"]["spr"] == 0.2 and ref["counts"].sum() == 5
    np.savez(tmp_path / "old.npz", meta_json=np.asarray(json.dumps({"data_format": 2})), counts=np.ones(5))
    with pytest.raises(ValueError, match="format"):
        load_dataset(tmp_path / "old.npz")
    with pytest.raises(ValueError):
        save_dataset({"x": np.ones(2)}, tmp_path / "bad")


def test_tuner_prefers_the_persistent_schedule_within_its_margin():
    from citrus_amd.scatter_solver.autotune import PERSISTENT_MARGIN, pick_config

    assert pick_config({("launches", 21): 118.0, ("persistent", 21): 119.3, ("persistent", 11): 214.6}) == \
        ("persistent", 21)
    far = 100.0 * PERSISTENT_MARGIN + 1.0
    assert pick_config({("launches", 11): 100.0, ("persistent", 21): far}) == ("launches", 11)
    assert pick_config({("launches", 7): 50.0, ("launches", 21): 40.0}) == ("launches", 21)
    with pytest.raises(ValueError):
        pick_config({})
