"""The density model on the GPU (utilities.py:710-753, citrus.py:466-560): the HIP kernel equals the float32 oracle
bit for bit, reproduces the attenuation, and ``material_fields_from_attenuation`` equals ``material_fields`` of the
model's densities."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from citrus_amd.density import DensityModel, _slot_tables, attenuation_to_densities, densities_to_attenuation
from citrus_amd.scatter_solver.cross_sections import seeded_library
from oracle import density as od
from tests.conftest import gpu_device

pytestmark = pytest.mark.gpu


def test_densities_equal_the_oracle_bit_for_bit():
    dev = gpu_device()
    lib = seeded_library(4, 3, 120.0, 20.0, seed=5, num_eta=33)
    model = DensityModel.from_library(lib)
    order, mu, nominal, mass = _slot_tables(lib, model.energy_kev)
    gen = torch.Generator().manual_seed(9)
    X = (torch.rand(7, 9, 11, generator=gen) * 0.8 - 0.1)
    X.view(-1)[:2] = torch.from_numpy(mu)  # exactly at the nominal attenuations
    got = model(X.to(dev))
    want = od.attenuation_to_densities(X.numpy(), order, mu, nominal, mass)
    assert np.array_equal(got.cpu().numpy(), want)
    assert abs(model.last_negative_fraction - float((X < 0).float().mean())) < 1e-12
    back = densities_to_attenuation(lib, got, model.energy_kev)
    assert torch.allclose(back, X.clamp_min(0).to(dev), rtol=3e-6, atol=1e-7)
    assert torch.equal(model.material(X.to(dev), 1), got[1])
    with pytest.raises(ValueError):
        attenuation_to_densities(lib, X)  # a CPU tensor: the product path has no CPU form


def test_fields_from_attenuation_equal_fields_of_the_densities():
    from citrus_amd.scatter_solver.config import SolverConfig
    from citrus_amd.scatter_solver.geometry import ConeGeometry, Grid
    from citrus_amd.scatter_solver.scatter_solver import ScatterSolver
    from citrus_amd.utilities import ellipsoid_phantom

    dev = gpu_device()
    lib = seeded_library(4, 3, 120.0, 20.0, seed=5, num_eta=33)
    config = SolverConfig(legendre_order=3, sweep_quadrature="level_symmetric", sn_order=4, transport_cell_cm=1.0,
                          quadrature_point_spacing_cm=0.5)
    solver = ScatterSolver(lib, config, ConeGeometry(30.0, 50.0, 16, 16, 0.5, 0.5), device=dev)
    grid = Grid((8, 8, 8), 0.5)
    model = DensityModel.from_library(lib)
    X = densities_to_attenuation(lib, ellipsoid_phantom(grid, seed=3).to(dev, torch.float32), model.energy_kev)
    fields = solver.material_fields_from_attenuation(X, grid, model)
    ref = solver.material_fields(model(X), grid)
    for name in ("rho_fine", "rho_coarse", "rho_quadrature", "mu_t_coarse"):
        assert torch.equal(getattr(fields, name), getattr(ref, name)), name
    assert fields.q == ref.q == 2
