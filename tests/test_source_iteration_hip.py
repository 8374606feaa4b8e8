"""Scattering-source moments and source iteration on the HIP path vs the FP64 oracle (GPU).

Mirrors tests/test_source_iteration.py of the reference: convergence and fixed point (:59-86), iteration cap (:88),
float32 vs float64 (:102), and the scattering source function (:114)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from citrus_amd.scatter_solver.angular_quadrature import product_quadrature
from citrus_amd.scatter_solver.cross_sections import mu_t_volume, synthetic_library, transfer_moments_volume
from citrus_amd.scatter_solver.geometry import Grid
from citrus_amd.scatter_solver.source_iteration import (MaterialTransfer, multiple_collision_source_moments,
                                                        scatter_moments, scatter_source_function, source_iteration)
from citrus_amd.scatter_solver.transport_sweep import make_sweep_context, sweep
from oracle import quadrature as oq, scattering as osc, sweep as osw
from tests.conftest import gpu_device
from tests.helpers import F32_REL, rel_err

pytestmark = pytest.mark.gpu


def _problem(device, n=5, h=2.0, L=3):
    """Small water/bone cube with the closed-form library; first-collision source in one corner, anisotropic."""
    library = synthetic_library(num_groups=4, legendre_order=5)
    grid = Grid((n, n, n), h)
    rho = torch.zeros(2, n, n, n, dtype=torch.float64)
    rho[0] = 1.0
    rho[1, 1:3, 2:4, 1:4] = 1.2
    rho[0, 1:3, 2:4, 1:4] = 0.0
    q = product_quadrature(4, 8)
    ctx = make_sweep_context(grid, q, L, device=device)
    G = library.num_groups
    iso = torch.zeros(G, 4, n, n, n, dtype=torch.float64)
    iso[:, 0, :2, :2, :2] = torch.linspace(1.0, 0.4, G, dtype=torch.float64)[:, None, None, None]
    iso[:, 1, :2, :2, :2] = 0.1
    omega = ctx.quadrature.omega.double().cpu()
    fc_all = (1.0 + 0.3 * omega[:, 0])[:, None, None, None, None, None] * iso[None]  # [M, G, 4, n, n, n]
    fc_dev = fc_all.to(device, torch.float32)

    def first_collision(index):
        return fc_dev.index_select(0, index)

    return library, grid, rho, q, ctx, fc_all, first_collision


def _oracle_inputs(ctx, q, library, rho):
    perm = ctx.permutation.cpu()
    om, w = q.omega[perm].numpy(), q.weight[perm].numpy()
    Y = oq.real_spherical_harmonics(ctx.L, om)
    mu_t = mu_t_volume(library, rho).numpy()
    mu_s = transfer_moments_volume(library, rho).numpy()
    return om, w, Y, mu_t, mu_s


@pytest.mark.parametrize("factored", [False, True])
def test_scatter_moments_match_oracle(factored):
    dev = gpu_device()
    library = synthetic_library(num_groups=5, legendre_order=3)
    gen = torch.Generator().manual_seed(9)
    rho = torch.rand(2, 4, 3, 5, generator=gen, dtype=torch.float64)
    phi = torch.randn(5, 16, 4, 4, 3, 5, generator=gen, dtype=torch.float64)
    ref = osc.scatter_moments(transfer_moments_volume(library, rho).numpy(), phi.numpy())
    if factored:
        transfer = MaterialTransfer(rho.to(dev, torch.float32), library.mu_s_l.to(dev, torch.float32))
        assert transfer.downscatter_only
    else:
        transfer = transfer_moments_volume(library, rho).to(dev, torch.float32)
    T = scatter_moments(transfer, phi.to(dev, torch.float32))
    assert T.shape == phi.shape
    assert rel_err(T, ref) < F32_REL


def test_scatter_source_function_matches_angular():
    dev = gpu_device()
    library, grid, rho, q, ctx, _, _ = _problem(dev, n=3)
    gen = torch.Generator().manual_seed(2)
    phi = torch.randn(4, ctx.num_moments, 4, *ctx.shape, generator=gen, dtype=torch.float64)
    T = scatter_moments(transfer_moments_volume(library, rho).to(dev, torch.float32), phi.to(dev, torch.float32))
    index = torch.tensor([0, 5, 17, 31, 30], device=dev)
    ref = torch.tensordot(ctx.basis.Y.double().cpu()[index.cpu()], T.double().cpu(), dims=([1], [1]))
    ref = ref.permute(0, 3, 4, 5, 1, 2)  # cell-major [D, R, C, S, G, 4]
    source = scatter_source_function(ctx, T)
    out = source(index)
    assert out.shape == (5,) + ctx.shape + (4, 4)
    assert rel_err(out, ref) < F32_REL
    buf = torch.zeros_like(out)
    assert source(index, out=buf) is buf and rel_err(buf, ref) < F32_REL
    source(index, out=buf, accumulate=True)
    assert rel_err(buf, 2 * ref) < F32_REL


def test_source_iteration_matches_oracle_and_is_fixed_point():
    dev = gpu_device()
    library, grid, rho, q, ctx, fc_all, first_collision = _problem(dev)
    mu_t = mu_t_volume(library, rho).to(dev, torch.float32)
    mu_s = transfer_moments_volume(library, rho).to(dev, torch.float32)
    tol = 1e-6
    seen = []
    result = source_iteration(ctx, mu_t, mu_s, first_collision, tolerance=tol, max_iterations=60,
                              callback=lambda n, phi_n, rel: seen.append((n, rel)))
    assert result.converged and result.num_iterations < 30
    assert len(result.history) == result.num_iterations == len(result.phi_norms) == len(seen)
    assert result.history[0] == 1.0 and result.history[-1] < tol
    om, w, Y, mu_t64, mu_s64 = _oracle_inputs(ctx, q, library, rho)
    phi_ref, hist_ref, n_ref, conv_ref = osc.source_iteration(ctx.shape, grid.deltas, om, w, Y, mu_t64, mu_s64,
                                                              fc_all.numpy(), tol, 60)
    assert conv_ref and abs(n_ref - result.num_iterations) <= 1
    err = rel_err(result.phi_total, phi_ref)
    print(f"source iteration: {result.num_iterations} sweeps, rel err vs FP64 oracle {err:.2e}")
    assert err < 1e-5
    T = multiple_collision_source_moments(mu_s, result.phi_total)
    fp = sweep(ctx, mu_t, scatter_source_function(ctx, T)) + sweep(ctx, mu_t, first_collision)
    assert float((fp - result.phi_total).norm() / result.phi_total.norm()) < 10 * tol
    assert float(result.phi_total[:, 0, 0].min()) > 0


def test_source_iteration_cap_and_thick_cells():
    dev = gpu_device()
    library, grid, rho, q, ctx, fc_all, first_collision = _problem(dev, h=4.0)
    mu_t = mu_t_volume(library, rho).to(dev, torch.float32)
    mu_s = MaterialTransfer(rho.to(dev, torch.float32), library.mu_s_l.to(dev, torch.float32))
    result = source_iteration(ctx, mu_t, mu_s, first_collision, tolerance=1e-12, max_iterations=3)
    assert result.num_iterations == 3 and not result.converged and len(result.history) == 3
    _, psi = sweep(ctx, mu_t, first_collision, return_psi=True)
    assert float(psi[:, :, 0].min()) < 0  # thick cells: negative LD cell means
    om, w, Y, mu_t64, mu_s64 = _oracle_inputs(ctx, q, library, rho)
    phi_ref, _, _, _ = osc.source_iteration(ctx.shape, grid.deltas, om, w, Y, mu_t64, mu_s64, fc_all.numpy(),
                                            1e-12, 3)
    assert rel_err(result.phi_total, phi_ref) < 1e-5


@pytest.mark.parametrize("G,shape", [(5, (3, 4, 6)), (8, (3, 4, 6)), (16, (5, 5, 5)), (16, (4, 4, 4))])
def test_scatter_moments_also_write_moment_major_cell_means(G, shape):
    """``scatter_moments_views(cell_means=True)``: the second output is the first's cell means, moment-major, bit for
    bit (the same accumulator written twice), and T is the float64 contraction over the input groups to float32
    rounding. G = 8 and 16 with downscatter only take the view step's form (a run of 4 groups of 4 cells per
    workgroup, the means staged in LDS): 72 and 125 cells leave a partial last run, 64 cells none."""
    from citrus_amd.scatter_solver.source_iteration import MaterialTransfer, scatter_moments_views

    dev = gpu_device()
    gen = torch.Generator().manual_seed(12)
    V, nsh, M, L = 2, 16, 2, 3
    rho = torch.rand(M, *shape, generator=gen).to(dev)
    sigma = (torch.rand(M, G, G, L + 1, generator=gen) * torch.triu(torch.ones(G, G))[None, :, :, None]).to(dev)
    phi = torch.randn(V, G, *shape, nsh, 4, generator=gen).to(dev)
    transfer = MaterialTransfer(rho, sigma)
    T = scatter_moments_views(transfer, phi)
    T2, means = scatter_moments_views(transfer, phi, cell_means=True)
    assert torch.equal(T, T2)
    assert torch.equal(means, T[..., 0].permute(0, 5, 1, 2, 3, 4))  # [V, G, R, C, S, nsh] -> [V, nsh, G, R, C, S]
    degree = torch.tensor([int(n ** 0.5) for n in range(nsh)])
    mu = torch.einsum("mrcs,mgho->rcsgho", rho.double().cpu(), sigma.double().cpu())  # [R, C, S, gi, go, l]
    ref = torch.einsum("rcsghn,vgrcsnk->vhrcsnk", mu[..., degree], phi.double().cpu())
    assert rel_err(T, ref) < F32_REL
