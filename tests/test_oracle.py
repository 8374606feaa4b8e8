"""Pins of the FP64 oracle against closed forms, brute force and invariants (CPU only).

Each oracle function is checked by something other than its own formula: scipy or closed-form spherical harmonics,
published quadrature moment conditions, numpy.linalg.solve of the dense cell and global systems, the textbook 1-D
slab recursion, particle balance, Beer's-law refinement and reflection symmetry (the reference's own test ideas:
tests/test_linear_discontinuous_cell.py:6, tests/test_transport_sweep.py:91-325, tests/test_angular_quadrature.py)."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import checks, ld, quadrature as oq, scattering, sweep as osw


def _octant_set(base=(0.3, 0.5, 0.81), replace=None):
    b = np.asarray(base, dtype=np.float64)
    b = b / np.linalg.norm(b)
    om = np.array([b * np.array(osw.octant_signs(o)) for o in range(8)])
    for o, v in (replace or {}).items():
        om[o] = v
    return om, np.full(8, 4 * math.pi / 8)


# ---- spherical harmonics and quadratures --------------------------------------------------------------------------

def test_sh_closed_form_degree_one():
    rng = np.random.default_rng(0)
    om = rng.standard_normal((50, 3))
    om /= np.linalg.norm(om, axis=1, keepdims=True)
    Y = oq.real_spherical_harmonics(1, om)
    assert np.allclose(Y[:, 0], 1 / math.sqrt(4 * math.pi))
    assert np.allclose(Y[:, 1:4], math.sqrt(3 / (4 * math.pi)) * om[:, [1, 2, 0]])  # (y, z, x) in index order


def test_sh_addition_theorem():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((20, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    b = rng.standard_normal((20, 3))
    b /= np.linalg.norm(b, axis=1, keepdims=True)
    L = 6
    Ya, Yb = oq.real_spherical_harmonics(L, a), oq.real_spherical_harmonics(L, b)
    P = oq.legendre(L, np.sum(a * b, axis=1))
    start = 0
    for l in range(L + 1):
        s = np.sum(Ya[:, start:start + 2 * l + 1] * Yb[:, start:start + 2 * l + 1], axis=1)
        assert np.allclose(s, (2 * l + 1) / (4 * math.pi) * P[:, l], atol=1e-12)
        start += 2 * l + 1


def test_product_set_orthonormal():
    om, w = oq.product_quadrature(8, 16)
    Y = oq.real_spherical_harmonics(7, om)
    assert np.abs((Y * w[:, None]).T @ Y - np.eye(64)).max() < 1e-12
    assert abs(w.sum() - 4 * math.pi) < 1e-12


@pytest.mark.parametrize("order", [2, 4, 6, 8, 12])
def test_level_symmetric_moments(order):
    om, w = oq.level_symmetric_quadrature(order)
    assert om.shape == (order * (order + 2), 3)
    assert abs(w.sum() - 4 * math.pi) < 1e-12
    assert np.allclose(np.linalg.norm(om, axis=1), 1.0)
    counts = np.bincount(oq.octant_of(om), minlength=8)
    assert (counts == order * (order + 2) // 8).all()
    for p in range(0, order + 1, 2):  # exact even moments of every axis to the table's 7 digits
        for a in range(3):
            assert abs((w * om[:, a] ** p).sum() / (4 * math.pi / (p + 1)) - 1) < 5e-7
    for p in range(1, order, 2):  # odd moments vanish by symmetry
        assert abs((w * om[:, 0] ** p).sum()) < 1e-12


def test_level_symmetric_s12_integrates_p5_basis():
    om, w = oq.level_symmetric_quadrature(12)
    Y = oq.real_spherical_harmonics(5, om)
    assert np.abs((Y * w[:, None]).T @ Y - np.eye(36)).max() < 1e-6


# ---- LD cell -------------------------------------------------------------------------------------------------------

def test_cell_closed_form_matches_dense_solve():
    rng = np.random.default_rng(0)
    c = rng.random((7, 3)) + 0.01
    s = np.where(rng.standard_normal((7, 3)) < 0, -1.0, 1.0)
    mu = rng.random(7) * 0.5 + 0.01
    src = rng.standard_normal((7, 4))
    up = rng.standard_normal((7, 3, 4))
    r0, ra = ld.right_hand_sides(c, s, src, up)
    x = ld.cell_solve(c, s, mu, r0, ra)
    for i in range(7):
        # right-hand sides rebuilt by hand from the upwind face integrals
        pb = [up[i, a, 0] + s[i, a] * up[i, a, 1 + a] for a in range(3)]
        R0 = src[i, 0] + sum(c[i, a] * pb[a] for a in range(3))
        Ra = [src[i, 1 + a] - 3 * s[i, a] * c[i, a] * pb[a]
              + sum(c[i, b] * up[i, b, 1 + a] for b in range(3) if b != a) for a in range(3)]
        ref = np.linalg.solve(ld.cell_matrix(c[i], s[i], mu[i]), np.array([R0] + Ra))
        assert np.allclose(x[i], ref, atol=1e-13)


# ---- sweep ---------------------------------------------------------------------------------------------------------

@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1, -1])
def test_sweep_slab_1d(axis, sign):
    rng = np.random.default_rng(10 + axis)
    K, G = 9, 2
    deltas = (0.7, 1.3, 0.9)
    shape = [3, 3, 3]
    vol_dim = {0: 1, 1: 0, 2: 2}[axis]
    shape[vol_dim] = K
    d = [0.0, 0.0, 0.0]
    d[axis] = float(sign)
    o = int(oq.octant_of(np.array(d)))
    om, _ = _octant_set(replace={o: d})
    mu1 = rng.random((G, K)) * 1.5 + 0.05
    S0 = rng.random((G, K))
    S1 = rng.standard_normal((G, K)) * 0.3

    def along(v):
        view = [G, 1, 1, 1]
        view[1 + vol_dim] = K
        return np.broadcast_to(v.reshape(view), (G, *shape)).copy()

    src = np.zeros((8, G, 4, *shape))
    src[:, :, 0] = along(S0)
    src[:, :, 1 + axis] = along(S1)
    psi = osw.sweep(tuple(shape), deltas, om, along(mu1), src)
    for g in range(G):
        p0, p1 = checks.slab_1d(mu1[g], S0[g], S1[g], deltas[axis], float(sign))
        assert np.allclose(psi[o, g, 0], along(np.broadcast_to(p0, (G, K)))[g], atol=1e-12)
        assert np.allclose(psi[o, g, 1 + axis], along(np.broadcast_to(p1, (G, K)))[g], atol=1e-12)


def test_sweep_balance_and_dense_assembly():
    rng = np.random.default_rng(3)
    shape, deltas, G = (3, 4, 3), (0.9, 1.2, 0.75), 2
    om, _ = _octant_set(base=(0.55, 0.38, 0.74))
    mu = rng.random((G, *shape)) + 0.1
    src = rng.standard_normal((8, G, 4, *shape))
    psi = osw.sweep(shape, deltas, om, mu, src)
    for m in range(8):
        lhs, rhs = checks.outflow_balance(psi[m], om[m], deltas, mu, src[m])
        assert np.allclose(lhs, rhs, rtol=1e-12, atol=1e-12)
        for g in range(G):
            assert np.allclose(psi[m, g], checks.dense_global_solve(om[m], deltas, mu[g], src[m, g]), atol=1e-12)


def test_sweep_beer_refinement():
    """Pure absorber, source cube: slab-integrated fluence vs the analytic exponential, error /3 per halving."""
    H, a, mu_val = 2.0, 0.5, 0.5
    om, _ = _octant_set()
    errors = []
    for n in (8, 16, 32):
        h = 2 * H / n
        kk = int(round(a / h))
        c0 = n // 2 - kk // 2
        src = np.zeros((8, 1, 4, n, n, n))
        src[:, 0, 0, c0:c0 + kk, c0:c0 + kk, c0:c0 + kk] = 1.0
        psi = osw.sweep((n, n, n), (h, h, h), om, np.full((1, n, n, n), mu_val), src)
        z = -H + h * np.arange(n + 1)
        err = 0.0
        for m in range(8):
            wz = abs(om[m, 2])
            for z0, z1 in ((0.5, 1.0), (1.0, 1.5), (1.5, 2.0)):
                lo, hi = (z0, z1) if om[m, 2] > 0 else (-z1, -z0)
                sel = (z[:-1] >= lo - 1e-12) & (z[1:] <= hi + 1e-12)
                num = psi[m, 0, 0][:, :, sel].sum() * h ** 3

                def edge(d):
                    return wz / mu_val * math.exp(-mu_val * d / wz) * 2 * math.sinh(mu_val * a / (2 * wz))

                exact = a * a * (edge(z0) - edge(z1)) / mu_val
                err = max(err, abs(num - exact) / exact)
        errors.append(err)
    assert errors[0] / errors[1] >= 3.0 and errors[1] / errors[2] >= 3.0, errors


def test_sweep_reflection_symmetry():
    n = 5
    om, w = oq.product_quadrature(8, 16)
    Y = oq.real_spherical_harmonics(3, om)
    src1 = np.zeros((1, 4, n, n, n))
    src1[0, 0, 1:4, 1:4, 1:4] = 0.5
    src1[0, 0, 2, 2, 2] = 2.0
    src = np.broadcast_to(src1[None], (om.shape[0], 1, 4, n, n, n))
    psi = osw.sweep((n, n, n), (1.0, 1.0, 1.0), om, np.full((1, n, n, n), 0.4), src)
    phi = osw.moments(psi, w, Y)
    p0 = phi[0, 0]
    tol = 1e-12 * abs(p0[0]).max()
    for a, dim in ((0, 2), (1, 1), (2, 3)):
        f = np.flip(p0, axis=dim)
        assert np.allclose(f[0], p0[0], atol=tol)
        for b in range(3):
            assert np.allclose(f[1 + b], (-1.0 if b == a else 1.0) * p0[1 + b], atol=tol)


# ---- scattering moments and source iteration ----------------------------------------------------------------------

def test_scatter_moments_brute_force():
    rng = np.random.default_rng(4)
    G, L, sp = 3, 2, (2, 3, 2)
    mu_s = rng.random((G, G, L + 1) + sp)
    phi = rng.standard_normal((G, (L + 1) ** 2, 4) + sp)
    T = scattering.scatter_moments(mu_s, phi)
    for go in range(G):
        for n in range((L + 1) ** 2):
            l = int(math.isqrt(n))
            ref = sum(mu_s[gi, go, l][None] * phi[gi, n] for gi in range(G))
            assert np.allclose(T[go, n], ref, atol=1e-14)


def test_source_iteration_fixed_point():
    rng = np.random.default_rng(5)
    n, G, L = 4, 3, 2
    shape, deltas = (n, n, n), (1.5, 1.5, 1.5)
    om, w = oq.product_quadrature(4, 8)
    Y = oq.real_spherical_harmonics(L, om)
    mu_t = rng.random((G, *shape)) * 0.3 + 0.3
    mu_s = np.zeros((G, G, L + 1) + shape)
    for gi in range(G):
        for go in range(gi, G):
            mu_s[gi, go, 0] = 0.08 * 0.6 ** (go - gi)
            mu_s[gi, go, 1] = 0.02 * 0.6 ** (go - gi)
            mu_s[gi, go, 2] = 0.005
    fc = np.zeros((om.shape[0], G, 4) + shape)
    fc[:, :, 0, :2, :2, :2] = 1.0
    phi, hist, nit, conv = scattering.source_iteration(shape, deltas, om, w, Y, mu_t, mu_s, fc, 1e-8, 60)
    assert conv and hist[0] == 1.0
    T = scattering.scatter_moments(mu_s, phi)
    fp = osw.moments(osw.sweep(shape, deltas, om, mu_t, osw.sources_from_moments(T, Y)), w, Y) \
        + osw.moments(osw.sweep(shape, deltas, om, mu_t, fc), w, Y)
    assert np.linalg.norm(fp - phi) / np.linalg.norm(phi) < 1e-7
