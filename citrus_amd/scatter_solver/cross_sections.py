"""Multigroup cross-section tables and per-voxel coefficients (host-side setup).

Keeps the layout of ``MultigroupLibrary`` (``citrus/scatter_solver/cross_sections.py:493``) for what the transport
path reads: ``mu_t [M, G]`` mass attenuation, ``mu_s_l [M, G_in, G_out, L+1]`` Legendre transfer moments with
``mu_s_l = 2 pi int mu_s(eta) P_l(eta) d eta`` (no (2l+1)/4pi factor), and the angular table ``mu_s_eta`` on
``eta_grid``. Library building from photon data is out of scope (it needs a network download); two offline
libraries are provided: ``synthetic_library`` (the reference's closed-form test library, cross_sections.py:1043,
kernel ``a (1 + b eta) / 4 pi``) and ``seeded_library`` (a seeded Henyey-Greenstein library of Compton-dominated
totals for the benchmark configurations).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import torch
from torch import Tensor


@dataclass(frozen=True)
class MultigroupLibrary:
    """Multigroup tables of M materials (float64, CPU)."""

    group_edges: Tensor        # [G+1] keV, descending
    mu_t: Tensor               # [M, G] cm^2/g
    mu_s_l: Tensor             # [M, G, G, L+1] cm^2/g
    eta_grid: Tensor           # [E]
    mu_s_eta: Tensor           # [M, G, G, E] cm^2/g/sr
    materials: tuple[str, ...]

    @property
    def num_groups(self) -> int:
        return int(self.mu_t.shape[1])

    @property
    def legendre_order(self) -> int:
        return int(self.mu_s_l.shape[3]) - 1

    @property
    def num_materials(self) -> int:
        return int(self.mu_t.shape[0])


def synthetic_library(num_groups: int = 4, legendre_order: int = 5, num_eta: int = 401) -> MultigroupLibrary:
    """The reference's two-material closed-form test library (cross_sections.py:1043-1100): kernel
    ``a (1 + b eta) / 4 pi`` with ``a = a_m 0.6^(go-gi)`` (a_m = 0.15, 0.12), ``b = 0.5`` in-group else 0.2, photo-
    electric ``0.05 / 0.2 * 1.5^g``; its Legendre moments are exactly ``a`` and ``a b / 3``."""
    G, L = num_groups, legendre_order
    edges = torch.linspace(125.0, 10.0, G + 1, dtype=torch.float64)
    eta = torch.linspace(-1.0, 1.0, num_eta, dtype=torch.float64)
    mu_s_eta = torch.zeros(2, G, G, num_eta, dtype=torch.float64)
    mu_s_l = torch.zeros(2, G, G, L + 1, dtype=torch.float64)
    for m, am in enumerate((0.15, 0.12)):
        for gi in range(G):
            for go in range(gi, G):
                a = am * 0.6 ** (go - gi)
                b = 0.5 if go == gi else 0.2
                mu_s_eta[m, gi, go] = a * (1.0 + b * eta) / (4.0 * math.pi)
                mu_s_l[m, gi, go, 0] = a
                if L >= 1:
                    mu_s_l[m, gi, go, 1] = a * b / 3.0
    mu_pe = torch.tensor([[0.05 * 1.5 ** g for g in range(G)], [0.2 * 1.5 ** g for g in range(G)]],
                         dtype=torch.float64)
    mu_t = mu_pe + mu_s_l[..., 0].sum(dim=2)
    return MultigroupLibrary(edges, mu_t, mu_s_l, eta, mu_s_eta, ("water", "bone_cortical_icrp"))


def seeded_library(num_groups: int, legendre_order: int, e_max_kev: float, e_min_kev: float, seed: int,
                   num_eta: int = 257) -> MultigroupLibrary:
    """Seeded synthetic two-material library (water-like, bone-like) of Compton-dominated totals.

    Water's total falls from about 0.6 to 0.15 cm^2/g across the groups (highest energy first); bone's is 1.25 to
    1.6 times higher. Scattering goes down in energy only, with Henyey-Greenstein kernels whose asymmetry grows
    with energy, so every Legendre moment is nonzero: ``mu_s_l = mu_s g^l`` (the HG expansion).
    """
    G, L = num_groups, legendre_order
    gen = torch.Generator().manual_seed(seed)
    edges = torch.linspace(e_max_kev, e_min_kev, G + 1, dtype=torch.float64)
    energy = 0.5 * (edges[:-1] + edges[1:])
    x = (energy - e_min_kev) / (e_max_kev - e_min_kev)  # 1 at the highest energy
    water_total = 0.6 - 0.45 * x + 0.01 * torch.rand(G, generator=gen, dtype=torch.float64)
    bone_factor = 1.25 + 0.35 * (1 - x) + 0.02 * torch.rand(G, generator=gen, dtype=torch.float64)
    totals = torch.stack([water_total, water_total * bone_factor])  # [2, G]
    scatter_fraction = torch.stack([0.9 - 0.2 * (1 - x), 0.75 - 0.35 * (1 - x)]).clamp(0.2, 0.95)
    asym = 0.15 + 0.35 * x  # [G] forward peaking grows with energy
    eta = torch.linspace(-1.0, 1.0, num_eta, dtype=torch.float64)
    mu_s_l = torch.zeros(2, G, G, L + 1, dtype=torch.float64)
    mu_s_eta = torch.zeros(2, G, G, num_eta, dtype=torch.float64)
    for m in range(2):
        for gi in range(G):
            share = torch.tensor([0.55 * 0.45 ** (go - gi) for go in range(gi, G)], dtype=torch.float64)
            share = share / share.sum()
            mu_s = totals[m, gi] * scatter_fraction[m, gi]
            for j, go in enumerate(range(gi, G)):
                g_hg = float(asym[gi]) * (1.0 if go == gi else 0.7)
                s = float(mu_s * share[j])
                mu_s_l[m, gi, go] = s * g_hg ** torch.arange(L + 1, dtype=torch.float64)
                mu_s_eta[m, gi, go] = s / (4 * math.pi) * (1 - g_hg ** 2) / (1 + g_hg ** 2 - 2 * g_hg * eta) ** 1.5
    return MultigroupLibrary(edges, totals, mu_s_l, eta, mu_s_eta, ("water", "bone"))


def mu_t_volume(library: MultigroupLibrary, rho: Tensor) -> Tensor:
    """[G, R, C, S] total attenuation ``sum_m mu_t[m, g] rho[m]`` in 1/cm (cross_sections.py:762)."""
    return torch.einsum("mg,mijk->gijk", library.mu_t.to(rho.device, rho.dtype), rho)


def transfer_moments_volume(library: MultigroupLibrary, rho: Tensor) -> Tensor:
    """Dense [G_in, G_out, L+1, R, C, S] transfer moments in 1/cm (cross_sections.py:886)."""
    return torch.einsum("mabl,mijk->ablijk", library.mu_s_l.to(rho.device, rho.dtype), rho)
