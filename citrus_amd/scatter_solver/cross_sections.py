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
from typing import Optional

import torch
from torch import Tensor


@dataclass(frozen=True)
class MultigroupLibrary:
    """Multigroup tables of M materials (float64, CPU), the fields of cross_sections.py:493 the hot path reads."""

    group_edges: Tensor        # [G+1] keV, descending
    mu_t: Tensor               # [M, G] cm^2/g
    mu_s_l: Tensor             # [M, G, G, L+1] cm^2/g
    eta_grid: Tensor           # [N_eta] angular-table nodes in eta = cos(scattering angle)
    mu_s_eta: Tensor           # [M, G, G, N_eta] cm^2/g/sr
    materials: tuple[str, ...]
    spectrum_fraction: Tensor  # [G] source photons per sr in each group (sum 1)
    fine_energy: Tensor        # [E] keV
    fine_mu_t: Tensor          # [M, E] cm^2/g
    fine_weight: Tensor        # [E] spectrum weight of each fine energy
    fine_group: Tensor         # [E] int64 group of each fine energy
    eta_map: Optional[Tensor] = None  # [K] node map of an adaptive angular table, None for a uniform one
    group_energy: Optional[Tensor] = None  # [G] keV, the group-edge midpoints when not given
    mu_pe: Optional[Tensor] = None  # [M, G] cm^2/g photoelectric part of mu_t (zeros when not given)
    nominal_density: Optional[Tensor] = None  # [M] g/cm^3 nominal densities of the materials (the density model's)

    def __post_init__(self) -> None:
        if self.group_energy is None:
            e = self.group_edges.to(torch.float64)
            object.__setattr__(self, "group_energy", 0.5 * (e[:-1] + e[1:]))
        if self.mu_pe is None:
            object.__setattr__(self, "mu_pe", torch.zeros_like(self.mu_t))

    # fields of the reference's library file, in its order (cross_sections.py:519-534)
    FILE_FIELDS = ("group_edges", "group_energy", "spectrum_fraction", "mu_t", "mu_pe", "mu_s_l", "eta_grid",
                   "mu_s_eta", "fine_energy", "fine_mu_t", "fine_weight", "fine_group", "materials", "nominal_density",
                   "eta_map")

    def save(self, path: str) -> None:
        """Write every field with ``torch.save`` as the dict of the reference's ``MultigroupLibrary.save``
        (cross_sections.py:624-631), so either implementation reads the other's files."""
        fields = {k: getattr(self, k) for k in self.FILE_FIELDS}
        fields["materials"] = tuple(self.materials)
        torch.save(fields, path)

    @staticmethod
    def load(path: str) -> "MultigroupLibrary":
        """Read a library file written by ``save`` here or by the reference (cross_sections.py:633-650), with
        ``torch.load(weights_only=True)`` (tensors, strings and None only; nothing in the file is executed). A file
        without ``eta_map`` loads as a uniform table; a file without another field raises ``KeyError``."""
        fields = torch.load(path, map_location="cpu", weights_only=True)
        if not isinstance(fields, dict):
            raise ValueError(f"library file {path!r} does not hold a field dict")
        fields.setdefault("eta_map", None)
        missing = [k for k in MultigroupLibrary.FILE_FIELDS if k not in fields]
        if missing:
            raise KeyError(f"library file {path!r} lacks fields {missing} (written by an older version)")
        return MultigroupLibrary(**{k: fields[k] for k in MultigroupLibrary.FILE_FIELDS})

    @property
    def num_groups(self) -> int:
        return int(self.mu_t.shape[1])

    @property
    def legendre_order(self) -> int:
        return int(self.mu_s_l.shape[3]) - 1

    @property
    def num_materials(self) -> int:
        return int(self.mu_t.shape[0])


def _fine_tables(edges: Tensor, mu_t: Tensor, n: int, spread: float):
    """Equal-weight fine bins per group whose weighted mean attenuation is the group's (cross_sections.py:1092-1099):
    (spectrum fraction [G], fine energy [E], fine mu_t [M, E], fine weight [E], fine group [E])."""
    G = int(mu_t.shape[1])
    if n < 1:
        raise ValueError("num_fine_per_group must be >= 1")
    fraction = torch.ones(G, dtype=torch.float64) / G
    t = (torch.arange(n, dtype=torch.float64) + 0.5) / n
    x = 1.0 - 2.0 * torch.arange(n, dtype=torch.float64) / (n - 1) if n > 1 else torch.zeros(1, dtype=torch.float64)
    fine_energy = torch.cat([edges[g + 1] + t * (edges[g] - edges[g + 1]) for g in range(G)])
    fine_group = torch.arange(G).repeat_interleave(n)
    fine_weight = fraction[fine_group] / n
    fine_mu_t = mu_t[:, fine_group] * (1.0 + float(spread) * x.repeat(G))[None, :]
    return fraction, fine_energy, fine_mu_t, fine_weight, fine_group


def synthetic_angular_map(num_knots: int = 512) -> Tensor:
    """[K] adaptive node placement dense near eta = 1 (cross_sections.py:1025-1040):
    ``u(q) = 1 - 0.6 log(1 + q/0.02)/log(51) - 0.4 q^2`` at ``q_k = k/(K-1)``, strictly decreasing from 1 to 0."""
    q = torch.arange(num_knots, dtype=torch.float64) / (num_knots - 1)
    u = 1.0 - 0.6 * torch.log1p(q / 0.02) / math.log(51.0) - 0.4 * q * q
    u[0], u[-1] = 1.0, 0.0
    return u


def angular_table_nodes(num_nodes: int, eta_map: Optional[Tensor] = None) -> Tensor:
    """[N] float64 node cosines, ascending from -1 to 1 (cross_sections.py:948-973): equally spaced, or node j at the
    map coordinate ``u = j/(N-1)``, ``eta_j = 1 - 2 q_j^2`` with q_j from the inverse of the piecewise-linear map."""
    n = int(num_nodes)
    if eta_map is None:
        return torch.linspace(-1.0, 1.0, n, dtype=torch.float64)
    uk = eta_map.detach().to("cpu", torch.float64)
    K = uk.numel()
    if not bool((uk[1:] < uk[:-1]).all()):
        raise ValueError("eta_map must decrease strictly from the first knot (eta = 1) to the last (eta = -1)")
    qk = torch.arange(K, dtype=torch.float64) / (K - 1)
    au, aq = uk.flip(0), qk.flip(0)
    u = torch.arange(n, dtype=torch.float64) / (n - 1)
    j = torch.searchsorted(au, u, right=True).clamp(1, K - 1)
    q = aq[j - 1] + (aq[j] - aq[j - 1]) * (u - au[j - 1]) / (au[j] - au[j - 1])
    eta = 1.0 - 2.0 * q * q
    eta[0], eta[-1] = -1.0, 1.0
    return eta


# (scale, bone share) of each default reference material's row in the seven-material synthetic library: the row is
# scale x ((1 - share) x water + share x bone) of the two closed-form rows (cross_sections.py:1112-1116)
_SEVEN_MATERIAL_BLENDS = {"air": (0.90, 0.0), "lung_icrp": (0.95, 0.02), "adipose_icrp": (0.97, 0.0),
                          "water": (1.0, 0.0), "muscle_icru44": (1.0, 0.02), "soft_tissue_icru44": (1.0, 0.04),
                          "bone_cortical_icrp": (1.0, 1.0)}
# nominal densities (g/cm^3) of the default reference materials (cross_sections.py:83-89)
_NOMINAL_DENSITIES = {"air": 0.0012, "lung_icrp": 0.26, "adipose_icrp": 0.95, "water": 1.0, "muscle_icru44": 1.05,
                      "soft_tissue_icru44": 1.06, "bone_cortical_icrp": 1.85}


def _blend_rows(rows: list, a: Tensor) -> Tensor:
    """[M, ...] material rows mixed from the two base rows ``a [2, ...]``: a row with share 0 or 1 and scale 1 is
    the base row itself (bit for bit), every other row ``scale ((1 - share) a0 + share a1)``."""
    out = []
    for scale, share in rows:
        if scale == 1.0 and share in (0.0, 1.0):
            out.append(a[int(share)].clone())
        else:
            out.append(scale * ((1.0 - share) * a[0] + share * a[1]))
    return torch.stack(out)


def synthetic_library(num_groups: int = 4, legendre_order: int = 5, num_eta: int = 401, num_fine_per_group: int = 4,
                      fine_spread: float = 0.0, num_materials: int = 2,
                      angular_table_grid: str = "uniform") -> MultigroupLibrary:
    """The reference's closed-form test library (cross_sections.py:1043-1131): kernel ``a (1 + b eta) / 4 pi`` with
    ``a = a_m 0.6^(go-gi)`` (a_m = 0.15, 0.12), ``b = 0.5`` in-group else 0.2, photoelectric ``0.05 / 0.2 * 1.5^g``;
    its Legendre moments are exactly ``a`` and ``a b / 3``. ``num_materials`` 2 gives (water, bone); 7 the seven
    default reference materials, each row a fixed blend of the two; any other M > 2 the rows "m0", "m1", ... (water,
    bone, then ``1 + 0.1 m`` times the blend at bone share ``(m - 1)/(M - 1)``). ``angular_table_grid`` "adaptive"
    places the table nodes by ``synthetic_angular_map``."""
    G, L = num_groups, legendre_order
    edges = torch.linspace(125.0, 10.0, G + 1, dtype=torch.float64)
    if angular_table_grid not in ("uniform", "adaptive"):
        raise ValueError(f"angular_table_grid must be 'uniform' or 'adaptive', got {angular_table_grid!r}")
    M = int(num_materials)
    if M < 2:
        raise ValueError("num_materials must be >= 2")
    eta_map = synthetic_angular_map() if angular_table_grid == "adaptive" else None
    eta = angular_table_nodes(num_eta, eta_map)
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
    fraction, fine_energy, fine_mu_t, fine_weight, fine_group = _fine_tables(edges, mu_t, num_fine_per_group,
                                                                             fine_spread)
    if M == 2:
        names, nominal = ("water", "bone_cortical_icrp"), torch.tensor([1.0, 1.85], dtype=torch.float64)
    else:
        if M == len(_SEVEN_MATERIAL_BLENDS):
            names = tuple(_SEVEN_MATERIAL_BLENDS)
            rows = [_SEVEN_MATERIAL_BLENDS[n] for n in names]
            nominal = torch.tensor([_NOMINAL_DENSITIES[n] for n in names], dtype=torch.float64)
        else:
            names = tuple(f"m{m}" for m in range(M))
            rows = [(1.0, 0.0), (1.0, 1.0)] + [(1.0 + 0.1 * m, (m - 1) / (M - 1)) for m in range(2, M)]
            nominal = torch.arange(1, M + 1, dtype=torch.float64)
        mu_t, mu_pe, mu_s_l, mu_s_eta, fine_mu_t = (_blend_rows(rows, a)
                                                    for a in (mu_t, mu_pe, mu_s_l, mu_s_eta, fine_mu_t))
    return MultigroupLibrary(edges, mu_t, mu_s_l, eta, mu_s_eta, names, fraction, fine_energy, fine_mu_t,
                             fine_weight, fine_group, eta_map=eta_map, mu_pe=mu_pe, nominal_density=nominal)


def seeded_library(num_groups: int, legendre_order: int, e_max_kev: float, e_min_kev: float, seed: int,
                   num_eta: int = 257, num_fine_per_group: int = 4, fine_spread: float = 0.15) -> MultigroupLibrary:
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
    return MultigroupLibrary(edges, totals, mu_s_l, eta, mu_s_eta, ("water", "bone"),
                             *_fine_tables(edges, totals, num_fine_per_group, fine_spread),
                             mu_pe=totals - mu_s_l[..., 0].sum(dim=2),
                             nominal_density=torch.tensor([1.0, 1.85], dtype=torch.float64))


def mu_t_volume(library: MultigroupLibrary, rho: Tensor) -> Tensor:
    """[G, R, C, S] total attenuation ``sum_m mu_t[m, g] rho[m]`` in 1/cm (cross_sections.py:762)."""
    return torch.einsum("mg,mijk->gijk", library.mu_t.to(rho.device, rho.dtype), rho)


def transfer_moments_volume(library: MultigroupLibrary, rho: Tensor) -> Tensor:
    """Dense [G_in, G_out, L+1, R, C, S] transfer moments in 1/cm (cross_sections.py:886)."""
    return torch.einsum("mabl,mijk->ablijk", library.mu_s_l.to(rho.device, rho.dtype), rho)


_FINE_TABLES: dict = {}


def fine_device_tables(library: "MultigroupLibrary", device, materials=None) -> tuple:
    """(fine mu_t [M', F] float32, fine group int32 [F], fine weight float32 [F]) of the library on the device for
    the materials listed (all when None), cached per library, materials and device: no host-to-device copy per
    call."""
    mats = None if materials is None else tuple(int(m) for m in materials)
    key = (id(library), mats, str(device))
    hit = _FINE_TABLES.get(key)
    if hit is not None and hit[0] is library:
        return hit[1]
    fine_mu = library.fine_mu_t if mats is None else library.fine_mu_t[list(mats)]
    tables = (fine_mu.to(device, torch.float32).contiguous(), library.fine_group.to(device, torch.int32).contiguous(),
              library.fine_weight.to(device, torch.float32).contiguous())
    _FINE_TABLES[key] = (library, tables)
    return tables


def group_transmission(library: MultigroupLibrary, path: Tensor, materials=None, filter_path: Optional[Tensor] = None,
                       filter_attenuation: Optional[Tensor] = None) -> Tensor:
    """[G, ...] group transmission from material line integrals ``path [M', ...]`` in g/cm^2 on the HIP kernel
    (cross_sections.py:797): ``T_g = sum_{e in g} w_e exp(-sum_m fine_mu_t[m, e] path[m] - mu_f[e] P_f)``."""
    from .. import _hip

    if (filter_path is None) != (filter_attenuation is None):
        raise ValueError("filter_path and filter_attenuation must be given together")
    device = path.device
    Mp = int(path.shape[0])
    spatial = tuple(path.shape[1:])
    mu, grp, w = fine_device_tables(library, device, materials)
    if int(mu.shape[0]) != Mp:
        raise ValueError(f"path has {Mp} materials, the library {int(mu.shape[0])}")
    paths = path.reshape(Mp, -1).contiguous()
    _hip.require_gpu_tensor(paths, "path")
    nrays = int(paths.shape[1])
    fp = fmu = None
    if filter_path is not None:
        fp = filter_path.reshape(-1).to(device, torch.float32).contiguous()
        fmu = filter_attenuation.reshape(-1).to(device, torch.float32).contiguous()
        if fp.numel() != nrays or fmu.numel() != int(mu.shape[1]):
            raise ValueError("filter_path must hold one value per ray and filter_attenuation one per fine energy")
    G = library.num_groups
    out = torch.empty(G, nrays, device=device, dtype=torch.float32)
    with torch.cuda.device(device):
        _hip.check(_hip.library().citrus_group_transmission(
            paths.data_ptr(), Mp, nrays, mu.data_ptr(), grp.data_ptr(), w.data_ptr(), int(mu.shape[1]), G,
            _hip.ptr(fp), _hip.ptr(fmu), out.data_ptr(), _hip.stream_of(device)), "citrus_group_transmission")
    return out.reshape((G,) + spatial)
