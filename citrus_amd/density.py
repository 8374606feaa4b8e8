"""Density model: the reconstruction's attenuation to the densities of the library's reference materials.

Keeps the interface of the reference's density model (``citrus/utilities.py:577-753``, ``citrus/citrus.py:466-560``):
``effective_energy`` (spectrum-mean energy), ``mass_attenuation_at`` (log-log interpolation of the fine table),
``nominal_densities``, ``nominal_attenuation``, ``material_order``, ``attenuation_to_densities`` (the interpolating
model, on the HIP kernel ``citrus_attenuation_to_densities``), ``densities_to_attenuation`` and ``DensityModel``.
``ScatterSolver.material_fields_from_attenuation`` builds the solver's fields from an attenuation volume with it.

The small per-material tables (energies, attenuations) are host float64, as in the reference; the per-voxel model
runs on the GPU in float32 with the reference's arithmetic.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch
from torch import Tensor

from . import _hip
from .scatter_solver.cross_sections import MultigroupLibrary

# two materials closer than this in nominal attenuation cannot be bracketed apart (relative)
MATERIAL_ORDER_GAP = 1e-6


def _fine_tables(library: MultigroupLibrary) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """(fine energies [E] ascending, their spectrum weights [E], mass attenuations [M, E]) as float64 numpy arrays
    (a stable sort keeps equal energies in library order)."""
    E = library.fine_energy.detach().cpu().double().numpy()
    order = np.argsort(E, kind="stable")
    return (E[order], library.fine_weight.detach().cpu().double().numpy()[order],
            library.fine_mu_t.detach().cpu().double().numpy()[:, order])


def effective_energy(library: MultigroupLibrary) -> float:
    """Spectrum-mean photon energy (keV): the fine energies averaged with their spectrum weights
    (utilities.py:577)."""
    E, w, _ = _fine_tables(library)
    return float(np.average(E, weights=w))


def mass_attenuation_at(library: MultigroupLibrary, energy_kev: float) -> Tensor:
    """[M] total mass attenuation (cm^2/g, float64 on the CPU) at one energy: the fine table's straight line in
    (log E, log mu) between the neighbouring fine energies; an energy equal to a fine energy gives that column (the
    last of equal ones), so a monoenergetic library works (utilities.py:590). ValueError outside the fine-energy
    range."""
    E, _, mu = _fine_tables(library)
    e = float(energy_kev)
    if not E[0] <= e <= E[-1]:
        raise ValueError(f"energy {e} keV outside the library's fine-energy range [{E[0]}, {E[-1]}]")
    same = np.flatnonzero(E == e)
    if same.size:
        return torch.from_numpy(mu[:, same[-1]].copy())
    j = int(np.searchsorted(E, e)) - 1  # E[j] < e < E[j + 1]
    frac = (math.log(e) - math.log(E[j])) / (math.log(E[j + 1]) - math.log(E[j]))
    log_mu = (1.0 - frac) * np.log(mu[:, j]) + frac * np.log(mu[:, j + 1])
    return torch.from_numpy(np.exp(log_mu))


def nominal_densities(library: MultigroupLibrary) -> Tensor:
    """[M] nominal densities, g/cm^3 (utilities.py:622); ValueError for a library without them."""
    if library.nominal_density is None:
        raise ValueError("the library holds no nominal densities (MultigroupLibrary.nominal_density)")
    return library.nominal_density.detach().to("cpu", torch.float64)


def nominal_attenuation(library: MultigroupLibrary, energy_kev: Optional[float] = None) -> Tensor:
    """[M] attenuation (1/cm) of each material at its nominal density (utilities.py:634)."""
    return nominal_densities(library) * mass_attenuation_at(
        library, effective_energy(library) if energy_kev is None else float(energy_kev))


def material_order(library: MultigroupLibrary, energy_kev: Optional[float] = None) -> list[int]:
    """Material indices in increasing nominal attenuation (utilities.py:648); ValueError when two materials cannot
    be told apart."""
    att = nominal_attenuation(library, energy_kev)
    order = torch.argsort(att).tolist()
    for a, b in zip(order, order[1:]):
        if float(att[b] - att[a]) <= MATERIAL_ORDER_GAP * float(att[b]):
            raise ValueError(f"materials {library.materials[a]!r} and {library.materials[b]!r} have nearly the "
                             f"same nominal attenuation ({float(att[a]):.6g}, {float(att[b]):.6g} 1/cm)")
    return order


def _slot_tables(library: MultigroupLibrary, energy_kev: Optional[float]):
    """float32 slot tables in bracketing order, as the reference forms them on the device (utilities.py:726-734)."""
    e = effective_energy(library) if energy_kev is None else float(energy_kev)
    order = material_order(library, e)
    mass = mass_attenuation_at(library, e)[order].to(torch.float32)
    nominal = nominal_densities(library)[order].to(torch.float32)
    mu = nominal * mass  # float32 product, the reference's `nominal * mass_attenuation`
    return (np.asarray(order, dtype=np.int32), mu.numpy().astype(np.float32), nominal.numpy().astype(np.float32),
            mass.numpy().astype(np.float32))


def attenuation_to_densities(library: MultigroupLibrary, X: Tensor, energy_kev: Optional[float] = None) -> Tensor:
    """[M, ...] densities (g/cm^3) of the interpolating model for an attenuation volume X [...] in 1/cm on the GPU
    (utilities.py:710): the two materials whose nominal attenuations bracket max(X, 0) share it linearly; below the
    first or above the last, that material alone with density X / m."""
    _hip.require_gpu_tensor(X, "X")
    order, mu, nominal, mass = _slot_tables(library, energy_kev)
    M = library.num_materials
    rho = torch.empty((M,) + tuple(X.shape), device=X.device, dtype=torch.float32)
    lib = _hip.library()
    keep = [np.ascontiguousarray(a) for a in (order, mu, nominal, mass)]
    with torch.cuda.device(X.device):
        _hip.check(lib.citrus_attenuation_to_densities(X.data_ptr(), X.numel(), M, keep[0].ctypes.data,
                                                       keep[1].ctypes.data, keep[2].ctypes.data,
                                                       keep[3].ctypes.data, rho.data_ptr(),
                                                       _hip.stream_of(X.device)),
                   "citrus_attenuation_to_densities")
    return rho


def densities_to_attenuation(library: MultigroupLibrary, rho: Tensor, energy_kev: Optional[float] = None) -> Tensor:
    """Attenuation ``sum_m rho[m] m_m(E)`` of a density volume [M, ...] (utilities.py:688)."""
    M = library.num_materials
    if int(rho.shape[0]) != M:
        raise ValueError(f"rho must be [{M}, ...] for the library's materials {tuple(library.materials)}")
    e = effective_energy(library) if energy_kev is None else float(energy_kev)
    m = mass_attenuation_at(library, e).to(rho.device, rho.dtype)
    return torch.tensordot(m, rho, dims=([0], [0]))


@dataclass
class DensityModel:
    """Attenuation X (1/cm at ``energy_kev``) to the densities of the library's materials (citrus.py:466): the
    interpolating model applied to X clamped at 0.

    Attributes:
        library: library whose materials and nominal densities define the model.
        energy_kev: effective energy of the attenuation, keV.
        last_negative_fraction: fraction of the voxels clamped at 0 in the last ``__call__`` or
            ``negative_fraction``.
    """

    library: MultigroupLibrary
    energy_kev: float
    last_negative_fraction: float = 0.0

    @classmethod
    def from_library(cls, library: MultigroupLibrary, energy_kev: Optional[float] = None) -> "DensityModel":
        """The model of a library at an energy (default: the spectrum-mean energy)."""
        return cls(library, effective_energy(library) if energy_kev is None else float(energy_kev))

    def __call__(self, X: Tensor) -> Tensor:
        """[M, ...] densities of X; sets ``last_negative_fraction``."""
        self.negative_fraction(X)
        return attenuation_to_densities(self.library, X, self.energy_kev)

    def material(self, X: Tensor, m: int) -> Tensor:
        """Density map of one material, equal to ``self(X)[m]``."""
        if not 0 <= int(m) < self.library.num_materials:
            raise ValueError(f"material {m} outside the library's {self.library.num_materials} materials")
        return attenuation_to_densities(self.library, X, self.energy_kev)[int(m)]

    def negative_fraction(self, X: Tensor) -> float:
        """Fraction of the voxels of X below 0 (also stored in ``last_negative_fraction``)."""
        self.last_negative_fraction = float((X < 0).float().mean()) if X.numel() else 0.0
        return self.last_negative_fraction
