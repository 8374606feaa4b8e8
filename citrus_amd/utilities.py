"""Seeded synthetic phantoms of the benchmark configurations (host-side, float64 CPU).

A water cylinder (1.0 g/cm^3, axis along z) holding 3-6 bone ellipsoids (density U[1.4, 1.9]) and 1-3 air
ellipsoids (0.001 g/cm^3), returned as two material channels (water-like, bone-like) of densities in g/cm^3 — the
``rho [M, R, C, S]`` input of the solver (the reference's WLAB phantom, citrus/utilities.py:104, plays the same role
for its clinical geometry). Also the WLAB phantom, its grid and its cone-beam geometry (``wlab_phantom``,
``wlab_grid``, ``wlab_geometry``: the reference's regression fixtures are solved on them) and the scan helpers
``view_angles`` and ``helical_scan``.
"""

from __future__ import annotations

import math

import numpy as np
import torch
from torch import Tensor

from .scatter_solver.config import MATERIALS
from .scatter_solver.geometry import ConeGeometry, Grid


def ellipsoid_phantom(grid: Grid, seed: int) -> Tensor:
    """[2, R, C, S] densities (water channel, bone channel) of a seeded water cylinder with ellipsoid inserts."""
    gen = torch.Generator().manual_seed(seed)

    def u(a, b):
        return a + (b - a) * float(torch.rand((), generator=gen, dtype=torch.float64))

    pos = grid.center_coordinates(dtype=torch.float64)  # [R, C, S, 3]
    x, y, z = pos[..., 0], pos[..., 1], pos[..., 2] - grid.recon_slice_offset
    ex, ey, ez = grid.extent
    radius = 0.42 * min(ex, ey)
    half_len = 0.45 * ez
    water = ((x * x + y * y <= radius * radius) & (z.abs() <= half_len)).to(torch.float64)
    bone = torch.zeros_like(water)

    def ellipsoid():
        c = torch.tensor([u(-0.5, 0.5) * radius, u(-0.5, 0.5) * radius, u(-0.6, 0.6) * half_len])
        ax = torch.tensor([u(0.08, 0.22) * radius, u(0.08, 0.22) * radius, u(0.1, 0.3) * half_len])
        t = u(0.0, math.pi)
        ct, st = math.cos(t), math.sin(t)
        dx, dy, dz = x - c[0], y - c[1], z - c[2]
        rx, ry = ct * dx + st * dy, -st * dx + ct * dy
        return (rx / ax[0]) ** 2 + (ry / ax[1]) ** 2 + (dz / ax[2]) ** 2 <= 1.0

    for _ in range(int(torch.randint(3, 7, (), generator=gen))):
        inside = ellipsoid() & (water > 0)
        bone[inside] = u(1.4, 1.9)
        water[inside] = 0.0
    for _ in range(int(torch.randint(1, 4, (), generator=gen))):
        inside = ellipsoid() & ((water > 0) | (bone > 0))
        water[inside] = 0.001
        bone[inside] = 0.0
    return torch.stack([water, bone])


def coarsen(rho: Tensor, factor: int) -> Tensor:
    """Cell means over ``factor^3`` blocks of [M, R, C, S] (exact conservative resampling for aligned grids)."""
    M, R, C, S = rho.shape
    f = int(factor)
    return rho.reshape(M, R // f, f, C // f, f, S // f, f).mean(dim=(2, 4, 6))


def view_angles(num_views: int, start_view_angle: float = 0.0, end_view_angle: float = 2.0 * math.pi):
    """[V] float64 view angles ``start + (end - start) i / V`` (utilities.py:1119; the end is excluded)."""
    n = int(num_views)
    if n < 1:
        raise ValueError("num_views must be >= 1")
    if not float(end_view_angle) > float(start_view_angle):
        raise ValueError("end_view_angle must be larger than start_view_angle")
    return float(start_view_angle) + (float(end_view_angle) - float(start_view_angle)) * \
        np.arange(n, dtype=np.float64) / n


def helical_scan(geometry, angles, pitch_cm: float, center_cm: float = 0.0):
    """Geometry of a helical scan with constant pitch (utilities.py:1153): the z shift of view angle a is
    ``center_cm + pitch_cm (a - a_mid) / (2 pi)``, a_mid the middle of the angular range."""
    a = np.asarray(angles.detach().cpu().numpy() if isinstance(angles, Tensor) else angles,
                   dtype=np.float64).reshape(-1)
    middle = 0.5 * (float(a.min()) + float(a.max())) if a.size else 0.0
    return geometry.with_helical_z_shifts(a, float(center_cm) + float(pitch_cm) * (a - middle) / (2.0 * math.pi))


# ---- the WLAB phantom (citrus/utilities.py:78-160, :495-525) --------------------------------------------------------
# A 30 x 20 x 40 cm (x, y, z) water prism about the isocentre with three 40 cm long square columns: lung (6 x 6 cm,
# water at 0.26 g/cm^3), air (2 x 2 cm, water at 1.205e-3) and cortical bone (2 x 2 cm at 1.85 in the bone row, no
# water). Membership is decided at the voxel centre with half-open intervals [centre - w/2, centre + w/2).
WLAB_SIZE_XYZ = (30.0, 20.0, 40.0)
WLAB_INSERTS = {"lung": ((6.0, 6.0), 0.26), "air": ((2.0, 2.0), 1.205e-3), "bone": ((2.0, 2.0), 1.85)}
_WLAB_TOLERANCE = 1e-9


def _row_of(materials, name: str) -> int:
    try:
        return list(materials).index(name)
    except ValueError:
        raise ValueError(f"the WLAB phantom needs {name!r} among the library materials, got {tuple(materials)}") \
            from None


def _interval(c: Tensor, centre: float, width: float) -> Tensor:
    lo = centre - 0.5 * width
    return (c >= lo) & (c < lo + width)


def wlab_phantom(grid: Grid, *, lung_center=(-5.0, 0.0), air_center=(3.0, 0.0), bone_center=(7.0, 0.0),
                 materials=None, device=None, dtype=torch.float32) -> Tensor:
    """[M, R, C, S] densities (g/cm^3) of the WLAB phantom rasterized on ``grid`` (utilities.py:104): the water row
    holds the prism with the lung and air columns at their reduced densities and none at the bone column, the bone
    row the bone column, every other row zero. ``materials`` defaults to the seven reference materials. Raises
    ``ValueError`` when water or bone is missing, the grid does not cover the prism, or a column leaves the prism or
    overlaps another."""
    names = MATERIALS if materials is None else tuple(materials)
    water, bone = _row_of(names, "water"), _row_of(names, "bone_cortical_icrp")
    half = [0.5 * e for e in WLAB_SIZE_XYZ]
    lo = grid.lower_corner()
    for axis in range(3):
        if lo[axis] > -half[axis] + _WLAB_TOLERANCE or lo[axis] + grid.extent[axis] < half[axis] - _WLAB_TOLERANCE:
            raise ValueError(f"grid box {lo} + {grid.extent} does not cover the WLAB prism {WLAB_SIZE_XYZ}")
    columns = [(tuple(map(float, lung_center)), WLAB_INSERTS["lung"]), (tuple(map(float, air_center)),
                                                                       WLAB_INSERTS["air"]),
               (tuple(map(float, bone_center)), WLAB_INSERTS["bone"])]
    for i, (ci, ((wx, wy), _)) in enumerate(columns):
        if abs(ci[0]) + wx / 2 > half[0] + _WLAB_TOLERANCE or abs(ci[1]) + wy / 2 > half[1] + _WLAB_TOLERANCE:
            raise ValueError(f"insert {i} at {ci} leaves the water prism")
        for cj, ((vx, vy), _) in columns[i + 1:]:
            if (abs(ci[0] - cj[0]) < (wx + vx) / 2 - _WLAB_TOLERANCE
                    and abs(ci[1] - cj[1]) < (wy + vy) / 2 - _WLAB_TOLERANCE):
                raise ValueError(f"inserts at {ci} and {cj} overlap")
    x, y, z = grid.axis_coordinates(device, torch.float64)
    X, Y, Z = x.view(1, -1, 1), y.view(-1, 1, 1), z.view(1, 1, -1)
    prism = _interval(X, 0.0, WLAB_SIZE_XYZ[0]) & _interval(Y, 0.0, WLAB_SIZE_XYZ[1]) & _interval(Z, 0.0,
                                                                                                   WLAB_SIZE_XYZ[2])
    water_map = prism.to(torch.float64)
    bone_map = torch.zeros_like(water_map)
    for name, (centre, ((wx, wy), density)) in zip(("lung", "air", "bone"), columns):
        inside = _interval(X, centre[0], wx) & _interval(Y, centre[1], wy) & prism
        if name == "bone":
            water_map = torch.where(inside, 0.0, water_map)
            bone_map = torch.where(inside, density, bone_map)
        else:
            water_map = torch.where(inside, density, water_map)
    rho = torch.zeros(len(names), *grid.shape, device=device, dtype=dtype)
    rho[water] = water_map.to(dtype)
    rho[bone] = bone_map.to(dtype)
    return rho


def wlab_grid(delta_voxel: float) -> Grid:
    """Grid that tiles the 30 x 20 x 40 cm WLAB box exactly (utilities.py:495); ``delta_voxel`` must divide every
    extent."""
    counts = []
    for length in (WLAB_SIZE_XYZ[1], WLAB_SIZE_XYZ[0], WLAB_SIZE_XYZ[2]):
        k = round(length / delta_voxel)
        if k < 1 or abs(k * delta_voxel - length) > 1e-9 * max(1.0, length):
            raise ValueError(f"delta_voxel={delta_voxel} does not divide the WLAB extent {length} cm")
        counts.append(int(k))
    return Grid(shape=tuple(counts), delta_voxel=float(delta_voxel))


def wlab_geometry() -> ConeGeometry:
    """The WLAB cone-beam setup (utilities.py:514): SID 100 cm, SDD 150 cm, a flat 100 x 75 panel of 4 mm pixels."""
    return ConeGeometry(source_iso_dist=100.0, source_detector_dist=150.0, num_det_rows=75, num_det_channels=100,
                        delta_det_row=0.4, delta_det_channel=0.4)


# ---- data sets (utilities.py:1480-1527) ------------------------------------------------------------------------------
# A data set is a dict of numpy arrays (sinograms, angles, images, ...) plus "meta", a JSON-serialisable dict. On disk
# it is <base>.npz holding the arrays and the meta as the JSON string "meta_json", and <base>.json with the meta alone
# (the reference's layout, so either implementation reads the other's files). Format 3 and later only.
DATASET_FORMAT = 3


def _dataset_base(path) -> str:
    base = str(path)
    for ext in (".npz", ".json"):
        if base.endswith(ext):
            return base[: -len(ext)]
    return base


def save_dataset(data: dict, out_path) -> list:
    """Write ``data`` (arrays plus "meta") as ``<base>.npz`` and ``<base>.json`` (a trailing .npz or .json of
    ``out_path`` is dropped); sets ``meta["files"]`` to the two paths and returns them."""
    import json
    import os

    if "meta" not in data or not isinstance(data["meta"], dict):
        raise ValueError("a data set needs a 'meta' dict")
    base = _dataset_base(out_path)
    folder = os.path.dirname(os.path.abspath(base))
    os.makedirs(folder, exist_ok=True)
    meta = data["meta"]
    meta.setdefault("data_format", DATASET_FORMAT)
    text = json.dumps(meta)
    arrays = {name: np.asarray(value) for name, value in data.items() if name != "meta"}
    if "meta_json" in arrays:
        raise ValueError("'meta_json' is reserved for the data set's meta")
    np.savez(base + ".npz", meta_json=np.asarray(text), **arrays)
    with open(base + ".json", "w") as f:
        f.write(json.dumps(meta, indent=2))
    meta["files"] = [base + ".npz", base + ".json"]
    return meta["files"]


def load_dataset(path) -> dict:
    """Read a data set written by ``save_dataset`` (here or by the reference): the arrays and "meta". Nothing in the
    file is unpickled (``numpy.load`` without pickles); a data set older than format 3 raises ``ValueError``."""
    import json

    base = _dataset_base(path)
    with np.load(base + ".npz", allow_pickle=False) as z:
        if "meta_json" not in z.files:
            raise ValueError(f"{base}.npz holds no 'meta_json': not a data set")
        out = {name: z[name] for name in z.files if name != "meta_json"}
        meta = json.loads(str(z["meta_json"][()]))
    if int(meta.get("data_format", 0)) < DATASET_FORMAT:
        raise ValueError(f"{base}.npz has data format {meta.get('data_format', 0)} (< {DATASET_FORMAT}): written "
                         "before the polychromatic primary; synthesize it again")
    out["meta"] = meta
    return out
