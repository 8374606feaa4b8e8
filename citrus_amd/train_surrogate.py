"""Train the Phi_{>2} surrogate of the 64^3 x 16-group x S12 x P5 view workload on the GPU and save its weights.

    python -m citrus_amd.train_surrogate --out citrus_amd/weights/surrogate_64cube_s12_p5_g16.npz

Training pairs come from the solver's own collision orders (plain source iteration, batched over four view angles per
phantom) on seeded ellipsoid phantoms of the 256^3 / 0.08 cm family, coarsened to 64^3 transport cells; the held-out
phantom of the benchmark (seed 1234) is never a training phantom. Every forward, backward and optimizer step runs on
the library's kernels. Prints one line per phase and the held-out error of phi1 + phi2 with and without the network.
"""

from __future__ import annotations

import argparse
import math
import time

import torch

from .scatter_solver.config import SolverConfig
from .scatter_solver.cross_sections import seeded_library
from .scatter_solver.detector import energy_weighted_response
from .scatter_solver.geometry import ConeGeometry, Grid
from .scatter_solver.scatter_solver import ScatterSolver
from .surrogate import (UNet3D, collision_orders_views, features, p1_means, sample_scale, scalar_flux_weight, train)
from .utilities import ellipsoid_phantom


def view_solver(device, cells: int = 64) -> tuple:
    """(solver, fine grid) of the view workload: 16 groups P5, S12, 0.32 cm cells, 16 x 32 cones, curved 256^2 panel."""
    fine = Grid((4 * cells, 4 * cells, 4 * cells), 0.08 * 64 / cells)
    config = SolverConfig(legendre_order=5, sweep_quadrature="level_symmetric", sn_order=12,
                          transport_cell_cm=0.32 * 64 / cells, quadrature_point_spacing_cm=0.5, cone_num_polar=16,
                          cone_num_azimuthal=32, direction_chunk=64)
    geometry = ConeGeometry(100.0, 150.0, 256, 256, 0.12, 0.12, use_curved_detector=True)
    solver = ScatterSolver(seeded_library(16, 5, 140.0, 20.0, seed=99), config, geometry,
                           detector=energy_weighted_response(), device=device)
    return solver, fine


def make_pairs(solver, fine: Grid, phantoms: int, seed0: int, tolerance: float, with_phi1: bool, log=print):
    xs, ys = [], []
    t0 = time.perf_counter()
    for i in range(phantoms):
        rho = ellipsoid_phantom(fine, seed=seed0 + i).to(solver.device, torch.float32)
        fields = solver.material_fields(rho, fine)
        angles = [2 * math.pi * i / phantoms + 0.5 * j * math.pi for j in range(4)]
        for o in collision_orders_views(solver, fields, angles, tolerance=tolerance):
            s = sample_scale(o.phi2)
            xs.append(features(fields.rho_coarse, o.phi2, s, o.phi1 if with_phi1 else None))
            ys.append(p1_means(o.total - o.phi1 - o.phi2).float() / s)
        del fields, rho
        log(f"phantom {i + 1}/{phantoms}: {time.perf_counter() - t0:.1f} s")
    return xs, ys


def held_out_errors(solver, fine: Grid, net: UNet3D, with_phi1: bool, seed: int = 1234):
    """(error of phi1 + phi2, error with the network) of the P1 cell means on the held-out phantom, four views."""
    fields = solver.material_fields(ellipsoid_phantom(fine, seed=seed).to(solver.device, torch.float32), fine)
    two, with_net = [], []
    for o in collision_orders_views(solver, fields, [2 * math.pi * k / 64 for k in range(4)], tolerance=1e-5):
        s = sample_scale(o.phi2)
        pred = net.forward(features(fields.rho_coarse, o.phi2, s, o.phi1 if with_phi1 else None)[None])[0] * s
        exact, base = p1_means(o.total), p1_means(o.phi1 + o.phi2)
        norm = torch.linalg.vector_norm(exact)
        two.append(float(torch.linalg.vector_norm(exact - base) / norm))
        with_net.append(float(torch.linalg.vector_norm(exact - base - pred) / norm))
    return two, with_net


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--phantoms", type=int, default=24)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--width", type=int, default=16)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--scalar-weight", type=float, default=4.0)
    ap.add_argument("--tolerance", type=float, default=1e-4)
    ap.add_argument("--with-phi1", action="store_true", help="inputs: densities, phi2 and phi1")
    args = ap.parse_args()
    device = torch.device("cuda:0")
    solver, fine = view_solver(device)
    with_phi1 = args.with_phi1
    xs, ys = make_pairs(solver, fine, args.phantoms, 3000, args.tolerance, with_phi1,
                        log=lambda m: print(m, flush=True))
    G = solver.num_groups
    net = UNet3D.create(int(xs[0].shape[0]), 4 * G, width=args.width, device=device, seed=0)
    print(f"{len(xs)} pairs, {net.num_parameters()} parameters", flush=True)
    weight = scalar_flux_weight(G, args.scalar_weight)
    t0 = time.perf_counter()
    stages = [(args.epochs // 2, args.lr), (args.epochs // 4, args.lr / 4),
              (args.epochs - args.epochs // 2 - args.epochs // 4, args.lr / 16)]
    for k, (epochs, lr) in enumerate(stages):
        hist = train(net, xs, ys, epochs=epochs, lr=lr, batch=args.batch, seed=k, channel_weight=weight,
                     log=lambda m: print(m, flush=True), log_every=10)
        print(f"stage {k}: lr {lr:g}, {epochs} epochs, loss {hist[0]:.5f} -> {hist[-1]:.5f} "
              f"({time.perf_counter() - t0:.1f} s)", flush=True)
    del xs, ys
    two, with_net = held_out_errors(solver, fine, net, with_phi1)
    print("held-out P1 error: phi1+phi2 " + ", ".join(f"{e:.4f}" for e in two) + "; with the network "
          + ", ".join(f"{e:.4f}" for e in with_net), flush=True)
    net.save(args.out)
    print(f"saved {args.out}", flush=True)


if __name__ == "__main__":
    main()
