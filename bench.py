#!/usr/bin/env python
"""Benchmark of the per-view transport step on one MI355X: CITRUS scatter view steps/s at 64^3 x 16 groups x S12.

Workload (the batched circular-scan configuration): a seeded 256^3 ellipsoid phantom at 0.08 cm (water cylinder, bone
and air ellipsoids), the 64^3 transport grid of 0.32 cm cells (one quadrature point per cell); 16 groups over 140-20
keV from a seeded two-material P5 library; the S12 level-symmetric sweep set (168 directions); the 16 x 32 cone set of
the scatter image on a 256 x 256 curved panel (0.12 cm pixels, SID 100 cm, SDD 150 cm); V views of a 64-view circle,
each with its own CITRUS state. One view step is the CITRUS step of one view: the uncollided trace, the corrected
first-collision source, one sweep of the summed first-collision and scattering sources (the accumulator update), the
signal-only collapsed-cone scatter image and the primary; the V views of a step run their accumulator updates in one
batched sweep. `value` counts completed view steps per second with every input resident in HBM; the sweep-only rate
(scattering moments plus one fused sweep per view, V views per launch sequence) and the sweep kernel's roofline are
reported beside it.

Prints one JSON line (rank 0) with the metric, the roofline of the sweep kernel (HIP events on the launch stream)
and the CPU baseline (the FP64 oracle's algorithm run in float32 on a sample of the same workload, on the host).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

METRIC = "scatter view-sweeps/s (64³×16g×S12), learned vs plain at ≤1% SPR error"
HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3  # dense float32 matrix-core peak of an MI355X


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--views", type=int, default=8, help="views per step (each with its own CITRUS state)")
    ap.add_argument("--sweep-only", action="store_true", help="time the sweep-only step alone")
    ap.add_argument("--cells", type=int, default=64, help="transport cells per axis")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-vs-plain comparison")
    ap.add_argument("--no-helical", action="store_true", help="skip the helical cropped-grid comparison")
    ap.add_argument("--cpu-directions", type=int, default=8, help="directions in the CPU-baseline sample")
    ap.add_argument("--schedule", choices=("auto", "persistent", "launches"), default="auto",
                    help="sweep schedule; 'auto' lets the autotuner pick it with the direction chunk")
    ap.add_argument("--chunk", type=int, default=0, help="directions per octant per sweep launch (0: the tuner's)")
    ap.add_argument("--profile", action="store_true",
                    help="the timed view steps and sweeps only (no bf16 form, scan stage, learned form, helical "
                         "configuration, graph or CPU baseline): the run to trace or count under rocprofv3")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a float32 buffer of MIB MiB three times on cuda:0 (the known-size kernel that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters), then exit")
    return ap.parse_args()


def copy_probe(mib: int) -> None:
    """Three plain device copies of ``mib`` MiB. Under ``rocprofv3 --pmc FETCH_SIZE`` (then ``WRITE_SIZE``) each copy
    dispatch moves exactly the buffer size each way, so its per-dispatch readings give the factors that turn the
    counters into bytes for the other kernels."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, device="cuda:0", dtype=torch.float32)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copy_probe: 3 copies of {mib} MiB ({n * 4} bytes each way)", flush=True)


def build_problem(args, device):
    import torch

    from citrus_amd.scatter_solver.angular_quadrature import level_symmetric_quadrature
    from citrus_amd.scatter_solver.cross_sections import mu_t_volume, seeded_library
    from citrus_amd.scatter_solver.geometry import Grid
    from citrus_amd.scatter_solver.source_iteration import MaterialTransfer
    from citrus_amd.scatter_solver.transport_sweep import make_sweep_context
    from citrus_amd.utilities import coarsen, ellipsoid_phantom

    n = args.cells
    fine = Grid((4 * n, 4 * n, 4 * n), 0.08 * 64 / n)
    coarse = Grid((n, n, n), 0.32 * 64 / n)
    rho = coarsen(ellipsoid_phantom(fine, seed=1234), 4)
    library = seeded_library(16, 5, 140.0, 20.0, seed=99)
    quadrature = level_symmetric_quadrature(12)
    ctx = make_sweep_context(coarse, quadrature, 5, device=device)
    mu_t = mu_t_volume(library, rho).to(device, torch.float32).contiguous()
    transfer = MaterialTransfer(rho.to(device, torch.float32).contiguous(),
                                library.mu_s_l.to(device, torch.float32).contiguous())
    G, nsh = 16, ctx.num_moments
    gen = torch.Generator(device=device).manual_seed(7)
    phi = torch.randn(args.views, G, n, n, n, nsh, 4, device=device, generator=gen) * 0.05
    phi[..., 0, 0] = phi[..., 0, 0].abs() + 1.0
    return ctx, mu_t, transfer, phi, library, rho, coarse


# ---- CPU baseline: the oracle's algorithm on the host's cores ------------------------------------------------------
# Runs before anything touches the GPU: each stage's bounded sample is split over a pool of forked worker processes
# (numpy, one thread each), timed by wall clock and scaled to one whole view step. Worker inputs live in _CPU, set
# before the pool forks.

_CPU: dict = {}
CPU_WORKERS_MAX = 16  # the host's share of cores per GPU


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def _w_uncollided(chunk):
    from oracle import rays as orays
    from oracle import uncollided as ounc

    c = _CPU
    pts = c["pts"][chunk]
    paths = orays.line_integrals(c["rho_c"], np.broadcast_to(c["src"], pts.shape), pts, c["lo"], c["delta"], c["nn"])
    return chunk, ounc.uncollided_fluence(paths, pts, c["fm"], c["fg"], c["fw"], c["G"], 100.0, 150.0, 256, 256,
                                          0.12, 0.12, 0.0, True)


def _w_first_collision(item):
    from oracle import first_collision as ofc

    chunk, phi_uc = item
    c = _CPU
    pts = c["pts"][chunk]
    rho_s = c["rho_c"].reshape(2, -1)[:, c["sample"][chunk]]
    ou = ofc.primary_directions(pts, c["src"])
    S = ofc.source_at_directions(c["mu_s_eta"], None, rho_s, phi_uc, ou, c["om"])
    I = ofc.quadrature_sums(S, ou, c["om"], c["w"])
    E = ofc.exact_moments(c["mu_s_l"], rho_s, phi_uc, 1)
    scale, beta = ofc.correction_factors(I, E)
    ofc.apply_correction(S, scale, beta, ou, c["om"])
    return S.shape


def _w_moments(rows):
    from oracle.scattering import scatter_moments

    c = _CPU
    mu_s = np.einsum("mabl,mijk->ablijk", c["mu_s_l"], c["rho32"][:, rows]).astype(np.float32)
    return scatter_moments(mu_s, c["phi"][:, :, :, rows]).shape


def _w_sweep(sel):
    from oracle import sweep as osw

    c = _CPU
    src = np.tensordot(c["Y"][sel].astype(np.float32), c["T"], axes=([1], [1])).astype(np.float32)
    psi = osw.sweep(c["shape"], c["deltas"], c["om"][sel], c["mu_t"], src, dtype=np.float32)
    out = np.tensordot((c["w"][sel, None] * c["Y"][sel]).astype(np.float32), psi, axes=([0], [0]))
    return out.shape


def _w_image(k):
    from oracle import image as oim
    from oracle import quadrature as oq
    from oracle import rays as orays

    c = _CPU
    wk = c["cone"][k]
    t0 = time.perf_counter()
    oim.downstream_tau(c["mu_coarse"], c["lo"], c["delta"], c["nn"], c["pts"], wk, 100.0, 150.0, 0.0, True)
    t1 = time.perf_counter()
    Y = oq.real_spherical_harmonics(5, wk[None])[0]
    S_mc = np.einsum("n,gnrcs->grcs", Y, c["Tm"])
    e = S_mc * np.exp(-np.abs(S_mc))  # the attenuated emission's arithmetic on the whole [G, 64^3] field
    t2 = time.perf_counter()
    a, b = orays.oblique_segments(c["pix"], wk, c["lo"], c["hi"])
    orays.line_integrals(e[:1], a, b, c["lo"], c["delta"], c["nn"])
    t3 = time.perf_counter()
    # one direction's image: depths on 1/64 of the cells and rays through 1/64 of the pixels, emission on every cell
    return (t1 - t0) * 64 + (t2 - t1) + (t3 - t2) * 64


def _w_primary(chunk):
    from oracle import rays as orays

    c = _CPU
    pix = c["pix"][chunk]
    return orays.line_integrals(c["rho_f"], np.broadcast_to(c["src"], pix.shape), pix, c["lof"], c["df"],
                                c["nf"]).shape


def cpu_baseline_parallel(args) -> dict:
    """The oracle's algorithm for one CITRUS view step of the benchmark (64^3 x 16 groups x S12 x P5, the 256^3
    phantom, the 16 x 32 cone set, the 256^2 curved panel) on P = min(16, host cores) worker processes: each stage
    on a bounded sample split over the workers, scaled to the whole view by its sample fraction."""
    import multiprocessing as mp

    import torch

    from citrus_amd.scatter_solver.angular_quadrature import hemisphere_quadrature, level_symmetric_quadrature
    from citrus_amd.scatter_solver.cross_sections import seeded_library
    from citrus_amd.scatter_solver.geometry import Grid
    from citrus_amd.utilities import coarsen, ellipsoid_phantom
    from oracle import quadrature as oq
    from oracle import rays as orays

    P = max(1, min(CPU_WORKERS_MAX, os.cpu_count() or 1))
    n = args.cells
    fine = Grid((4 * n, 4 * n, 4 * n), 0.08 * 64 / n)
    coarse = Grid((n, n, n), 0.32 * 64 / n)
    rho_f = ellipsoid_phantom(fine, seed=1234)
    rho_c = coarsen(rho_f, 4).numpy()
    library = seeded_library(16, 5, 140.0, 20.0, seed=99)
    q = level_symmetric_quadrature(12)
    om, w = q.omega.numpy(), q.weight.numpy()
    G, M = 16, int(om.shape[0])
    rng = np.random.default_rng(11)
    pts_all = coarse.center_coordinates(dtype=torch.float64).reshape(-1, 3).numpy()
    sample = rng.choice(pts_all.shape[0], pts_all.shape[0] // 64, replace=False)
    pixels = orays.pixel_centers(100.0, 150.0, 256, 256, 0.12, 0.12, 0.0, True)[:, :, 0, 0].reshape(-1, 3)
    pix = pixels[rng.choice(pixels.shape[0], pixels.shape[0] // 64, replace=False)]
    cone = hemisphere_quadrature(torch.tensor([0.0, -1.0, 0.0], dtype=torch.float64), 16, 32, 0.3,
                                 tangent_axes=([1.0, 0.0, 0.0], [0.0, 0.0, 1.0])).omega.numpy()
    phi = (rng.standard_normal((G, 36, 4) + coarse.shape) * 0.05).astype(np.float32)
    phi[:, 0, 0] = np.abs(phi[:, 0, 0]) + 1.0
    lo, delta = np.array(coarse.lower_corner()), np.array(coarse.deltas)
    nn = (coarse.num_cols, coarse.num_rows, coarse.num_slices)
    _CPU.clear()
    _CPU.update(
        pts=pts_all[sample], sample=sample, rho_c=rho_c, rho32=rho_c.astype(np.float32), src=orays.source_position(100.0, 0.0),
        lo=lo, hi=lo + delta * np.array(nn), delta=delta, nn=nn, fm=library.fine_mu_t.numpy(),
        fg=library.fine_group.numpy(), fw=library.fine_weight.numpy(), G=G, mu_s_eta=library.mu_s_eta.numpy(),
        mu_s_l=library.mu_s_l.numpy(), om=om, w=w, Y=oq.real_spherical_harmonics(5, om), phi=phi,
        T=(rng.standard_normal((G, 36, 4) + coarse.shape) * 0.05).astype(np.float32),
        Tm=rng.standard_normal((G, 36) + coarse.shape).astype(np.float32) * 1e-3,
        mu_t=np.einsum("mg,mijk->gijk", library.mu_t.numpy(), rho_c).astype(np.float32),
        mu_coarse=np.einsum("mg,mrcs->grcs", library.mu_t.numpy(), rho_c), shape=coarse.shape, deltas=coarse.deltas,
        cone=cone, pix=pix, rho_f=rho_f.numpy(), lof=np.array(fine.lower_corner()), df=np.array(fine.deltas),
        nf=(fine.num_cols, fine.num_rows, fine.num_slices))
    octant = oq.octant_of(om)
    per_octant = [np.nonzero(octant == o)[0] for o in range(8)]
    sweep_sets = [np.concatenate([d[i % d.size:i % d.size + 1] for d in per_octant]) for i in range(P)]
    chunks = np.array_split(np.arange(sample.size), P)
    rows = np.array_split(np.arange(n), P)
    pix_chunks = np.array_split(np.arange(pix.shape[0]), P)
    times = {}
    ctx = mp.get_context("fork")
    with ctx.Pool(P) as pool:
        t0 = time.perf_counter()
        unc = pool.map(_w_uncollided, chunks)
        times["uncollided"] = (time.perf_counter() - t0) * 64
        t0 = time.perf_counter()
        pool.map(_w_first_collision, [(ch, phi_uc) for ch, phi_uc in unc])
        times["first_collision"] = (time.perf_counter() - t0) * 64
        t0 = time.perf_counter()
        pool.map(_w_moments, rows)
        times["scatter_moments"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        pool.map(_w_sweep, sweep_sets)
        t_sweep = time.perf_counter() - t0
        times["sweep"] = t_sweep * M / (8 * P)
        per_direction = pool.map(_w_image, list(range(P)))  # each timed inside its worker, all P running at once
        times["image"] = float(np.mean(per_direction)) * cone.shape[0] / P
        t0 = time.perf_counter()
        pool.map(_w_primary, pix_chunks)
        times["primary"] = (time.perf_counter() - t0) * 64
    _CPU.clear()
    per_view = sum(times.values())
    sweep_only = times["scatter_moments"] + times["sweep"]
    return {"value": 1.0 / per_view, "unit": "view-sweeps/s", "cores": P, "kind": "port",
            "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "sweep_only_value": 1.0 / sweep_only,
            "sample": f"one view step of the oracle's algorithm (numpy float32/float64) on {P} worker processes of "
                      "one thread each, every stage's sample split over them and timed by wall clock: "
                      + ", ".join(f"{k} {v:.1f} s" for k, v in times.items())
                      + f" = {per_view:.1f} s per view step (uncollided, first collision and primary on 1/64 of the "
                        f"points or pixel rays; scattering moments on every cell; the sweep (source expansion, LD "
                        f"sweep, moment reduction) on {8 * P} of {M} directions x 16 groups x {n}^3 cells; the image "
                        f"(downstream depths and oblique rays on 1/64 of the cells and pixels, the emission on every "
                        f"cell) on {P} of {cone.shape[0]} cone directions, one per worker, each timed in its worker "
                        f"while all run at once; each stage scaled to the whole view)"}


def sweep_launches(grid_cells: int, block: int = 4) -> int:
    """Launches of one view-sweep: K + 1 block planes per octant pair (K = sum of blocks - 1 over the axes), the
    middle plane split in two when K is even, four octant pairs."""
    K = 3 * ((grid_cells + block - 1) // block - 1)
    return 4 * (K + 1 + (1 if K % 2 == 0 else 0))


def profiled_traffic(kernel: str, views: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter summaries of this workload
    (profiles/ld_sweep_*_64cube_s12_p5_g16.txt), or None (also when the summary was taken at another number of views
    per launch).

    A summary lists FETCH_SIZE and WRITE_SIZE per dispatch scaled by the same counters' reading of a plain copy of
    known size (its `traffic_bytes_per_launch` line)."""
    folder = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")
    try:
        names = sorted(f for f in os.listdir(folder) if f.startswith("ld_sweep_") and f.endswith(".txt"))
    except OSError:
        return None
    for name in names:
        with open(os.path.join(folder, name)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch") and kernel in line and f"{views} views" in line:
                    return float(line.split("=")[1].split()[0]), f"profiles/{name}"
    return None


def build_view_solver(args, device, rho_fine):
    """Solver, material fields and view angles of the view-step workload (the grids of ``build_problem``)."""
    from citrus_amd.scatter_solver.config import SolverConfig
    from citrus_amd.scatter_solver.cross_sections import seeded_library
    from citrus_amd.scatter_solver.detector import energy_weighted_response
    from citrus_amd.scatter_solver.geometry import ConeGeometry, Grid
    from citrus_amd.scatter_solver.scatter_solver import ScatterSolver

    n = args.cells
    fine = Grid((4 * n, 4 * n, 4 * n), 0.08 * 64 / n)
    config = SolverConfig(legendre_order=5, sweep_quadrature="level_symmetric", sn_order=12,
                          transport_cell_cm=0.32 * 64 / n, quadrature_point_spacing_cm=0.5, cone_num_polar=16,
                          cone_num_azimuthal=32, direction_chunk=64)
    geometry = ConeGeometry(100.0, 150.0, 256, 256, 0.12, 0.12, use_curved_detector=True)
    solver = ScatterSolver(seeded_library(16, 5, 140.0, 20.0, seed=99), config, geometry,
                           detector=energy_weighted_response(), device=device)
    fields = solver.material_fields(rho_fine, fine)
    angles = [2.0 * math.pi * k / 64 for k in range(args.views)]
    return solver, fields, angles


REPEATS = 3  # timed repeats of each helical step
WARM = 2  # untimed repeats before them


def helical_views(args, device, rho_fine, views: int = 8):
    """A turn of a helical scan (pitch 1.0: table feed per turn = beam width at the isocentre) on a 64-row curved panel
    with a bowtie filter (0-3 cm aluminium-like, quadratic in the channel) and 5% off-focal radiation: the views' CITRUS
    steps on their beam-cropped transport grids against the same steps on the covering grid."""
    import torch

    from citrus_amd.citrus import CitrusState, citrus_cropped_views_step, citrus_views_step
    from citrus_amd.scatter_solver.config import SolverConfig
    from citrus_amd.scatter_solver.cross_sections import seeded_library
    from citrus_amd.scatter_solver.detector import energy_weighted_response
    from citrus_amd.scatter_solver.geometry import ConeGeometry, Grid
    from citrus_amd.scatter_solver.scatter_solver import ScatterSolver
    from citrus_amd.scatter_solver.source import SourceModel
    from citrus_amd.utilities import helical_scan, view_angles

    n = args.cells
    fine = Grid((4 * n, 4 * n, 4 * n), 0.08 * 64 / n)
    library = seeded_library(16, 5, 140.0, 20.0, seed=99)
    base = ConeGeometry(100.0, 150.0, 64, 256, 0.12, 0.12, use_curved_detector=True)
    beam_at_iso = base.num_det_rows * base.delta_det_row / base.magnification
    angles = [float(a) for a in view_angles(views, 0.0, 2 * math.pi)]
    geometry = helical_scan(base, angles, pitch_cm=1.0 * beam_at_iso)
    u = (torch.arange(256, dtype=torch.float64) - 127.5) * 0.12
    al_mu = 0.15 + 5.0 * (30.0 / library.fine_energy.double()) ** 2.5  # cm^2/g, aluminium-like
    source = SourceModel(off_focal_distance_cm=(0.5,), off_focal_fraction=(0.05,),
                         filter_thickness_cm=3.0 * (u / u.abs().max()) ** 2, filter_density=2.7,
                         filter_attenuation_table=al_mu)
    config = SolverConfig(legendre_order=5, sweep_quadrature="level_symmetric", sn_order=12,
                          transport_cell_cm=0.32 * 64 / n, quadrature_point_spacing_cm=0.5, cone_num_polar=16,
                          cone_num_azimuthal=32, direction_chunk=64)
    solver = ScatterSolver(library, config, geometry, detector=energy_weighted_response(), source=source,
                           device=device)
    rho = rho_fine.to(device, torch.float32)
    cones = [solver.cone_quadrature(fine, a) for a in angles]
    fields = solver.material_fields(rho, fine)

    def cropped(states):
        return citrus_cropped_views_step(solver, states, rho, fine, angles)

    def covering(states):
        return citrus_views_step(solver, states, fields, angles)

    out = {}
    for name, step in (("cropped", cropped), ("covering", covering)):
        states = res = None
        torch.cuda.empty_cache()
        states = [CitrusState(cone_quadrature=c) for c in cones]
        res = step(states)
        states = res[1]
        for _ in range(WARM):  # the caching allocator reaches its steady state (fresh device memory is cleared by
            res = step(states)  # the driver, about 20 ms per GB, so a step that allocates new segments is slower)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(REPEATS):  # the same step from the same states: the mean of a few repeats
            res = step(states)
        torch.cuda.synchronize()
        out[name] = {"views_per_s": REPEATS * views / (time.perf_counter() - t0)}
        if name == "cropped":
            cells = [(k1 - k0) // 4 for k0, k1 in res[2]]
            out[name]["transport_slices_per_view"] = cells
    out["cropped_speedup"] = out["cropped"]["views_per_s"] / out["covering"]["views_per_s"]
    # the two-stage schedule on the cropped grids: views 0-3 take their coarse step first, then one step of all
    # eight (0-3 entering the fine stage with prolonged accumulators, 4-7 at the coarse stage) with every view's
    # accumulator update in one persistent launch over the uneven grids
    from dataclasses import replace as dc_replace

    from citrus_amd.multigrid import MultigridSchedule, citrus_multigrid_views_step

    coarse_solver = ScatterSolver(library, dc_replace(config, transport_cell_cm=2 * config.transport_cell_cm),
                                  geometry, detector=energy_weighted_response(), source=source, device=device)
    sched = MultigridSchedule(coarse=coarse_solver, fine=solver, coarse_steps=1)
    half = views // 2
    fresh = [CitrusState(cone_quadrature=c) for c in cones]
    _, early, _ = citrus_multigrid_views_step(sched, fresh[:half], rho, fine, angles[:half], crop=True)
    mixed = early + fresh[half:]
    for _ in range(WARM):  # warm-up of this configuration
        res = citrus_multigrid_views_step(sched, mixed, rho, fine, angles, crop=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(REPEATS):
        res = citrus_multigrid_views_step(sched, mixed, rho, fine, angles, crop=True)
    torch.cuda.synchronize()
    out["multigrid_cropped"] = {
        "views_per_s": REPEATS * views / (time.perf_counter() - t0),
        "grids": [list(s.collided_moments.shape[-3:]) for s in res[1]],
        "stages": [s for s, _ in res[2]],
        "note": "one step of 8 helical views on beam-cropped grids, 4 entering the fine stage (64x64 cells at "
                "0.32 cm, accumulators prolonged from the coarse stage) and 4 at the coarse stage (32x32 at 0.64 cm); "
                "all accumulator updates in one persistent sweep launch; wall clock, mean of 3 repeats of the step"}
    out["note"] = (f"{views} views of one helical turn (pitch 1.0 = {beam_at_iso:.2f} cm per turn), 64 x 256 curved "
                   "panel, bowtie 0-3 cm at 2.7 g/cm^3, 5% off-focal at 0.5 cm; second step of each view (with its "
                   "accumulator), wall clock, mean of 3 repeats; the covering grid is 64 slices")
    return out


def scan_stage(solver, fields, device, views: int, batch: int):
    """One outer iteration's scatter stage over a scan of `views` views (the whole 64-view circle) with the
    accumulators in the compact state store (scaled float16 cell means, resident in HBM): per batch of `batch` views,
    the accumulators from the store, one batched view step, the new accumulators back; scatter, primary and SPR of
    every view. The second stage is timed (the first starts from empty accumulators)."""
    import torch

    from citrus_amd.citrus import CitrusState
    from citrus_amd.state_store import StateStore, citrus_scan_step

    angles = [2.0 * math.pi * k / 64 for k in range(views)]
    ctx = solver.sweep_context(fields.grid_coarse)
    store = StateStore(views, (solver.num_groups, ctx.num_moments, 4) + tuple(fields.grid_coarse.shape), device)
    states = [CitrusState() for _ in angles]
    _, _, _, states = citrus_scan_step(solver, store, states, fields, angles, views_per_batch=batch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, _, ratio, states = citrus_scan_step(solver, store, states, fields, angles, views_per_batch=batch)
    torch.cuda.synchronize()
    seconds = time.perf_counter() - t0
    out = {"views_per_s": views / seconds, "views": views, "views_per_batch": batch,
           "store": store.describe(), "store_seconds_per_stage": store.transfer_seconds / 2,
           "spr_mean": float(ratio.mean()),
           "note": "second scatter stage of a scan: accumulators decoded from and packed into the state store "
                   "(float16 cell means with per-(group, moment) scales) around each batched view step; wall clock"}
    store.close()
    return out


SURROGATE_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "citrus_amd", "weights",
                                 "surrogate_64cube_s12_p5_g16.npz")


def learned_vs_plain(solver, fields, angles, device):
    """Converged scatter images of the step's views from scratch: the plain source iteration (stopping at the
    configured tolerance) against the learned form (two sweeps, the surrogate's Phi_{>2}, one judging sweep at a 1%
    threshold, plain fallback), both batched over the views and timed with the uncollided and first-collision work
    and the signal images; errors against a source iteration converged to 1e-5."""
    import torch

    from citrus_amd.surrogate import UNet3D, collision_orders_views, learned_views, plain_views

    if not os.path.exists(SURROGATE_WEIGHTS):
        return {"error": f"surrogate weights missing ({SURROGATE_WEIGHTS})"}
    net = UNet3D.load(SURROGATE_WEIGHTS, device)
    cones = [solver.cone_quadrature(fields.grid_fine, a) for a in angles]

    def images(moments):
        out = []
        for a, cone, m in zip(angles, cones, moments):
            phi_uc = solver.uncollided_fluence(fields, a)
            T = solver.multiple_collision_source_moments(fields, m)
            out.append(solver.scatter_image(fields, a, phi_uc, T, cone, signal_only=True)["signal"])
        return out

    ref = [o.total for o in collision_orders_views(solver, fields, angles, tolerance=1e-5)]
    ref_images = images(ref)

    def errors(moments, imgs):
        img = [float(torch.linalg.vector_norm(i - r) / torch.linalg.vector_norm(r)) for i, r in zip(imgs, ref_images)]
        flux = [float(torch.linalg.vector_norm((m - x)[:, 0, 0]) / torch.linalg.vector_norm(x[:, 0, 0]))
                for m, x in zip(moments, ref)]
        return max(img), max(flux)

    tol = solver.config.source_iteration_tolerance
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    plain, plain_sweeps = plain_views(solver, fields, angles, tol)
    plain_images = images(plain)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    lv = learned_views(solver, fields, angles, net, threshold=0.01, fallback_tolerance=tol)
    learned_images = images(lv.collided_moments)
    torch.cuda.synchronize()
    t_learned = time.perf_counter() - t0
    plain_img, plain_flux = errors(plain, plain_images)
    learned_img, learned_flux = errors(lv.collided_moments, learned_images)
    V = len(angles)
    first_step = citrus_first_step(solver, fields, angles, net, ref, ref_images)
    return {"citrus_first_step": first_step,
            "plain_views_per_s": V / t_plain, "learned_views_per_s": V / t_learned,
            "learned_speedup": t_plain / t_learned, "plain_sweeps": plain_sweeps, "learned_sweeps": lv.sweeps,
            "learned_accepted": sum(lv.accepted), "judging_residuals": [round(r, 5) for r in lv.residuals],
            "plain_tolerance": tol, "plain_image_error": plain_img, "learned_image_error": learned_img,
            "plain_flux_error": plain_flux, "learned_flux_error": learned_flux,
            "surrogate_parameters": net.num_parameters(),
            "note": "errors: max over views of the relative L2 error of the signal-only scatter image (the SPR "
                    "numerator) and of the scalar-flux cell means, against a source iteration converged to 1e-5"}


def citrus_first_step(solver, fields, angles, net, ref, ref_images) -> dict:
    """The first CITRUS step of the views from empty states, plain (its accumulator is the first collision order)
    against the learned start (``citrus_views_step(learned=net)``: two more batched sweeps and one inference, judged
    at 1%): wall-clock rate and the errors of the step's scatter images and accumulators against the converged
    source iteration, then the relative change the second (plain) step makes."""
    import torch

    from citrus_amd.citrus import CitrusState, citrus_views_step

    out = {}
    for name, kw in (("plain", {}), ("learned", {"learned": net, "learned_threshold": 0.01})):
        citrus_views_step(solver, [CitrusState() for _ in angles], fields, angles, **kw)  # warm the allocator
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        results, states = citrus_views_step(solver, [CitrusState() for _ in angles], fields, angles, **kw)
        torch.cuda.synchronize()
        seconds = time.perf_counter() - t0
        img = max(float(torch.linalg.vector_norm(r[0] - x) / torch.linalg.vector_norm(x))
                  for r, x in zip(results, ref_images))
        flux = max(float(torch.linalg.vector_norm((st.collided_moments - x)[:, 0, 0]) /
                         torch.linalg.vector_norm(x[:, 0, 0])) for st, x in zip(states, ref))
        residuals = [st.learned_residual for st in states]
        _, states = citrus_views_step(solver, states, fields, angles)
        out[name] = {"views_per_s": len(angles) / seconds, "image_error": img, "flux_error": flux,
                     "second_step_change": max(st.history[1] for st in states)}
        if name == "learned":
            out[name]["judging_residuals"] = [round(r, 5) for r in residuals]
        del results, states
        torch.cuda.empty_cache()
    out["note"] = ("first CITRUS step from empty states; errors are max over views against a source iteration "
                   "converged to 1e-5; second_step_change is the relative change of the scalar-flux cell means the "
                   "next plain step makes")
    return out


def mixed_precision_steps(view_step, states, images, V: int, args, stream) -> dict:
    """The same view steps with the mixed-precision source contraction (degrees l <= 1 in float32, l >= 2 as a
    bfloat16 matrix product, float32 accumulation) from empty states: rate, and errors at the last step against the
    float32 steps (which the oracle tests pin to the FP64 restatement at 1e-5)."""
    import torch

    from citrus_amd.citrus import CitrusState

    states16 = [CitrusState() for _ in range(V)]
    for _ in range(max(1, args.warmup)):
        view_step("bf16", states16)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(args.steps):
        images16 = view_step("bf16", states16)
    stop.record(stream)
    torch.cuda.synchronize()
    ms16 = start.elapsed_time(stop)
    rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    return {"bf16": {
        "view_steps_per_s": V * args.steps / (ms16 * 1e-3), "ms_per_step": ms16 / args.steps,
        "image_error_vs_fp32": max(rel(a[0], b[0]) for a, b in zip(images16, images)),
        "spr_error_vs_fp32": max(rel(a[0] / a[1], b[0] / b[1]) for a, b in zip(images16, images)),
        "accumulator_error_vs_fp32": max(rel(a.collided_moments, b.collided_moments)
                                         for a, b in zip(states16, states)),
        "note": "the step's scattering-source expansion in mixed precision: degrees l <= 1 in float32, l >= 2 as a "
                "bfloat16 matrix product with float32 accumulation, the operand written by the scattering-moment "
                "kernel; errors are relative L2 against the float32 steps from the same empty states after "
                f"{max(1, args.warmup) + args.steps} steps"}}


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): the per-view "
              "transport step is sharded by views across GPUs by design but this build measures one GPU.",
              file=sys.stderr)
        sys.exit(2)
    if args.profile:
        args.no_cpu_baseline = args.no_learned = args.no_helical = True
    cpu = None
    if not args.no_cpu_baseline:  # on the host's cores, before anything touches the GPU (workers are forked)
        cpu = cpu_baseline_parallel(args)
    import torch

    from citrus_amd.citrus import CitrusState, citrus_views_step
    from citrus_amd.scatter_solver.geometry import Grid
    from citrus_amd.scatter_solver.source_iteration import scatter_moments_views
    from citrus_amd.scatter_solver.transport_sweep import sweep_views
    from citrus_amd.utilities import ellipsoid_phantom

    device = torch.device("cuda:0")
    V = args.views
    stream = torch.cuda.current_stream(device)
    n = args.cells
    rho_fine = ellipsoid_phantom(Grid((4 * n, 4 * n, 4 * n), 0.08 * 64 / n), seed=1234)

    # ---- launch shape: the autotuner's direction chunk for V views of this grid (measured, outside the timing) --
    from citrus_amd.scatter_solver.autotune import default_tuner

    ctx_t, mu_t_t = build_problem(argparse.Namespace(**{**vars(args), "views": 1}), device)[:2]
    if args.schedule != "auto" and args.chunk > 0:
        schedule, chunk, tuned = args.schedule, min(args.chunk, ctx_t.directions_per_octant), {}
    else:
        schedule, chunk = default_tuner().best_config(ctx_t, mu_t_t, V)
        tuned = {f"{s}:{c}": round(v, 3) for (s, c), v in
                 default_tuner().timings[("config",) + default_tuner().key(ctx_t, 16, V, "fp32")].items()}
        if schedule == "persistent":  # resident waves per CU at the chosen chunk (0: the kernel's occupancy)
            default_tuner().best_waves(ctx_t, mu_t_t, V, chunk)
            tuned.update({f"persistent:{chunk}:waves{w}": round(v, 3) for w, v in
                          default_tuner().timings[("waves", chunk) + default_tuner().key(ctx_t, 16, V, "fp32")]
                          .items()})
    del ctx_t, mu_t_t

    # ---- CITRUS view steps (the metric) --------------------------------------------------------------------
    view = None
    if not args.sweep_only:
        solver, fields, angles = build_view_solver(args, device, rho_fine.to(device, torch.float32))
        states = [CitrusState() for _ in range(V)]

        def view_step(contraction="fp32", st=states):
            results, st[:] = citrus_views_step(solver, st, fields, angles, contraction=contraction,
                                               direction_chunk=chunk, schedule=schedule)
            return results

        for _ in range(max(1, args.warmup)):  # the first step of a view has no accumulator yet
            view_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(args.steps):
            images = view_step()
        stop.record(stream)
        torch.cuda.synchronize()
        view = {"wall": time.perf_counter() - t0, "ms": start.elapsed_time(stop),
                # HBM of the view steps: the peak within a step, and what stays allocated between steps (states,
                # material fields, solver tables and the last step's images)
                "memory_gib": {"step_peak": torch.cuda.max_memory_allocated(device) / 2**30,
                               "between_steps": torch.cuda.memory_allocated(device) / 2**30}}
        scatter, primary = images[0]
        spr = float((scatter.sum() / primary.sum()).item())
        assert all(bool(torch.isfinite(s.collided_moments).all()) for s in states), "non-finite accumulator"
        view["spr"] = spr
        view["history"] = [round(h, 6) for h in states[0].history]
        # the same steps with the mixed-precision source contraction: rate and error against the float32 steps
        if not args.profile:
            view.update(mixed_precision_steps(view_step, states, images, V, args, stream))
        del states, images, view_step  # the closure holds the states: without it their 6 GB per view stay allocated
        torch.cuda.empty_cache()
        if not args.profile:
            view["scan_stage"] = scan_stage(solver, fields, device, views=64, batch=V)
            torch.cuda.empty_cache()
        if not args.no_learned:
            try:
                view["learned"] = learned_vs_plain(solver, fields, angles, device)
            except Exception as e:  # reported, not fatal: the view-step rate above is the metric
                view["learned"] = {"error": f"{type(e).__name__}: {e}"}
        del solver, fields
        torch.cuda.empty_cache()
        if not args.no_helical:
            try:
                view["helical"] = helical_views(args, device, rho_fine)
            except Exception as e:  # reported, not fatal
                view["helical"] = {"error": f"{type(e).__name__}: {e}"}
            torch.cuda.empty_cache()

    # ---- sweep only: scattering moments + one fused sweep per view, V views per launch sequence -------------
    args_sweep = argparse.Namespace(**vars(args))
    ctx, mu_t, transfer, phi_a, library, rho, coarse = build_problem(args_sweep, device)
    phi_b = torch.empty_like(phi_a)
    T = torch.empty_like(phi_a)
    for _ in range(1):
        scatter_moments_views(transfer, phi_a, out=T)
        sweep_views(ctx, mu_t, T, out=phi_b, direction_chunk=chunk, schedule=schedule)
        phi_a, phi_b = phi_b, phi_a
    torch.cuda.synchronize()
    sweep_steps = 2 if view is not None else args.steps
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(sweep_steps)]
    s_start, s_stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s_start.record(stream)
    for s in range(sweep_steps):
        scatter_moments_views(transfer, phi_a, out=T)
        ev[s][0].record(stream)
        sweep_views(ctx, mu_t, T, out=phi_b, direction_chunk=chunk, schedule=schedule)
        ev[s][1].record(stream)
        phi_a, phi_b = phi_b, phi_a
    s_stop.record(stream)
    torch.cuda.synchronize()
    sweep_wall = time.perf_counter() - t0
    sweep_total_ms = s_start.elapsed_time(s_stop)
    sweep_ms = sum(a.elapsed_time(b) for a, b in ev) / sweep_steps  # per step = V view-sweeps
    assert bool(torch.isfinite(phi_a).all()), "non-finite moments"
    # the same batched sweep captured once in a HIP graph and replayed
    from citrus_amd.scatter_solver.transport_sweep import SweepGraph

    graph_ms = None
    if not args.profile:
        graph = SweepGraph(ctx, mu_t, T, direction_chunk=chunk, schedule=schedule)
        torch.cuda.synchronize()
        g_start, g_stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        g_start.record(stream)
        for _ in range(sweep_steps):
            graph.replay()
        g_stop.record(stream)
        torch.cuda.synchronize()
        graph_ms = g_start.elapsed_time(g_stop) / sweep_steps
        del graph
    G, nsh = 16, ctx.num_moments
    N = ctx.num_cells
    # compulsory bytes of one view-sweep in the sweep kernel: moments T in and phi out (float32) plus mu_t. The
    # launch sequence splits a batched sweep into block-plane launches, the persistent schedule runs one launch per
    # direction chunk for all V views; either way bytes per launch / launch time averages to the same ratio
    alg_bytes = 2 * G * N * nsh * 4 * 4 + G * N * 4
    # algorithmic matrix-core work of one view-sweep: the source expansion S = T Y^T and the moment reduction
    # phi = psi Yw, each 2 flops x 4 LD components x nsh moments x M directions per cell and group (float32)
    alg_flops = 2 * 2 * 4 * nsh * ctx.num_directions * G * N
    n_chunks = -(-ctx.directions_per_octant // chunk)
    if schedule == "persistent":
        launches = n_chunks  # per batched sweep of V views
        kernel = f"ld_sweep_persistent_kernel<{nsh}, {(chunk + 3) // 4}>"
    else:
        launches = sweep_launches(coarse.shape[0]) * n_chunks * 1  # per batched sweep (V views per launch)
        kernel = f"ld_sweep_mfma_kernel<{nsh}, {(chunk + 3) // 4}>"
    achieved = alg_bytes * V / (sweep_ms * 1e-3) / 1e9
    achieved_tflops = alg_flops * V / (sweep_ms * 1e-3) / 1e12
    traffic, traffic_source = profiled_traffic(kernel, V) or (None, None)
    sweep_rate = V * sweep_steps / (sweep_total_ms * 1e-3)
    if view is not None:
        value, ms_per_step, wall = V * args.steps / (view["ms"] * 1e-3), view["ms"] / args.steps, view["wall"]
        workload = ("CITRUS view step: uncollided trace + corrected first-collision source + one fused sweep of the "
                    "summed sources (S12, P5) + signal-only collapsed-cone image (16x32 cone directions, 256x256 "
                    "curved panel) + primary")
    else:
        value, ms_per_step, wall = sweep_rate, sweep_total_ms / sweep_steps, sweep_wall
        workload = "sweep only: HIP scattering moments + fused LD sweep (S12, P5)"
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "view-sweeps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 256^3 ellipsoid phantom, seeded 2-material P5 library)",
        "config": {"workload": workload,
                   "grid": f"{coarse.shape[0]}^3 cells at {coarse.delta_voxel:.2f} cm", "groups": G,
                   "directions": ctx.num_directions, "moments": nsh, "views_per_step": V,
                   "sweep_only_view_sweeps_per_s": sweep_rate, "sweep_ms_per_view": sweep_ms / V,
                   "sweep_graph_ms_per_view": None if graph_ms is None else graph_ms / V,
                   "schedule": schedule, "direction_chunk": chunk, "tuner_ms_per_schedule_chunk": tuned,
                   "wall_s": wall, "peak_allocated_gib": torch.cuda.max_memory_allocated(device) / 2**30,
                   **({"spr_view0": view["spr"], "accumulator_change_view0": view["history"],
                       "view_step_memory_gib": view["memory_gib"]} if view else {}),
                   **({"learned_vs_plain": view["learned"]} if view and "learned" in view else {}),
                   **({"bf16_contraction": view["bf16"]} if view and "bf16" in view else {}),
                   **({"scan_stage": view["scan_stage"]} if view and "scan_stage" in view else {}),
                   **({"helical_cropped": view["helical"]} if view and "helical" in view else {})},
        # the binding roof of the sweep is the float32 matrix cores: its algorithmic flops at the fp32 MFMA peak take
        # alg_flops / 157.3 TFLOP/s, 4.2x its algorithmic bytes at 8 TB/s; both are reported
        "roofline": {"bound": "mfma", "achieved": achieved_tflops, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP32_MFMA_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": kernel, "alg_flops_per_view_sweep": alg_flops,
                     "alg_flops_per_launch": alg_flops * V / launches,
                     "avg_launch_ms": sweep_ms / launches, "launches_per_batched_sweep": launches,
                     "views_per_batched_sweep": V,
                     "roof_ms_per_view_sweep": {"mfma": alg_flops / FP32_MFMA_PEAK_TFLOPS / 1e9,
                                                "hbm": alg_bytes / HBM_PEAK_GBS / 1e6},
                     "hbm": {"achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": achieved / HBM_PEAK_GBS, "alg_bytes_per_view_sweep": alg_bytes,
                             "alg_bytes_per_launch": alg_bytes * V / launches,
                             "traffic_rate_gbs": (traffic / (sweep_ms / launches * 1e-3) / 1e9) if traffic else None,
                             "traffic_frac": ((traffic / (sweep_ms / launches * 1e-3) / 1e9 / HBM_PEAK_GBS)
                                              if traffic else None)},
                     "traffic_source": traffic_source},
    }
    if cpu is not None:
        if view is None:  # sweep-only unit: the scattering moments and the sweep of the same run
            cpu = {**cpu, "value": cpu["sweep_only_value"]}
        result["cpu_baseline"] = cpu
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
