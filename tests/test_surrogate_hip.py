"""The Phi_{>2} surrogate on the HIP kernels (GPU): every kernel and the whole network's forward and backward against
the FP64 oracle (oracle/surrogate.py) with the same bfloat16-rounded weights, training, and what the trained network
gains over the method without it, with the judging sweep and its fallback. FP32 bar ``F32_REL``."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from citrus_amd import _hip
from citrus_amd.surrogate import (LAYERS, UNet3D, collision_orders, features, learned_collided_moments, mse_gradient,
                                  p1_means, sample_scale, train, training_pair)
from oracle import surrogate as osg
from tests.conftest import gpu_device
from tests.helpers import F32_REL, rel_err

pytestmark = pytest.mark.gpu


def _bf16(rng, shape, scale=0.3):
    w = (rng.standard_normal(shape) * scale).astype(np.float32)
    return torch.from_numpy(w).to(torch.bfloat16), osg.bf16_round(w)


def _stream(dev):
    return _hip.stream_of(dev)


@pytest.mark.parametrize("k", [3, 1])
def test_conv_kernels_match_oracle(k):
    dev = gpu_device()
    lib = _hip.library()
    rng = np.random.default_rng(11)
    B, Ci, Co, D, H, W = 2, 5, 20, 4, 6, 10  # 20 output channels: one full block of 16 and a partial one
    x = rng.standard_normal((B, Ci, D, H, W))
    w16, w64 = _bf16(rng, (Co, Ci, k ** 3))
    b = rng.standard_normal(Co).astype(np.float32)
    xd = torch.from_numpy(x.astype(np.float32)).to(dev)
    wd, bd = w16.to(dev), torch.from_numpy(b).to(dev)
    y = torch.empty(B, Co, D, H, W, device=dev)
    with torch.cuda.device(dev):
        _hip.check(lib.citrus_nn_conv3d_forward(xd.data_ptr(), Ci * D * H * W, B, Ci, D, H, W, wd.data_ptr(),
                                                bd.data_ptr(), Co, k, 1, y.data_ptr(), Co * D * H * W,
                                                _stream(dev)), "forward")
    x32 = x.astype(np.float32).astype(np.float64)
    pre = osg.conv3d(x32, w64.reshape(Co, Ci, k, k, k), b.astype(np.float64))
    assert rel_err(y, osg.leaky(pre)) < F32_REL
    g = rng.standard_normal((B, Co, D, H, W)).astype(np.float32)
    gd = torch.from_numpy(g).to(dev)
    dx = torch.empty(B, Ci, D, H, W, device=dev)
    dw = torch.empty(Co, Ci, k ** 3, device=dev)
    db = torch.empty(Co, device=dev)
    ws = torch.empty(int(lib.citrus_nn_wgrad_workspace_floats(B, Ci, Co, D, H, W, k)), device=dev)
    with torch.cuda.device(dev):
        _hip.check(lib.citrus_nn_conv3d_backward_data(gd.data_ptr(), Co * D * H * W, B, Co, D, H, W, wd.data_ptr(),
                                                      Ci, k, dx.data_ptr(), Ci * D * H * W, _stream(dev)), "data")
        _hip.check(lib.citrus_nn_conv3d_backward_weight(xd.data_ptr(), Ci * D * H * W, gd.data_ptr(), Co * D * H * W,
                                                        B, Ci, Co, D, H, W, k, dw.data_ptr(), db.data_ptr(),
                                                        ws.data_ptr(), _stream(dev)), "weight")
    rdx, rdw, rdb = osg.conv3d_backward(x32, w64.reshape(Co, Ci, k, k, k), g.astype(np.float64))
    assert rel_err(dx, rdx) < F32_REL
    assert rel_err(dw, rdw.reshape(Co, Ci, -1)) < F32_REL
    assert rel_err(db, rdb) < F32_REL


def test_pool_upsample_activation_and_adam_match_oracle():
    dev = gpu_device()
    lib = _hip.library()
    st = _stream(dev)
    rng = np.random.default_rng(12)
    B, C, D, H, W = 2, 3, 4, 6, 8
    x = rng.standard_normal((B, C, D, H, W)).astype(np.float32)
    xd = torch.from_numpy(x).to(dev)
    p = torch.empty(B, C, D // 2, H // 2, W // 2, device=dev)
    u = torch.empty(B, C, 2 * D, 2 * H, 2 * W, device=dev)
    n = C * D * H * W
    with torch.cuda.device(dev):
        _hip.check(lib.citrus_nn_avg_pool2(xd.data_ptr(), n, B, C, D, H, W, p.data_ptr(), n // 8, st), "pool")
        _hip.check(lib.citrus_nn_upsample2(xd.data_ptr(), n, B, C, D, H, W, u.data_ptr(), 8 * n, st), "up")
    assert rel_err(p, osg.avg_pool2(x.astype(np.float64))) < F32_REL
    assert torch.equal(u.cpu(), torch.from_numpy(osg.upsample2(x)))
    gp = rng.standard_normal((B, C, D // 2, H // 2, W // 2)).astype(np.float32)
    skip = rng.standard_normal((B, C, D, H, W)).astype(np.float32)
    dx = torch.from_numpy(skip.copy()).to(dev)
    gu = rng.standard_normal((B, C, 2 * D, 2 * H, 2 * W)).astype(np.float32)
    du = torch.empty(B, C, D, H, W, device=dev)
    with torch.cuda.device(dev):
        _hip.check(lib.citrus_nn_avg_pool2_backward(torch.from_numpy(gp).to(dev).data_ptr(), n // 8, B, C, D, H, W,
                                                    dx.data_ptr(), n, 1, st), "pool back")
        _hip.check(lib.citrus_nn_upsample2_backward(torch.from_numpy(gu).to(dev).data_ptr(), 8 * n, B, C, D, H, W,
                                                    du.data_ptr(), n, st), "up back")
    assert rel_err(dx, osg.avg_pool2_backward(gp.astype(np.float64)) + skip) < F32_REL
    assert rel_err(du, osg.upsample2_backward(gu.astype(np.float64))) < F32_REL
    # activation gradient from the output
    y = rng.standard_normal((B, n)).astype(np.float32)
    g = rng.standard_normal((B, n)).astype(np.float32)
    gd = torch.from_numpy(g).to(dev)
    with torch.cuda.device(dev):
        _hip.check(lib.citrus_nn_leaky_backward(gd.data_ptr(), n, torch.from_numpy(y).to(dev).data_ptr(), n, n, B, st),
                   "leaky")
    slope = osg.leaky_grad_from_output(y).astype(np.float32)  # 1 or 0.01 rounded to float32, as the kernel has it
    assert torch.equal(gd.cpu(), torch.from_numpy(g * slope))
    # Adam, two steps, and the bfloat16 copy
    p0 = rng.standard_normal(1000).astype(np.float32)
    pd, md, vd = torch.from_numpy(p0.copy()).to(dev), torch.zeros(1000, device=dev), torch.zeros(1000, device=dev)
    p16 = torch.empty(1000, device=dev, dtype=torch.bfloat16)
    pr, mr, vr = p0.astype(np.float64), np.zeros(1000), np.zeros(1000)
    for step in (1, 2):
        gr = rng.standard_normal(1000).astype(np.float32)
        with torch.cuda.device(dev):
            _hip.check(lib.citrus_nn_adam(pd.data_ptr(), torch.from_numpy(gr).to(dev).data_ptr(), md.data_ptr(),
                                          vd.data_ptr(), p16.data_ptr(), 1000, 1e-2, 0.9, 0.999, 1e-8, step, st),
                       "adam")
        # the hyperparameters as the kernel holds them (float32: 1 - 0.999f is 0.00099998712 exactly)
        f = lambda a: float(np.float32(a))
        pr, mr, vr = osg.adam(pr, gr.astype(np.float64), mr, vr, step, f(1e-2), f(0.9), f(0.999), f(1e-8))
    assert rel_err(pd, pr) < F32_REL and rel_err(md, mr) < F32_REL and rel_err(vd, vr) < F32_REL
    assert np.array_equal(p16.double().cpu().numpy(), osg.bf16_round(pd.cpu().numpy()))


def _oracle_params(net: UNet3D):
    params = {}
    for name in LAYERS:
        w = net.weights16[name].double().cpu().numpy()
        co, ci, kk = w.shape
        k = round(kk ** (1 / 3))
        params[name] = (w.reshape(co, ci, k, k, k), net.biases[name].double().cpu().numpy())
    return params


def test_network_forward_backward_match_oracle():
    dev = gpu_device()
    net = UNet3D.create(3, 2, width=4, device=dev, seed=5)
    gen = torch.Generator().manual_seed(8)
    for name in LAYERS:  # nonzero biases so their gradients and effect are exercised
        net.biases[name] = (torch.randn(net.biases[name].shape, generator=gen) * 0.1).to(dev)
    x = torch.randn(2, 3, 8, 8, 8, generator=gen)
    target = torch.randn(2, 2, 8, 8, 8, generator=gen)
    out, cache = net.forward(x.to(dev), keep=True)
    params = _oracle_params(net)
    ref, acts = osg.forward(params, x.double().numpy())
    assert rel_err(out, ref) < F32_REL
    loss, g = mse_gradient(out, target.to(dev))
    rloss, rg = osg.mse(ref, target.double().numpy())
    assert abs(loss - rloss) < 1e-5 * rloss
    grads = net.backward(cache, g)
    rgrads = osg.backward(params, acts, rg)
    for name in LAYERS:
        dw, db = grads[name]
        assert rel_err(dw, rgrads[name][0].reshape(dw.shape)) < F32_REL, name
        assert rel_err(db, rgrads[name][1]) < F32_REL, name
    assert net.num_parameters() == sum(co * ci * k ** 3 + co for co, ci, k in osg.layer_shapes(3, 2, 4).values())


def test_training_reduces_the_loss_and_weights_round_trip(tmp_path):
    dev = gpu_device()
    gen = torch.Generator().manual_seed(9)
    xs = [torch.randn(2, 8, 8, 8, generator=gen) for _ in range(8)]
    # target: a smooth linear functional of the input (a local average of the channels)
    ys = [torch.nn.functional.avg_pool3d(x[None], 3, 1, 1, count_include_pad=False)[0].sum(0, keepdim=True)
          for x in xs]
    net = UNet3D.create(2, 1, width=4, device=dev, seed=1)
    hist = train(net, [x.to(dev) for x in xs], [y.to(dev) for y in ys], epochs=40, lr=3e-3)
    assert hist[-1] < 0.1 * hist[0], hist
    path = str(tmp_path / "net.npz")
    net.save(path)
    again = UNet3D.load(path, dev)
    x = xs[0][None].to(dev)
    assert torch.equal(again.forward(x), net.forward(x))


def _small_solver(dev):
    from citrus_amd.scatter_solver.config import SolverConfig
    from citrus_amd.scatter_solver.cross_sections import seeded_library
    from citrus_amd.scatter_solver.geometry import ConeGeometry
    from citrus_amd.scatter_solver.scatter_solver import ScatterSolver

    library = seeded_library(4, 3, 120.0, 20.0, seed=31, num_eta=129)
    config = SolverConfig(legendre_order=3, sweep_quadrature="level_symmetric", sn_order=4, transport_cell_cm=1.0,
                          quadrature_point_spacing_cm=1.0, max_source_iterations=60,
                          source_iteration_tolerance=1e-4)
    return ScatterSolver(library, config, ConeGeometry(30.0, 50.0, 16, 16, 0.5, 0.5), device=dev)


@pytest.fixture(scope="module")
def trained_small():
    """(solver, fields_of(seed), net, loss history): the small solver's network trained on the collision orders of 12
    seeded 16^3 phantoms (held-out seeds 500+ are not among them)."""
    from citrus_amd.scatter_solver.geometry import Grid
    from citrus_amd.utilities import ellipsoid_phantom

    dev = gpu_device()
    solver = _small_solver(dev)
    grid = Grid((16, 16, 16), 1.0)

    def fields_of(seed):
        return solver.material_fields(ellipsoid_phantom(grid, seed=seed).to(dev), grid)

    xs, ys = [], []
    for seed in range(12):
        x, y = training_pair(solver, fields_of(100 + seed), 0.3 * seed)
        xs.append(x)
        ys.append(y)
    G = solver.num_groups
    net = UNet3D.create(2 + 4 * G, 4 * G, width=8, device=dev, seed=3)
    hist = train(net, xs, ys, epochs=60, lr=2e-3)
    return solver, fields_of, net, hist


def test_trained_surrogate_gains_over_two_orders_and_falls_back(trained_small):
    """Trained on the solver's own collision orders of seeded phantoms, the network's Phi_{>2} brings phi1 + phi2
    much closer to the converged moments on held-out phantoms; the judging sweep accepts or, at threshold 0, falls
    back to the plain source iteration."""
    solver, fields_of, net, hist = trained_small
    assert hist[-1] < 0.2 * hist[0], hist
    gains = []
    for seed in (500, 501, 502):
        fields = fields_of(seed)
        orders = collision_orders(solver, fields, 0.7)
        s = sample_scale(orders.phi2)
        pred = net.forward(features(fields.rho_coarse, orders.phi2, s)[None])[0] * s
        exact = p1_means(orders.total)
        two = p1_means(orders.phi1 + orders.phi2)
        err_two = float(torch.linalg.vector_norm(exact - two) / torch.linalg.vector_norm(exact))
        err_net = float(torch.linalg.vector_norm(exact - two - pred) / torch.linalg.vector_norm(exact))
        gains.append(err_two / err_net)
    assert min(gains) > 2.0, gains
    fields = fields_of(501)
    res = learned_collided_moments(solver, fields, 0.7, net, threshold=0.05)
    orders = collision_orders(solver, fields, 0.7)
    assert res.sweeps == 3 and res.accepted and res.residual < 0.05
    assert rel_err(res.collided_moments[:, 0, 0], orders.total[:, 0, 0]) < 0.05
    forced = learned_collided_moments(solver, fields, 0.7, net, threshold=0.0, fallback_tolerance=1e-4)
    assert not forced.accepted and forced.sweeps > 3
    assert rel_err(forced.collided_moments[:, 0, 0], orders.total[:, 0, 0]) < 1e-3


def test_batched_views_equal_single_views():
    """Collision orders and the learned form of several views in batched sweeps equal the per-view functions."""
    from citrus_amd.scatter_solver.geometry import Grid
    from citrus_amd.surrogate import collision_orders_views, learned_views, plain_views
    from citrus_amd.utilities import ellipsoid_phantom

    dev = gpu_device()
    solver = _small_solver(dev)
    grid = Grid((16, 16, 16), 1.0)
    fields = solver.material_fields(ellipsoid_phantom(grid, seed=77).to(dev), grid)
    angles = [0.2, 2.5]
    batched = collision_orders_views(solver, fields, angles, tolerance=1e-4)
    for o, a in zip(batched, angles):
        single = collision_orders(solver, fields, a, tolerance=1e-4)
        assert rel_err(o.phi1, single.phi1) < F32_REL and rel_err(o.phi2, single.phi2) < F32_REL
        assert rel_err(o.total, single.total) < 1e-4  # the batch stops when its slowest view has converged
    totals, sweeps = plain_views(solver, fields, angles, 1e-4)
    assert sweeps == batched[0].sweeps and all(torch.equal(t, o.total) for t, o in zip(totals, batched))
    G = solver.num_groups
    net = UNet3D.create(2 + 4 * G, 4 * G, width=4, device=dev, seed=2)
    lv = learned_views(solver, fields, angles, net, threshold=1.0)
    assert lv.accepted == [True, True] and lv.sweeps == 3
    for v, a in enumerate(angles):
        single = learned_collided_moments(solver, fields, a, net, threshold=1.0)
        assert rel_err(lv.collided_moments[v], single.collided_moments) < F32_REL
        assert abs(lv.residuals[v] - single.residual) < 1e-5
    forced = learned_views(solver, fields, angles, net, threshold=0.0, fallback_tolerance=1e-4)
    assert forced.accepted == [False, False] and forced.sweeps == 3 + batched[0].sweeps
    assert all(torch.equal(m, o.total) for m, o in zip(forced.collided_moments, batched))


def test_learned_start_of_the_citrus_step(trained_small):
    """The learned form as the first CITRUS step of views with empty states: the step's sweep is phi^(1), two more
    batched sweeps and one inference give the judged estimate, so the first accumulator of a held-out phantom is
    within a few percent of the converged collided moments where the plain first step (phi^(1) alone) is far off; the
    next step changes it little. At threshold 0 the judge rejects and the views keep the plain first step bit for
    bit; the residual is recorded either way."""
    from citrus_amd.citrus import CitrusState, citrus_views_step

    solver, fields_of, net, _ = trained_small
    fields = fields_of(501)
    angles = [0.7, 2.1]
    print("learned start: 16^3, 4 groups, S4, P3, two views", flush=True)
    _, plain = citrus_views_step(solver, [CitrusState() for _ in angles], fields, angles)
    _, learned = citrus_views_step(solver, [CitrusState() for _ in angles], fields, angles, learned=net,
                                   learned_threshold=0.05)
    _, forced = citrus_views_step(solver, [CitrusState() for _ in angles], fields, angles, learned=net,
                                  learned_threshold=0.0)
    for v, a in enumerate(angles):
        total = collision_orders(solver, fields, a).total
        err_plain = rel_err(plain[v].collided_moments[:, 0, 0], total[:, 0, 0])
        err_learned = rel_err(learned[v].collided_moments[:, 0, 0], total[:, 0, 0])
        assert learned[v].learned_residual < 0.05 and err_learned < 0.05, (learned[v].learned_residual, err_learned)
        assert err_plain > 5.0 * err_learned, (err_plain, err_learned)
        assert plain[v].learned_residual is None and forced[v].learned_residual > 0.0
        assert torch.equal(forced[v].collided_moments, plain[v].collided_moments)
    _, after_learned = citrus_views_step(solver, learned, fields, angles)
    _, after_plain = citrus_views_step(solver, plain, fields, angles)
    for v in range(len(angles)):
        assert after_learned[v].history[1] < 0.05 < after_plain[v].history[1], \
            (after_learned[v].history, after_plain[v].history)
