"""Compact per-view accumulators on the GPU (the reference's StateStore, citrus/citrus.py:623-757): the HIP packing
equals the oracle's float16 format bit for bit, the store round-trips through GPU and pinned host slots, and the
scan-level step carries every view's accumulator through the store."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from citrus_amd.state_store import (StateStore, citrus_scan_step, pack_accumulators, scatter_corrected_counts,
                                    unpack_accumulators)
from oracle import state_store as oss
from tests.conftest import gpu_device

pytestmark = pytest.mark.gpu


def _moments(dev, V=2, G=3, shape=(5, 6, 7), nsh=16, seed=3):
    gen = torch.Generator().manual_seed(seed)
    m = torch.randn(V, G, *shape, nsh, 4, generator=gen)
    m *= torch.tensor([1e-6, 1.0, 3e3])[None, :G, None, None, None, None, None]
    m[0, 1, ..., 5, 0] = 0.0  # an all-zero (group, moment): scale 1
    m[1, 2, 1, 2, 3, 7, 0] = 2e-5  # far below its moment's maximum: a float16 subnormal
    return m.to(dev).contiguous()


def test_pack_unpack_equal_the_oracle_bit_for_bit():
    dev = gpu_device()
    m = _moments(dev)
    q, scale = pack_accumulators(m)
    back = unpack_accumulators(q, scale)
    torch.cuda.synchronize()
    for v in range(m.shape[0]):
        cm = m[v, ..., 0].permute(0, 4, 1, 2, 3).cpu().numpy()  # [G, nsh, R, C, S]
        wq, ws = oss.pack(cm)
        assert np.array_equal(q[v].cpu().numpy().view(np.uint16), wq.view(np.uint16))
        assert np.array_equal(scale[v].cpu().numpy(), ws)
        want = oss.unpack(wq, ws)
        got = back[v].cpu().numpy()
        assert np.array_equal(got[..., 0].transpose(0, 4, 1, 2, 3), want)
        assert not got[..., 1:].any()


@pytest.mark.parametrize("location", ["gpu", "cpu"])
def test_store_round_trip(location):
    dev = gpu_device()
    m = _moments(dev, V=3)
    G, R, C, S, nsh = 3, 5, 6, 7, 16
    store = StateStore(4, (G, nsh, 4, R, C, S), dev, location=location)
    assert store.location == location and store.get(2) is None
    store.put_views([0, 2], m[:2])  # batched group-major moments
    store.put(3, m[2].permute(0, 4, 5, 1, 2, 3))  # one accumulator [G, nsh, 4, R, C, S]
    for slot, v in ((0, 0), (2, 1), (3, 2)):
        got = store.get(slot)
        assert tuple(got.shape) == (G, nsh, 4, R, C, S) and got.device == dev
        cm = m[v, ..., 0].permute(0, 4, 1, 2, 3).cpu().numpy()
        want = oss.unpack(*oss.pack(cm))
        assert np.array_equal(got[:, :, 0].cpu().numpy(), want)
        assert not got[:, :, 1:].any()
    assert store.get(1) is None
    d = store.describe()
    assert d["stored_shape"] == [G, nsh, 1, R, C, S] and d["bytes_per_view"] == G * nsh * R * C * S * 2 + G * nsh * 4
    assert store.transfer_seconds > 0
    with pytest.raises(IndexError):
        store.get(4)
    with pytest.raises(ValueError):
        store.put(0, m[0])
    store.close()
    assert store.get(0) is None


def test_scan_step_carries_accumulators_through_the_store():
    """Five views in batches of two: the stored accumulators are the packed results of the batched step, and the next
    stage starts from their decoded form; the ratio is scatter / primary; Y / (1 + SPR)."""
    from dataclasses import replace

    from citrus_amd.citrus import CitrusState, citrus_views_step
    from tests.test_citrus_hip import _solver

    dev = gpu_device()
    solver, grid, rho = _solver(dev, 1.0)
    fields = solver.material_fields(rho, grid)
    angles = [0.2, 1.1, 2.3, 3.9, 5.0]
    ctx = solver.sweep_context(fields.grid_coarse)
    G, nsh = solver.num_groups, ctx.num_moments
    store = StateStore(len(angles), (G, nsh, 4) + tuple(fields.grid_coarse.shape), dev)
    states = [CitrusState() for _ in angles]
    scatter, primary, ratio, states = citrus_scan_step(solver, store, states, fields, angles, views_per_batch=2)
    assert tuple(scatter.shape) == (5,) + tuple(primary.shape[1:]) and all(s.collided_moments is None for s in states)
    assert torch.equal(ratio, scatter / primary.clamp_min(torch.finfo(torch.float32).tiny))
    # the stored accumulators are the batch steps' results, packed
    ref_res, ref_states = citrus_views_step(solver, [CitrusState() for _ in angles[:2]], fields, angles[:2])
    for v in range(2):
        assert torch.equal(scatter[v], ref_res[v][0])
        cm = ref_states[v].collided_moments[:, :, 0].cpu().numpy()
        assert np.array_equal(store.get(v)[:, :, 0].cpu().numpy(), oss.unpack(*oss.pack(cm)))
    # the second stage starts from the decoded accumulators
    decoded = [replace(s, collided_moments=store.get(v)) for v, s in enumerate(states)]
    scatter2, _, _, states2 = citrus_scan_step(solver, store, states, fields, angles, views_per_batch=2)
    want, want_states = citrus_views_step(solver, decoded[2:4], fields, angles[2:4])
    assert torch.equal(scatter2[2], want[0][0]) and torch.equal(scatter2[3], want[1][0])
    # the stages form the step's moments from the stored means directly: the same change norms as the decoded form
    assert states2[2].history == want_states[0].history and states2[3].history == want_states[1].history
    assert [s.num_steps for s in states2] == [2] * 5
    y = torch.rand_like(primary) + 1.0
    assert torch.equal(scatter_corrected_counts(y, ratio), y / (1.0 + ratio))
    assert scatter_corrected_counts(y, None) is y


@pytest.mark.parametrize("G,shape", [(16, (5, 6, 7)), (5, (4, 4, 3))])
def test_scattering_moments_from_stored_accumulators(G, shape):
    """scatter_moments_stored forms T from the stored float16 cell means and scales (zero slopes) bit for bit as
    scatter_moments_views forms it from their unpacked float32 form; G = 16 downscatter-only takes the view step's
    run form, G = 5 the general one."""
    from citrus_amd.scatter_solver.source_iteration import MaterialTransfer, scatter_moments_stored, \
        scatter_moments_views
    from citrus_amd.state_store import pack_accumulators, unpack_accumulators

    dev = gpu_device()
    gen = torch.Generator().manual_seed(23)
    V, nsh, M, L = 3, 36, 2, 5
    rho = torch.rand(M, *shape, generator=gen).to(dev)
    sigma = (torch.rand(M, G, G, L + 1, generator=gen) * torch.triu(torch.ones(G, G))[None, :, :, None]).to(dev)
    transfer = MaterialTransfer(rho, sigma)
    phi = torch.randn(V, G, *shape, nsh, 4, generator=gen).to(dev)
    q, scale = pack_accumulators(phi)
    want = scatter_moments_views(transfer, unpack_accumulators(q, scale))
    got = scatter_moments_stored(transfer, q.contiguous(), scale.contiguous())
    assert torch.equal(got, want)
    assert not bool(got[..., 1:].any())  # zero slopes in, zero slopes out
