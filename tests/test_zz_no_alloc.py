"""The allocation-free pack ``citrus_state_pack_static`` (no workspace argument; its scratch is a file-scope device
array) called through the C ABI: two calls in a row, the second on a smaller input, each equal bit for bit to
``citrus_state_pack`` with a caller workspace and to the oracle's float16 format; a size beyond the array is
rejected before anything is launched."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from citrus_amd import _hip
from oracle import state_store as oss
from tests.conftest import gpu_device

pytestmark = pytest.mark.gpu

SCRATCH_FLOATS = 1 << 23  # CITRUS_STATE_PACK_SCRATCH_FLOATS


def _moments(dev, V, G, shape, nsh, seed):
    gen = torch.Generator().manual_seed(seed)
    m = torch.randn(V, G, *shape, nsh, 4, generator=gen) * torch.logspace(-3, 3, G)[None, :, None, None, None, None,
                                                                                     None]
    m[0, 0, ..., 1, 0] = 0.0
    return m.to(dev).contiguous()


def _pack(lib, m, static: bool):
    V, G, R, C, S, nsh, _ = (int(v) for v in m.shape)
    N = R * C * S
    scale = torch.full((V, G, nsh), -1.0, device=m.device)
    q = torch.full((V, G, nsh, N), -7.0, device=m.device, dtype=torch.float16)
    stream = _hip.stream_of(m.device)
    with torch.cuda.device(m.device):
        if static:
            rc = lib.citrus_state_pack_static(m.data_ptr(), V, G, nsh, N, scale.data_ptr(), q.data_ptr(), stream)
            _hip.check(rc, "citrus_state_pack_static")
        else:
            ws = torch.empty(int(lib.citrus_state_pack_workspace_floats(V, G, nsh, N)), device=m.device)
            rc = lib.citrus_state_pack(m.data_ptr(), V, G, nsh, N, scale.data_ptr(), q.data_ptr(), ws.data_ptr(),
                                       ws.numel(), stream)
            _hip.check(rc, "citrus_state_pack")
    torch.cuda.synchronize()
    return scale, q


def test_static_pack_equals_the_workspace_pack_on_two_calls():
    dev = gpu_device()
    lib = _hip.library()
    for V, G, shape, nsh, seed in ((3, 4, (17, 9, 33), 16, 1), (1, 2, (5, 6, 7), 9, 2)):  # the second call is smaller
        print(f"no-alloc pack: V={V} G={G} shape={shape} nsh={nsh}", flush=True)
        m = _moments(dev, V, G, shape, nsh, seed)
        s_ref, q_ref = _pack(lib, m, static=False)
        s, q = _pack(lib, m, static=True)
        assert torch.equal(s, s_ref)
        assert torch.equal(q.view(torch.int16), q_ref.view(torch.int16))
        for v in range(V):
            wq, ws = oss.pack(m[v, ..., 0].permute(0, 4, 1, 2, 3).cpu().numpy())
            assert np.array_equal(q[v].cpu().numpy().view(np.uint16), wq.reshape(G, nsh, -1).view(np.uint16))
            assert np.array_equal(s[v].cpu().numpy(), ws.reshape(G, nsh))


def test_static_pack_rejects_sizes_beyond_its_scratch():
    dev = gpu_device()
    lib = _hip.library()
    m = torch.zeros(4, device=dev)
    V, G, nsh = 64, 16, 36
    N = (SCRATCH_FLOATS // (V * G * nsh) + 1) * 256  # one block more per slab than the array holds
    assert int(lib.citrus_state_pack_workspace_floats(V, G, nsh, N)) > SCRATCH_FLOATS
    rc = lib.citrus_state_pack_static(m.data_ptr(), V, G, nsh, N, m.data_ptr(), m.data_ptr(), _hip.stream_of(dev))
    assert rc == 1 and "CITRUS_STATE_PACK_SCRATCH_FLOATS" in lib.citrus_last_error().decode()
