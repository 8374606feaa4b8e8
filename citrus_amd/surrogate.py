"""The Phi_{>2} surrogate: a 3D U-Net on the HIP kernels that predicts the collided moments of collision order three
and higher, checked by one plain sweep.

The surrogate has no counterpart in the reference. Its place in the CITRUS step (``citrus/citrus.py:192-366``) is the
accumulator: the plain method reaches the converged collided moments ``Phi = sum_n phi^(n)`` by repeated sweeps
``Phi <- phi^(1) + K Phi``; the learned form sweeps twice (``phi^(1)``, ``phi^(2) = K phi^(1)``), predicts
``Phi_{>2} = sum_{n>=3} phi^(n)`` from the densities and ``phi^(2)``, and lets one more plain sweep judge the result:
``Phi~ = phi^(1) + K (phi^(1) + phi^(2) + Phi_{>2})``, accepted when its relative change in the scalar-flux cell
means (the quantity the source iteration's stopping rule measures, source_iteration.py:126-204) is at most the
threshold, else the plain source iteration runs instead.

Network (``UNet3D``): two pooling levels, base width c, leaky ReLU (slope 0.01), a 1x1x1 head; inputs are the M
material densities of the transport cells and the P1 cell means of ``phi^(2)`` (4 moments x G groups), optionally
also those of ``phi^(1)``, divided by the sample's scale (RMS of the scalar-flux cell means of ``phi^(2)``); outputs
the P1 cell means of ``Phi_{>2}`` in the same scale (the transport problem is linear in the source, so the scale carries through). The LD slopes and the
moments l >= 2 of ``Phi_{>2}`` are left to the judging sweep, which forms them from the predicted part.

Every layer runs on the library's kernels (``citrus_nn_*``): bfloat16 weights, float32 activations and accumulation;
float32 master weights and Adam moments are updated by ``citrus_nn_adam``, which also re-rounds the bfloat16 copy.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch
from torch import Tensor

from . import _hip

LAYERS = ("e0a", "e0b", "e1a", "e1b", "b0", "b1", "d1a", "d1b", "d0a", "d0b", "head")


def layer_shapes(cin: int, cout: int, c: int) -> dict:
    """{name: (out channels, in channels, kernel size)} of the U-Net."""
    return {
        "e0a": (c, cin, 3), "e0b": (c, c, 3),
        "e1a": (2 * c, c, 3), "e1b": (2 * c, 2 * c, 3),
        "b0": (4 * c, 2 * c, 3), "b1": (4 * c, 4 * c, 3),
        "d1a": (2 * c, 6 * c, 3), "d1b": (2 * c, 2 * c, 3),
        "d0a": (c, 3 * c, 3), "d0b": (c, c, 3),
        "head": (cout, c, 1),
    }


def _stream(device) -> int:
    return _hip.stream_of(device)


def _check(rc: int, what: str) -> None:
    _hip.check(rc, what)


class _Act:
    """A [B, C, D, H, W] activation that may be a channel slice of a larger buffer: (tensor view, batch stride)."""

    def __init__(self, t: Tensor):
        if t.dtype != torch.float32 or t.device.type != "cuda":
            raise ValueError("surrogate activations are float32 GPU tensors")
        if t.dim() != 5 or t.stride(4) != 1 or t.stride(3) != t.shape[4] or t.stride(2) != t.shape[3] * t.shape[4] \
                or t.stride(1) != t.shape[2] * t.shape[3] * t.shape[4]:
            raise ValueError("activation channels must be contiguous within a sample")
        self.t = t

    @property
    def ptr(self) -> int:
        return self.t.data_ptr()

    @property
    def bs(self) -> int:
        return self.t.stride(0)

    @property
    def shape(self):
        return tuple(self.t.shape)


@dataclass
class UNet3D:
    """The U-Net's parameters on one device: float32 master weights, their bfloat16 copies and the Adam moments."""

    cin: int
    cout: int
    width: int
    device: torch.device
    weights: dict = field(default_factory=dict)      # name -> float32 [Co, Ci, k^3]
    biases: dict = field(default_factory=dict)       # name -> float32 [Co]
    weights16: dict = field(default_factory=dict)    # name -> bfloat16 [Co, Ci, k^3]
    moments: dict = field(default_factory=dict)      # name -> (m_w, v_w, m_b, v_b)
    steps: int = 0

    @classmethod
    def create(cls, cin: int, cout: int, width: int = 16, device="cuda", seed: int = 0) -> "UNet3D":
        """He-initialised weights (seeded, on the CPU then copied) and zero biases; the head starts at zero scale."""
        net = cls(cin, cout, width, torch.device(device))
        gen = torch.Generator().manual_seed(seed)
        for name, (co, ci, k) in layer_shapes(cin, cout, width).items():
            fan_in = ci * k ** 3
            std = math.sqrt(2.0 / fan_in) * (0.1 if name == "head" else 1.0)
            w = torch.randn(co, ci, k ** 3, generator=gen, dtype=torch.float32) * std
            net.weights[name] = w.to(net.device).contiguous()
            net.biases[name] = torch.zeros(co, device=net.device, dtype=torch.float32)
        net._round_all()
        for name in LAYERS:
            w, b = net.weights[name], net.biases[name]
            net.moments[name] = (torch.zeros_like(w), torch.zeros_like(w), torch.zeros_like(b), torch.zeros_like(b))
        return net

    def num_parameters(self) -> int:
        return sum(w.numel() + self.biases[n].numel() for n, w in self.weights.items())

    def _round_all(self) -> None:
        lib = _hip.library()
        for name, w in self.weights.items():
            w16 = torch.empty(w.shape, device=self.device, dtype=torch.bfloat16)
            with torch.cuda.device(self.device):
                _check(lib.citrus_nn_to_bf16(w.data_ptr(), w16.data_ptr(), w.numel(), _stream(self.device)),
                       "citrus_nn_to_bf16")
            self.weights16[name] = w16

    # ---- layers ------------------------------------------------------------------------------------------------
    def _conv(self, name: str, x: _Act, out: Optional[Tensor] = None, act: bool = True) -> _Act:
        co, ci, k = layer_shapes(self.cin, self.cout, self.width)[name]
        B, C, D, H, W = x.shape
        if C != ci:
            raise ValueError(f"layer {name} takes {ci} channels, got {C}")
        y = _Act(torch.empty(B, co, D, H, W, device=self.device, dtype=torch.float32) if out is None else out)
        with torch.cuda.device(self.device):
            _check(_hip.library().citrus_nn_conv3d_forward(
                x.ptr, x.bs, B, C, D, H, W, self.weights16[name].data_ptr(), self.biases[name].data_ptr(), co, k,
                int(act), y.ptr, y.bs, _stream(self.device)), "citrus_nn_conv3d_forward")
        return y

    def _pool(self, x: _Act) -> _Act:
        B, C, D, H, W = x.shape
        y = _Act(torch.empty(B, C, D // 2, H // 2, W // 2, device=self.device, dtype=torch.float32))
        with torch.cuda.device(self.device):
            _check(_hip.library().citrus_nn_avg_pool2(x.ptr, x.bs, B, C, D, H, W, y.ptr, y.bs, _stream(self.device)),
                   "citrus_nn_avg_pool2")
        return y

    def _up(self, x: _Act, out: Tensor) -> None:
        B, C, D, H, W = x.shape
        y = _Act(out)
        with torch.cuda.device(self.device):
            _check(_hip.library().citrus_nn_upsample2(x.ptr, x.bs, B, C, D, H, W, y.ptr, y.bs, _stream(self.device)),
                   "citrus_nn_upsample2")

    # ---- forward / backward ------------------------------------------------------------------------------------
    def forward(self, x: Tensor, keep: bool = False):
        """Output [B, cout, D, H, W] of x [B, cin, D, H, W] (D, H, W divisible by 4); with ``keep`` also the
        activations the backward pass needs."""
        if x.dim() != 5 or int(x.shape[1]) != self.cin:
            raise ValueError(f"input must be [B, {self.cin}, D, H, W], got {tuple(x.shape)}")
        B, _, D, H, W = (int(s) for s in x.shape)
        if D % 4 or H % 4 or W % 4:
            raise ValueError(f"grid sizes must be divisible by 4, got {(D, H, W)}")
        c = self.width
        f32 = dict(device=self.device, dtype=torch.float32)
        xa = _Act(x.to(self.device, torch.float32).contiguous())
        a = {"x": xa}
        cat0 = torch.empty(B, 3 * c, D, H, W, **f32)
        cat1 = torch.empty(B, 6 * c, D // 2, H // 2, W // 2, **f32)
        a["cat0"], a["cat1"] = _Act(cat0), _Act(cat1)
        a["e0a"] = self._conv("e0a", xa)
        a["e0b"] = self._conv("e0b", a["e0a"], out=cat0[:, 2 * c:])
        a["p0"] = self._pool(a["e0b"])
        a["e1a"] = self._conv("e1a", a["p0"])
        a["e1b"] = self._conv("e1b", a["e1a"], out=cat1[:, 4 * c:])
        a["p1"] = self._pool(a["e1b"])
        a["b0"] = self._conv("b0", a["p1"])
        a["b1"] = self._conv("b1", a["b0"])
        self._up(a["b1"], cat1[:, :4 * c])
        a["d1a"] = self._conv("d1a", a["cat1"])
        a["d1b"] = self._conv("d1b", a["d1a"])
        self._up(a["d1b"], cat0[:, :2 * c])
        a["d0a"] = self._conv("d0a", a["cat0"])
        a["d0b"] = self._conv("d0b", a["d0a"])
        out = self._conv("head", a["d0b"], act=False)
        return (out.t, a) if keep else out.t

    def backward(self, a: dict, g_out: Tensor) -> dict:
        """Parameter gradients {name: (dw [Co, Ci, k^3], db [Co])} of sum(g_out * output), from ``forward(keep=True)``."""
        lib = _hip.library()
        st = _stream(self.device)
        shapes = layer_shapes(self.cin, self.cout, self.width)
        grads = {}
        c = self.width
        f32 = dict(device=self.device, dtype=torch.float32)

        def mask(g: _Act, y: _Act) -> None:
            B, C, D, H, W = g.shape
            _check(lib.citrus_nn_leaky_backward(g.ptr, g.bs, y.ptr, y.bs, C * D * H * W, B, st),
                   "citrus_nn_leaky_backward")

        def layer(name: str, inp: _Act, g: _Act, need_dx: bool = True, dx_out: Optional[Tensor] = None):
            co, ci, k = shapes[name]
            B, _, D, H, W = g.shape
            dw = torch.empty(co, ci, k ** 3, **f32)
            db = torch.empty(co, **f32)
            ws = torch.empty(int(lib.citrus_nn_wgrad_workspace_floats(B, ci, co, D, H, W, k)), **f32)
            _check(lib.citrus_nn_conv3d_backward_weight(inp.ptr, inp.bs, g.ptr, g.bs, B, ci, co, D, H, W, k,
                                                        dw.data_ptr(), db.data_ptr(), ws.data_ptr(), st),
                   "citrus_nn_conv3d_backward_weight")
            grads[name] = (dw, db)
            if not need_dx:
                return None
            dx = _Act(torch.empty(B, ci, D, H, W, **f32) if dx_out is None else dx_out)
            _check(lib.citrus_nn_conv3d_backward_data(g.ptr, g.bs, B, co, D, H, W, self.weights16[name].data_ptr(),
                                                      ci, k, dx.ptr, dx.bs, st), "citrus_nn_conv3d_backward_data")
            return dx

        def up_back(g: _Act, dims_small) -> _Act:
            B, C, D, H, W = g.shape
            dx = _Act(torch.empty(B, C, *dims_small, **f32))
            _check(lib.citrus_nn_upsample2_backward(g.ptr, g.bs, B, C, *dims_small, dx.ptr, dx.bs, st),
                   "citrus_nn_upsample2_backward")
            return dx

        def pool_back_into(g: _Act, dx: _Act) -> None:
            B, C, D, H, W = dx.shape
            _check(lib.citrus_nn_avg_pool2_backward(g.ptr, g.bs, B, C, D, H, W, dx.ptr, dx.bs, 1, st),
                   "citrus_nn_avg_pool2_backward")

        with torch.cuda.device(self.device):
            g = _Act(g_out.to(self.device, torch.float32).contiguous())
            g = layer("head", a["d0b"], g)
            mask(g, a["d0b"])
            g = layer("d0b", a["d0a"], g)
            mask(g, a["d0a"])
            g_cat0 = layer("d0a", a["cat0"], g)  # [B, 3c]: upsampled d1b, then the skip e0b
            g = up_back(_Act(g_cat0.t[:, :2 * c]), a["d1b"].shape[2:])
            mask(g, a["d1b"])
            g = layer("d1b", a["d1a"], g)
            mask(g, a["d1a"])
            g_cat1 = layer("d1a", a["cat1"], g)  # [B, 6c]: upsampled b1, then the skip e1b
            g = up_back(_Act(g_cat1.t[:, :4 * c]), a["b1"].shape[2:])
            mask(g, a["b1"])
            g = layer("b1", a["b0"], g)
            mask(g, a["b0"])
            g = layer("b0", a["p1"], g)
            g_e1b = _Act(g_cat1.t[:, 4 * c:])  # skip gradient, plus the pooling's adjoint
            pool_back_into(g, g_e1b)
            mask(g_e1b, a["e1b"])
            g = layer("e1b", a["e1a"], g_e1b)
            mask(g, a["e1a"])
            g = layer("e1a", a["p0"], g)
            g_e0b = _Act(g_cat0.t[:, 2 * c:])
            pool_back_into(g, g_e0b)
            mask(g_e0b, a["e0b"])
            g = layer("e0b", a["e0a"], g_e0b)
            mask(g, a["e0a"])
            layer("e0a", a["x"], g, need_dx=False)
        return grads

    def adam_step(self, grads: dict, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        """One Adam step on every parameter (``citrus_nn_adam``), re-rounding the bfloat16 weights."""
        self.steps += 1
        lib = _hip.library()
        st = _stream(self.device)
        with torch.cuda.device(self.device):
            for name in LAYERS:
                dw, db = grads[name]
                mw, vw, mb, vb = self.moments[name]
                w, w16, b = self.weights[name], self.weights16[name], self.biases[name]
                _check(lib.citrus_nn_adam(w.data_ptr(), dw.data_ptr(), mw.data_ptr(), vw.data_ptr(), w16.data_ptr(),
                                          w.numel(), lr, beta1, beta2, eps, self.steps, st), "citrus_nn_adam")
                _check(lib.citrus_nn_adam(b.data_ptr(), db.data_ptr(), mb.data_ptr(), vb.data_ptr(), None, b.numel(),
                                          lr, beta1, beta2, eps, self.steps, st), "citrus_nn_adam")

    # ---- persistence -------------------------------------------------------------------------------------------
    def save(self, path: str) -> None:
        """float32 master weights and biases as a plain .npz (no pickled objects)."""
        arrays = {"shape": np.array([self.cin, self.cout, self.width], dtype=np.int64)}
        for name in LAYERS:
            arrays[f"{name}.w"] = self.weights[name].cpu().numpy()
            arrays[f"{name}.b"] = self.biases[name].cpu().numpy()
        np.savez(path, **arrays)

    @staticmethod
    def read_arrays(path: str) -> tuple[tuple[int, int, int], dict]:
        """((cin, cout, width), {array name: numpy array}) of a weight file written by ``save``, every array checked
        against ``layer_shapes``: weights ``[Co, Ci, k^3]`` and biases ``[Co]``, float32. The convolution kernels
        index the weights by those shapes, so a truncated, reshaped or float64 file raises ``ValueError`` here instead
        of being read out of bounds on the device."""
        with np.load(path, allow_pickle=False) as f:
            if "shape" not in f.files:
                raise ValueError(f"{path}: no 'shape' entry (not a surrogate weight file)")
            shape = f["shape"]
            if shape.shape != (3,) or shape.dtype.kind not in "iu":
                raise ValueError(f"{path}: 'shape' must be 3 integers, got {shape.dtype} {shape.shape}")
            cin, cout, width = (int(v) for v in shape)
            if min(cin, cout, width) < 1:
                raise ValueError(f"{path}: channel counts must be positive, got {(cin, cout, width)}")
            arrays = {}
            for name, (co, ci, k) in layer_shapes(cin, cout, width).items():
                for key, want in ((f"{name}.w", (co, ci, k ** 3)), (f"{name}.b", (co,))):
                    if key not in f.files:
                        raise ValueError(f"{path}: missing array {key!r}")
                    a = f[key]
                    if a.dtype != np.float32:
                        raise ValueError(f"{path}: {key} is {a.dtype}, expected float32")
                    if tuple(a.shape) != want:
                        raise ValueError(f"{path}: {key} has shape {tuple(a.shape)}, expected {want}")
                    arrays[key] = np.ascontiguousarray(a)
        return (cin, cout, width), arrays

    @classmethod
    def load(cls, path: str, device="cuda") -> "UNet3D":
        (cin, cout, width), arrays = cls.read_arrays(path)
        net = cls.create(cin, cout, width, device)
        for name in LAYERS:
            net.weights[name] = torch.from_numpy(arrays[f"{name}.w"]).to(net.device).contiguous()
            net.biases[name] = torch.from_numpy(arrays[f"{name}.b"]).to(net.device).contiguous()
        net._round_all()
        return net


def mse_gradient(pred: Tensor, target: Tensor, channel_weight: Optional[Tensor] = None) -> tuple[float, Tensor]:
    """(mean squared error, its gradient 2 (pred - target) / n); with ``channel_weight`` [C] each channel's squared
    error is weighted."""
    d = pred - target
    if channel_weight is None:
        return float((d * d).mean()), d * (2.0 / d.numel())
    w = channel_weight.to(d.device, d.dtype).reshape(1, -1, *([1] * (d.dim() - 2)))
    return float((w * d * d).mean()), w * d * (2.0 / d.numel())


def scalar_flux_weight(groups: int, weight: float) -> Tensor:
    """[4 G] channel weights of the P1 outputs: ``weight`` on the scalar-flux channel (l = 0) of every group, 1 on
    the three l = 1 channels."""
    w = torch.ones(groups, 4)
    w[:, 0] = weight
    return w.reshape(-1)


# ---- features and targets from the solver ----------------------------------------------------------------------

def p1_means(phi: Tensor) -> Tensor:
    """[4 G, R, C, S] P1 (l <= 1) cell means of moments [G, nsh, 4, R, C, S]."""
    G = int(phi.shape[0])
    return phi[:, :4, 0].reshape(4 * G, *phi.shape[3:])


def sample_scale(phi2: Tensor) -> float:
    """RMS of the scalar-flux cell means of phi^(2): the normalisation of one sample."""
    s = float(torch.sqrt((phi2[:, 0, 0].double() ** 2).mean()))
    return s if s > 0 else 1.0


def features(rho_coarse: Tensor, phi2: Tensor, scale: float, phi1: Optional[Tensor] = None) -> Tensor:
    """[M + 4 G (+ 4 G), R, C, S] network input: densities, the scaled P1 cell means of phi^(2) and, when given,
    those of phi^(1)."""
    parts = [rho_coarse.to(phi2.device, torch.float32), p1_means(phi2).float() / scale]
    if phi1 is not None:
        parts.append(p1_means(phi1).float() / scale)
    return torch.cat(parts, dim=0)


def _uses_phi1(net: "UNet3D", fields) -> bool:
    """Whether the network's input carries phi^(1) besides phi^(2) (told by its input width)."""
    M = int(fields.rho_coarse.shape[0])
    extra = net.cin - M
    if extra == net.cout:
        return False
    if extra == 2 * net.cout:
        return True
    raise ValueError(f"network input width {net.cin} fits neither {M + net.cout} nor {M + 2 * net.cout}")


@dataclass
class CollisionOrders:
    """phi^(1), phi^(2) and the converged sum of one view, from the plain source iteration."""

    phi1: Tensor
    phi2: Tensor
    total: Tensor
    sweeps: int


def collision_orders(solver, fields, angle: float, tolerance: float = 1e-4, max_sweeps: int = 200) -> CollisionOrders:
    """Sweep collision order by collision order (``phi^(n+1) = K phi^(n)``) until the order's scalar-flux cell-mean
    norm falls below ``tolerance`` of the total's (the plain source iteration, source_iteration.py:138-204)."""
    ctx = solver.sweep_context(fields.grid_coarse)
    phi_uc = solver.uncollided_fluence(fields, angle)
    fc = solver.first_collision_moment_source(fields, angle, phi_uc, ctx)
    phi1 = solver.transport_sweep(fields, fc)
    del fc
    total = phi1.clone()
    phi_n = phi1
    phi2 = None
    sweeps = 1
    while sweeps < max_sweeps:
        phi_n = solver.transport_sweep(fields, solver.scatter_moment_source(fields, phi_n, ctx))
        sweeps += 1
        if phi2 is None:
            phi2 = phi_n
        total.add_(phi_n)
        n, t = torch.stack([torch.linalg.vector_norm(phi_n[:, 0, 0].double()),
                            torch.linalg.vector_norm(total[:, 0, 0].double())]).tolist()
        if n <= tolerance * t:
            break
    return CollisionOrders(phi1=phi1, phi2=phi2, total=total, sweeps=sweeps)


def training_pair(solver, fields, angle: float, tolerance: float = 1e-4, with_phi1: bool = False):
    """(input [M + 4G (+ 4G), R, C, S], target [4G, R, C, S]) of one view: the scaled P1 cell means of Phi_{>2}."""
    orders = collision_orders(solver, fields, angle, tolerance)
    s = sample_scale(orders.phi2)
    higher = orders.total - orders.phi1 - orders.phi2
    return (features(fields.rho_coarse, orders.phi2, s, orders.phi1 if with_phi1 else None),
            p1_means(higher).float() / s)


def train(net: UNet3D, inputs: Sequence[Tensor], targets: Sequence[Tensor], epochs: int, lr: float = 1e-3,
          batch: int = 1, seed: int = 0, channel_weight: Optional[Tensor] = None, log=None,
          log_every: int = 10) -> list[float]:
    """Adam on the (channel-weighted) mean squared error over the pairs, shuffled per epoch (seeded); returns the
    mean loss per epoch. ``log(message)``, when given, receives the epoch and its loss every ``log_every`` epochs."""
    gen = torch.Generator().manual_seed(seed)
    n = len(inputs)
    history = []
    for epoch in range(epochs):
        order = torch.randperm(n, generator=gen).tolist()
        total = 0.0
        for i in range(0, n, batch):
            idx = order[i:i + batch]
            x = torch.stack([inputs[j] for j in idx]).to(net.device)
            y = torch.stack([targets[j] for j in idx]).to(net.device)
            out, cache = net.forward(x, keep=True)
            loss, g = mse_gradient(out, y, channel_weight)
            net.adam_step(net.backward(cache, g), lr)
            total += loss * len(idx)
        history.append(total / n)
        if log is not None and (epoch + 1) % max(1, int(log_every)) == 0:
            log(f"  epoch {epoch + 1}/{epochs}: loss {history[-1]:.5f}")
    return history


# ---- the learned step -----------------------------------------------------------------------------------------

@dataclass
class LearnedResult:
    """Collided moments of one view from the learned form, the judging sweep's relative change, whether it was
    accepted, and the sweeps spent."""

    collided_moments: Tensor
    residual: float
    accepted: bool
    sweeps: int


def _embed_p1(pred: Tensor, like: Tensor, scale: float) -> Tensor:
    """Moments shaped like ``like`` [G, nsh, 4, R, C, S] holding the P1 cell means pred [4G, R, C, S] * scale."""
    G = int(like.shape[0])
    out = torch.zeros(like.shape[0], like.shape[1], like.shape[2], *like.shape[3:], device=like.device,
                      dtype=like.dtype)
    out[:, :4, 0] = pred.reshape(G, 4, *like.shape[3:]).to(like.dtype) * scale
    return out


def learned_collided_moments(solver, fields, angle: float, net: UNet3D, threshold: float = 0.01,
                             fallback_tolerance: float = 1e-3) -> LearnedResult:
    """Collided moments of one view by the learned form: phi^(1), phi^(2), the predicted Phi_{>2}, then one judging
    sweep ``Phi~ = phi^(1) + K Phi^``; Phi~ is returned when its relative scalar-flux change is <= ``threshold``,
    else the plain source iteration's result (stopping at ``fallback_tolerance``)."""
    ctx = solver.sweep_context(fields.grid_coarse)
    phi_uc = solver.uncollided_fluence(fields, angle)
    fc = solver.first_collision_moment_source(fields, angle, phi_uc, ctx)
    phi1 = solver.transport_sweep(fields, fc)
    phi2 = solver.transport_sweep(fields, solver.scatter_moment_source(fields, phi1, ctx))
    s = sample_scale(phi2)
    x = features(fields.rho_coarse, phi2, s, phi1 if _uses_phi1(net, fields) else None)[None]
    pred = net.forward(x)[0]
    estimate = _embed_p1(pred, phi1, s).add_(phi1).add_(phi2)
    del pred, x, phi2
    from .scatter_solver.source_iteration import summed_source_function

    judged = solver.transport_sweep(fields, summed_source_function(solver.scatter_moment_source(fields, estimate, ctx),
                                                                   fc))
    n, d = torch.stack([torch.linalg.vector_norm(judged[:, 0, 0].double()),
                        torch.linalg.vector_norm((judged[:, 0, 0] - estimate[:, 0, 0]).double())]).tolist()
    residual = d / n if n > 0 else 0.0
    if residual <= threshold:
        return LearnedResult(judged, residual, True, 3)
    del judged, estimate
    from .scatter_solver.source_iteration import source_iteration

    plain = source_iteration(ctx, fields.mu_t_coarse, fields.mu_s_l_coarse, fc, tolerance=fallback_tolerance,
                             max_iterations=max(1, solver.config.max_source_iterations),
                             direction_chunk=solver.sweep_direction_chunk(fields))
    return LearnedResult(plain.phi_total, residual, False, 3 + plain.num_iterations)


# ---- several views of one volume at once (batched sweeps) ---------------------------------------------------------

def _views_setup(solver, fields, angles):
    """The sweep context, the views' first-collision source and zero scattering moments. With one quadrature point
    per cell the source stays uncorrected in the sweep's block order (``DeferredFirstCollision``, as the CITRUS view
    step keeps it: the persistent sweep corrects it as it reads it, bit for bit the corrected source), else the
    group-major corrected LD source."""
    from .scatter_solver.uncollided import DeferredFirstCollision, first_collision_deferred_samples

    ctx = solver.sweep_context(fields.grid_coarse)
    V = len(angles)
    G, nsh, M = solver.num_groups, ctx.num_moments, ctx.num_directions
    R, C, S = fields.grid_coarse.shape
    dev = solver.device
    if fields.q == 1:
        source = DeferredFirstCollision.empty(V, G, (R, C, S), M, dev)
        omega = ctx.quadrature.omega.to(dev, torch.float32)
        for v, a in enumerate(angles):
            phi_uc = solver.uncollided_fluence(fields, a)
            first_collision_deferred_samples(solver.library, fields.rho_quadrature, phi_uc, fields.grid_quadrature,
                                             solver.geometry, a, omega, ctx.quadrature.weight, source, v,
                                             return_correction=False)
            del phi_uc
    else:
        source = torch.empty((V, G, R, C, S, M, 4), device=dev, dtype=torch.float32)
        for v, a in enumerate(angles):
            phi_uc = solver.uncollided_fluence(fields, a)
            solver.first_collision_group_major(fields, a, phi_uc, ctx, out=source[v])
            del phi_uc
    T = torch.zeros(V, G, R, C, S, nsh, 4, device=dev, dtype=torch.float32)
    return ctx, source, T


def _api(phi_v: Tensor) -> Tensor:
    """[G, nsh, 4, R, C, S] view of one view's [G, R, C, S, nsh, 4] moments."""
    return phi_v.permute(0, 4, 5, 1, 2, 3)


def _scalar_norms_views(*phis: Tensor) -> list:
    """Per view, float64 norms of the scalar-flux cell means of [V, G, R, C, S, nsh, 4] moments (one sync)."""
    from .scatter_solver.kernels.sweep_kernels import decode_status, fold_status

    rows = torch.stack([torch.linalg.vector_norm(p[..., 0, 0].reshape(p.shape[0], -1).double(), dim=1)
                        for p in phis])
    flat = decode_status(fold_status(rows).tolist(), rows.device)  # raises if a persistent sweep aborted
    V = int(rows.shape[1])
    return [flat[i * V:(i + 1) * V] for i in range(len(phis))]


def collision_orders_views(solver, fields, angles: Sequence[float], tolerance: float = 1e-4,
                           max_sweeps: int = 200) -> list[CollisionOrders]:
    """``collision_orders`` of several views of one volume, each collision order of all views in one batched sweep
    (``sweep_views``); stops when every view's order norm is below ``tolerance`` of its total."""
    from .scatter_solver.source_iteration import scatter_moments_views
    from .scatter_solver.transport_sweep import sweep_views

    ctx, source, T = _views_setup(solver, fields, angles)
    phi1 = sweep_views(ctx, fields.mu_t_coarse, T, source=source)
    del source
    total = phi1.clone()
    phi_n, phi2, sweeps = phi1, None, 1
    while sweeps < max_sweeps:
        T = scatter_moments_views(fields.transfer_coarse, phi_n, out=T)
        phi_n = sweep_views(ctx, fields.mu_t_coarse, T)
        sweeps += 1
        if phi2 is None:
            phi2 = phi_n
        total.add_(phi_n)
        n, t = _scalar_norms_views(phi_n, total)
        if all(a <= tolerance * b for a, b in zip(n, t)):
            break
    return [CollisionOrders(phi1=_api(phi1[v]), phi2=_api(phi2[v]), total=_api(total[v]), sweeps=sweeps)
            for v in range(len(angles))]


def plain_views(solver, fields, angles: Sequence[float], tolerance: float, max_sweeps: int = 200):
    """Plain source iteration of several views (batched sweeps) until every view's last collision order is below
    ``tolerance`` of its total: ([G, nsh, 4, R, C, S] per view, sweeps)."""
    orders = collision_orders_views(solver, fields, angles, tolerance, max_sweeps)
    return [o.total for o in orders], orders[0].sweeps


@dataclass
class LearnedViews:
    """The learned form of several views: collided moments per view, the judging sweep's relative change per view,
    which views it accepted (the others hold the plain source iteration's result) and the sweeps spent."""

    collided_moments: list
    residuals: list
    accepted: list
    sweeps: int


def learned_judged(solver, fields, ctx, source, phi1: Tensor, T: Tensor, net: UNet3D):
    """The learned form's estimate of V views from their first collision order, judged: phi^(2) (one batched sweep
    of K phi^(1)), the network on all views at once, then one batched judging sweep ``Phi~ = G(S_FC + K Phi^)`` of the
    estimate ``Phi^ = phi^(1) + phi^(2) + Phi^_{>2}``. phi1 [V, G, R, C, S, nsh, 4] is left as it was; T is float32
    scratch of the same shape (or None: allocated). Returns (judged [V, ...], residuals: per view the relative change
    of the scalar-flux cell means from the estimate to the judged moments)."""
    from .scatter_solver.source_iteration import scatter_moments_views
    from .scatter_solver.transport_sweep import sweep_views

    V = int(phi1.shape[0])
    if not (isinstance(T, Tensor) and T.shape == phi1.shape and T.dtype == torch.float32 and T.is_contiguous()):
        T = torch.empty_like(phi1)
    T = scatter_moments_views(fields.transfer_coarse, phi1, out=T)
    phi2 = sweep_views(ctx, fields.mu_t_coarse, T)
    scales = [sample_scale(_api(phi2[v])) for v in range(V)]
    with1 = _uses_phi1(net, fields)
    x = torch.stack([features(fields.rho_coarse, _api(phi2[v]), scales[v], _api(phi1[v]) if with1 else None)
                     for v in range(V)])
    pred = net.forward(x)
    G = solver.num_groups
    estimate = phi2.add_(phi1)  # phi1 + phi2, in place of phi2
    for v in range(V):  # predicted P1 cell means of Phi_{>2}: moments 0..3, LD component 0
        est = estimate[v]
        est[..., :4, 0] += (pred[v].reshape(G, 4, *est.shape[1:4]) * scales[v]).permute(0, 2, 3, 4, 1)
    del pred, x
    T = scatter_moments_views(fields.transfer_coarse, estimate, out=T)
    judged = sweep_views(ctx, fields.mu_t_coarse, T, source=source)
    n, d = _scalar_norms_views(judged, judged - estimate)
    return judged, [di / ni if ni > 0 else 0.0 for ni, di in zip(n, d)]


def learned_views(solver, fields, angles: Sequence[float], net: UNet3D, threshold: float = 0.01,
                  fallback_tolerance: float = 1e-3) -> LearnedViews:
    """``learned_collided_moments`` of several views of one volume with batched sweeps and one batched inference:
    phi^(1) and phi^(2) of every view (two batched sweeps), the network on all views at once, one batched judging
    sweep (``learned_judged``); views whose relative change exceeds ``threshold`` are recomputed by the plain source
    iteration."""
    from .scatter_solver.transport_sweep import sweep_views

    V = len(angles)
    ctx, source, T = _views_setup(solver, fields, angles)
    phi1 = sweep_views(ctx, fields.mu_t_coarse, T, source=source)
    judged, residuals = learned_judged(solver, fields, ctx, source, phi1, T, net)
    del phi1
    accepted = [r <= threshold for r in residuals]
    result = [_api(judged[v]) for v in range(V)]
    sweeps = 3
    rejected = [v for v in range(V) if not accepted[v]]
    if rejected:
        del source, T
        totals, extra = plain_views(solver, fields, [angles[v] for v in rejected], fallback_tolerance)
        for v, tot in zip(rejected, totals):
            result[v] = tot
        sweeps += extra
    return LearnedViews(result, residuals, accepted, sweeps)
