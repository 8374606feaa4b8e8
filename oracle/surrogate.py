"""FP64 oracle of the Phi_{>2} surrogate network (test infrastructure only).

The surrogate has no counterpart in the reference: it is the learned part of this path (a 3D U-Net that predicts the
collided moments of collision order three and higher from the densities and the second-order moments). This module
restates every operation of the network in numpy float64 so the HIP forward, backward and optimizer kernels can be
checked element by element:

- ``conv3d`` / ``conv3d_backward``: 3D cross-correlation with a cubic kernel of odd size k (k = 3 or 1), zero
  padding (k - 1)/2, stride 1, bias; the input gradient is the correlation of the output gradient with the flipped,
  transposed kernel; the weight gradient the correlation of the input with the output gradient.
- ``leaky`` and its derivative (from the output, which has the sign of the input).
- ``avg_pool2`` / ``upsample2`` and their adjoints (pooling: mean of each 2x2x2 block; upsampling: nearest).
- ``UNet`` forward and backward over a parameter dict, and the mean-squared loss.
- ``adam`` one step, and ``bf16_round`` (round to nearest even of the float32 bit pattern, as the kernels store
  weights).

Pins (tests/test_oracle.py): the convolution against a brute-force loop, both gradients and the full network's
gradient against central finite differences in float64, the pooling/upsampling adjoint identity, Adam against its
closed form on the first step, and bf16 rounding against the bit pattern of known values.
"""

from __future__ import annotations

import numpy as np

SLOPE = 0.01  # leaky-ReLU slope for negative inputs


def conv3d(x, w, b):
    """x [B, Ci, D, H, W], w [Co, Ci, k, k, k], b [Co] -> [B, Co, D, H, W] (zero padding (k-1)/2)."""
    k = w.shape[-1]
    p = (k - 1) // 2
    B, Ci, D, H, W = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (p, p), (p, p), (p, p)))
    out = np.broadcast_to(b[None, :, None, None, None], (B, w.shape[0], D, H, W)).astype(np.float64).copy()
    for dz in range(k):
        for dy in range(k):
            for dx in range(k):
                sl = xp[:, :, dz:dz + D, dy:dy + H, dx:dx + W]
                out += np.einsum("oi,bizyx->bozyx", w[:, :, dz, dy, dx], sl)
    return out


def conv3d_backward(x, w, g):
    """Gradients (dx, dw, db) of sum(g * conv3d(x, w, b)) for g [B, Co, D, H, W]."""
    k = w.shape[-1]
    p = (k - 1) // 2
    B, Ci, D, H, W = x.shape
    xp = np.pad(x, ((0, 0), (0, 0), (p, p), (p, p), (p, p)))
    dxp = np.zeros_like(xp, dtype=np.float64)
    dw = np.zeros_like(w, dtype=np.float64)
    for dz in range(k):
        for dy in range(k):
            for dx in range(k):
                sl = xp[:, :, dz:dz + D, dy:dy + H, dx:dx + W]
                dw[:, :, dz, dy, dx] = np.einsum("bozyx,bizyx->oi", g, sl)
                dxp[:, :, dz:dz + D, dy:dy + H, dx:dx + W] += np.einsum("oi,bozyx->bizyx", w[:, :, dz, dy, dx], g)
    dx_ = dxp[:, :, p:p + D, p:p + H, p:p + W] if p else dxp
    return dx_, dw, g.sum(axis=(0, 2, 3, 4))


def leaky(x):
    return np.where(x > 0, x, SLOPE * x)


def leaky_grad_from_output(y):
    """Derivative of the leaky ReLU at the input, read from the output (same sign)."""
    return np.where(y > 0, 1.0, SLOPE)


def avg_pool2(x):
    B, C, D, H, W = x.shape
    return x.reshape(B, C, D // 2, 2, H // 2, 2, W // 2, 2).mean(axis=(3, 5, 7))


def avg_pool2_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, axis=2), 2, axis=3), 2, axis=4) / 8.0


def upsample2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, axis=2), 2, axis=3), 2, axis=4)


def upsample2_backward(g):
    B, C, D, H, W = g.shape
    return g.reshape(B, C, D // 2, 2, H // 2, 2, W // 2, 2).sum(axis=(3, 5, 7))


# network: two levels of pooling; layer names and shapes
def layer_shapes(cin: int, cout: int, c: int):
    return {
        "e0a": (c, cin, 3), "e0b": (c, c, 3),
        "e1a": (2 * c, c, 3), "e1b": (2 * c, 2 * c, 3),
        "b0": (4 * c, 2 * c, 3), "b1": (4 * c, 4 * c, 3),
        "d1a": (2 * c, 6 * c, 3), "d1b": (2 * c, 2 * c, 3),
        "d0a": (c, 3 * c, 3), "d0b": (c, c, 3),
        "head": (cout, c, 1),
    }


def forward(params, x):
    """Returns (output, cache of activations). params[name] = (w, b)."""
    a = {}

    def conv(name, inp, act=True):
        w, b = params[name]
        y = conv3d(inp, w, b)
        return leaky(y) if act else y

    a["x"] = x
    a["e0a"] = conv("e0a", x)
    a["e0b"] = conv("e0b", a["e0a"])
    a["p0"] = avg_pool2(a["e0b"])
    a["e1a"] = conv("e1a", a["p0"])
    a["e1b"] = conv("e1b", a["e1a"])
    a["p1"] = avg_pool2(a["e1b"])
    a["b0"] = conv("b0", a["p1"])
    a["b1"] = conv("b1", a["b0"])
    a["c1"] = np.concatenate([upsample2(a["b1"]), a["e1b"]], axis=1)
    a["d1a"] = conv("d1a", a["c1"])
    a["d1b"] = conv("d1b", a["d1a"])
    a["c0"] = np.concatenate([upsample2(a["d1b"]), a["e0b"]], axis=1)
    a["d0a"] = conv("d0a", a["c0"])
    a["d0b"] = conv("d0b", a["d0a"])
    a["out"] = conv("head", a["d0b"], act=False)
    return a["out"], a


def backward(params, a, g_out):
    """Parameter gradients {name: (dw, db)} of sum(g_out * output)."""
    grads = {}

    def back(name, inp, out, g, act=True):
        if act:
            g = g * leaky_grad_from_output(out)
        dx, dw, db = conv3d_backward(inp, params[name][0], g)
        grads[name] = (dw, db)
        return dx

    g = back("head", a["d0b"], a["out"], g_out, act=False)
    g = back("d0b", a["d0a"], a["d0b"], g)
    g = back("d0a", a["c0"], a["d0a"], g)
    n_up = a["d1b"].shape[1]
    g_skip0 = g[:, n_up:]
    g = upsample2_backward(g[:, :n_up])
    g = back("d1b", a["d1a"], a["d1b"], g)
    g = back("d1a", a["c1"], a["d1a"], g)
    n_up = a["b1"].shape[1]
    g_skip1 = g[:, n_up:]
    g = upsample2_backward(g[:, :n_up])
    g = back("b1", a["b0"], a["b1"], g)
    g = back("b0", a["p1"], a["b0"], g)
    g = avg_pool2_backward(g) + g_skip1
    g = back("e1b", a["e1a"], a["e1b"], g)
    g = back("e1a", a["p0"], a["e1a"], g)
    g = avg_pool2_backward(g) + g_skip0
    g = back("e0b", a["e0a"], a["e0b"], g)
    back("e0a", a["x"], a["e0a"], g)
    return grads


def mse(pred, target):
    """(loss, d loss / d pred) of the mean squared error."""
    d = pred - target
    return float((d * d).mean()), 2.0 * d / d.size


def adam(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (bias-corrected); returns (p, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16_round(x):
    """float32 values rounded to bfloat16 (nearest, ties to even), returned as float64."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)
