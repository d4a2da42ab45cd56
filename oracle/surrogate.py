"""FP64 restatement of the learned warm-start step (forward, loss, gradient, Adam).

TEST INFRASTRUCTURE ONLY: imported by tests/ (and bench.py's checks) as the
checker of the HIP kernels in lcdrsimulator_amd/csrc/qsr_surr.hip.  The product
path never imports this module.

The network has no counterpart in the reference (QSR.cu has only the plain
relaxation loop, QSR.cu:3126-3476); its definition is the one stated in
include/qsr_amd.h and lcdrsimulator_amd/csrc/qsr_surr.hpp:

    z1 = conv(W1, x) + b1,  a1 = gelu(z1)
    z2 = conv(W2, a1) + b2, a2 = gelu(z2)
    y  = conv(W3, a2) + b3
    conv(W, a)[v, co] = sum_{t, ci} W[t][ci][co] a[v + off(t), ci], zero outside the box
    t = (di+1)*9 + (dj+1)*3 + (dk+1), off(t) = (di, dj, dk) along (i, j, k)
    gelu(z) = z/2 (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))
    loss = sum_v m_v sum_c (y - target)^2 / (5 V)

Under increment and gain input the network's output modulates a given field g
(the normalised increment of one explicit step): the predicted change is
p = g (1 + y), and

    loss = sum_v m_v sum_c (g (1 + y) - target)^2 / (5 V),  dL/dy = 2 m (p - target) g / (5 V)

or, under step-size input, one gain per voxel: p = g (1 + y_0) in every component and
dL/dy_0 = 2 m sum_c (p_c - target_c) g_c / (5 V).

With ``emulate=True`` the values the GPU keeps in BF16 are rounded the same way
(round to nearest even): x, the weights, a1, a2 and the three output gradients
dy, dz2, dz1.  Everything else is FP64.

Pins (tests/test_surrogate_oracle.py): ``conv`` against scipy.ndimage.correlate,
``bf16`` against torch's bfloat16 conversion, the gradient against central
finite differences, the input-gradient convolution by the adjoint identity, and
``adam`` against torch.optim.Adam.
"""
from __future__ import annotations

import numpy as np

TAPS = 27
CQ, CH = 5, 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + TAPS * CQ * CH
OFF_W2 = OFF_B1 + CH
OFF_B2 = OFF_W2 + TAPS * CH * CH
OFF_W3 = OFF_B2 + CH
OFF_B3 = OFF_W3 + TAPS * CH * CQ
NPARAM = OFF_B3 + CQ
GELU_C = np.sqrt(2.0 / np.pi)
GELU_A = 0.044715


def tap_offsets():
    """(di, dj, dk) of tap t = (di+1)*9 + (dj+1)*3 + (dk+1)."""
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(TAPS)]


def bf16(x):
    """Round to the nearest BF16 (ties to even), via FP32 as the GPU does; returns float64."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(TAPS, CQ, CH), b1=p[OFF_B1:OFF_W2],
                W2=p[OFF_W2:OFF_B2].reshape(TAPS, CH, CH), b2=p[OFF_B2:OFF_W3],
                W3=p[OFF_W3:OFF_B3].reshape(TAPS, CH, CQ), b3=p[OFF_B3:NPARAM])


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3]).astype(np.float64)


def shifted(a, di, dj, dk):
    """s[n, i, j, k] = a[n, i+di, j+dj, k+dk], zero outside the box."""
    n, X, Y, Z, c = a.shape
    p = np.zeros((n, X + 2, Y + 2, Z + 2, c))
    p[:, 1:-1, 1:-1, 1:-1] = a
    return p[:, 1 + di:1 + di + X, 1 + dj:1 + dj + Y, 1 + dk:1 + dk + Z]


def conv(a, W, b=None):
    """a: (n, X, Y, Z, ci); W: (27, ci, co) -> (n, X, Y, Z, co)."""
    out = np.zeros(a.shape[:4] + (W.shape[2],))
    for t, (di, dj, dk) in enumerate(tap_offsets()):
        out += shifted(a, di, dj, dk) @ W[t]
    if b is not None:
        out += b
    return out


def conv_input_grad(dz, W):
    """Adjoint of conv in its input: the convolution of dz with W'[t] = W[26 - t]^T."""
    Wt = np.ascontiguousarray(W[::-1].transpose(0, 2, 1))
    return conv(dz, Wt)


def conv_weight_grad(a, dz):
    """dW[t][ci][co] = sum_v a[v + off(t), ci] dz[v, co]; db = sum_v dz[v]."""
    dW = np.zeros((TAPS, a.shape[-1], dz.shape[-1]))
    dz2 = dz.reshape(-1, dz.shape[-1])
    for t, (di, dj, dk) in enumerate(tap_offsets()):
        dW[t] = shifted(a, di, dj, dk).reshape(-1, a.shape[-1]).T @ dz2
    return dW, dz2.sum(axis=0)


def gelu(z):
    return 0.5 * z * (1.0 + np.tanh(GELU_C * (z + GELU_A * z ** 3)))


def gelu_grad(z):
    t = np.tanh(GELU_C * (z + GELU_A * z ** 3))
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z)


def forward(params, x, emulate=True):
    """x: (n, X, Y, Z, 5).  Returns the intermediates and the output y."""
    P = unpack(params)
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    xb = r(x)
    W1, W2, W3 = r(P["W1"]), r(P["W2"]), r(P["W3"])
    b1, b2, b3 = (np.float32(P[k]).astype(np.float64) if emulate else P[k] for k in ("b1", "b2", "b3"))
    z1 = conv(xb, W1, b1)
    a1 = r(gelu(z1))
    z2 = conv(a1, W2, b2)
    a2 = r(gelu(z2))
    y = conv(a2, W3, b3)
    return dict(x=xb, W1=W1, W2=W2, W3=W3, z1=z1, a1=a1, z2=z2, a2=a2, y=y)


def gain_prediction(g, y):
    """Modulated prediction g (1 + y)."""
    return np.asarray(g, dtype=np.float64) * (1.0 + y)


def loss_and_grad(params, x, target, mask=None, emulate=True, mod=None, scalar=False):
    """loss = sum m |y - target|^2 / (5 V) (or, with a modulator ``mod`` = g, of g (1 + y) - target;
    with ``scalar`` of g (1 + y_0) - target, one gain per voxel) and its gradient in the flat layout."""
    f = forward(params, x, emulate)
    r = bf16 if emulate else (lambda v: v)
    V = np.prod(x.shape[:4])
    m = np.ones(x.shape[:4]) if mask is None else np.asarray(mask, dtype=np.float64)
    g = None if mod is None else np.asarray(mod, dtype=np.float64)
    y = f["y"][..., :1] if scalar else f["y"]
    d = (f["y"] if g is None else gain_prediction(g, y)) - target
    loss = float(np.sum(m[..., None] * d * d) / (5.0 * V))
    scale = np.float32(2.0 / (5.0 * V)) if emulate else 2.0 / (5.0 * V)
    if scalar:  # dL/dy_0 = sum_c dL/dp_c g_c; the other outputs do not enter the loss
        dy = np.zeros_like(f["y"])
        dy[..., 0] = np.sum(d * g, axis=-1)
        dy = r(np.float64(scale) * m[..., None] * dy)
    else:
        dy = r(np.float64(scale) * m[..., None] * d * (1.0 if g is None else g))
    dW3, db3 = conv_weight_grad(f["a2"], dy)
    dz2 = r(conv_input_grad(dy, f["W3"]) * gelu_grad(f["z2"]))
    dW2, db2 = conv_weight_grad(f["a1"], dz2)
    dz1 = r(conv_input_grad(dz2, f["W2"]) * gelu_grad(f["z1"]))
    dW1, db1 = conv_weight_grad(f["x"], dz1)
    return loss, pack(dW1, db1, dW2, db2, dW3, db3), f


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) with bias correction, step index t >= 1.
    Returns (p, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** t)
    vh = v / (1.0 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def random_params(seed=0, scale=1.0):
    """Uniform(-1/sqrt(fan_in), +) weights and biases (test inputs)."""
    rng = np.random.default_rng(seed)
    def u(n, fan):
        return rng.uniform(-1, 1, n) * scale / np.sqrt(fan)
    return np.concatenate([u(TAPS * CQ * CH, TAPS * CQ), u(CH, TAPS * CQ), u(TAPS * CH * CH, TAPS * CH),
                           u(CH, TAPS * CH), u(TAPS * CH * CQ, TAPS * CH), u(CQ, TAPS * CH)])
