"""ORACLE - test infrastructure only.

NumPy restatement of the learned-initial-guess network that libbcmhip trains
on the GPU: nin -> nh (tanh) -> nh (tanh) -> nout (linear), loss
mean((y - t)^2), parameters in one array [W1 b1 W2 b2 W3 b3] with W_l
row-major (fan_in x fan_out). Only tests/ may import it, as the checker.

``bf16=True`` rounds every matrix-product operand to bfloat16 (round to
nearest even) the way the device kernels feed the matrix cores, and keeps
everything else in float64; with ``bf16=False`` the functions are the exact
float64 mathematics, which tests/test_oracle.py pins against central finite
differences. This network has no counterpart in the reference (parity for it
is pinned by the finite-difference check, not by reference vectors).
"""

import numpy as np


def bf16_round(a):
    """float64 -> nearest bfloat16 value (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def layout(nin, nh, nout):
    """Offsets of W1 b1 W2 b2 W3 b3 in the flat parameter array."""
    shapes = [(nin, nh), (nh,), (nh, nh), (nh,), (nh, nout), (nout,)]
    offs, o = [], 0
    for s in shapes:
        offs.append((o, s))
        o += int(np.prod(s))
    return offs, o


def unpack(p, nin, nh, nout):
    offs, _ = layout(nin, nh, nout)
    return [np.asarray(p[o:o + int(np.prod(s))], dtype=np.float64).reshape(s) for o, s in offs]


def pack(parts):
    return np.concatenate([np.asarray(a, dtype=np.float64).ravel() for a in parts])


def _mm(a, b, bf16):
    if bf16:
        a, b = bf16_round(a), bf16_round(b)
    return a @ b


def forward(p, x, nin, nh, nout, bf16=False):
    w1, b1, w2, b2, w3, b3 = unpack(p, nin, nh, nout)
    h1 = np.tanh(_mm(x, w1, bf16) + b1)
    h2 = np.tanh(_mm(h1, w2, bf16) + b2)
    y = _mm(h2, w3, bf16) + b3
    return y, (h1, h2)


def loss_grad(p, x, t, nin, nh, nout, bf16=False):
    """(loss, flat gradient) of mean((y - t)^2)."""
    x = np.asarray(x, dtype=np.float64)
    w1, b1, w2, b2, w3, b3 = unpack(p, nin, nh, nout)
    y, (h1, h2) = forward(p, x, nin, nh, nout, bf16)
    d = y - t
    loss = float(np.mean(d * d))
    dy = 2.0 * d / d.size
    gw3 = _mm(h2.T, dy, bf16)
    gb3 = dy.sum(axis=0)
    dz2 = _mm(dy, w3.T, bf16) * (1.0 - h2 * h2)
    gw2 = _mm(h1.T, dz2, bf16)
    gb2 = dz2.sum(axis=0)
    dz1 = _mm(dz2, w2.T, bf16) * (1.0 - h1 * h1)
    gw1 = _mm(x.T, dz1, bf16)
    gb1 = dz1.sum(axis=0)
    return loss, pack([gw1, gb1, gw2, gb2, gw3, gb3])


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (step counts from 1); returns new (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v
