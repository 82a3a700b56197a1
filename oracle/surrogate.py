"""TEST INFRASTRUCTURE: the FP64 oracle of the barostat screen's network.

The 16 -> 64 -> 64 -> 1 tanh MLP of rainmd_amd/csrc/surrogate.hip in plain
numpy FP64: forward, the mean-squared-error loss and its gradient by
backpropagation, plus a forward that rounds every matrix-product operand to
BF16 exactly where the device does (inputs, W1, layer-1 activations, W2), so
the device's forward can be held to FP32 accumulation error.

There is no counterpart in the reference (its barostat has no learned part).
Pinned by: central differences of the loss (test_surrogate.py), and the
closed form of a network whose hidden layers are switched off (W1 = W2 = 0
gives y = W3 . tanh(b2) + b3).

Only tests/ may import this module; the product path never does.
"""
import numpy as np

INPUTS = 16
HIDDEN = 64
PARAMS = INPUTS * HIDDEN + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1  # 5313


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    assert p.shape == (PARAMS,)
    o = 0
    W1 = p[o:o + INPUTS * HIDDEN].reshape(INPUTS, HIDDEN); o += INPUTS * HIDDEN
    b1 = p[o:o + HIDDEN]; o += HIDDEN
    W2 = p[o:o + HIDDEN * HIDDEN].reshape(HIDDEN, HIDDEN); o += HIDDEN * HIDDEN
    b2 = p[o:o + HIDDEN]; o += HIDDEN
    W3 = p[o:o + HIDDEN]; o += HIDDEN
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.reshape(-1), b1, W2.reshape(-1), b2, W3, [b3]])


def bf16(a):
    """Round FP32 values to BF16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, x, emulate_bf16=False):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64).reshape(-1, INPUTS)
    if emulate_bf16:
        z1 = bf16(x) @ bf16(W1) + np.float32(b1)
        a1 = np.tanh(z1)
        z2 = bf16(a1) @ bf16(W2) + np.float32(b2)
    else:
        z1 = x @ W1 + b1
        a1 = np.tanh(z1)
        z2 = a1 @ W2 + b2
    a2 = np.tanh(z2)
    return a2 @ W3 + b3


def loss_and_gradient(p, x, t):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    x = np.asarray(x, dtype=np.float64).reshape(-1, INPUTS)
    t = np.asarray(t, dtype=np.float64)
    n = len(x)
    a1 = np.tanh(x @ W1 + b1)
    a2 = np.tanh(a1 @ W2 + b2)
    y = a2 @ W3 + b3
    e = y - t
    loss = np.mean(e * e)
    dy = 2.0 * e / n
    gW3 = a2.T @ dy
    gb3 = dy.sum()
    dz2 = np.outer(dy, W3) * (1.0 - a2 * a2)
    gW2 = a1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * (1.0 - a1 * a1)
    gW1 = x.T @ dz1
    gb1 = dz1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def loss(p, x, t):
    y = forward(p, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))
