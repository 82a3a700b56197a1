"""TEST INFRASTRUCTURE: the FP64 oracle of the barostat screen's network.

The 16 -> 64 -> 64 -> 1 tanh MLP of rainmd_amd/csrc/surrogate.hip in plain
numpy FP64: forward, the mean-squared-error loss and its gradient by
backpropagation, and the same forward and backward with every matrix-product
operand rounded to BF16 exactly where the device rounds it (inputs, W1,
layer-1 activations, W2; the back-propagated dZ2 and dZ1), so the device's
forward and gradient can be held to FP32 accumulation error.

There is no counterpart in the reference (its barostat has no learned part).
Pinned by: central differences of the loss (test_surrogate.py), the
closed form of a network whose hidden layers are switched off (W1 = W2 = 0
gives y = W3 . tanh(b2) + b3), and, for the BF16 variant, agreement with the
exact path on BF16-representable parameters and inputs where only the
rounded intermediates differ (within BF16's 2^-9 relative step) plus the
identity of the two paths when rounding is switched off.

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


def _rounding(emulate_bf16):
    return bf16 if emulate_bf16 else (lambda a: np.asarray(a, dtype=np.float64))


def forward(p, x, emulate_bf16=False):
    """y for every row of x. With emulate_bf16 the operands of the two hidden
    layers' matrix products (x, W1, layer-1 activations, W2) are rounded to
    BF16 as the device rounds them; everything else stays FP64."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    rb = _rounding(emulate_bf16)
    x = np.asarray(x, dtype=np.float64).reshape(-1, INPUTS)
    a1 = np.tanh(rb(x) @ rb(W1) + b1)
    a2 = np.tanh(rb(a1) @ rb(W2) + b2)
    return a2 @ W3 + b3


def loss_and_gradient(p, x, t, emulate_bf16=False):
    """The mean-squared-error loss and its gradient by backpropagation.

    With emulate_bf16, every operand the device feeds its BF16 matrix
    products is rounded as it is there: the forward's (as in forward()), the
    back-propagated dZ2 = dy w3 (1 - a2^2) and dZ1 = (dZ2 W2^T)(1 - a1^2)
    before they enter the weight-gradient and dA1 products, and the bias
    gradients b1, b2 summed from those rounded values. Without it this is
    the exact FP64 gradient (the same code with rounding the identity)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    rb = _rounding(emulate_bf16)
    x = np.asarray(x, dtype=np.float64).reshape(-1, INPUTS)
    t = np.asarray(t, dtype=np.float64)
    n = len(x)
    xr, W1r, W2r = rb(x), rb(W1), rb(W2)
    a1 = np.tanh(xr @ W1r + b1)
    a1r = rb(a1)
    a2 = np.tanh(a1r @ W2r + b2)
    y = a2 @ W3 + b3
    e = y - t
    loss = np.mean(e * e)
    dy = 2.0 * e / n
    gW3 = a2.T @ dy
    gb3 = dy.sum()
    dz2 = rb(np.outer(dy, W3) * (1.0 - a2 * a2))
    gW2 = a1r.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = rb((dz2 @ W2r.T) * (1.0 - a1 * a1))
    gW1 = xr.T @ dz1
    gb1 = dz1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def loss(p, x, t):
    y = forward(p, x)
    return float(np.mean((y - np.asarray(t, dtype=np.float64)) ** 2))
