"""The barostat screen's network: the FP64 oracle pinned by central
differences and a closed form (CPU), and the BF16-MFMA device kernels held to
the oracle: forward, gradient and training (GPU)."""
import numpy as np
import pytest

import surrogate as S
import rainmd_amd as R


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    W1 = rng.uniform(-1, 1, (16, 64)) * np.sqrt(6 / 80) * scale
    W2 = rng.uniform(-1, 1, (64, 64)) * np.sqrt(6 / 128) * scale
    W3 = rng.uniform(-1, 1, 64) * np.sqrt(6 / 65) * scale
    return S.pack(W1, rng.normal(0, 0.1, 64), W2, rng.normal(0, 0.1, 64), W3, 0.3)


def test_oracle_gradient_matches_central_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    x = rng.normal(size=(37, 16))
    t = rng.normal(size=37)
    _, g = S.loss_and_gradient(p, x, t)
    h = 1e-6
    for k in rng.choice(S.PARAMS, 60, replace=False).tolist() + [0, 1024, 1088, 5184, 5248, 5312]:
        e = np.zeros(S.PARAMS)
        e[k] = h
        fd = (S.loss(p + e, x, t) - S.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(fd)), (k, fd, g[k])


def test_oracle_closed_form_with_hidden_layers_off():
    p = _params(3)
    W1, b1, W2, b2, W3, b3 = S.unpack(p)
    p0 = S.pack(0 * W1, b1, 0 * W2, b2, W3, b3)
    x = np.random.default_rng(4).normal(size=(5, 16))
    assert np.allclose(S.forward(p0, x), W3 @ np.tanh(b2) + b3, rtol=0, atol=1e-14)


def test_bf16_oracle_against_central_differences_of_its_own_loss():
    """The BF16-emulating gradient against central differences of the
    BF16-emulating loss where the loss is smooth in the parameter: W3 and b3
    sit after every rounding (exact agreement), b2 before the rounding of
    dZ2 only (agreement within the BF16 step of the summed terms)."""
    rng = np.random.default_rng(21)
    p = _params(22)
    x = rng.normal(size=(41, 16))
    t = rng.normal(size=41)
    _, g = S.loss_and_gradient(p, x, t, emulate_bf16=True)
    _, exact = S.loss_and_gradient(p, x, t)
    h = 1e-6

    def fd(k):
        e = np.zeros(S.PARAMS)
        e[k] = h
        lp = np.mean((S.forward(p + e, x, emulate_bf16=True) - t) ** 2)
        lm = np.mean((S.forward(p - e, x, emulate_bf16=True) - t) ** 2)
        return (lp - lm) / (2 * h)

    for k in list(range(5248, 5313, 7)) + [5312]:
        assert abs(fd(k) - g[k]) <= 1e-7 * max(1.0, abs(fd(k))), k
    # b2: the device sums BF16-rounded dZ2 terms, each within 2^-9 of its value
    W1, b1, W2, b2, W3, b3 = S.unpack(p)
    a1 = S.bf16(np.tanh(S.bf16(x) @ S.bf16(W1) + b1))
    a2 = np.tanh(a1 @ S.bf16(W2) + b2)
    e_ = a2 @ W3 + b3 - t
    bound = (np.abs(np.outer(2 * e_ / len(x), W3) * (1 - a2 * a2)).sum(0)) * 2 ** -8
    for o in range(0, 64, 5):
        assert abs(fd(5184 + o) - g[5184 + o]) <= bound[o] + 1e-9, o
    # the rounding moves the gradient, by what BF16 explains and no more
    r = np.linalg.norm(g - exact) / np.linalg.norm(exact)
    assert 1e-5 < r < 2e-2, r


def test_bf16_rounding_is_nearest_even():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 3.14159], dtype=np.float32)
    out = S.bf16(vals)
    assert out[0] == 1.0 and out[2] == 1.0
    # exact ties go to the even mantissa: 1 + 2^-8 (halfway to 1 + 2^-7) rounds down to 1
    assert out[1] == 1.0
    assert out[3] == 1.0 + 2 ** -7
    assert out[4] == -2.5
    assert abs(out[5] - 3.14159) < 2 ** -7 * 4


@pytest.mark.gpu
def test_device_forward_matches_the_bf16_oracle():
    net = R.Surrogate(seed=11)
    p = _params(5).astype(np.float32)
    net.params = p
    x = np.random.default_rng(6).normal(size=(200, 16)).astype(np.float32)  # 3 full tiles and a ragged one
    y = net.forward(x)
    ref = S.forward(p.astype(np.float64), x, emulate_bf16=True)
    full = S.forward(p.astype(np.float64), x)
    scale = np.abs(full).max()
    print("forward vs bf16 oracle", np.abs(y - ref).max() / scale, "vs fp64", np.abs(y - full).max() / scale)
    assert np.abs(y - ref).max() <= 2e-5 * scale
    assert np.abs(y - full).max() <= 3e-2 * scale
    net.release()


@pytest.mark.gpu
def test_device_gradient_matches_the_oracle():
    """Held to the BF16-emulating oracle (the same roundings, FP64
    elsewhere) at FP32 accumulation error, every block on its own; the
    exact FP64 gradient only bounds what BF16 costs."""
    net = R.Surrogate(seed=12)
    p = _params(7).astype(np.float32)
    net.params = p
    rng = np.random.default_rng(8)
    x = rng.normal(size=(150, 16)).astype(np.float32)  # two full tiles and a ragged one
    t = rng.normal(size=150).astype(np.float32)
    loss, g = net.gradient(x, t)
    ref_loss, ref = S.loss_and_gradient(p.astype(np.float64), x, t, emulate_bf16=True)
    exact_loss, exact = S.loss_and_gradient(p.astype(np.float64), x, t)
    rel = float(np.linalg.norm(g - ref) / np.linalg.norm(ref))
    rel_exact = float(np.linalg.norm(g - exact) / np.linalg.norm(exact))
    print("loss", loss, ref_loss, exact_loss, "gradient vs bf16 oracle", rel, "vs fp64", rel_exact)
    assert abs(loss - ref_loss) <= 1e-6 * ref_loss
    assert rel < 1e-5
    assert abs(loss - exact_loss) <= 2e-2 * exact_loss and rel_exact < 3e-2
    for name, sl in [("W1", slice(0, 1024)), ("b1", slice(1024, 1088)), ("W2", slice(1088, 5184)),
                     ("b2", slice(5184, 5248)), ("W3", slice(5248, 5312)), ("b3", slice(5312, 5313))]:
        e = np.linalg.norm(g[sl] - ref[sl]) / max(np.linalg.norm(ref[sl]), 1e-12)
        print(name, e)
        assert e < 1e-4, name
    net.release()


@pytest.mark.gpu
def test_device_training_fits_a_smooth_function():
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, size=(2048, 16)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 3] ** 2 + 0.2 * x[:, 4]).astype(np.float32)
    t = (t - t.mean()) / t.std()
    net = R.Surrogate(seed=13)
    before = S.loss(net.params.astype(np.float64), x, t)
    last = net.train(x, t, steps=3000, learning_rate=3e-3)
    after = S.loss(net.params.astype(np.float64), x, t)
    held = rng.uniform(-1, 1, size=(512, 16)).astype(np.float32)
    th = np.sin(2 * held[:, 0]) + 0.5 * held[:, 1] * held[:, 2] - 0.3 * held[:, 3] ** 2 + 0.2 * held[:, 4]
    raw = (np.sin(2 * x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 3] ** 2 + 0.2 * x[:, 4])
    th = (th - raw.mean()) / raw.std()
    held_loss = S.loss(net.params.astype(np.float64), held, th)
    print("loss before", before, "after", after, "last minibatches", last, "held out", held_loss)
    assert after < 0.03 * before and after < 0.03 and held_loss < 0.05
    net.release()
