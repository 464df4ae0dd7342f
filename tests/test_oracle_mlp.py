"""Pins of the FP64 network oracle: its gradient against central finite differences of its own
loss (so a wrong backward term, transposed weight or dropped GELU factor fails), the loss of
a network whose outputs are known in closed form, and the parameter layout."""
import numpy as np

from oracle import oracle_py as O


def _random_params(rng, scale=1.0):
    P = O.mlp_param_count()
    dims = [(256, 5), (256, 256), (256, 256), (256, 256), (6, 256)]
    out, off = np.zeros(P, np.float32), 0
    for o, i in dims:
        out[off:off + o * i] = rng.uniform(-1, 1, o * i) * scale * np.sqrt(6.0 / i)
        off += o * i
        out[off:off + o] = rng.normal(size=o) * 0.1
        off += o
    assert off == P
    return out


def test_param_count():
    assert O.mlp_param_count() == 256 * 5 + 256 + 3 * (256 * 256 + 256) + 6 * 256 + 6


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(11)
    p = _random_params(rng)
    n = 6
    x = rng.uniform(-1, 1, (n, 5)).astype(np.float32)
    cls = np.array([0, 1, 2, 0, 0, 1], np.int32)
    d = rng.normal(scale=0.3, size=(n, 3)).astype(np.float32)
    loss, g = O.mlp_loss_grad(p, x, cls, d)
    # every layer's weights and biases, including the first and last
    P = len(p)
    idx = np.r_[rng.choice(1280, 20, replace=False), 1280 + rng.choice(256, 5, replace=False),
                rng.choice(np.arange(1536, P - 1542), 40, replace=False), P - 1542 + rng.choice(1536, 20, replace=False),
                P - 6 + np.arange(6)]
    pd = p.astype(np.float64)
    for j in idx:
        eps = 1e-3 * max(1.0, abs(pd[j]))
        # the oracle takes float32 parameters, so step in a way float32 represents exactly
        hi, lo = p.copy(), p.copy()
        hi[j] = np.float32(pd[j] + eps)
        lo[j] = np.float32(pd[j] - eps)
        fd = (O.mlp_loss_grad(hi, x, cls, d)[0] - O.mlp_loss_grad(lo, x, cls, d)[0]) / (float(hi[j]) - float(lo[j]))
        assert abs(fd - g[j]) < 1e-5 * max(1.0, abs(g[j])) + 1e-7, (j, fd, g[j])


def test_zero_network_loss_is_closed_form():
    """All-zero parameters: logits 0, residual 0, so the loss is ln 3 + w |target|^2 for escaped
    samples; the output-bias gradient is (1/3 - onehot, -2 w target) averaged."""
    P = O.mlp_param_count()
    p = np.zeros(P, np.float32)
    x = np.random.default_rng(2).uniform(-1, 1, (4, 5)).astype(np.float32)
    cls = np.array([0, 1, 2, 0], np.int32)
    d = np.array([[0.1, -0.2, 0.3], [1, 1, 1], [1, 1, 1], [0.0, 0.5, 0.0]], np.float32)
    loss, g = O.mlp_loss_grad(p, x, cls, d)
    w = 100.0
    esc = cls == 0
    expect = np.log(3.0) + w * (d[esc].astype(np.float64) ** 2).sum() / 4
    assert abs(loss - expect) < 1e-12
    gb = g[-6:]
    onehot = np.eye(3)[cls]
    assert np.allclose(gb[:3], (1.0 / 3 - onehot).mean(0), atol=1e-15)
    assert np.allclose(gb[3:], (-2 * w * d * esc[:, None]).mean(0), atol=1e-12)
    out = O.mlp_forward(p, x)
    assert np.all(out == 0)


def test_bf16_rounding_is_small_and_deterministic():
    rng = np.random.default_rng(5)
    p = _random_params(rng)
    x = rng.uniform(-1, 1, (16, 5)).astype(np.float32)
    a = O.mlp_forward(p, x, round_bf16=False)
    b = O.mlp_forward(p, x, round_bf16=True)
    assert np.array_equal(b, O.mlp_forward(p, x, round_bf16=True))
    rel = np.abs(a - b).max() / np.abs(a).max()
    assert 1e-5 < rel < 5e-2, rel
