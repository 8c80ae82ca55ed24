"""Pins the float64 network oracle (oracle/cnn.py) by closed forms and finite
differences: it is the checker of the gfx950 network kernels."""
import numpy as np

from oracle import cnn


def test_conv_identity_shift_and_linearity():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 16, 16, 32))
    w = np.zeros((32, 3, 3, 32))
    w[np.arange(32), 1, 1, np.arange(32)] = 1.0          # centre tap: identity
    assert np.allclose(cnn.conv3x3(x, w), x)
    w2 = np.zeros((32, 3, 3, 32))
    w2[np.arange(32), 0, 2, np.arange(32)] = 1.0         # y[i, j] = x[i - 1, j + 1]
    y = cnn.conv3x3(x, w2)
    assert np.allclose(y[:, 1:, :-1], x[:, :-1, 1:])
    assert np.allclose(y[:, 0], 0.0) and np.allclose(y[:, :, -1], 0.0)  # zero padding
    wa, wb = rng.standard_normal((2, 32, 3, 3, 32))
    b = rng.standard_normal(32)
    assert np.allclose(cnn.conv3x3(x, wa + 2 * wb, b), cnn.conv3x3(x, wa, b) + 2 * cnn.conv3x3(x, wb))


def test_conv_backward_matches_finite_differences():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((1, 16, 16, 32))
    w = rng.standard_normal((32, 3, 3, 32)) * 0.1
    b = rng.standard_normal(32)
    dy = rng.standard_normal((1, 16, 16, 32))
    dw, db, dx = cnn.conv3x3_backward(x, w, dy)
    loss = lambda x_, w_, b_: float((cnn.conv3x3(x_, w_, b_) * dy).sum())
    h = 1e-6
    for idx in [(0, 0, 0, 0), (5, 2, 1, 17), (31, 1, 2, 3)]:
        e = np.zeros_like(w)
        e[idx] = h
        assert abs((loss(x, w + e, b) - loss(x, w - e, b)) / (2 * h) - dw[idx]) < 1e-5
    for idx in [(0, 0, 0, 0), (0, 7, 15, 9), (0, 15, 0, 31)]:
        e = np.zeros_like(x)
        e[idx] = h
        assert abs((loss(x + e, w, b) - loss(x - e, w, b)) / (2 * h) - dx[idx]) < 1e-5
    e = np.zeros(32)
    e[4] = h
    assert abs((loss(x, w, b + e) - loss(x, w, b - e)) / (2 * h) - db[4]) < 1e-5


def test_softmax_xent_gradient():
    rng = np.random.default_rng(3)
    z = rng.standard_normal(50)
    loss, g = cnn.softmax_xent(z, 7)
    assert abs(loss - (np.log(np.exp(z).sum()) - z[7])) < 1e-12
    h = 1e-6
    for i in (0, 7, 49):
        e = np.zeros(50)
        e[i] = h
        fd = (cnn.softmax_xent(z + e, 7)[0] - cnn.softmax_xent(z - e, 7)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7


def test_bf16_round():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], dtype=np.float32)
    r = cnn.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6  # ties to even
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2 ** -8
