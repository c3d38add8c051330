"""Pins the numpy oracle of the learned-seed network (oracle/qpnet_ref.py):
the convolution against scipy.signal.correlate2d, every gradient against
central finite differences, pooling/upsampling as adjoint pairs."""
import numpy as np
from scipy.signal import correlate2d

from oracle import qpnet_ref as nr


def test_conv3x3_is_a_sum_of_correlations():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((7, 9, 3))
    w = rng.standard_normal((9, 2, 3))
    b = rng.standard_normal(2)
    y = nr.conv3x3(x, w, b, relu=False)
    for co in range(2):
        want = np.full(x.shape[:2], b[co])
        for ci in range(3):
            k = w[:, co, ci].reshape(3, 3)        # k[dy+1][dx+1]
            want += correlate2d(x[:, :, ci], k, mode="same")
        assert np.allclose(y[:, :, co], want, atol=1e-12)


def _numgrad(f, a, eps=1e-6):
    g = np.zeros_like(a)
    it = np.nditer(a, flags=["multi_index"])
    for _ in it:
        i = it.multi_index
        old = a[i]
        a[i] = old + eps
        fp = f()
        a[i] = old - eps
        fm = f()
        a[i] = old
        g[i] = (fp - fm) / (2 * eps)
    return g


def test_conv3x3_gradients_match_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((5, 6, 2))
    w = rng.standard_normal((9, 3, 2)) * 0.5
    b = rng.standard_normal(3) * 0.1
    gy = rng.standard_normal((5, 6, 3))
    loss = lambda: float((nr.conv3x3(x, w, b) * gy).sum())   # noqa: E731
    y = nr.conv3x3(x, w, b)
    gx, gw, gb = nr.conv3x3_bwd(x, w, y, gy)
    assert np.allclose(gx, _numgrad(loss, x), atol=1e-6)
    assert np.allclose(gw, _numgrad(loss, w), atol=1e-6)
    assert np.allclose(gb, _numgrad(loss, b), atol=1e-6)


def test_pool_up_head_mse():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 6, 3))
    g = rng.standard_normal((2, 3, 3))
    # adjoint pairs: <pool x, g> = <x, pool_bwd g>, <up g, x> = <g, up_bwd x>
    assert np.isclose((nr.pool2(x) * g).sum(), (x * nr.pool2_bwd(g)).sum())
    assert np.isclose((nr.up2(g) * x).sum(), (g * nr.up2_bwd(x)).sum())
    assert np.allclose(nr.pool2(nr.up2(g)), g)
    h = rng.standard_normal((4, 6, 5))
    w = rng.standard_normal((2, 5))
    b = rng.standard_normal(2)
    o = nr.head(h, w, b)
    assert o.shape == (2, 24) and np.isclose(o[1, 7], b[1] + h.reshape(-1, 5)[7] @ w[1])
    t = rng.standard_normal(o.shape)
    loss, grad = nr.mse(o, t)
    num = _numgrad(lambda: nr.mse(o, t)[0], o)
    assert np.allclose(grad, num, atol=1e-7)
