"""CPU oracle of the learned-seed network's layers -- TEST INFRASTRUCTURE ONLY.

Plain numpy FP64 restatements of what each kernel of include/qpnet.h
computes, used by tests/test_gpu_net.py as the checker: 3x3 'same'
convolution (NHWC, weights [tap][cout][cin], tap = 3 (dy+1) + dx + 1) with
bias and ReLU, its data and weight gradients through the ReLU, 2x2 average
pooling and nearest upsampling with their gradients, the 1x1 output layer
and the mean squared error.  The convolution is also checked against
scipy.signal.correlate2d (a textbook routine) in the same tests.
"""
import numpy as np

OFFS = [(t // 3 - 1, t % 3 - 1) for t in range(9)]


def shift(x, dy, dx):
    """s[y, x] = x[y + dy, x + dx] with zeros outside (x: H x W x C)."""
    H, W = x.shape[:2]
    s = np.zeros_like(x)
    ys, ye = max(0, -dy), min(H, H - dy)
    xs, xe = max(0, -dx), min(W, W - dx)
    s[ys:ye, xs:xe] = x[ys + dy:ye + dy, xs + dx:xe + dx]
    return s


def conv3x3(x, w, b=None, relu=True):
    """y[p][co] = b[co] + sum_{tap, ci} x[p + off(tap)][ci] w[tap][co][ci]"""
    H, W, _ = x.shape
    y = np.zeros((H, W, w.shape[1]))
    for t, (dy, dx) in enumerate(OFFS):
        y += shift(x, dy, dx) @ w[t].T
    if b is not None:
        y += b
    return np.maximum(y, 0.0) if relu else y


def conv3x3_bwd(x, w, y, gy):
    """Gradients of conv3x3 with ReLU output y: (gx, gw, gb)."""
    g = gy * (y > 0)
    gb = g.sum(axis=(0, 1))
    gw = np.stack([g.reshape(-1, g.shape[2]).T @ shift(x, dy, dx).reshape(-1, x.shape[2])
                   for dy, dx in OFFS])
    gx = np.zeros_like(x)
    for t, (dy, dx) in enumerate(OFFS):
        gx += shift(g @ w[t], -dy, -dx)
    return gx, gw, gb


def pool2(x):
    return 0.25 * (x[0::2, 0::2] + x[0::2, 1::2] + x[1::2, 0::2] + x[1::2, 1::2])


def pool2_bwd(g):
    return 0.25 * np.repeat(np.repeat(g, 2, axis=0), 2, axis=1)


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=0), 2, axis=1)


def up2_bwd(g):
    return g[0::2, 0::2] + g[0::2, 1::2] + g[1::2, 0::2] + g[1::2, 1::2]


def head(h, w, b):
    """o[c][p] = b[c] + sum_k w[c][k] h[p][k]; h: H x W x cin."""
    return (h.reshape(-1, h.shape[2]) @ w.T + b).T


def mse(o, t):
    d = o - t
    return float((d * d).mean()), 2.0 * d / d.size
