"""The learned-seed network's BF16-MFMA kernels against the numpy oracle
(oracle/qpnet_ref.py) on the same BF16 inputs.

Bars: a kernel that accumulates exact BF16 products in FP32 and stores FP32
(weight, bias and head gradients) agrees to 1e-5 relative; one that stores
BF16 agrees to the BF16 rounding of its output (2^-8 relative) plus FP32
accumulation.  The whole network's gradients are checked against the oracle
composed layer by layer with the same BF16 roundings.
"""
import numpy as np
import pytest

from oracle import qpnet_ref as nr

torch = pytest.importorskip("torch")
qpnet = pytest.importorskip("qpnet")

pytestmark = pytest.mark.gpu


def banner(name):
    print(f"[gpu-test] {name}", flush=True)


def bf(a):
    """numpy float array rounded to BF16 (round to nearest even)."""
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))
    return t.to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)


def dev_bits(a):
    return qpnet.bf16_bits(torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).cuda()).flatten()


def host(bits, shape):
    return qpnet.bits_to_f32(bits).cpu().numpy().astype(np.float64).reshape(shape)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("H,W,cin,cout", [(32, 32, 48, 32), (8, 80, 16, 192), (16, 16, 192, 64)])
def test_conv3x3_forward_and_gradients(H, W, cin, cout):
    banner(f"conv3x3 {H}x{W} {cin}->{cout}")
    rng = np.random.default_rng(H + cin + cout)
    x = bf(rng.standard_normal((H, W, cin)))
    w = bf(rng.standard_normal((9, cout, cin)) * 0.1)
    b = rng.standard_normal(cout).astype(np.float32).astype(np.float64) * 0.1
    want = nr.conv3x3(x, w, b)
    xd, wd = dev_bits(x), dev_bits(w)
    bd = torch.from_numpy(b.astype(np.float32)).cuda()
    yd = torch.zeros(H * W * cout, dtype=torch.int16, device="cuda")
    qpnet.conv3x3(xd, H, W, cin, cin, wd, bd, yd, cout, cout)
    y = host(yd, (H, W, cout))
    assert rel(y, want) < 8e-3
    # gradients through the ReLU, with the GPU's own BF16 output as y
    gy = bf(rng.standard_normal((H, W, cout)))
    gyd = dev_bits(gy)
    gx_want, gw_want, gb_want = nr.conv3x3_bwd(x, w, y, gy)
    part = torch.zeros(32 * 9 * cout * cin, device="cuda")
    gw = torch.zeros(9 * cout * cin, device="cuda")
    gb = torch.zeros(cout, device="cuda")
    qpnet.conv3x3_wgrad(xd, cin, cin, gyd, yd, cout, cout, H, W, part, gw, gb)
    assert rel(gw.cpu().numpy().reshape(9, cout, cin), gw_want) < 1e-5
    assert rel(gb.cpu().numpy(), gb_want) < 1e-5
    wt = torch.zeros_like(wd)
    qpnet.flip_weights(wd, cout, cin, wt)
    gxd = torch.zeros(H * W * cin, dtype=torch.int16, device="cuda")
    qpnet.conv3x3(gyd, H, W, cout, cout, wt, None, gxd, cin, cin, relu=False, mask=yd)
    assert rel(host(gxd, (H, W, cin)), gx_want) < 8e-3


def test_pool_up_head():
    banner("pool / up / head")
    L = qpnet._lib()
    rng = np.random.default_rng(5)
    H, W, C = 8, 12, 16
    x = bf(rng.standard_normal((H, W, C)))
    xd = dev_bits(x)
    pd = torch.zeros(H * W * C // 4, dtype=torch.int16, device="cuda")
    qpnet._call(L.qp_net_pool2, qpnet.stream(), qpnet.ptr(xd), C, qpnet.ptr(pd), C, H, W, C)
    assert rel(host(pd, (H // 2, W // 2, C)), nr.pool2(x)) < 4e-3
    g = bf(rng.standard_normal((H // 2, W // 2, C)))
    gd = dev_bits(g)
    bd = torch.zeros(H * W * C, dtype=torch.int16, device="cuda")
    qpnet._call(L.qp_net_pool2_bwd, qpnet.stream(), qpnet.ptr(gd), C, qpnet.ptr(bd), C, H, W, C, 0)
    assert np.array_equal(host(bd, (H, W, C)), nr.pool2_bwd(g))   # exact: times 1/4
    ud = torch.zeros(H * W * C, dtype=torch.int16, device="cuda")
    qpnet._call(L.qp_net_up2, qpnet.stream(), qpnet.ptr(gd), C, qpnet.ptr(ud), C, H // 2, W // 2, C)
    assert np.array_equal(host(ud, (H, W, C)), nr.up2(g))
    ubd = torch.zeros(H * W * C // 4, dtype=torch.int16, device="cuda")
    qpnet._call(L.qp_net_up2_bwd, qpnet.stream(), qpnet.ptr(xd), C, qpnet.ptr(ubd), C, H // 2, W // 2, C)
    assert rel(host(ubd, (H // 2, W // 2, C)), nr.up2_bwd(x)) < 4e-3
    hw = rng.standard_normal((2, C)).astype(np.float32)
    hb = rng.standard_normal(2).astype(np.float32)
    o = torch.zeros(2 * H * W, device="cuda")
    hwd, hbd = torch.from_numpy(hw).cuda(), torch.from_numpy(hb).cuda()   # held past the launch
    qpnet._call(L.qp_net_head, qpnet.stream(), qpnet.ptr(xd), C, C, qpnet.ptr(hwd),
                qpnet.ptr(hbd), 2, H * W, qpnet.ptr(o))
    assert rel(o.cpu().numpy().reshape(2, -1), nr.head(x, hw, hb)) < 1e-5


def _numpy_unet(net, x):
    """The network composed from the oracle's layers with the GPU's BF16
    roundings; returns the activations and the output."""
    c1, c2, c3, c4 = qpnet.CH
    W = lambda cv: bf(cv.w.cpu().numpy())   # noqa: E731
    B = lambda cv: cv.b.cpu().numpy().astype(np.float64)   # noqa: E731
    e, d = net.enc, net.dec
    A = {}
    A["x0"] = x
    A["e1a"] = bf(nr.conv3x3(x, W(e[0]), B(e[0])))
    A["e1"] = bf(nr.conv3x3(A["e1a"], W(e[1]), B(e[1])))
    A["p1"] = bf(nr.pool2(A["e1"]))
    A["e2a"] = bf(nr.conv3x3(A["p1"], W(e[2]), B(e[2])))
    A["e2"] = bf(nr.conv3x3(A["e2a"], W(e[3]), B(e[3])))
    A["p2"] = bf(nr.pool2(A["e2"]))
    A["e3a"] = bf(nr.conv3x3(A["p2"], W(e[4]), B(e[4])))
    A["e3"] = bf(nr.conv3x3(A["e3a"], W(e[5]), B(e[5])))
    A["p3"] = bf(nr.pool2(A["e3"]))
    A["ba"] = bf(nr.conv3x3(A["p3"], W(e[6]), B(e[6])))
    A["bb"] = bf(nr.conv3x3(A["ba"], W(e[7]), B(e[7])))
    A["cat3"] = np.concatenate([nr.up2(A["bb"]), A["e3"]], axis=2)
    A["d3a"] = bf(nr.conv3x3(A["cat3"], W(d[0]), B(d[0])))
    A["d3b"] = bf(nr.conv3x3(A["d3a"], W(d[1]), B(d[1])))
    A["cat2"] = np.concatenate([nr.up2(A["d3b"]), A["e2"]], axis=2)
    A["d2a"] = bf(nr.conv3x3(A["cat2"], W(d[2]), B(d[2])))
    A["d2b"] = bf(nr.conv3x3(A["d2a"], W(d[3]), B(d[3])))
    A["cat1"] = np.concatenate([nr.up2(A["d2b"]), A["e1"]], axis=2)
    A["d1a"] = bf(nr.conv3x3(A["cat1"], W(d[4]), B(d[4])))
    A["d1b"] = bf(nr.conv3x3(A["d1a"], W(d[5]), B(d[5])))
    out = nr.head(A["d1b"], net.hw.cpu().numpy().astype(np.float64),
                  net.hb.cpu().numpy().astype(np.float64))
    return A, out


def test_unet_forward_backward_against_oracle():
    banner("unet forward/backward")
    H = W = 32
    rng = np.random.default_rng(9)
    net = qpnet.UNet(seed=3)
    for cv in net.enc + net.dec:   # non-zero biases so every ReLU branch is taken
        cv.b.copy_(torch.from_numpy(rng.standard_normal(cv.cout).astype(np.float32) * 0.05))
    planes = rng.standard_normal((qpnet.NIN, H, W)).astype(np.float32)
    pd = torch.from_numpy(planes).cuda().flatten()
    out = net.forward(pd, H * W, W, H, W).cpu().numpy().astype(np.float64)
    x = np.zeros((H, W, 16))
    x[:, :, :qpnet.NIN] = bf(np.transpose(planes, (1, 2, 0)))
    A, want = _numpy_unet(net, x)
    assert rel(out, want) < 5e-2
    tgt = rng.standard_normal(out.shape).astype(np.float32)
    loss = net.backward(torch.from_numpy(tgt).cuda())
    torch.cuda.synchronize()
    ref_loss, gout = nr.mse(out, tgt.astype(np.float64))
    assert abs(loss / ref_loss - 1) < 1e-6
    # the last two layers' gradients from the GPU's own activations
    c1 = qpnet.CH[0]
    hw = net.hw.cpu().numpy().astype(np.float64)
    d1b = host(net.a["d1b"], (H, W, c1))
    g_d1b = (gout.T @ hw).reshape(H, W, c1)
    gw_head = net.ghwb.cpu().numpy().reshape(2, c1 + 1)
    assert rel(gw_head[:, :c1], gout @ d1b.reshape(-1, c1)) < 1e-5
    d1a = host(net.a["d1a"], (H, W, c1))
    _, gw5, gb5 = nr.conv3x3_bwd(d1a, bf(net.dec[5].w.cpu().numpy()), d1b, bf(g_d1b))
    assert rel(net.dec[5].gw.cpu().numpy(), gw5) < 2e-2
    assert rel(net.dec[5].gb.cpu().numpy(), gb5) < 2e-2
    # every parameter received a finite, non-zero gradient
    for cv in net.enc + net.dec:
        g = cv.gw.cpu().numpy()
        assert np.isfinite(g).all() and np.abs(g).max() > 0


def test_unet_backward_every_layer_against_oracle():
    """Every layer's backward from the GPU's own activations and incoming
    gradient: the weight and bias gradients of all 14 convolutions and the
    data gradients through each convolution, pooling and upsampling, with
    the skip connections' sums (the gradient of an encoder output is its
    skip part plus the pooled path's)."""
    banner("unet backward, every layer")
    H = W = 32
    rng = np.random.default_rng(21)
    net = qpnet.UNet(seed=5)
    for cv in net.enc + net.dec:
        cv.b.copy_(torch.from_numpy(rng.standard_normal(cv.cout).astype(np.float32) * 0.05))
    planes = torch.from_numpy(rng.standard_normal((qpnet.NIN, H, W)).astype(np.float32)).cuda()
    out = net.forward(planes.flatten(), H * W, W, H, W).cpu().numpy().astype(np.float64)
    tgt = rng.standard_normal(out.shape).astype(np.float32)
    net.backward(torch.from_numpy(tgt).cuda())
    torch.cuda.synchronize()
    c1, c2, c3, c4 = qpnet.CH
    dims = {"x0": (H, W, 16), "e1a": (H, W, c1), "cat1": (H, W, c2 + c1),
            "p1": (H // 2, W // 2, c1), "e2a": (H // 2, W // 2, c2),
            "cat2": (H // 2, W // 2, c3 + c2), "p2": (H // 4, W // 4, c2),
            "e3a": (H // 4, W // 4, c3), "cat3": (H // 4, W // 4, c4 + c3),
            "p3": (H // 8, W // 8, c3), "ba": (H // 8, W // 8, c4), "bb": (H // 8, W // 8, c4),
            "d3a": (H // 4, W // 4, c3), "d3b": (H // 4, W // 4, c3),
            "d2a": (H // 2, W // 2, c2), "d2b": (H // 2, W // 2, c2),
            "d1a": (H, W, c1), "d1b": (H, W, c1)}
    A = {k: host(net.a[k], v) for k, v in dims.items()}
    G = {k: host(net.g[k], v) for k, v in dims.items()}
    e, d = net.enc, net.dec
    worst = {}

    def layer(name, cv, x, y, gy):
        """numpy gradients of conv cv from the GPU's x, y and gy; checks
        the GPU's weight and bias gradients; returns the data gradient."""
        gx, gw, gb = nr.conv3x3_bwd(x, bf(cv.w.cpu().numpy()), y, gy)
        ew = rel(cv.gw.cpu().numpy(), gw)
        eb = rel(cv.gb.cpu().numpy(), gb)
        worst[name] = max(ew, eb)
        assert ew < 2e-2 and eb < 2e-2, (name, ew, eb)
        return gx

    def same(name, got, want, bar=3e-2):
        err = rel(got, want)
        worst[name] = max(worst.get(name, 0.0), err)
        assert err < bar, (name, err)

    gx = layer("dec5", d[5], A["d1a"], A["d1b"], G["d1b"])
    same("g_d1a", G["d1a"], gx)
    gx = layer("dec4", d[4], A["cat1"], A["d1a"], G["d1a"])
    skip1 = gx[:, :, c2:]
    same("g_d2b", G["d2b"], nr.up2_bwd(G["cat1"][:, :, :c2]))
    same("g_cat1_up", G["cat1"][:, :, :c2], gx[:, :, :c2])
    gx = layer("dec3", d[3], A["d2a"], A["d2b"], G["d2b"])
    same("g_d2a", G["d2a"], gx)
    gx = layer("dec2", d[2], A["cat2"], A["d2a"], G["d2a"])
    skip2 = gx[:, :, c3:]
    same("g_cat2_up", G["cat2"][:, :, :c3], gx[:, :, :c3])
    same("g_d3b", G["d3b"], nr.up2_bwd(G["cat2"][:, :, :c3]))
    gx = layer("dec1", d[1], A["d3a"], A["d3b"], G["d3b"])
    same("g_d3a", G["d3a"], gx)
    gx = layer("dec0", d[0], A["cat3"], A["d3a"], G["d3a"])
    skip3 = gx[:, :, c4:]
    same("g_cat3_up", G["cat3"][:, :, :c4], gx[:, :, :c4])
    same("g_bb", G["bb"], nr.up2_bwd(G["cat3"][:, :, :c4]))
    gx = layer("enc7", e[7], A["ba"], A["bb"], G["bb"])
    same("g_ba", G["ba"], gx)
    gx = layer("enc6", e[6], A["p3"], A["ba"], G["ba"])
    same("g_p3", G["p3"], gx)
    # encoder level 3 output: skip gradient plus the pooled path's
    same("g_e3", G["cat3"][:, :, c4:], skip3 + nr.pool2_bwd(G["p3"]))
    gx = layer("enc5", e[5], A["e3a"], A["cat3"][:, :, c4:], G["cat3"][:, :, c4:])
    same("g_e3a", G["e3a"], gx)
    gx = layer("enc4", e[4], A["p2"], A["e3a"], G["e3a"])
    same("g_p2", G["p2"], gx)
    same("g_e2", G["cat2"][:, :, c3:], skip2 + nr.pool2_bwd(G["p2"]))
    gx = layer("enc3", e[3], A["e2a"], A["cat2"][:, :, c3:], G["cat2"][:, :, c3:])
    same("g_e2a", G["e2a"], gx)
    gx = layer("enc2", e[2], A["p1"], A["e2a"], G["e2a"])
    same("g_p1", G["p1"], gx)
    same("g_e1", G["cat1"][:, :, c2:], skip1 + nr.pool2_bwd(G["p1"]))
    gx = layer("enc1", e[1], A["e1a"], A["cat1"][:, :, c2:], G["cat1"][:, :, c2:])
    same("g_e1a", G["e1a"], gx)
    layer("enc0", e[0], A["x0"], A["e1a"], G["e1a"])
    print("worst relative error per layer:",
          " ".join(f"{k}={v:.1e}" for k, v in worst.items()), flush=True)


def test_network_fits_a_fixed_set_of_input_dependent_targets():
    """Overfitting check of the whole training path (forward, backward,
    Adam): 4 fixed random inputs whose targets are a local function of
    each input (a 3x3 stencil of two channels and a product of two others,
    so the output depends on the input, not only on the biases).  Adam on
    the 4 samples in turn must bring the mean loss below 10 % of where it
    started."""
    banner("unet overfit")
    H = W = 32
    rng = np.random.default_rng(31)
    net = qpnet.UNet(seed=6)
    xs, ys = [], []
    for _ in range(4):
        x = rng.standard_normal((qpnet.NIN, H, W))
        # smooth inputs (the seed's inputs are smooth fields)
        for _ in range(2):
            x = 0.25 * (np.roll(x, 1, 1) + np.roll(x, -1, 1) + np.roll(x, 1, 2) + np.roll(x, -1, 2))
        x /= x.std(axis=(1, 2), keepdims=True)
        t0 = 0.5 * x[0] + 0.25 * (np.roll(x[2], 1, 0) - np.roll(x[2], -1, 0))
        t1 = 0.5 * x[1] * x[4] - 0.3 * x[7]
        xs.append(torch.from_numpy(x.astype(np.float32)).cuda().flatten())
        ys.append(torch.from_numpy(np.stack([t0, t1]).reshape(2, -1).astype(np.float32)).cuda())
    losses = []
    for it in range(600):
        k = it % 4
        net.forward(xs[k], H * W, W, H, W)
        losses.append(net.backward(ys[k]))
        net.adam(lr=1e-3)
    first, last = float(np.mean(losses[:4])), float(np.mean(losses[-4:]))
    print(f"overfit: mean loss {first:.4g} -> {last:.4g}", flush=True)
    assert last < 0.1 * first


def test_training_reduces_the_loss():
    banner("unet training")
    H = W = 32
    rng = np.random.default_rng(11)
    net = qpnet.UNet(seed=4)
    planes = torch.from_numpy(rng.standard_normal((qpnet.NIN, H, W)).astype(np.float32)).cuda()
    yy, xx = np.mgrid[0:H, 0:W] / H
    tgt = np.stack([np.sin(3 * xx) * np.cos(2 * yy), np.cos(4 * yy) * xx]).astype(np.float32)
    tgt_d = torch.from_numpy(tgt.reshape(2, -1)).cuda()
    losses = []
    for _ in range(60):
        net.forward(planes.flatten(), H * W, W, H, W)
        losses.append(net.backward(tgt_d))
        net.adam(lr=3e-3)
    print("loss", losses[0], "->", losses[-1], flush=True)
    assert losses[-1] < 0.3 * losses[0]


def _bilinear_up(c, f, ny1, nx1):
    """numpy restatement of the seed's upsampling: bilinear between coarse
    pixel centres f*T + (f-1)/2, clamped at the edges, zero beyond f*H, f*W."""
    H, W = c.shape
    out = np.zeros((ny1, nx1))

    def w(t, n):
        u = (t - 0.5 * (f - 1)) / f
        if u <= 0:
            return 0, 0, 0.0
        if u >= n - 1:
            return n - 1, n - 1, 0.0
        i0 = int(np.floor(u))
        return i0, i0 + 1, u - i0

    for y in range(min(ny1, f * H)):
        y0, y1, wy = w(y, H)
        for x in range(min(nx1, f * W)):
            x0, x1, wx = w(x, W)
            top = (1 - wx) * c[y0, x0] + wx * c[y0, x1]
            bot = (1 - wx) * c[y1, x0] + wx * c[y1, x1]
            out[y, x] = (1 - wy) * top + wy * bot
    return out


@pytest.mark.parametrize("f", [1, 2, 4])
def test_block_mean_and_seed_apply(f):
    banner(f"block mean / seed apply f={f}")
    rng = np.random.default_rng(20 + f)
    ny, nx = 32, 48
    ld32, ld64 = nx + 7, nx + 3
    a32 = rng.standard_normal((ny + 1, ld32)).astype(np.float32)
    a64 = rng.standard_normal((ny + 1, ld64))
    d32 = torch.from_numpy(a32).cuda()
    d64 = torch.from_numpy(a64).cuda()
    H, W = ny // f, nx // f
    out = torch.zeros((2, H, W), device="cuda")
    qpnet.block_mean([(d32.data_ptr(), ld32, False), (d64.data_ptr(), ld64, True)], f, H, W, out)
    got = out.cpu().numpy()
    for c, src in enumerate((a32.astype(np.float64), a64)):
        want = src[: f * H, : f * W].reshape(H, f, W, f).mean(axis=(1, 3))
        assert np.abs(got[c] - want).max() < 1e-5 * max(1.0, np.abs(want).max())
    # seed: 2 a - b + s up(corr), into an FP64 and an FP32 destination
    hist = torch.from_numpy(rng.standard_normal((2, 2, ny + 1, nx + 1)).astype(np.float32)).cuda()
    corr = torch.from_numpy(rng.standard_normal((2, H, W)).astype(np.float32)).cuda()
    s = 0.37
    h = hist.cpu().numpy().astype(np.float64)
    cc = corr.cpu().numpy().astype(np.float64)
    want = np.stack([2 * h[0, c] - h[1, c] + s * _bilinear_up(cc[c], f, ny + 1, nx + 1)
                     for c in range(2)])
    for dbl, ldd in ((True, nx + 5), (False, nx + 9)):
        pl = (ny + 1) * ldd + 11
        dst = torch.zeros(2 * pl, dtype=torch.float64 if dbl else torch.float32, device="cuda")
        qpnet.seed_apply(corr, H, W, f, s, hist, dst.data_ptr(), dbl, ldd, pl, ny + 1, nx + 1)
        d = dst.cpu().numpy().astype(np.float64)
        got = np.stack([d[c * pl: c * pl + (ny + 1) * ldd].reshape(ny + 1, ldd)[:, : nx + 1]
                        for c in range(2)])
        assert np.abs(got - want).max() < 2e-5 * np.abs(want).max()


def test_seed_combine():
    banner("seed combine")
    rng = np.random.default_rng(31)
    n = 1000
    for dbl in (False, True):
        dt = torch.float64 if dbl else torch.float32
        e0, h0, c0 = (rng.standard_normal(n) for _ in range(3))
        e = torch.tensor(e0, dtype=dt, device="cuda")
        h = torch.tensor(h0, dtype=dt, device="cuda")
        c = torch.tensor(c0, dtype=dt, device="cuda")
        qpnet.seed_combine(e.data_ptr(), h.data_ptr(), c.data_ptr(), n, True, dbl)
        tol = 1e-12 if dbl else 1e-5
        assert np.abs(e.cpu().numpy() - (2 * e0 - h0 + c0)).max() < tol * 10
        assert np.abs(h.cpu().numpy() - e0).max() < tol
        qpnet.seed_combine(e.data_ptr(), h.data_ptr(), None, n, False, dbl)   # record only
        assert np.array_equal(h.cpu().numpy(), e.cpu().numpy())


def test_forward_graph_replays_match_eager():
    """The captured forward replays to the eager result, with eager work and
    synchronisations between replays, and sees weight updates (Adam writes
    the BF16 weights in place)."""
    banner("forward graph")
    H = W = 32
    rng = np.random.default_rng(12)
    net = qpnet.UNet(seed=5)
    x = torch.from_numpy(rng.standard_normal((qpnet.NIN, H, W)).astype(np.float32)).cuda()
    want = net.forward(x, H * W, W, H, W).clone()
    for r in range(4):
        got = net.forward_graph(x, H * W, W, H, W).clone()
        torch.cuda.synchronize()
        assert torch.equal(got, want), r
        y = torch.randn(1000, device="cuda").sum()   # eager work between replays
        torch.cuda.synchronize()
        del y
    tgt = torch.zeros(2, H * W, device="cuda")
    net.backward(tgt)
    net.adam(lr=1e-2)
    want2 = net.forward(x, H * W, W, H, W).clone()
    assert not torch.equal(want2, want)
    got2 = net.forward_graph(x, H * W, W, H, W).clone()
    torch.cuda.synchronize()
    assert torch.equal(got2, want2)
    net.drop_graph()
