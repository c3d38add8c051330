"""The learned corrector seed: a 4-level U-Net (3x3 convolutions, 16-32-64-128
channels, average pooling, nearest upsampling, skip concatenations, a 1x1
FP32 output layer) whose forward and backward passes are the library's own
BF16-MFMA kernels (include/qpnet.h).  PyTorch only provides device memory and
the stream.

Input channels (8): B_perp of the two previous slices (real space, 4), and of
the current slice the plasma charge, the predictor current J_x, J_y and psi.
Output (2): the correction of B_perp over its linear extrapolation
2 B(xi-1) - B(xi-2).  The network only seeds the corrector; the plain
iteration to the residual tolerance decides the answer.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

import qpslice as qps

CH = (16, 32, 64, 128)
NIN, NOUT = 8, 2

_NET = None


def _lib():
    global _NET
    if _NET is None:
        L = qps.lib()
        vp, i, sz, f, u16p, fp, dp = (C.c_void_p, C.c_int, C.c_size_t, C.c_float, C.c_void_p,
                                      C.c_void_p, C.c_void_p)
        sig = {
            "qp_net_conv3x3": [vp, u16p, u16p, i, i, i, i, u16p, fp, u16p, i, i, i, i],
            "qp_net_conv3x3_wgrad": [vp, u16p, i, i, u16p, u16p, i, i, i, i, i, fp, fp, fp],
            "qp_net_flip_weights": [vp, u16p, i, i, u16p],
            "qp_net_pool2": [vp, u16p, i, u16p, i, i, i, i],
            "qp_net_pool2_bwd": [vp, u16p, i, u16p, i, i, i, i, i],
            "qp_net_up2": [vp, u16p, i, u16p, i, i, i, i],
            "qp_net_up2_bwd": [vp, u16p, i, u16p, i, i, i, i],
            "qp_net_pack_input": [vp, fp, i, sz, i, fp, i, i, u16p, i],
            "qp_net_head": [vp, u16p, i, i, fp, fp, i, sz, fp],
            "qp_net_head_bwd": [vp, u16p, i, i, fp, i, sz, fp, u16p, i, fp, fp],
            "qp_net_mse": [vp, fp, fp, sz, fp, dp, i],
            "qp_net_adam": [vp, fp, fp, fp, fp, sz, f, f, f, f, i, u16p],
            "qp_net_to_bf16": [vp, fp, sz, u16p],
            "qp_net_block_mean": [vp, i, C.POINTER(C.c_void_p), C.POINTER(C.c_int), i, i, i, i,
                                  fp],
            "qp_net_seed_apply": [vp, fp, i, i, i, f, fp, fp, i, sz, vp, i, i, sz, i, i],
            "qp_net_seed_combine": [vp, vp, vp, vp, sz, i, i],
            "qp_net_graph_begin": [vp],
            "qp_net_graph_end": [vp, C.POINTER(C.c_void_p)],
            "qp_net_graph_launch": [vp, vp],
            "qp_net_graph_destroy": [vp],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _NET = L
    return _NET


def _call(fn, *args):
    rc = fn(*args)
    if rc != 0:
        raise qps.QPError(rc, qps.lib().qp_last_error().decode())


def ptr(t, offset=0):
    """Device address of a tensor, plus an element offset."""
    return C.c_void_p(t.data_ptr() + offset * t.element_size()) if t is not None else None


def stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def block_mean(planes, f, H, W, out):
    """planes: list of (device address, row pitch, is_fp64); out: FP32
    tensor [len(planes)][H][W] of f x f block means."""
    n = len(planes)
    ps = (C.c_void_p * n)(*[p for p, _, _ in planes])
    lds = (C.c_int * n)(*[ld for _, ld, _ in planes])
    dbl = sum(1 << c for c, (_, _, d) in enumerate(planes) if d)
    _call(_lib().qp_net_block_mean, stream(), n, ps, lds, dbl, f, H, W, ptr(out))


def seed_apply(corr, H, W, f, s, hist, dst, dst_double, ld_d, pl_d, ny1, nx1):
    """dst <- 2 hist[0] - hist[1] + s up(corr) (see qpnet.h), or s up(corr)
    when hist is None; hist: FP32 tensor [2][2][ny1][nx1]; dst: device address
    of the 2-component plane."""
    a, b = (ptr(hist[0]), ptr(hist[1])) if hist is not None else (None, None)
    _call(_lib().qp_net_seed_apply, stream(), ptr(corr), H, W, f, s, a, b, nx1, ny1 * nx1,
          C.c_void_p(dst), 1 if dst_double else 0, ld_d, pl_d, ny1, nx1)


def seed_combine(epw, hist, corr, n, extrap, dbl):
    """k space: e = epw; epw = 2 e - hist (+ corr) if extrap; hist = e
    (device addresses; corr may be None)."""
    _call(_lib().qp_net_seed_combine, stream(), C.c_void_p(epw), C.c_void_p(hist),
          C.c_void_p(corr) if corr else None, n, 1 if extrap else 0, 1 if dbl else 0)


def bf16_bits(a: torch.Tensor) -> torch.Tensor:
    """float32 -> BF16 bit patterns (int16 tensor), round to nearest even."""
    return a.to(torch.bfloat16).view(torch.int16)


def bits_to_f32(b: torch.Tensor) -> torch.Tensor:
    return b.view(torch.bfloat16).to(torch.float32)


# ---------------------------------------------------------------------------
# layer-level entry points (also used by the tests)
# ---------------------------------------------------------------------------
def conv3x3(x, H, W, cs_in, cin, w16, bias, out, cs_out, cout, relu=True, accum=False,
            mask=None, x_off=0, out_off=0):
    _call(_lib().qp_net_conv3x3, stream(), ptr(x, x_off), ptr(mask, x_off) if mask is not None
          else None, H, W, cs_in, cin, ptr(w16), ptr(bias), ptr(out, out_off), cs_out, cout,
          1 if relu else 0, 1 if accum else 0)


def conv3x3_wgrad(x, cs_in, cin, dy, yout, cs_out, cout, H, W, part, dw, db, x_off=0,
                  y_off=0, nchunk=32):
    _call(_lib().qp_net_conv3x3_wgrad, stream(), ptr(x, x_off), cs_in, cin, ptr(dy, y_off),
          ptr(yout, y_off), cs_out, cout, H, W, nchunk, ptr(part), ptr(dw), ptr(db))


def flip_weights(w16, cout, cin, wt):
    _call(_lib().qp_net_flip_weights, stream(), ptr(w16), cout, cin, ptr(wt))


class Conv:
    """3x3 convolution with ReLU: FP32 master weights [9][cout][cin], their
    BF16 copy, FP32 bias, gradients and Adam moments."""

    def __init__(self, cin, cout, gen, dev):
        self.cin, self.cout = cin, cout
        fan_in = 9 * cin
        w = torch.randn((9, cout, cin), generator=gen) * np.sqrt(2.0 / fan_in)
        self.w = w.to(dev)
        self.b = torch.zeros(cout, device=dev)
        self.w16 = bf16_bits(self.w)
        self.wt16 = torch.empty_like(self.w16)
        self.gw = torch.zeros_like(self.w)
        self.gb = torch.zeros_like(self.b)
        self.m = [torch.zeros_like(self.w), torch.zeros_like(self.b)]
        self.v = [torch.zeros_like(self.w), torch.zeros_like(self.b)]

    def params(self):
        return [(self.w, self.gw, self.m[0], self.v[0], self.w16), (self.b, self.gb, self.m[1],
                                                                     self.v[1], None)]


class UNet:
    """The 4-level U-Net.  Buffers are allocated per input size (H, W
    divisible by 8)."""

    def __init__(self, seed=0, device="cuda"):
        gen = torch.Generator().manual_seed(seed)
        self.dev = torch.device(device)
        c1, c2, c3, c4 = CH
        self.enc = [Conv(16, c1, gen, self.dev), Conv(c1, c1, gen, self.dev),
                    Conv(c1, c2, gen, self.dev), Conv(c2, c2, gen, self.dev),
                    Conv(c2, c3, gen, self.dev), Conv(c3, c3, gen, self.dev),
                    Conv(c3, c4, gen, self.dev), Conv(c4, c4, gen, self.dev)]
        self.dec = [Conv(c4 + c3, c3, gen, self.dev), Conv(c3, c3, gen, self.dev),
                    Conv(c3 + c2, c2, gen, self.dev), Conv(c2, c2, gen, self.dev),
                    Conv(c2 + c1, c1, gen, self.dev), Conv(c1, c1, gen, self.dev)]
        hw = torch.randn((NOUT, c1), generator=gen) * np.sqrt(1.0 / c1) * 0.1
        self.hw = hw.to(self.dev)
        self.hb = torch.zeros(NOUT, device=self.dev)
        self.ghwb = torch.zeros(NOUT * (c1 + 1), device=self.dev)
        self.hm = [torch.zeros_like(self.hw), torch.zeros_like(self.hb)]
        self.hv = [torch.zeros_like(self.hw), torch.zeros_like(self.hb)]
        self.in_scale = torch.ones(NIN, device=self.dev)
        self.out_scale = 1.0
        self.step = 0
        self._shape = None
        self._graph = None     # (key, executable graph) of the captured forward

    # -- buffers ------------------------------------------------------------
    def _alloc(self, H, W):
        if self._shape == (H, W):
            return
        assert H % 8 == 0 and W % 8 == 0
        c1, c2, c3, c4 = CH
        z = lambda h, w, c: torch.zeros((h * w * c,), dtype=torch.int16, device=self.dev)  # noqa
        s = {}
        s["x0"] = z(H, W, 16)
        s["e1a"] = z(H, W, c1)
        s["cat1"] = z(H, W, c2 + c1)
        s["p1"] = z(H // 2, W // 2, c1)
        s["e2a"] = z(H // 2, W // 2, c2)
        s["cat2"] = z(H // 2, W // 2, c3 + c2)
        s["p2"] = z(H // 4, W // 4, c2)
        s["e3a"] = z(H // 4, W // 4, c3)
        s["cat3"] = z(H // 4, W // 4, c4 + c3)
        s["p3"] = z(H // 8, W // 8, c3)
        s["ba"] = z(H // 8, W // 8, c4)
        s["bb"] = z(H // 8, W // 8, c4)
        s["d3a"] = z(H // 4, W // 4, c3)
        s["d3b"] = z(H // 4, W // 4, c3)
        s["d2a"] = z(H // 2, W // 2, c2)
        s["d2b"] = z(H // 2, W // 2, c2)
        s["d1a"] = z(H, W, c1)
        s["d1b"] = z(H, W, c1)
        self.a = s
        self.g = {k: torch.zeros_like(v) for k, v in s.items()}
        self.out = torch.zeros((NOUT, H * W), device=self.dev)
        self.gout = torch.zeros_like(self.out)
        npmax = H * W * 192
        self.part = torch.zeros(max(32 * 9 * 192 * 128, 64 * NOUT * (c1 + 1)), device=self.dev)
        self.mse_part = torch.zeros(256, dtype=torch.float64, device=self.dev)
        self._shape = (H, W)
        del npmax

    # -- forward --------------------------------------------------------------
    def forward(self, planes, plane, ld, H, W):
        """planes: FP32 device tensor holding NIN planes (plane elements apart,
        row pitch ld); returns the FP32 output [NOUT][H*W] (in output units)."""
        self._alloc(H, W)
        a = self.a
        c1, c2, c3, c4 = CH
        _call(_lib().qp_net_pack_input, stream(), ptr(planes), NIN, plane, ld, ptr(self.in_scale),
              H, W, ptr(a["x0"]), 16)
        e, d = self.enc, self.dec
        conv3x3(a["x0"], H, W, 16, 16, e[0].w16, e[0].b, a["e1a"], c1, c1)
        conv3x3(a["e1a"], H, W, c1, c1, e[1].w16, e[1].b, a["cat1"], c2 + c1, c1, out_off=c2)
        _call(_lib().qp_net_pool2, stream(), ptr(a["cat1"], c2), c2 + c1, ptr(a["p1"]), c1, H, W, c1)
        H2, W2 = H // 2, W // 2
        conv3x3(a["p1"], H2, W2, c1, c1, e[2].w16, e[2].b, a["e2a"], c2, c2)
        conv3x3(a["e2a"], H2, W2, c2, c2, e[3].w16, e[3].b, a["cat2"], c3 + c2, c2, out_off=c3)
        _call(_lib().qp_net_pool2, stream(), ptr(a["cat2"], c3), c3 + c2, ptr(a["p2"]), c2, H2, W2,
              c2)
        H4, W4 = H // 4, W // 4
        conv3x3(a["p2"], H4, W4, c2, c2, e[4].w16, e[4].b, a["e3a"], c3, c3)
        conv3x3(a["e3a"], H4, W4, c3, c3, e[5].w16, e[5].b, a["cat3"], c4 + c3, c3, out_off=c4)
        _call(_lib().qp_net_pool2, stream(), ptr(a["cat3"], c4), c4 + c3, ptr(a["p3"]), c3, H4, W4,
              c3)
        H8, W8 = H // 8, W // 8
        conv3x3(a["p3"], H8, W8, c3, c3, e[6].w16, e[6].b, a["ba"], c4, c4)
        conv3x3(a["ba"], H8, W8, c4, c4, e[7].w16, e[7].b, a["bb"], c4, c4)
        _call(_lib().qp_net_up2, stream(), ptr(a["bb"]), c4, ptr(a["cat3"]), c4 + c3, H8, W8, c4)
        conv3x3(a["cat3"], H4, W4, c4 + c3, c4 + c3, d[0].w16, d[0].b, a["d3a"], c3, c3)
        conv3x3(a["d3a"], H4, W4, c3, c3, d[1].w16, d[1].b, a["d3b"], c3, c3)
        _call(_lib().qp_net_up2, stream(), ptr(a["d3b"]), c3, ptr(a["cat2"]), c3 + c2, H4, W4, c3)
        conv3x3(a["cat2"], H2, W2, c3 + c2, c3 + c2, d[2].w16, d[2].b, a["d2a"], c2, c2)
        conv3x3(a["d2a"], H2, W2, c2, c2, d[3].w16, d[3].b, a["d2b"], c2, c2)
        _call(_lib().qp_net_up2, stream(), ptr(a["d2b"]), c2, ptr(a["cat1"]), c2 + c1, H2, W2, c2)
        conv3x3(a["cat1"], H, W, c2 + c1, c2 + c1, d[4].w16, d[4].b, a["d1a"], c1, c1)
        conv3x3(a["d1a"], H, W, c1, c1, d[5].w16, d[5].b, a["d1b"], c1, c1)
        _call(_lib().qp_net_head, stream(), ptr(a["d1b"]), c1, c1, ptr(self.hw), ptr(self.hb), NOUT,
              H * W, ptr(self.out))
        return self.out

    def forward_graph(self, planes, plane, ld, H, W):
        """forward() replayed from a HIP graph captured on the first call
        with these buffers and sizes (the weights are read at replay, so
        training between replays is seen); returns the output buffer.  The
        legacy null stream cannot be captured: from it, the graph runs on a
        side stream ordered after and before the caller's work."""
        cur = torch.cuda.current_stream()
        if cur.cuda_stream == 0:
            if getattr(self, "_side", None) is None:
                self._side = torch.cuda.Stream()
            self._side.wait_stream(cur)
            with torch.cuda.stream(self._side):
                out = self.forward_graph(planes, plane, ld, H, W)
            cur.wait_stream(self._side)
            return out
        key = (planes.data_ptr(), plane, ld, H, W, cur.cuda_stream)
        if self._graph is None or self._graph[0] != key:
            self.drop_graph()
            self.forward(planes, plane, ld, H, W)   # buffers allocated outside the capture
            st = stream()
            _call(_lib().qp_net_graph_begin, st)
            try:
                self.forward(planes, plane, ld, H, W)
            finally:
                ge = C.c_void_p()
                _call(_lib().qp_net_graph_end, st, C.byref(ge))
            self._graph = (key, ge)
            return self.out
        _call(_lib().qp_net_graph_launch, self._graph[1], stream())
        return self.out

    def drop_graph(self):
        if self._graph is not None:
            _call(_lib().qp_net_graph_destroy, self._graph[1])
            self._graph = None

    # -- backward -------------------------------------------------------------
    def _conv_bwd(self, cv, x, cs_in, H, W, y, gy, cs_out, gx, x_off=0, y_off=0, gx_off=0,
                  accum=False):
        """Gradients of one conv: weights and bias into cv.gw/gb, data into gx
        (if given), from gy (the gradient of its ReLU output y)."""
        conv3x3_wgrad(x, cs_in, cv.cin, gy, y, cs_out, cv.cout, H, W, self.part, cv.gw, cv.gb,
                      x_off=x_off, y_off=y_off)
        if gx is not None:
            flip_weights(cv.w16, cv.cout, cv.cin, cv.wt16)
            conv3x3(gy, H, W, cs_out, cv.cout, cv.wt16, None, gx, cs_in, cv.cin, relu=False,
                    accum=accum, mask=y, x_off=y_off, out_off=gx_off)

    def backward(self, target):
        """MSE against target [NOUT][H*W] (output units); fills every
        gradient; returns the loss."""
        H, W = self._shape
        a, g = self.a, self.g
        c1, c2, c3, c4 = CH
        n = NOUT * H * W
        _call(_lib().qp_net_mse, stream(), ptr(self.out), ptr(target), n, ptr(self.gout),
              ptr(self.mse_part), 256)
        _call(_lib().qp_net_head_bwd, stream(), ptr(a["d1b"]), c1, c1, ptr(self.hw), NOUT, H * W,
              ptr(self.gout), ptr(g["d1b"]), 64, ptr(self.part), ptr(self.ghwb))
        e, d = self.enc, self.dec
        H2, W2, H4, W4, H8, W8 = H // 2, W // 2, H // 4, W // 4, H // 8, W // 8
        # decoder level 1
        self._conv_bwd(d[5], a["d1a"], c1, H, W, a["d1b"], g["d1b"], c1, g["d1a"])
        self._conv_bwd(d[4], a["cat1"], c2 + c1, H, W, a["d1a"], g["d1a"], c1, g["cat1"])
        _call(_lib().qp_net_up2_bwd, stream(), ptr(g["cat1"]), c2 + c1, ptr(g["d2b"]), c2, H2, W2,
              c2)
        # decoder level 2
        self._conv_bwd(d[3], a["d2a"], c2, H2, W2, a["d2b"], g["d2b"], c2, g["d2a"])
        self._conv_bwd(d[2], a["cat2"], c3 + c2, H2, W2, a["d2a"], g["d2a"], c2, g["cat2"])
        _call(_lib().qp_net_up2_bwd, stream(), ptr(g["cat2"]), c3 + c2, ptr(g["d3b"]), c3, H4, W4,
              c3)
        # decoder level 3
        self._conv_bwd(d[1], a["d3a"], c3, H4, W4, a["d3b"], g["d3b"], c3, g["d3a"])
        self._conv_bwd(d[0], a["cat3"], c4 + c3, H4, W4, a["d3a"], g["d3a"], c3, g["cat3"])
        _call(_lib().qp_net_up2_bwd, stream(), ptr(g["cat3"]), c4 + c3, ptr(g["bb"]), c4, H8, W8,
              c4)
        # bottleneck
        self._conv_bwd(e[7], a["ba"], c4, H8, W8, a["bb"], g["bb"], c4, g["ba"])
        self._conv_bwd(e[6], a["p3"], c3, H8, W8, a["ba"], g["ba"], c4, g["p3"])
        # encoder level 3: e3 = cat3[c4:], its gradient = skip part + pooled path
        _call(_lib().qp_net_pool2_bwd, stream(), ptr(g["p3"]), c3, ptr(g["cat3"], c4), c4 + c3,
              H4, W4, c3, 1)
        self._conv_bwd(e[5], a["e3a"], c3, H4, W4, a["cat3"], g["cat3"], c4 + c3, g["e3a"],
                       y_off=c4)
        self._conv_bwd(e[4], a["p2"], c2, H4, W4, a["e3a"], g["e3a"], c3, g["p2"])
        _call(_lib().qp_net_pool2_bwd, stream(), ptr(g["p2"]), c2, ptr(g["cat2"], c3), c3 + c2,
              H2, W2, c2, 1)
        self._conv_bwd(e[3], a["e2a"], c2, H2, W2, a["cat2"], g["cat2"], c3 + c2, g["e2a"],
                       y_off=c3)
        self._conv_bwd(e[2], a["p1"], c1, H2, W2, a["e2a"], g["e2a"], c2, g["p1"])
        _call(_lib().qp_net_pool2_bwd, stream(), ptr(g["p1"]), c1, ptr(g["cat1"], c2), c2 + c1, H,
              W, c1, 1)
        self._conv_bwd(e[1], a["e1a"], c1, H, W, a["cat1"], g["cat1"], c2 + c1, g["e1a"],
                       y_off=c2)
        self._conv_bwd(e[0], a["x0"], 16, H, W, a["e1a"], g["e1a"], c1, None)
        return float(self.mse_part.sum().item()) / n

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        L = _lib()
        for cv in self.enc + self.dec:
            for w, gw, m, v, w16 in cv.params():
                _call(L.qp_net_adam, stream(), ptr(w), ptr(gw), ptr(m), ptr(v), w.numel(), lr, b1,
                      b2, eps, self.step, ptr(w16))
        c1 = CH[0]
        g = self.ghwb.view(NOUT, c1 + 1)
        gw = g[:, :c1].contiguous()
        gb = g[:, c1].contiguous()
        for w, gg, m, v in ((self.hw, gw, self.hm[0], self.hv[0]),
                            (self.hb, gb, self.hm[1], self.hv[1])):
            _call(L.qp_net_adam, stream(), ptr(w), ptr(gg), ptr(m), ptr(v), w.numel(), lr, b1, b2,
                  eps, self.step, None)

    # -- persistence ------------------------------------------------------------
    def state(self):
        d = {f"enc{i}_w": c.w.cpu() for i, c in enumerate(self.enc)}
        d.update({f"enc{i}_b": c.b.cpu() for i, c in enumerate(self.enc)})
        d.update({f"dec{i}_w": c.w.cpu() for i, c in enumerate(self.dec)})
        d.update({f"dec{i}_b": c.b.cpu() for i, c in enumerate(self.dec)})
        d.update({"head_w": self.hw.cpu(), "head_b": self.hb.cpu(),
                  "in_scale": self.in_scale.cpu(),
                  "out_scale": torch.tensor([self.out_scale])})
        return d

    def load(self, d):
        for i, c in enumerate(self.enc):
            c.w.copy_(d[f"enc{i}_w"])
            c.b.copy_(d[f"enc{i}_b"])
            c.w16.copy_(bf16_bits(c.w))
        for i, c in enumerate(self.dec):
            c.w.copy_(d[f"dec{i}_w"])
            c.b.copy_(d[f"dec{i}_b"])
            c.w16.copy_(bf16_bits(c.w))
        self.hw.copy_(d["head_w"])
        self.hb.copy_(d["head_b"])
        self.in_scale.copy_(d["in_scale"])
        self.out_scale = float(d["out_scale"][0])
