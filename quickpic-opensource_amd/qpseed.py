"""Learned corrector seed for the slice loop, and the training of its network.

Between qp_slice_prepare and qp_slice_finish of slice j the network
(qpnet.UNet) reads, in real space on its (coarsened) grid:
  B_perp of slice j-1 and its change B(j-1) - B(j-2) (4 channels), the
  plasma charge of slice j, the current J_x, J_y of the slice before and psi
  of slice j,
each group divided by its own root mean square over the slice, and predicts
the correction of B_perp over 2 B(j-1) - B(j-2) in units of the root mean
square of B(j-1) - B(j-2).  So the network sees fields of order one whatever
the driver's strength, and its output scales back with the slice.  The seed is
that extrapolation, formed in k space as the library's own extrapolated seed
is, plus the correction's kind-1 transform (QP_TUNE_SEED = 2); the plain
iteration to the residual tolerance is unchanged, so the network only changes
the number of passes.  B_perp here is the corrector's field epw (in real space
through the kind-1 transform of two components).

Every device operation here runs on the context's stream
(torch.cuda.ExternalStream over qp_ctx_stream), so torch's element-wise
steps, the network's kernels and the library's slice calls are ordered
without host synchronisation.
"""
from __future__ import annotations

import numpy as np
import torch

import qpnet
import qpslice as qps


def _wrap(addr, n, dtype):
    """A torch tensor over n elements of existing device memory."""
    class _Holder:
        pass
    h = _Holder()
    h.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f4" if dtype == torch.float32
                                  else "<f8", "data": (addr, False), "version": 2}
    return torch.as_tensor(h, device="cuda")


def ctx_stream(ctx: qps.Ctx):
    """A torch stream object for the context's HIP stream."""
    return torch.cuda.ExternalStream(ctx.stream())


# the normalisation group of each input channel: B(j-1), B(j-1) - B(j-2),
# charge, current, psi
GROUP = [0, 0, 1, 1, 2, 3, 3, 4]


class SeedFeatures:
    """Features, targets and seeds of one simulation, on the device, on a
    grid coarsened f times (the network's grid: (ny / f) x (nx / f), both
    multiples of 8).  Call its methods inside
    `with torch.cuda.stream(ctx_stream(ctx))`."""

    def __init__(self, ctx: qps.Ctx, sim: qps.Sim, f=1):
        self.ctx, self.sim, self.f = ctx, sim, f
        self.nx, self.ny = ctx.nx, ctx.ny
        self.H, self.W = self.ny // f, self.nx // f
        if self.H * f != self.ny or self.W * f != self.nx or self.H % 8 or self.W % 8:
            raise ValueError(f"grid {self.nx}x{self.ny} does not coarsen by {f} to multiples of 8")
        self.fp64 = ctx.precision == qps.FP64
        self.dtype = torch.float64 if self.fp64 else torch.float32
        self.tmp = qps.Field2d(ctx, 2, state=1)
        self.hk = qps.Field2d(ctx, 2, state=1)   # k-space B_perp of the slice before
        self.nk = 0
        self.hist = torch.zeros((2, 2, self.ny + 1, self.nx + 1), dtype=torch.float32,
                                device="cuda")
        self.nhist = 0
        self.feat = torch.zeros((qpnet.NIN, self.H, self.W), dtype=torch.float32, device="cuda")
        self.xn = torch.zeros_like(self.feat)
        self.sd = self.sb = None
        self.targ = torch.zeros((2, self.H, self.W), dtype=torch.float32, device="cuda")
        self.diff = torch.zeros((2, self.ny + 1, self.nx + 1), dtype=torch.float32,
                                device="cuda")

    def view(self, field, comp, space=0):
        """Component comp of a field as a (ny+1, nx+1) view of its storage
        (k space: (nx+1, ny+1))."""
        p, ld, cs = field.device(space)
        es = 8 if self.fp64 else 4
        rows, cols = (self.ny + 1, self.nx + 1) if space == 0 else (self.nx + 1, self.ny + 1)
        buf = _wrap(p + comp * cs * es, rows * ld, self.dtype)
        return buf.view(rows, ld)[:, :cols]

    def _real_plane(self, field, comp):
        p, ld, cs = field.device(0)
        return (p + comp * cs * (8 if self.fp64 else 4), ld, self.fp64)

    def _hist_plane(self, t):
        return (t.data_ptr(), self.nx + 1, False)

    def gather(self):
        """After qp_slice_prepare: the 8 normalised input planes on the
        network's grid (self.xn); self.sd is the output unit (a device scalar,
        the root mean square of B(j-1) - B(j-2))."""
        h = self.hist
        cu = self.sim.field(qps.F_CU)
        planes = [self._hist_plane(h[0, 0]), self._hist_plane(h[0, 1]),
                  self._hist_plane(h[1, 0]), self._hist_plane(h[1, 1]),
                  self._real_plane(self.sim.field(qps.F_QE), 0),
                  self._real_plane(cu, 0), self._real_plane(cu, 1),
                  self._real_plane(self.sim.field(qps.F_PSI), 0)]
        f = self.feat
        qpnet.block_mean(planes, self.f, self.H, self.W, f)
        f[2:4] = f[0:2] - f[2:4]
        ms = (f * f).mean(dim=(1, 2))
        g = torch.stack([ms[0:2].mean(), ms[2:4].mean(), ms[4], ms[5:7].mean(), ms[7]])
        g = g.sqrt().clamp(min=1e-30)
        torch.div(f, g[GROUP][:, None, None], out=self.xn)
        self.sd = g[1]
        self.sb = g[0]
        return self.xn

    def record(self, target=False):
        """After qp_slice_finish: the converged B_perp (real space) enters
        the history; with target (gather was called for this slice), first
        the network's target, the block means of converged minus
        extrapolated B_perp in units of self.sd, into self.targ."""
        self.tmp.asa(self.sim.field(qps.F_EPW))
        self.tmp.fftkr(1)
        e = [self.view(self.tmp, c) for c in range(2)]
        if target:
            for c in range(2):
                torch.sub(e[c].float(), 2.0 * self.hist[0, c] - self.hist[1, c],
                          out=self.diff[c])
            qpnet.block_mean([self._hist_plane(self.diff[0]), self._hist_plane(self.diff[1])],
                             self.f, self.H, self.W, self.targ)
            self.targ /= self.sd
        self.hist[1].copy_(self.hist[0])
        for c in range(2):
            self.hist[0, c].copy_(e[c])
        self.nhist += 1

    def seed(self, corr=None, scale=1.0):
        """Between prepare and finish: epw <- 2 B(j-1) - B(j-2) + scale up(corr)
        in k space once two slices are recorded (the correction through the
        kind-1 forward transform; corr None: no correction), else epw is kept
        (the previous slice's field); B(j-1) enters the k-space history."""
        epw = self.sim.field(qps.F_EPW)
        pe, _, cs = epw.device(1)
        ph = self.hk.device(1)[0]
        extrap = self.nk >= 2
        pc = None
        if extrap and corr is not None:
            pt, ld, cst = self.tmp.device(0)
            qpnet.seed_apply(corr, self.H, self.W, self.f, scale, None, pt, self.fp64, ld, cst,
                             self.ny + 1, self.nx + 1)
            self.tmp.fftrk(1)
            pc = self.tmp.device(1)[0]
        qpnet.seed_combine(pe, ph, pc, 2 * cs, extrap, self.fp64)
        self.nk += 1

    def reset(self):
        self.nhist = 0
        self.nk = 0


class LearnedSeed:
    """Runs slices with the network's seed (QP_TUNE_SEED = 2)."""

    def __init__(self, ctx: qps.Ctx, sim: qps.Sim, net: qpnet.UNet, f=1):
        self.ctx, self.sim, self.net = ctx, sim, net
        ctx.tune(qps.TUNE_SEED, 2)
        net.in_scale.fill_(1.0)   # the features arrive normalised
        self.stream = ctx_stream(ctx)
        with torch.cuda.stream(self.stream):
            self.fs = SeedFeatures(ctx, sim, f)

    def slice(self, j, iter_max=20, tol=1e-6):
        fs = self.fs
        with torch.cuda.stream(self.stream):
            self.sim.prepare(j)
            out = None
            if fs.nhist >= 2:
                x = fs.gather()
                out = self.net.forward_graph(x, fs.H * fs.W, fs.W, fs.H, fs.W)
                out *= fs.sd
            fs.seed(out)
            it, res = self.sim.finish(j, iter_max=iter_max, tol=tol)
            fs.record()
        return it, res

    def begin_step(self):
        self.sim.begin_step()
        self.fs.reset()


def collect(ctx, sim, j0, j1, f=1, iter_max=20, tol=1e-6, every=1, sink=None):
    """Runs slices j0..j1-1 with the extrapolated seed (QP_TUNE_SEED = 1);
    for every `every`-th slice with two slices of history, calls
    sink(features, target) (device tensors, valid during the call).
    Returns the passes of every slice."""
    ctx.tune(qps.TUNE_SEED, 1)
    passes = []
    with torch.cuda.stream(ctx_stream(ctx)):
        fs = SeedFeatures(ctx, sim, f)
        for j in range(j0, j1):
            sim.prepare(j)
            take = sink is not None and fs.nhist >= 2 and (j - j0) % every == 0
            if take:
                fs.gather()
            it, _ = sim.finish(j, iter_max=iter_max, tol=tol)
            passes.append(it)
            fs.record(target=take)
            # slices where B has not started to change carry no signal
            if take and float(fs.sd) > 1e-6 * float(fs.sb):
                sink(fs.xn, fs.targ)
    ctx.tune(qps.TUNE_SEED, -1)
    return passes


class Trainer:
    """Online training of the network on normalised (features, target)
    pairs as the plain solver produces them: every pair enters a replay
    buffer (replacing a random entry once it is full) and is followed by
    `per_pair` Adam steps on random entries, after the first `warm` pairs."""

    def __init__(self, net: qpnet.UNet, H, W, buf=64, warm=16, per_pair=1, lr=1e-3, seed=0):
        self.net, self.H, self.W = net, H, W
        net.in_scale.fill_(1.0)
        net.out_scale = 1.0
        self.x = torch.zeros((buf, qpnet.NIN, H, W), device="cuda")
        self.y = torch.zeros((buf, 2, H, W), device="cuda")
        self.n, self.warm, self.per_pair, self.lr = 0, warm, per_pair, lr
        self.rng = np.random.default_rng(seed)
        self.losses = []
        self.target_ms = []   # mean square of the target of each step's pair

    def add(self, x, y):
        b = len(self.x)
        i = self.n if self.n < b else int(self.rng.integers(b))
        self.x[i].copy_(x)
        self.y[i].copy_(y)
        self.n += 1
        if self.n >= self.warm:
            for _ in range(self.per_pair):
                self.step()

    def step(self):
        k = min(self.n, len(self.x))
        i = int(self.rng.integers(k))
        H, W = self.H, self.W
        self.net.forward(self.x[i], H * W, W, H, W)
        self.losses.append(self.net.backward(self.y[i].reshape(2, -1)))
        self.target_ms.append(float((self.y[i] * self.y[i]).mean()))
        self.net.adam(lr=self.lr)

    def normalised_loss(self, last=200):
        """Mean loss over the last steps divided by their targets' mean
        square: 1 is what predicting zero (the plain extrapolation) scores."""
        n = min(last, len(self.losses))
        if n == 0:
            return float("nan")
        return float(np.sum(self.losses[-n:]) / max(np.sum(self.target_ms[-n:]), 1e-300))


def save(net, path):
    np.savez_compressed(path, **{k: v.numpy() for k, v in net.state().items()})


def load(net, path):
    d = np.load(path, allow_pickle=False)
    net.load({k: torch.from_numpy(d[k]) for k in d.files})
