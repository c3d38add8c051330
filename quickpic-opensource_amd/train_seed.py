"""Trains the learned corrector seed on slices that the plain solver produces,
and measures what it gains.

Each training run is one xi sweep of a bi-Gaussian driver through uniform
plasma with parameters drawn from nb/n0 ~ U[1, 20], sigma_r ~ U[0.2, 1.0],
sigma_z ~ U[0.3, 1.5] (per-run beam seed); the plain predictor-corrector
iterates every slice to the residual tolerance with the extrapolated seed,
and each slice's (features, converged - extrapolated B_perp) pair trains the
network online (qpseed.Trainer).  evaluate() runs a held-out beam with the
previous-slice seed (the reference's), the extrapolated seed and the learned
seed and reports passes per slice and slices/s for each.

usage: python quickpic-opensource_amd/train_seed.py --ind 9 --runs 200 \
           --out quickpic-opensource_amd/weights/seed_512.npz
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

import torch  # noqa: E402

import qpnet  # noqa: E402
import qpseed  # noqa: E402
import qpslice as qps  # noqa: E402


def draw_beam(rng):
    """(peak nb/n0, sigma_r, sigma_z) of one run."""
    return float(rng.uniform(1.0, 20.0)), float(rng.uniform(0.2, 1.0)), float(rng.uniform(0.3, 1.5))


class Run:
    """One plasma and one beam on a context; a new beam per run."""

    def __init__(self, ctx, ppc, beam_np, center_z=2.0):
        self.ctx, self.beam_np, self.center_z = ctx, beam_np, center_z
        self.sp = qps.Species(ctx, ppc=(ppc, ppc))
        self.bm = None
        self.sim = None

    def new_beam(self, peak, sr, sz, seed):
        if self.sim is not None:
            self.sim.close()
        if self.bm is not None:
            self.bm.close()
        self.bm = qps.Beam(self.ctx, center=(0.0, 0.0, self.center_z), sigma=(sr, sr, sz),
                           sigma_v=(1.0, 1.0, 0.0), np_=self.beam_np, gamma=2.0e4, peak=peak,
                           dt=10.0, seed=seed)
        self.sp.renew(0.0)
        self.sim = qps.Sim(self.ctx, [self.sp], [self.bm], iter_=2)
        self.sim.begin_step()

    def slices(self, sz, span):
        """Slices from 3 sigma_z ahead of the beam centre, `span` of them."""
        j0 = max(1, int((self.center_z - 3.0 * sz) / self.ctx.dxi))
        return j0, min(self.ctx.nz, j0 + span)


def train(args, log=print):
    ctx = qps.Ctx(args.ind, args.ind, args.indz, precision=qps.FP32)
    run = Run(ctx, args.ppc, tuple(args.beam_np))
    net = qpnet.UNet(seed=args.seed)
    H, W = ctx.ny // args.f, ctx.nx // args.f
    tr = qpseed.Trainer(net, H, W, buf=args.buf, warm=args.warm, per_pair=args.per_pair,
                        lr=args.lr, seed=args.seed)
    rng = np.random.default_rng(args.seed)
    t0 = time.time()
    stopped = 0
    for r in range(args.runs):
        peak, sr, sz = draw_beam(rng)
        try:
            run.new_beam(peak, sr, sz, seed=1000 + r)
        except qps.QPError:
            # the move check of a run that stopped is read at the next
            # reload; the reload after it is clean
            run.new_beam(peak, sr, sz, seed=1000 + r)
        j0, j1 = run.slices(sz, args.span)
        try:
            # slices ahead of the sampled range only advance the plasma
            for j in range(1, j0):
                run.sim.slice(j, iter_max=args.iter_max, tol=args.tol)
            passes = qpseed.collect(ctx, run.sim, j0, j1, f=args.f, iter_max=args.iter_max,
                                    tol=args.tol, every=args.every, sink=tr.add)
        except qps.QPError as e:
            # a plasma electron crossed more than one tile in a slice: the
            # reference stops such a run (PPPCHECK2L); its slices up to there
            # have trained the network
            stopped += 1
            log(f"[train] run {r}: nb {peak:.1f} sr {sr:.2f} sz {sz:.2f} stops: {e}")
            continue
        if r % max(1, args.runs // 20) == 0 or r == args.runs - 1:
            loss = np.mean(tr.losses[-50:]) if tr.losses else float("nan")
            log(f"[train] run {r}: nb {peak:.1f} sr {sr:.2f} sz {sz:.2f}  passes "
                f"{np.mean(passes):.2f}  pairs {tr.n}  loss {loss:.4g}  {time.time() - t0:.0f}s")
    ctx.sync()
    log(f"[train] {args.runs} runs, {stopped} stopped by the move check, {tr.n} pairs")
    return net, tr


def evaluate(net, ind, indz, ppc, beam_np, f, peak, sr, sz, seed, iter_max=20, tol=1e-6,
             span=None, log=print):
    """Passes per slice and slices/s over the same slices of one beam with
    seed 0 (previous slice), 1 (extrapolation) and the learned seed."""
    ctx = qps.Ctx(ind, ind, indz, precision=qps.FP32)
    run = Run(ctx, ppc, tuple(beam_np))
    out = {}
    for mode in ("previous", "extrapolated", "learned"):
        run.new_beam(peak, sr, sz, seed)
        j0, j1 = run.slices(sz, span or ctx.nz)
        ctx.tune(qps.TUNE_SEED, {"previous": 0, "extrapolated": 1, "learned": 2}[mode])
        ls = qpseed.LearnedSeed(ctx, run.sim, net, f) if mode == "learned" else None
        for j in range(1, j0):
            (ls.slice if ls else run.sim.slice)(j, iter_max=iter_max, tol=tol)
        ctx.sync()
        torch.cuda.synchronize()
        t = time.perf_counter()
        its = [(ls.slice if ls else run.sim.slice)(j, iter_max=iter_max, tol=tol)[0]
               for j in range(j0, j1)]
        ctx.sync()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        out[mode] = {"passes": float(np.mean(its)), "slices_per_s": (j1 - j0) / dt,
                     "slices": j1 - j0}
        log(f"[eval] {mode}: {out[mode]['passes']:.2f} passes/slice, "
            f"{out[mode]['slices_per_s']:.1f} slices/s over {j1 - j0} slices")
        ctx.tune(qps.TUNE_SEED, -1)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--ind", type=int, default=9)
    ap.add_argument("--indz", type=int, default=9)
    ap.add_argument("--ppc", type=int, default=4, help="particles per cell per dimension")
    ap.add_argument("--beam-np", type=int, nargs=3, default=[64, 64, 256])
    ap.add_argument("--f", type=int, default=4, help="coarsening of the network's grid")
    ap.add_argument("--runs", type=int, default=200)
    ap.add_argument("--span", type=int, default=160, help="sampled slices per run")
    ap.add_argument("--every", type=int, default=2)
    ap.add_argument("--buf", type=int, default=256)
    ap.add_argument("--warm", type=int, default=32)
    ap.add_argument("--per-pair", type=int, default=1)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--iter-max", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=None)
    ap.add_argument("--eval", type=float, nargs=3, default=None,
                    metavar=("PEAK", "SIGMA_R", "SIGMA_Z"), help="held-out beam to evaluate")
    ap.add_argument("--eval-ind", type=int, nargs="*", default=None,
                    help="log2 grids to evaluate on (default: the training grid)")
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    net, tr = train(args, log=log)
    res = {"runs": args.runs, "pairs": tr.n, "grid": 2 ** args.ind,
           "normalised_loss_last_200": tr.normalised_loss(200)}
    log(f"[train] normalised loss over the last 200 steps: {res['normalised_loss_last_200']:.4g}")
    if args.out:
        os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
        qpseed.save(net, args.out)
        log(f"[train] weights -> {args.out}")
    if args.eval:
        res["eval"] = {}
        for ind in (args.eval_ind or [args.ind]):
            res["eval"][2 ** ind] = evaluate(net, ind, ind, args.ppc, args.beam_np, args.f,
                                             *args.eval, seed=424242, iter_max=args.iter_max,
                                             tol=args.tol, span=args.span * 2 ** (ind - args.ind),
                                             log=log)
    import json
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
