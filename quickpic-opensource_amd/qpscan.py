"""Batched parameter scan: many independent slice-solver instances on one GPU.

Each instance is a context with its own HIP stream, plasma species and
bi-Gaussian driver; instances differ in transverse grid, particles per cell
and driver parameters (nb/n0 ~ U[1, 20], sigma_r ~ U[0.2, 1.0], sigma_z ~
U[0.3, 1.5], per-instance beam seed).  The transverse grids are powers of
two, as the reference's are (indx, indy are log2 sizes,
simulation_class.f03:151-224), so a scan draws them from 2^indx choices.

run_graph captures one slice of every instance into one HIP graph
(qpslice.Scan, qp_scan_*: the instances are independent branches, each slice
index a device word) and replays it once per slice, so the GPU works on all
instances at once from a single launch per slice; the corrector runs either
the reference's fixed number of passes or to a residual tolerance (the
passes after convergence return at once inside the graph).  run_batched
enqueues slice j of every instance before slice j+1 of any, each on its
instance's stream, with fixed passes.  run_sequential runs the instances one
after the other.  All three give bitwise-equal fields per instance (every
kernel is deterministic and an instance's work stays on its own stream),
which tests/test_gpu_scan.py checks.

autotune() picks, per instance, the fastest launch shape of each deposit
kernel (qp_ctx_tune) by timing a few slices of that instance with the
library's event profiler.

usage: python quickpic-opensource_amd/qpscan.py --n 64 --grids 8 9 --slices 256
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

import qpslice as qps  # noqa: E402

BOX = (-10.0, 10.0, -10.0, 10.0, 0.0, 10.0)


class Instance:
    def __init__(self, ind, indz, ppc, peak, sr, sz, seed, beam_np, precision=qps.FP32,
                 iter_=2, tol=0.0, seed_mode=0):
        self.params = {"grid": 2 ** ind, "ppc": ppc * ppc, "peak": peak, "sigma_r": sr,
                       "sigma_z": sz, "seed": seed}
        self.ctx = qps.Ctx(ind, ind, indz, box=BOX, precision=precision)
        self.sp = qps.Species(self.ctx, ppc=(ppc, ppc))
        self.bm = qps.Beam(self.ctx, center=(0.0, 0.0, 2.0), sigma=(sr, sr, sz),
                           sigma_v=(1.0, 1.0, 0.0), np_=beam_np, gamma=2.0e4, peak=peak,
                           dt=10.0, seed=seed)
        self.sim = qps.Sim(self.ctx, [self.sp], [self.bm], iter_=iter_)
        self.iter, self.tol = iter_, tol
        if tol > 0.0:
            self.ctx.tune(qps.TUNE_SEED, seed_mode)

    def begin(self):
        self.sp.renew(0.0)
        self.sim.begin_step()

    def slice(self, j):
        return self.sim.slice(j, iter_max=self.iter, tol=self.tol)[0]

    def fields(self):
        F = self.sim.field
        return [F(qps.F_PSIT).download(0), F(qps.F_BXYZ).download(0),
                F(qps.F_FXY).download(0), F(qps.F_CU).download(1)]

    def close(self):
        self.ctx.close()


def draw_one(i, grids, ppcs, indz, beam_np, seed=0, attempt=0, precision=qps.FP32, iter_=2,
             tol=0.0, seed_mode=0):
    """A replacement for instance i of a scan (attempt >= 1): grid 2^g (g
    from grids), ppc^2 particles per cell (ppc from ppcs), driver drawn from
    the ranges above, with a stream and beam seed of its own."""
    rng = np.random.default_rng([seed, i, attempt])
    g = int(rng.choice(grids))
    p = int(rng.choice(ppcs))
    peak, sr, sz = (float(rng.uniform(1, 20)), float(rng.uniform(0.2, 1.0)),
                    float(rng.uniform(0.3, 1.5)))
    return Instance(g, indz, p, peak, sr, sz, seed=1 + 7919 * seed + 104729 * attempt + i,
                    beam_np=beam_np, precision=precision, iter_=iter_, tol=tol,
                    seed_mode=seed_mode)


def draw(n, grids, ppcs, indz, beam_np, seed=0, precision=qps.FP32, iter_=2, tol=0.0,
         seed_mode=0):
    """n instances with grids 2^g (g from grids), ppc^2 particles per cell
    (ppc from ppcs) and drivers drawn from the ranges above, from one stream
    seeded by `seed` (draw_one makes the replacements)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        g = int(rng.choice(grids))
        p = int(rng.choice(ppcs))
        peak, sr, sz = (float(rng.uniform(1, 20)), float(rng.uniform(0.2, 1.0)),
                        float(rng.uniform(0.3, 1.5)))
        out.append(Instance(g, indz, p, peak, sr, sz, seed=1 + 7919 * seed + i,
                            beam_np=beam_np, precision=precision, iter_=iter_, tol=tol,
                            seed_mode=seed_mode))
    return out


def screen(insts, nslices, redraw, log=None, rounds=8):
    """Runs every instance once over nslices slices with the strict particle
    move check (PPPCHECK2L: a particle crossing more than one tile in a
    slice stops the reference's run); an instance that stops is replaced by
    redraw(i, attempt), a fresh draw, until all pass.  Returns the count of
    replaced draws."""
    replaced = 0
    for attempt in range(rounds):
        bad = []
        for i, x in enumerate(insts):
            x.ctx.tune(qps.TUNE_STRICT, 1)
            try:
                x.begin()
                for j in range(1, nslices + 1):
                    x.slice(j)
            except qps.QPError as e:
                bad.append(i)
                if log:
                    log(f"[scan] instance {i} {x.params} stops: {e}")
            finally:
                x.ctx.tune(qps.TUNE_STRICT, -1)
        if not bad:
            return replaced
        for i in bad:
            insts[i].close()
            insts[i] = redraw(i, attempt + 1)
            replaced += 1
    raise RuntimeError(f"scan: draws still stop after {rounds} rounds")


def autotune(inst, j0, reps=2):
    """Fastest variant of each deposit kernel for this instance, timed over
    `reps` slices per variant starting at slice j0 (the slices are computed
    and the instance must be restarted afterwards); returns {knob: value}."""
    ctx, best = inst.ctx, {}
    for knob, kind in ((qps.TUNE_AMJDP, "amjdp"), (qps.TUNE_DEPOSIT_Q, "deposit_q")):
        _, nv = ctx.tune_get(knob)
        k = qps.PROF_NAMES.index(kind)
        times = []
        j = j0
        for v in range(nv):
            ctx.tune(knob, v)
            ctx.profile(True)
            for _ in range(reps):
                inst.slice(j)
                j += 1
            t, cnt, _ = ctx.profile_get(k)
            ctx.profile(False)
            times.append(t / max(cnt, 1))
        best[knob] = int(np.argmin(times))
        ctx.tune(knob, best[knob])
    return best


def run_batched(insts, nslices):
    for x in insts:
        x.begin()
    for x in insts:
        x.ctx.sync()
    t = time.perf_counter()
    for j in range(1, nslices + 1):
        for x in insts:
            x.slice(j)
    for x in insts:
        x.ctx.sync()
    return time.perf_counter() - t


def run_sequential(insts, nslices):
    """Returns the wall time; each instance's passes per slice go to
    x.passes."""
    for x in insts:
        x.begin()
    for x in insts:
        x.ctx.sync()
    t = time.perf_counter()
    for x in insts:
        x.passes = [x.slice(j) for j in range(1, nslices + 1)]
        x.ctx.sync()
    return time.perf_counter() - t


def run_graph(insts, nslices, chunk=None):
    """All instances from one captured slice graph, replayed nslices times
    (in runs of `chunk` slices); returns the wall time of the replays.  Each
    instance's passes per slice go to x.passes."""
    for x in insts:
        x.begin()
    scan = qps.Scan([x.sim for x in insts], iter_max=insts[0].iter, tol=insts[0].tol)
    try:
        for x in insts:
            x.ctx.sync()
        t = time.perf_counter()
        done = 0
        while done < nslices:
            k = min(chunk or nslices, nslices - done)
            scan.run(k)
            done += k
        el = time.perf_counter() - t
        for i, x in enumerate(insts):
            x.passes = [int(p) for p in scan.passes(i, nslices)]
    finally:
        scan.close()
    return el


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--grids", type=int, nargs="+", default=[8, 9],
                    help="log2 transverse sizes to draw from")
    ap.add_argument("--ppcs", type=int, nargs="+", default=[2, 3, 4],
                    help="particles per cell per dimension to draw from")
    ap.add_argument("--indz", type=int, default=8)
    ap.add_argument("--slices", type=int, default=256)
    ap.add_argument("--beam-np", type=int, nargs=3, default=[32, 32, 256])
    ap.add_argument("--tol", type=float, default=1e-6,
                    help="corrector residual tolerance (0: the reference's fixed passes)")
    ap.add_argument("--iter-max", type=int, default=20)
    ap.add_argument("--seed-mode", type=int, default=1, choices=(0, 1),
                    help="corrector seed with a tolerance: 0 previous slice, 1 extrapolated")
    ap.add_argument("--fp64", action="store_true", help="FP64 instances (default FP32)")
    ap.add_argument("--tune", action="store_true", help="autotune each instance first")
    ap.add_argument("--no-batched", dest="batched", action="store_false",
                    help="skip the round-robin stream run (fixed passes only)")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.time()
    iter_ = args.iter_max if args.tol > 0 else 2
    insts = draw(args.n, args.grids, args.ppcs, args.indz, tuple(args.beam_np), args.seed,
                 precision=qps.FP64 if args.fp64 else qps.FP32, iter_=iter_, tol=args.tol,
                 seed_mode=args.seed_mode)
    log(f"[scan] {args.n} instances set up in {time.time() - t0:.1f}s")
    replaced = screen(insts, args.slices,
                      lambda i, a: draw_one(i, args.grids, args.ppcs, args.indz,
                                            tuple(args.beam_np), args.seed, a,
                                            qps.FP64 if args.fp64 else qps.FP32, iter_, args.tol,
                                            args.seed_mode), log=log)
    log(f"[scan] screened: {replaced} draws replaced ({time.time() - t0:.1f}s)")
    if args.tune:
        for x in insts:
            x.begin()
            x.params["tune"] = autotune(x, 1)
    total = args.n * args.slices
    ts = run_sequential(insts, args.slices)
    seq_passes = [list(x.passes) for x in insts]
    log(f"[scan] sequential {total / ts:.1f} slices/s")
    tg = run_graph(insts, args.slices)
    same = all(list(x.passes) == p for x, p in zip(insts, seq_passes))
    log(f"[scan] graph {total / tg:.1f} slices/s ({ts / tg:.2f}x), passes as sequential: {same}")
    res = {"instances": args.n, "slices_each": args.slices,
           "precision": "fp64" if args.fp64 else "fp32",
           "corrector": (f"tolerance {args.tol:g} (max {iter_} passes), "
                         + ("extrapolated" if args.seed_mode else "previous-slice") + " seed")
           if args.tol > 0 else "2 fixed passes",
           "sequential_slices_per_s": total / ts, "graph_slices_per_s": total / tg,
           "graph_over_sequential": ts / tg, "graph_passes_equal_sequential": same,
           "passes_mean": float(np.mean(seq_passes)),
           "grids": sorted({x.params["grid"] for x in insts}),
           "ppc": sorted({x.params["ppc"] for x in insts}),
           "draws_replaced_by_move_check": replaced}
    if args.batched and args.tol <= 0:
        tb = run_batched(insts, args.slices)
        res["batched_streams_slices_per_s"] = total / tb
        log(f"[scan] round-robin streams {total / tb:.1f} slices/s")
    for x in insts:
        x.close()
    import json
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
