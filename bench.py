"""xi-slice throughput of the quasi-static solver on one MI355X.

Workload (the precision-form configuration): 1024 x 1024 transverse grid over
x, y in [-10, 10] c/wp, xi in [0, 10] c/wp with 1024 slices, uniform plasma with
4 x 4 particles per cell, a bi-Gaussian electron driver of 2^25 particles
(nb/n0 = 10, sigma_r = 0.3, sigma_z = 0.5, gamma = 2e4, centred at xi = 2).
A step is one xi slice: beam-charge copy, charge deposit, psi, E_perp, the
B_perp predictor-corrector to relative change 1e-6 (at most 20 passes), the
plasma push with its cell sort, and the slice writes into the 3-D wake.  The
slices timed lie in the beam (the xi range where the corrector works); the
slices before it are run untimed.  Inputs are resident on the GPU when timing
starts; the clock is bracketed by stream synchronisations.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "quickpic-opensource_amd")
for _p in (REPO, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = ("ξ-slices/s at 1024², 16 ppc; CNN-seeded FP32 vs plain FP64 at same B "
          "residual")
HBM_PEAK_GBS = 8000.0
PROFILE = os.path.join(REPO, "profiles", "slice_loop_fp32_1024sq_16ppc.md")
KERNEL_SYMBOL = {"amjdp": "k_amjdp<{p}", "push": "k_push<{p},1>", "extpsi": "k_extpsi<{p}",
                 "deposit_q": "k_deposit_q<{p}", "sort": "k_push<{p},2>"}


def profiled_traffic(kind, precision_name):
    """HBM bytes per launch of a kernel class as rocprofv3 --pmc measured it
    (FETCH_SIZE and WRITE_SIZE scaled by a copy of known size), from the
    committed profile summary: the call-weighted mean over the kernels of
    the class (e.g. the deposit with and without the psi gather); None when
    that summary has no such entry."""
    if not os.path.exists(PROFILE):
        return None
    want = "qp::" + KERNEL_SYMBOL[kind].format(p=precision_name)
    tot = calls = 0.0
    with open(PROFILE) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) in (3, 4) and parts[0] == "traffic" and parts[1].startswith(want):
                n = float(parts[3]) if len(parts) == 4 else 1.0
                tot += float(parts[2]) * n
                calls += n
    return tot / calls if calls else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_gpu(args, precision, beam_part=None):
    import numpy as np

    import qpslice as qps
    ind = args.ind
    ctx = qps.Ctx(ind, ind, args.indz, precision=precision)
    # corrector seed: the FP32 leg uses --fp32-seed, the plain FP64 leg the
    # reference's previous-slice field; both iterate to the same tolerance
    ctx.tune(qps.TUNE_SEED, args.fp32_seed if precision == qps.FP32 else 0)
    # corrector iteration: the FP32 leg mixes (--fp32-mix), the FP64 leg is plain
    ctx.tune(qps.TUNE_MIX, args.fp32_mix if precision == qps.FP32 else 0)
    # enqueue the next pass's deposits before each residual read (results
    # are bit-identical either way)
    ctx.tune(qps.TUNE_SPECULATE, args.speculate)
    n = ctx.nx
    t0 = time.time()
    sp = qps.Species(ctx, ppc=(args.ppc, args.ppc))
    bm = qps.Beam(ctx, center=(0.0, 0.0, 2.0), sigma=(0.3, 0.3, 0.5), sigma_v=(1.0, 1.0, 0.0),
                  np_=tuple(args.beam_np) if beam_part is None else (0, 0, 0), gamma=2.0e4,
                  peak=args.peak, dt=10.0,
                  npmax=0 if beam_part is None else beam_part.shape[0] + 2)
    if beam_part is not None:
        bm.upload(beam_part)
    sim = qps.Sim(ctx, [sp], [bm], iter_=2)
    # the same beam for the FP64 leg
    part = bm.download() if beam_part is None and args.fp64 else None
    log(f"[bench] {'FP32' if precision == qps.FP32 else 'FP64'} setup {time.time() - t0:.1f}s "
        f"(plasma {sp.np}, beam {bm.np})")
    sim.begin_step()
    # first timed slice: a little ahead of the beam centre (xi = 2)
    j_first = args.j_first or max(1, int(round((2.0 - 2.0 * 0.5) / ctx.dxi)))
    t0 = time.time()
    for j in range(1, j_first - args.warmup):
        sim.slice(j, iter_max=args.iter_max, tol=args.tol)
    log(f"[bench] fast-forward to slice {j_first - args.warmup}: {time.time() - t0:.1f}s")
    for j in range(max(1, j_first - args.warmup), j_first):
        sim.slice(j, iter_max=args.iter_max, tol=args.tol)
    # timed region: no per-launch events
    passes = []
    ctx.sync()
    t0 = time.perf_counter()
    for j in range(j_first, j_first + args.steps):
        it, _ = sim.slice(j, iter_max=args.iter_max, tol=args.tol)
        passes.append(it)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    # per-kernel-class durations (HIP events around each launch on the
    # library's stream) over the next slices, outside the timed region
    nprof = min(args.prof_steps, ctx.nz - j_first - args.steps)
    ctx.profile(True)
    for j in range(j_first + args.steps, j_first + args.steps + nprof):
        sim.slice(j, iter_max=args.iter_max, tol=args.tol)
    ctx.sync()
    prof = {qps.PROF_NAMES[k]: ctx.profile_get(k) for k in range(len(qps.PROF_NAMES))}
    ctx.profile(False)
    sweep = {}
    j_cont = j_first + args.steps + nprof   # next slice of the step
    if args.sweep:
        # the deposit launch shapes, each over the same profiled slices
        j0 = j_cont
        for knob, name in ((qps.TUNE_AMJDP, "amjdp"), (qps.TUNE_DEPOSIT_Q, "deposit_q")):
            _, nv = ctx.tune_get(knob)
            kind = qps.PROF_NAMES.index(name)
            for v in range(nv):
                ctx.tune(knob, v)
                ctx.profile(True)
                for j in range(j0, j0 + 3):
                    sim.slice(j, iter_max=args.iter_max, tol=args.tol)
                t, cnt, _ = ctx.profile_get(kind)
                ctx.profile(False)
                sweep[f"{name}[{v}]"] = round(t / max(cnt, 1), 4)
                j0 += 3
            ctx.tune(knob, -1)
        # corrector speculation off / on: wall time per slice
        for v in (0, 1):
            ctx.tune(qps.TUNE_SPECULATE, v)
            ctx.sync()
            t0 = time.perf_counter()
            for j in range(j0, j0 + 5):
                sim.slice(j, iter_max=args.iter_max, tol=args.tol)
            ctx.sync()
            sweep[f"speculate[{v}] ms/slice"] = round((time.perf_counter() - t0) / 5 * 1e3, 4)
            j0 += 5
        ctx.tune(qps.TUNE_SPECULATE, args.speculate)
        # corrector seed: previous slice (the reference) / linear extrapolation
        for v in (0, 1):
            ctx.tune(qps.TUNE_SEED, v)
            ctx.sync()
            t0 = time.perf_counter()
            its = [sim.slice(j, iter_max=args.iter_max, tol=args.tol)[0]
                   for j in range(j0, j0 + 10)]
            ctx.sync()
            sweep[f"seed[{v}] ms/slice"] = round((time.perf_counter() - t0) / 10 * 1e3, 4)
            sweep[f"seed[{v}] passes"] = float(np.mean(its[2:]))
            j0 += 10
        ctx.tune(qps.TUNE_SEED, args.fp32_seed if precision == qps.FP32 else 0)
        # corrector iteration: plain / Anderson-mixed, wall time and passes
        for v in (0, 1):
            ctx.tune(qps.TUNE_MIX, v)
            ctx.sync()
            t0 = time.perf_counter()
            its = [sim.slice(j, iter_max=args.iter_max, tol=args.tol)[0]
                   for j in range(j0, j0 + 10)]
            ctx.sync()
            sweep[f"mix[{v}] ms/slice"] = round((time.perf_counter() - t0) / 10 * 1e3, 4)
            sweep[f"mix[{v}] passes"] = float(np.mean(its))
            j0 += 10
        ctx.tune(qps.TUNE_MIX, args.fp32_mix if precision == qps.FP32 else 0)
        # kernel time per slice of the classes a knob changes (HIP events on
        # the library's stream), so the pass count of the slices drops out
        def class_ms(knob, v, names, nsl=5):
            nonlocal j0
            ctx.tune(knob, v)
            ctx.profile(True)
            passes = 0
            for j in range(j0, j0 + nsl):
                passes += sim.slice(j, iter_max=args.iter_max, tol=args.tol)[0]
            t = sum(ctx.profile_get(qps.PROF_NAMES.index(n))[0] for n in names)
            work = ctx.profile_get(qps.PROF_NAMES.index(names[0]))[2]
            ctx.profile(False)
            j0 += nsl
            return round(t / nsl, 4), passes / nsl, t / max(work, 1.0) * 1e9
        # transforms: transpose kernels / column-pair y pass / LDS paddings;
        # ms per GB of row-transform work does not depend on the pass count
        for v in range(ctx.tune_get(qps.TUNE_FFT)[1]):
            t, ps, per_gb = class_ms(qps.TUNE_FFT, v, ("fft", "transpose"))
            sweep[f"fft[{v}] ms/slice"] = t
            sweep[f"fft[{v}] passes"] = ps
            sweep[f"fft[{v}] ms/GB"] = round(per_gb, 4)
        ctx.tune(qps.TUNE_FFT, -1)
        # the push's re-sort: count then push into place / push then scatter
        for v in range(ctx.tune_get(qps.TUNE_PUSH)[1]):
            sweep[f"push[{v}] ms/slice"] = class_ms(qps.TUNE_PUSH, v, ("push", "sort"))[0]
        ctx.tune(qps.TUNE_PUSH, -1)
        log(f"[bench] sweep (ms per launch): {sweep}")
        j_cont = j0
    # every kernel class: time, launches and algorithmic bytes (operands read
    # or written once) counted by the library at each launch
    classes = {}
    for name, (t, c, b) in prof.items():
        if c:
            classes[name] = {"ms_total": round(t, 3), "launches": c, "ms_avg": round(t / c, 4),
                             "GBs": round(b / (t / 1e3) / 1e9, 1) if t > 0 else None}
    # the roofline object names the dominant particle kernel (the path's
    # bandwidth-bound hot spot) by accumulated time
    particle = ("amjdp", "push", "extpsi", "deposit_q", "sort")
    dom = max(particle, key=lambda k: prof[k][0])
    t, cnt, b = prof[dom]
    roof = None
    if cnt:
        avg_s = t / cnt / 1e3
        ach = (b / cnt) / avg_s / 1e9
        traffic = profiled_traffic(dom, "float" if precision == qps.FP32 else "double")
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "avg_launch_ms": round(t / cnt, 4), "algorithmic_bytes_per_launch": b / cnt}
    total_ms = sum(v[0] for v in prof.values())
    nprof = max(nprof, 1)
    result = {
        "slices_per_s": args.steps / elapsed,
        "ms_per_slice": elapsed / args.steps * 1e3,
        "passes_mean": float(np.mean(passes)),
        "passes": passes,
        "first_slice": j_first,
        "roofline": roof,
        "kernels": classes,
        "kernel_ms_per_slice": total_ms / nprof,
        "np_plasma": sp.np,
        "np_beam": bm.np,
        "sweep": sweep,
    }
    # finish the 3-D step (untimed) through the slices the beam occupies (to
    # 5 sigma_z behind its centre; the beam push reads only those), then the
    # beam push; sample the wake on the way for the FP32-vs-FP64 comparison.
    # A particle that crosses more than one tile in a slice stops the step,
    # as PPPCHECK2L stops the reference (part2d_class.f03:582-588); the
    # comparison then covers what was computed.
    j0 = j_cont
    j_end = min(ctx.nz, int(np.ceil((2.0 + 5 * 0.5) / ctx.dxi)) + 2)
    samples = sorted({j0 + k * (j_end - j0) // 4 for k in range(5)})
    psi = {}
    g0 = bm.moments()[0]
    stop = None
    try:
        for j in range(j0, j_end + 1):
            sim.slice(j, iter_max=args.iter_max, tol=args.tol)
            if j in samples:
                psi[j] = sim.field(qps.F_PSI).download(0)[0]
        sim.end_step()
    except qps.QPError as e:
        stop = f"slice {j}: {e}"
        log(f"[bench] 3-D step stopped at {stop}")
    e3, b3 = sim.bexyz, sim.bbxyz
    # per sampled slice j (completed): the 3-D wake plane lpos = j + 1 and psi
    result["wake"] = {j: {"Ez": e3.plane(3, j + 1),
                          "Ex-By": e3.plane(1, j + 1) - b3.plane(2, j + 1),
                          "Bz": b3.plane(3, j + 1),
                          "psi": psi[j]} for j in samples if j in psi}
    result["beam_energy"] = None if stop else (g0, bm.moments()[0])
    result["step_stop"] = stop
    ctx.close()
    return result, part


def calib_copy(args):
    """A plain copy of known size (1 GiB read, 1 GiB written) for scaling the
    HBM byte counters of a profiler run."""
    import torch
    n = (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    log(f"[bench] copy calibration: 3 x {n * 4} B read and written")


def cpu_baseline(args):
    """The oracle (FP64, one thread) on a bounded sample of the same workload:
    --cpu-slices xi slices (about 10 s) at the full transverse size and ppc,
    2 corrector passes, no beam (the per-slice CPU cost does not depend on
    the beam)."""
    from oracle import qpo
    t0 = time.time()
    grid = qpo.Grid(args.ind, args.ind, 4)
    sim = qpo.Simulation(grid, [qpo.SpeciesSpec(ppc=(args.ppc, args.ppc))], [], dt=10.0,
                         iter_=2)
    sim.begin_step()
    log(f"[bench] cpu oracle setup {time.time() - t0:.1f}s")
    nsl = args.cpu_slices
    t0 = time.perf_counter()
    for j in range(1, nsl + 1):
        sim.slice(j)
    el = time.perf_counter() - t0
    return {"value": round(nsl / el, 4), "unit": "xi-slices/s", "cores": 1, "kind": "port",
            "sample": f"{nsl} xi slices of the FP64 C oracle at {grid.nx}x{grid.ny}, "
                      f"{args.ppc}x{args.ppc} ppc, 2 corrector passes, one thread "
                      f"({el:.1f}s)"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--ind", type=int, default=10, help="log2 of the transverse grid")
    ap.add_argument("--indz", type=int, default=10)
    ap.add_argument("--ppc", type=int, default=4, help="particles per cell per dimension")
    ap.add_argument("--beam-np", type=int, nargs=3, default=[128, 128, 2048])
    ap.add_argument("--peak", type=float, default=10.0)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--fp32-seed", type=int, default=1, choices=(0, 1),
                    help="corrector seed of the FP32 leg: 0 previous slice, 1 extrapolated")
    ap.add_argument("--fp32-mix", type=int, default=0, choices=(0, 1),
                    help="corrector iteration of the FP32 leg: 0 plain, 1 Anderson-mixed")
    ap.add_argument("--speculate", type=int, default=1, choices=(0, 1),
                    help="enqueue the next corrector pass's deposits before reading a residual")
    ap.add_argument("--iter-max", type=int, default=20)
    ap.add_argument("--no-fp64", dest="fp64", action="store_false")
    ap.add_argument("--no-cpu", dest="cpu", action="store_false")
    ap.add_argument("--cpu-slices", type=int, default=4)
    ap.add_argument("--prof-steps", type=int, default=10,
                    help="slices after the timed ones run with per-launch events")
    ap.add_argument("--sweep", action="store_true",
                    help="also time every compiled deposit launch shape")
    ap.add_argument("--j-first", type=int, default=0,
                    help="first timed slice (default: ahead of the beam centre)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a 1 GiB copy kernel (profiler counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)
    if args.calib_copy:
        calib_copy(args)
        return
    import numpy as np

    import qpslice as qps
    r32, beam = run_gpu(args, qps.FP32)
    log(f"[bench] FP32: {r32['slices_per_s']:.1f} slices/s, passes {r32['passes_mean']:.2f}")
    r64 = None
    if args.fp64:
        r64, _ = run_gpu(args, qps.FP64, beam_part=beam)
        log(f"[bench] FP64: {r64['slices_per_s']:.1f} slices/s, passes {r64['passes_mean']:.2f}")
    errs = None
    if r64 is not None:
        # relative error of the precision form against the FP64 path on the
        # same beam: max |a32 - a64| / max |a64| over the sampled wake slices
        common = sorted(set(r32["wake"]) & set(r64["wake"]))
        errs, peaks = {}, {}
        for k in ("Ez", "Ex-By", "Bz", "psi"):
            if common:
                a = np.stack([r32["wake"][j][k] for j in common])
                b = np.stack([r64["wake"][j][k] for j in common])
                errs[k] = float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
                peaks[k] = float(np.abs(b).max())
        if r32["beam_energy"] and r64["beam_energy"]:
            (a0, a1), (b0, b1) = r32["beam_energy"], r64["beam_energy"]
            errs["beam energy"] = abs(a1 - b1) / abs(b1)
            errs["beam energy change"] = abs((a1 - a0) - (b1 - b0)) / max(abs(b1 - b0), 1e-300)
        log(f"[bench] FP32 vs FP64 relative error: {errs}")
    cpu = cpu_baseline(args) if args.cpu else None
    out = {
        "metric": METRIC,
        "value": round(r32["slices_per_s"], 3),
        "unit": "xi-slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_slice"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded beam, quiet-start plasma)",
        "config": {
            "workload": f"precision form: {2 ** args.ind}^2 x {2 ** args.indz} slices, "
                        f"{args.ppc * args.ppc} ppc, beam {r32['np_beam']} particles, "
                        f"nb/n0={args.peak}, B residual {args.tol:g} (max {args.iter_max} "
                        "passes), FP32 corrector seed: "
                        + {0: "previous slice", 1: "linear extrapolation"}[args.fp32_seed]
                        + {0: "", 1: ", Anderson-mixed passes"}[args.fp32_mix]
                        + "; FP64 leg: previous slice (plain)",
            "fp32_seed": {0: "previous", 1: "extrapolated"}[args.fp32_seed],
            "fp32_corrector": {0: "plain", 1: "anderson1"}[args.fp32_mix],
            "speculative_deposits": bool(args.speculate),
            "grid": [2 ** args.ind, 2 ** args.ind, 2 ** args.indz],
            "ppc": args.ppc * args.ppc,
            "plasma_particles": r32["np_plasma"],
            "beam_particles": r32["np_beam"],
            "passes_per_slice_mean": r32["passes_mean"],
            "first_timed_slice": r32["first_slice"],
        },
        "roofline": r32["roofline"],
        "cpu_baseline": cpu,
        "kernels": r32["kernels"],
        "kernel_ms_per_step": round(r32["kernel_ms_per_slice"], 4),
    }
    if r64 is not None:
        out["fp64_plain"] = {"value": round(r64["slices_per_s"], 3),
                             "ms_per_step": round(r64["ms_per_slice"], 4),
                             "passes_per_slice_mean": r64["passes_mean"],
                             "fp32_over_fp64": round(r32["slices_per_s"] / r64["slices_per_s"], 3),
                             "fp32_rel_err": {k: float(f"{v:.3g}") for k, v in errs.items()},
                             "fp32_rel_err_slices": common,
                             "fp64_peak": {k: float(f"{v:.4g}") for k, v in peaks.items()},
                             "step_stop": {"fp32": r32["step_stop"], "fp64": r64["step_stop"]},
                             "roofline": r64["roofline"]}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
