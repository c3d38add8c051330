"""xi-slice throughput of the quasi-static solver on one MI355X.

Workload (the precision-form configuration): 1024 x 1024 transverse grid over
x, y in [-10, 10] c/wp, xi in [0, 10] c/wp with 1024 slices, uniform plasma with
4 x 4 particles per cell, a bi-Gaussian electron driver of 2^25 particles
(nb/n0 = 10, sigma_r = 0.3, sigma_z = 0.5, gamma = 2e4, centred at xi = 2).

A step is one 3-D step of the reference's loop (simulation_class.f03:
611-765): the beam charge deposit, the xi slices (beam-charge copy, charge
deposit, psi, E_perp, the B_perp predictor-corrector to relative change 1e-6
with at most 20 passes, the plasma push with its cell sort, the slice writes
into the 3-D wake), the beam push and the plasma renewal.  The slices are all
1024 where the reference completes the step.  Behind a strong driver a plasma
electron eventually crosses more than one tile in one slice (the quasi-static
push's dx/dxi = p_perp / (1 + psi) grows without bound as psi approaches -1),
and the reference stops there (PPPCHECK2L, part2d_class.f03:582-588); an
untimed probe step with the strict move check finds that slice for each
precision, and the step then covers the slices before the earlier of the two
stops.  `value` is those slices x K steps / the timed wall clock, so the beam
kernels are amortised in it.  Inputs are resident on the GPU when timing
starts; the clock is bracketed by stream synchronisations.

Legs: the FP32 precision form (the headline, nb/n0 = 10) and the plain FP64
path on the same beam, each also at nb/n0 = 1 and 20.  The first warm-up step
of every leg samples the wake for the FP32-vs-FP64 error report.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import math
import os
import platform
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
# one profiled run of the bench command holds both legs' kernels
PROFILES = {p: os.path.join(REPO, "profiles", "step_1024sq_16ppc_nb10.md")
            for p in ("float", "double")}
KERNEL_SYMBOL = {"amjdp": "k_amjdp<{p}", "push": "k_push<{p},1>", "extpsi": "k_extpsi<{p}",
                 "deposit_q": "k_deposit_q<{p}", "sort": "k_push<{p},2>"}
# slices whose wake planes the first step of each leg keeps (FP32 vs FP64)
SAMPLES = (16, 32, 48, 64, 80, 96, 128, 192, 256, 384, 512, 768, 1024)
HEAD = 96   # samples up to here: the beam's front, ahead of its centre (slice 205)


def profiled_traffic(kind, precision_name):
    """HBM bytes per launch of a kernel class as rocprofv3 --pmc measured it
    (FETCH_SIZE and WRITE_SIZE scaled by a copy of known size), from the
    committed profile summary of this workload: the call-weighted mean over
    the kernels of the class (e.g. the deposit with and without the psi
    gather); None when that summary has no such entry."""
    path = PROFILES[precision_name]
    if not os.path.exists(path):
        return None
    want = "qp::" + KERNEL_SYMBOL[kind].format(p=precision_name)
    tot = calls = 0.0
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) in (3, 4) and parts[0] == "traffic" and parts[1].startswith(want):
                n = float(parts[3]) if len(parts) == 4 else 1.0
                tot += float(parts[2]) * n
                calls += n
    return tot / calls if calls else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_step(sim, args, nslices, samples=None, psi=None):
    """One 3-D step over slices 1..nslices; returns the corrector passes of
    every slice.  With samples, psi[j] is downloaded after each sampled
    slice j (untimed steps only)."""
    import qpslice as qps
    sim.begin_step()
    passes = []
    for j in range(1, nslices + 1):
        it, _ = sim.slice(j, iter_max=args.iter_max, tol=args.tol)
        passes.append(it)
        if samples and j in samples:
            psi[j] = sim.field(qps.F_PSI).download(0)[0]
    sim.end_step()
    return passes


def run_leg(args, precision, peak, nwarm, nsteps, nslices, beam_part=None, breakdown=False):
    """One context: W warm-up steps (the first samples the wake), K timed
    steps of nslices slices, optionally one more step with every kernel
    class timed."""
    import numpy as np

    import qpslice as qps
    name = "FP32" if precision == qps.FP32 else "FP64"
    ctx = qps.Ctx(args.ind, args.ind, args.indz, precision=precision)
    try:
        # corrector seed: the FP32 leg uses --fp32-seed, the plain FP64 leg
        # the reference's previous-slice field; both iterate to the same
        # tolerance
        ctx.tune(qps.TUNE_SEED, args.fp32_seed if precision == qps.FP32 else 0)
        ctx.tune(qps.TUNE_MIX, args.fp32_mix if precision == qps.FP32 else 0)
        ctx.tune(qps.TUNE_SPECULATE, args.speculate)
        ctx.tune(qps.TUNE_FFT, args.fft if precision == qps.FP32 else -1)
        ctx.tune(qps.TUNE_AMJDP, args.amjdp)
        ctx.tune(qps.TUNE_DEPOSIT_Q, args.deposit_q)
        t0 = time.time()
        sp = qps.Species(ctx, ppc=(args.ppc, args.ppc))
        bm = qps.Beam(ctx, center=(0.0, 0.0, 2.0), sigma=(args.sigma_r, args.sigma_r, 0.5),
                      sigma_v=(1.0, 1.0, 0.0),
                      np_=tuple(args.beam_np) if beam_part is None else (0, 0, 0),
                      gamma=2.0e4, peak=peak, dt=10.0,
                      npmax=0 if beam_part is None else beam_part.shape[0] + 2)
        if beam_part is not None:
            bm.upload(beam_part)
        sim = qps.Sim(ctx, [sp], [bm], iter_=2)
        part = bm.download() if beam_part is None else None
        log(f"[bench] {name} nb/n0={peak:g}: set-up {time.time() - t0:.1f}s "
            f"(plasma {sp.np}, beam {bm.np})")
        nz = nslices
        samples = [j for j in SAMPLES if j <= nz]
        wake, energy, first_passes = None, None, None
        for w in range(nwarm):
            t0 = time.time()
            if w == 0:
                psi = {}
                g0 = bm.moments()[0]
                first_passes = run_step(sim, args, nz, samples, psi)
                e3, b3 = sim.bexyz, sim.bbxyz
                wake = {j: {"Ez": e3.plane(3, j + 1), "Ex-By": e3.plane(1, j + 1) - b3.plane(2, j + 1),
                            "Ey+Bx": e3.plane(2, j + 1) + b3.plane(1, j + 1),
                            "Bz": b3.plane(3, j + 1), "psi": psi[j]} for j in samples}
                energy = (g0, bm.moments()[0])
            else:
                run_step(sim, args, nz)
            log(f"[bench] {name} nb/n0={peak:g}: warm-up step {w + 1} {time.time() - t0:.2f}s")
        # timed steps: HIP events on one current-deposit launch in
        # --events-every, spread evenly over the steps' passes and slices
        if args.events_every > 0:
            ctx.profile(True, classes=[qps.PROF_AMJDP])
            ctx.profile_every(args.events_every)
        passes = []
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(nsteps):
            passes += run_step(sim, args, nz)
        ctx.sync()
        elapsed = time.perf_counter() - t0
        am_ms, am_n, am_b = ctx.profile_get(qps.PROF_AMJDP)
        ctx.profile(False)
        ctx.profile_every(1)
        res = {"slices_per_s": nz * nsteps / elapsed, "ms_per_step": elapsed / nsteps * 1e3,
               "passes_mean": float(np.mean(passes)), "passes_max": int(np.max(passes)),
               "np_plasma": sp.np, "np_beam": bm.np, "wake": wake, "beam_energy": energy,
               "first_step_passes_mean": float(np.mean(first_passes)) if first_passes else None}
        log(f"[bench] {name} nb/n0={peak:g}: {res['slices_per_s']:.1f} slices/s "
            f"({res['ms_per_step']:.0f} ms per step), passes {res['passes_mean']:.2f} "
            f"(max {res['passes_max']})")
        if am_n:
            avg_s = am_ms / am_n / 1e3
            ach = am_b / am_n / avg_s / 1e9
            pname = "float" if precision == qps.FP32 else "double"
            res["roofline"] = {"kernel": "amjdp", "bound": "hbm", "achieved": round(ach, 1),
                               "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(ach / HBM_PEAK_GBS, 4),
                               "traffic": profiled_traffic("amjdp", pname),
                               "avg_launch_ms": round(am_ms / am_n, 4),
                               "launches": am_n,
                               "algorithmic_bytes_per_launch": am_b / am_n}
        if breakdown:
            # every kernel class over one more step (events on every launch;
            # not part of the timed value)
            ctx.profile(True)
            bp = run_step(sim, args, nz)
            ctx.sync()
            classes = {}
            total = 0.0
            for k, kname in enumerate(qps.PROF_NAMES):
                t, c, b = ctx.profile_get(k)
                total += t
                if c:
                    classes[kname] = {"ms_total": round(t, 3), "launches": c,
                                      "ms_avg": round(t / c, 4),
                                      "GBs": round(b / (t / 1e3) / 1e9, 1) if t > 0 else None}
            ctx.profile(False)
            res["kernels"] = classes
            res["kernel_ms_per_step"] = round(total, 3)
            res["breakdown_passes_mean"] = float(np.mean(bp))
        if args.sweep:
            res["sweep"] = sweep(ctx, sim, args)
        return res, part
    finally:
        ctx.close()


def sweep(ctx, sim, args):
    """Launch-shape and corrector variants, each over slices of the beam
    region of a fresh step (per-launch events or wall clock per slice)."""
    import numpy as np

    import qpslice as qps
    out = {}
    sim.begin_step()
    j0 = max(1, int(round(1.0 / ctx.dxi)))
    for j in range(1, j0):
        sim.slice(j, iter_max=args.iter_max, tol=args.tol)
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
            out[f"{name}[{v}]"] = round(t / max(cnt, 1), 4)
            j0 += 3
        ctx.tune(knob, -1)

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
    for v in range(ctx.tune_get(qps.TUNE_FFT)[1]):
        t, ps, per_gb = class_ms(qps.TUNE_FFT, v, ("fft", "transpose"))
        out[f"fft[{v}] ms/slice"] = t
        out[f"fft[{v}] ms/GB"] = round(per_gb, 4)
    ctx.tune(qps.TUNE_FFT, -1)
    for v in range(ctx.tune_get(qps.TUNE_PUSH)[1]):
        out[f"push[{v}] ms/slice"] = class_ms(qps.TUNE_PUSH, v, ("push", "sort"))[0]
    ctx.tune(qps.TUNE_PUSH, -1)
    for knob, name in ((qps.TUNE_SEED, "seed"), (qps.TUNE_MIX, "mix")):
        for v in (0, 1):
            ctx.tune(knob, v)
            ctx.sync()
            t0 = time.perf_counter()
            its = [sim.slice(j, iter_max=args.iter_max, tol=args.tol)[0]
                   for j in range(j0, j0 + 10)]
            ctx.sync()
            out[f"{name}[{v}] ms/slice"] = round((time.perf_counter() - t0) / 10 * 1e3, 4)
            out[f"{name}[{v}] passes"] = float(np.mean(its))
            j0 += 10
        ctx.tune(knob, -1)
    log(f"[bench] sweep: {out}")
    return out


def wake_errors(r32, r64):
    """max |a32 - a64| / max |a64| per field, over the sampled slices of the
    first step, with the FP64 field's peak over every sample as the scale:
    over the head of the beam (slices <= HEAD) and over the whole step."""
    import numpy as np
    if not (r32.get("wake") and r64.get("wake")):
        return None
    common = sorted(set(r32["wake"]) & set(r64["wake"]))
    out = {"head": {}, "step": {}, "peak": {}, "per_slice": {}}
    for k in ("Ez", "Ex-By", "Ey+Bx", "Bz", "psi"):
        peak = max(float(np.abs(r64["wake"][j][k]).max()) for j in common)
        errs = {j: float(np.abs(r32["wake"][j][k] - r64["wake"][j][k]).max()) / max(peak, 1e-300)
                for j in common}
        out["peak"][k] = float(f"{peak:.4g}")
        out["head"][k] = float(f"{max(v for j, v in errs.items() if j <= HEAD):.3g}")
        out["step"][k] = float(f"{max(errs.values()):.3g}")
        out["per_slice"][k] = {j: float(f"{v:.3g}") for j, v in errs.items()}
    if r32["beam_energy"] and r64["beam_energy"]:
        (a0, a1), (b0, b1) = r32["beam_energy"], r64["beam_energy"]
        out["step"]["beam energy"] = float(f"{abs(a1 - b1) / abs(b1):.3g}")
        out["step"]["beam energy change"] = float(
            f"{abs((a1 - a0) - (b1 - b0)) / max(abs(b1 - b0), 1e-300):.3g}")
    return out


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(args, passes):
    """The reference's own kernels (its F77 routines compiled from source with
    their OpenMP directives, oracle/_ref/libqpref_omp.so), composed into the
    reference's slice loop, on a bounded sample of the same workload: xi
    slices at the full transverse size and ppc with `passes` corrector passes
    each (the FP64 GPU leg's mean, rounded up), no beam (the per-slice cost
    does not depend on it).  Timed with every thread OpenMP is given
    (OMP_NUM_THREADS) and with one.  Without that build, the oracle's C port
    on one thread."""
    import ctypes

    from oracle import qpo
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count() or 1
    omp_so = os.path.join(REPO, "oracle", "_ref", "libqpref_omp.so")
    grid = qpo.Grid(args.ind, args.ind, 4)
    spec = [qpo.SpeciesSpec(ppc=(args.ppc, args.ppc))]
    info = {"unit": "xi-slices/s", "cpu_model": cpu_model(), "nproc": os.cpu_count()}
    if os.path.exists(omp_so):
        from oracle import refharness, refslice
        ref = refharness.Ref(omp_so)
        omp = ctypes.CDLL("libomp.so")
        omp.omp_set_num_threads(threads)
        t0 = time.time()
        sim = refslice.RefSimulation(ref, grid, spec, [], dt=10.0, iter_=passes)
        sim.begin_step()
        log(f"[bench] cpu reference set-up {time.time() - t0:.1f}s ({threads} threads)")
        t0 = time.perf_counter()
        n_all = args.cpu_slices
        for j in range(1, n_all + 1):
            sim.slice(j)
        el_all = time.perf_counter() - t0
        omp.omp_set_num_threads(1)
        t0 = time.perf_counter()
        sim.slice(n_all + 1)
        el_one = time.perf_counter() - t0
        omp.omp_set_num_threads(threads)
        info.update({"value": round(n_all / el_all, 4), "cores": threads, "kind": "reference",
                     "value_1_thread": round(1.0 / el_one, 4),
                     "sample": f"the reference's F77 kernels (OpenMP) in its slice loop: "
                               f"{n_all} xi slices with {threads} threads ({el_all:.1f}s) "
                               f"and 1 slice with 1 thread ({el_one:.1f}s), "
                               f"{grid.nx}x{grid.ny}, {args.ppc}x{args.ppc} ppc, "
                               f"{passes} corrector passes, no beam"})
        return info
    t0 = time.time()
    sim = qpo.Simulation(grid, spec, [], dt=10.0, iter_=passes)
    sim.begin_step()
    t0 = time.perf_counter()
    for j in range(1, args.cpu_slices + 1):
        sim.slice(j)
    el = time.perf_counter() - t0
    info.update({"value": round(args.cpu_slices / el, 4), "cores": 1, "kind": "port",
                 "sample": f"{args.cpu_slices} xi slices of the FP64 C oracle at "
                           f"{grid.nx}x{grid.ny}, {args.ppc}x{args.ppc} ppc, {passes} "
                           f"corrector passes, one thread ({el:.1f}s), no beam"})
    return info


def probe_leg(args, pk, prec):
    """One untimed 3-D step with the strict move check: the slices computed
    before the step stops (a particle crossing more than one tile in a
    slice, PPPCHECK2L, where the reference stops), the reason, and the
    corrector passes per slice."""
    import numpy as np

    import qpslice as qps
    name = "fp32" if prec == qps.FP32 else "fp64"
    ctx = qps.Ctx(args.ind, args.ind, args.indz, precision=prec)
    try:
        ctx.tune(qps.TUNE_STRICT, 1)
        ctx.tune(qps.TUNE_SEED, args.fp32_seed if prec == qps.FP32 else 0)
        sp = qps.Species(ctx, ppc=(args.ppc, args.ppc))
        bm = qps.Beam(ctx, center=(0.0, 0.0, 2.0), sigma=(args.sigma_r, args.sigma_r, 0.5),
                      sigma_v=(1.0, 1.0, 0.0), np_=tuple(args.beam_np), gamma=2.0e4, peak=pk,
                      dt=10.0)
        sim = qps.Sim(ctx, [sp], [bm], iter_=2)
        sim.begin_step()
        passes, stop = [], None
        t0 = time.time()
        try:
            for j in range(1, ctx.nz + 1):
                passes.append(sim.slice(j, iter_max=args.iter_max, tol=args.tol)[0])
            sim.end_step()
        except qps.QPError as e:
            stop = str(e)
        p = np.array(passes)
        out = {"slices_done": len(passes), "stop": stop, "seconds": round(time.time() - t0, 2),
               "passes_mean": float(p.mean()) if len(p) else None,
               "passes_max": int(p.max()) if len(p) else None,
               "passes_per_128": [round(float(p[i:i + 128].mean()), 2)
                                  for i in range(0, len(p), 128)]}
        log(f"[bench] probe nb/n0={pk:g} {name}: {out}")
        return out
    finally:
        ctx.close()


def probe(args):
    """One strict-checked step per precision and nb/n0 (probe_leg)."""
    import qpslice as qps
    out = {}
    for pk in [args.peak] + list(args.other_peaks):
        for prec, name in ((qps.FP32, "fp32"), (qps.FP64, "fp64")):
            out[f"nb/n0={pk:g} {name}"] = probe_leg(args, pk, prec)
    print(json.dumps(out), flush=True)


def leg_slices(args, pk):
    """The slices a step of this nb/n0 covers: all of them if neither
    precision's probe step stops, else those before the earlier stop."""
    import qpslice as qps
    res = {"fp32": probe_leg(args, pk, qps.FP32)}
    if args.fp64:
        res["fp64"] = probe_leg(args, pk, qps.FP64)
    n = min(r["slices_done"] for r in res.values())
    stops = {k: r["stop"] for k, r in res.items() if r["stop"]}
    return n, stops


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed 3-D steps of the headline leg")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--ind", type=int, default=10, help="log2 of the transverse grid")
    ap.add_argument("--indz", type=int, default=10)
    ap.add_argument("--ppc", type=int, default=4, help="particles per cell per dimension")
    ap.add_argument("--beam-np", type=int, nargs=3, default=[128, 128, 2048])
    ap.add_argument("--peak", type=float, default=10.0, help="nb/n0 of the headline leg")
    ap.add_argument("--other-peaks", type=float, nargs="*", default=[1.0, 20.0],
                    help="nb/n0 of the other legs (one warm-up and one timed step each)")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--fp32-seed", type=int, default=1, choices=(0, 1, 3),
                    help="corrector seed of the FP32 leg: 0 previous slice, 1 linear, "
                         "3 quadratic extrapolation")
    ap.add_argument("--fp32-mix", type=int, default=0, choices=(0, 1),
                    help="corrector iteration of the FP32 leg: 0 plain, 1 Anderson-mixed")
    ap.add_argument("--speculate", type=int, default=1, choices=(0, 1, 2),
                    help="enqueue the next corrector pass's deposits before reading a residual")
    ap.add_argument("--fft", type=int, default=-1, choices=(-1, 0, 1, 2, 3, 4, 5, 6),
                    help="transform form of the FP32 leg (QP_TUNE_FFT): 0 row passes with a "
                         "transpose between them, 1 column-pair passes, 5 / 6 column pairs "
                         "only for k -> real / real -> k, -1 the library's default")
    ap.add_argument("--amjdp", type=int, default=-1,
                    help="launch shape of the current deposit (QP_TUNE_AMJDP), -1 the default")
    ap.add_argument("--deposit-q", type=int, default=-1,
                    help="launch shape of the charge deposit (QP_TUNE_DEPOSIT_Q), -1 the default")
    ap.add_argument("--events-every", type=int, default=8,
                    help="time one current-deposit launch in this many during the timed "
                         "steps, spread evenly over them (0: none, no roofline); every "
                         "launch's events cost the timed steps 2-3 %%")
    ap.add_argument("--iter-max", type=int, default=20)
    ap.add_argument("--no-fp64", dest="fp64", action="store_false")
    ap.add_argument("--no-cpu", dest="cpu", action="store_false")
    ap.add_argument("--no-breakdown", dest="breakdown", action="store_false",
                    help="skip the step with every kernel class timed")
    ap.add_argument("--cpu-slices", type=int, default=2)
    ap.add_argument("--sweep", action="store_true",
                    help="also time the compiled launch shapes and corrector variants")
    ap.add_argument("--sigma-r", type=float, default=0.3, help="driver's transverse sigma")
    ap.add_argument("--probe", action="store_true",
                    help="only run one strict-checked step per precision and nb/n0 and report "
                         "where it stops")
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
    if args.probe:
        probe(args)
        return
    import qpslice as qps
    legs, span = {}, {}
    for pk in [args.peak] + [p for p in args.other_peaks if p != args.peak]:
        span[pk] = leg_slices(args, pk)
    n0 = span[args.peak][0]
    r32, beam = run_leg(args, qps.FP32, args.peak, args.warmup, args.steps, n0,
                        breakdown=args.breakdown)
    legs[args.peak] = {"fp32": r32}
    if args.fp64:
        r64, _ = run_leg(args, qps.FP64, args.peak, max(1, min(args.warmup, 1)),
                         max(1, min(args.steps, 1)), n0, beam_part=beam,
                         breakdown=args.breakdown)
        legs[args.peak]["fp64"] = r64
    del beam
    for pk in args.other_peaks:
        if pk == args.peak:
            continue
        a, b = run_leg(args, qps.FP32, pk, 1, 1, span[pk][0])
        legs[pk] = {"fp32": a}
        if args.fp64:
            legs[pk]["fp64"], _ = run_leg(args, qps.FP64, pk, 1, 1, span[pk][0], beam_part=b)
    for pk, lg in legs.items():
        if "fp64" in lg:
            lg["err"] = wake_errors(lg["fp32"], lg["fp64"])
            log(f"[bench] nb/n0={pk:g} FP32 vs FP64: head {lg['err']['head']}, "
                f"step {lg['err']['step']}")
    cpu = None
    if args.cpu:
        r64 = legs[args.peak].get("fp64")
        passes = math.ceil(r64["passes_mean"]) if r64 else 2
        cpu = cpu_baseline(args, passes)
    out = {
        "metric": METRIC,
        "value": round(r32["slices_per_s"], 3),
        "unit": "xi-slices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded beam, quiet-start plasma)",
        "config": {
            "workload": f"precision form: {2 ** args.ind}^2 x {2 ** args.indz} slices, "
                        f"{args.ppc * args.ppc} ppc, beam {r32['np_beam']} particles, "
                        f"nb/n0={args.peak:g}, sigma_r={args.sigma_r:g}, B residual {args.tol:g} "
                        f"(max {args.iter_max} passes); a step is one 3-D step (beam deposit, "
                        f"slices 1..{n0}, beam push, plasma renewal): "
                        + ("every slice" if n0 == 2 ** args.indz else
                           f"the slices before the reference's move check (PPPCHECK2L) stops "
                           f"the step at slice {n0 + 1} (a plasma electron crossing more than "
                           "one tile)")
                        + "; FP32 corrector seed: "
                        + {0: "previous slice", 1: "linear extrapolation",
                           3: "quadratic extrapolation"}[args.fp32_seed]
                        + {0: "", 1: ", Anderson-mixed passes"}[args.fp32_mix]
                        + "; FP64 leg: previous slice (plain)",
            "fp32_seed": {0: "previous", 1: "extrapolated", 3: "quadratic"}[args.fp32_seed],
            "fp32_corrector": {0: "plain", 1: "anderson1"}[args.fp32_mix],
            "speculative_deposits": bool(args.speculate),
            "grid": [2 ** args.ind, 2 ** args.ind, 2 ** args.indz],
            "ppc": args.ppc * args.ppc,
            "plasma_particles": r32["np_plasma"],
            "beam_particles": r32["np_beam"],
            "slices_per_step": n0,
            "step_stop": span[args.peak][1] or None,
            "passes_per_slice_mean": round(r32["passes_mean"], 3),
            "passes_per_slice_max": r32["passes_max"],
        },
        "roofline": r32.get("roofline"),
        "cpu_baseline": cpu,
        "kernels": r32.get("kernels"),
        "kernel_ms_per_step": r32.get("kernel_ms_per_step"),
    }
    r64 = legs[args.peak].get("fp64")
    if r64 is not None:
        out["fp64_plain"] = {"value": round(r64["slices_per_s"], 3),
                             "ms_per_step": round(r64["ms_per_step"], 3),
                             "passes_per_slice_mean": round(r64["passes_mean"], 3),
                             "fp32_over_fp64": round(r32["slices_per_s"] / r64["slices_per_s"], 3),
                             "roofline": r64.get("roofline"),
                             "kernels": r64.get("kernels")}
    out["legs"] = {}
    for pk, lg in legs.items():
        e = {"slices_per_step": span[pk][0], "step_stop": span[pk][1] or None}
        for prec in ("fp32", "fp64"):
            if prec in lg:
                r = lg[prec]
                e[prec] = {"value": round(r["slices_per_s"], 3),
                           "ms_per_step": round(r["ms_per_step"], 3),
                           "passes_per_slice_mean": round(r["passes_mean"], 3),
                           "passes_per_slice_max": r["passes_max"]}
        if "fp64" in lg:
            e["fp32_over_fp64"] = round(lg["fp32"]["slices_per_s"] / lg["fp64"]["slices_per_s"], 3)
            err = lg["err"]
            e["fp32_rel_err_head"] = err["head"]
            e["fp32_rel_err_step"] = err["step"]
            e["fp64_peak"] = err["peak"]
            if pk == args.peak:
                e["fp32_rel_err_per_slice"] = err["per_slice"]
        out["legs"][f"nb/n0={pk:g}"] = e
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
