"""GPU parity at the benchmark configurations, against the oracle's
restatement of the step (oracle/qpo.py Simulation, simulation_class.f03:611-765)
on the same seeded inputs.

* 128^3, 2x2 ppc, 2^20-particle driver (nb/n0 = 4, sigma = 0.5, gamma = 2e4,
  sigma_p = 1), 2 corrector passes: the whole 3-D step (128 slices and the
  beam push) in FP64, every slice's stored fields within 1e-10 relative, the
  per-cell plasma counts bit-exact at every 16th slice, the pushed beam's
  survivors within 1e-10.
* 512^2, 2x2 ppc, 2^24-particle driver (nb/n0 = 10, sigma_perp = 0.3,
  sigma_z = 0.5): FP64 slices 1..112, compared at every slice of the window
  96..112 that crosses the driver's centre (slice 102.4).
* 1024^2, 4x4 ppc, 2^25-particle driver (nb/n0 = 10): the first FP64 slices
  against the oracle; and FP32 properties at the full size: two runs give
  bitwise-equal fields and particles, the deposited plasma charge equals the
  particles' total charge.

The large beams are loaded on the GPU and handed to the oracle (its own
loader, held bit-exact to the GPU loader by test_gpu_parity's loader test,
runs at about 5 us per particle, minutes for 2^24).  Progress lines are
printed as each stage ends.
"""
import time

import numpy as np
import pytest

from oracle import qpo

qps = pytest.importorskip("qpslice")

pytestmark = pytest.mark.gpu


def rel(a, b):
    s = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(np.asarray(a) - b).max()) / s


def say(msg):
    print(f"[gpu-configs {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def build(ind, indz, ppc, center, sigma, np_, peak, precision, iter_=2, with_oracle=True):
    grid = qpo.Grid(ind, ind, indz)
    ctx = qps.Ctx(ind, ind, indz, precision=precision)
    sp = qps.Species(ctx, ppc=ppc)
    bm = qps.Beam(ctx, center=center, sigma=sigma, sigma_v=(1.0, 1.0, 0.0), np_=np_,
                  gamma=2.0e4, peak=peak, dt=10.0)
    gsim = qps.Sim(ctx, [sp], [bm], iter_=iter_)
    osim = None
    if with_oracle:
        bspec = qpo.BeamSpec(center=center, sigma=sigma, sigma_v=(1.0, 1.0, 0.0), np_=np_,
                             gamma=2.0e4, peak=peak)
        osim = qpo.Simulation(grid, [qpo.SpeciesSpec(ppc=ppc)], [bspec], dt=10.0,
                              iter_=iter_, beam_parts=[bm.download()])
    return grid, ctx, sp, bm, gsim, osim


def plane_errors(gsim, osim, j):
    errs = {}
    for c in range(3):
        errs[f"e{c}"] = rel(gsim.bexyz.plane(c + 1, j + 1), osim.e3[c, j])
        errs[f"b{c}"] = rel(gsim.bbxyz.plane(c + 1, j + 1), osim.b3[c, j])
    return errs


def check_counts(grid, sp, osim):
    want = np.bincount(qpo.cell_of(grid.nx, grid.ny, osim.species[0].part),
                       minlength=grid.nx * grid.ny)
    assert np.array_equal(sp.cell_counts(), want)


@pytest.mark.timeout(900)
def test_config1_full_step_fp64():
    say("config 1: 128^3, 2x2 ppc, 2^20 beam, FP64, full step")
    grid, ctx, sp, bm, gsim, osim = build(7, 7, (2, 2), (0.0, 0.0, 2.0), (0.5, 0.5, 0.5),
                                          (64, 64, 256), 4.0, qps.FP64)
    try:
        assert bm.np == osim.beams[0][0].shape[0] == 2 ** 20
        osim.begin_step()
        gsim.begin_step()
        worst = 0.0
        for j in range(1, grid.nz + 1):
            osim.slice(j)
            gsim.slice(j)
            errs = plane_errors(gsim, osim, j)
            worst = max(worst, max(errs.values()))
            assert max(errs.values()) < 1e-10, (j, errs)
            if j % 16 == 0:
                check_counts(grid, sp, osim)
                say(f"slice {j}: worst plane error so far {worst:.2e}")
        # beam push (simulation_class.f03:752-757), the drops compacted
        osim.push_beams()
        gsim.end_step()
        got, want = bm.download(), osim.beams[0][0]
        assert got.shape == want.shape, (got.shape, want.shape)
        key = lambda a: a[np.lexsort(a[:, :3].T[::-1])]   # noqa: E731
        err = rel(key(got), key(want))
        say(f"beam push: {got.shape[0]} survivors, rel err {err:.2e}")
        assert err < 1e-10
    finally:
        ctx.close()


@pytest.mark.timeout(900)
def test_config2_window_through_driver_centre_fp64():
    say("config 2: 512^2, 2x2 ppc, 2^24 beam, nb/n0 = 10, sigma_perp = 0.3, FP64")
    t0 = time.time()
    grid, ctx, sp, bm, gsim, osim = build(9, 9, (2, 2), (0.0, 0.0, 2.0), (0.3, 0.3, 0.5),
                                          (128, 128, 1024), 10.0, qps.FP64)
    try:
        say(f"set-up {time.time() - t0:.1f} s, beam {bm.np}")
        osim.begin_step()
        gsim.begin_step()
        worst = 0.0
        for j in range(1, 113):
            osim.slice(j)
            gsim.slice(j)
            if j >= 96:
                errs = plane_errors(gsim, osim, j)
                worst = max(worst, max(errs.values()))
                assert max(errs.values()) < 1e-10, (j, errs)
                check_counts(grid, sp, osim)
            if j % 16 == 0:
                say(f"slice {j}: worst plane error {worst:.2e}")
        say(f"window 96..112 worst {worst:.2e}")
    finally:
        ctx.close()


@pytest.mark.timeout(900)
def test_config3_first_slices_fp64():
    say("config 3: 1024^2, 4x4 ppc, 2^25 beam, nb/n0 = 10, FP64")
    t0 = time.time()
    grid, ctx, sp, bm, gsim, osim = build(10, 10, (4, 4), (0.0, 0.0, 2.0), (0.3, 0.3, 0.5),
                                          (128, 128, 2048), 10.0, qps.FP64)
    try:
        say(f"set-up {time.time() - t0:.1f} s, plasma {sp.np}, beam {bm.np}")
        osim.begin_step()
        gsim.begin_step()
        for j in range(1, 4):
            osim.slice(j)
            gsim.slice(j)
            errs = plane_errors(gsim, osim, j)
            say(f"slice {j}: " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()))
            assert max(errs.values()) < 1e-10, (j, errs)
        check_counts(grid, sp, osim)
    finally:
        ctx.close()


@pytest.mark.timeout(900)
def test_config3_fp32_properties():
    """FP32 at 1024^2, 16 ppc, 2^25 beam, 16 slices from the front of the box:
    two runs bitwise equal; the plasma charge deposit conserves charge."""
    say("config 3: FP32 properties")
    runs = []
    for r in range(2):
        grid, ctx, sp, bm, gsim, _ = build(10, 10, (4, 4), (0.0, 0.0, 2.0), (0.3, 0.3, 0.5),
                                           (128, 128, 2048), 10.0, qps.FP32, with_oracle=False)
        try:
            gsim.begin_step()
            for j in range(1, 17):
                gsim.slice(j)
            # charge conservation of the deposit with the guard fold:
            # species2d%qdp into a zero field gives deposit + qn, qn the
            # negated deposit of the initial plasma, so its node sum is the
            # change of the plasma's total charge: zero
            q = qps.Field2d(ctx, 1)
            q.as_(0.0)
            sp.qdp(q)
            part = sp.download()
            total = float(np.abs(part[:, 7]).sum())
            dq = float(q.download(0).sum())
            say(f"run {r}: node sum of deposit + qn {dq:.3e} against total |q| {total:.3e}")
            assert abs(dq) < 1e-6 * total
            planes = [gsim.bexyz.plane(c + 1, 17) for c in range(3)] + \
                     [gsim.bbxyz.plane(c + 1, 17) for c in range(3)]
            runs.append((planes, part))
        finally:
            ctx.close()
    (pa, xa), (pb, xb) = runs
    for a, b in zip(pa, pb):
        assert np.array_equal(a, b)
    assert np.array_equal(xa, xb)
