"""The batched scan (quickpic-opensource_amd/qpscan.py): instances run
round-robin on their own streams give, instance by instance, bitwise the
fields of running them one after the other; the autotuner returns a valid
launch shape per deposit kernel and leaves the results within the oracle's
rounding of the default shape."""
import numpy as np
import pytest

qps = pytest.importorskip("qpslice")
qpscan = pytest.importorskip("qpscan")

pytestmark = pytest.mark.gpu


def banner(name):
    print(f"[gpu-test] {name}", flush=True)


def test_batched_equals_sequential_bitwise():
    banner("scan batched vs sequential")
    insts = qpscan.draw(4, grids=[6, 7], ppcs=[2, 3], indz=6, beam_np=(16, 16, 64), seed=3)
    try:
        qpscan.run_sequential(insts, 24)
        seq = [x.fields() for x in insts]
        qpscan.run_batched(insts, 24)
        bat = [x.fields() for x in insts]
        for i, (a, b) in enumerate(zip(seq, bat)):
            for fa, fb in zip(a, b):
                assert np.array_equal(fa, fb), i
            assert np.abs(a[0]).max() > 0   # the wake is there
    finally:
        for x in insts:
            x.close()


def test_autotune_picks_valid_shapes():
    banner("scan autotune")
    (x,) = qpscan.draw(1, grids=[7], ppcs=[4], indz=6, beam_np=(16, 16, 64), seed=5)
    try:
        x.begin()
        best = qpscan.autotune(x, 1)
        for knob, v in best.items():
            cur, nv = x.ctx.tune_get(knob)
            assert 0 <= v < nv and cur == v
        # a tuned run and a default run of the same instance agree to rounding
        x.begin()
        for j in range(1, 9):
            x.slice(j)
        tuned = x.fields()
        for knob in best:
            x.ctx.tune(knob, -1)
        x.begin()
        for j in range(1, 9):
            x.slice(j)
        plain = x.fields()
        for a, b in zip(tuned, plain):
            assert np.abs(a - b).max() <= 1e-4 * max(np.abs(b).max(), 1e-30)
    finally:
        x.close()


def _fields_and_parts(insts):
    return [x.fields() + [x.sp.download()] for x in insts]


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
def test_graph_scan_equals_sequential_fixed_passes(precision):
    """One captured slice graph replayed for every slice (qp_scan) gives,
    instance by instance, bitwise the fields and particles of running the
    instances one after the other, also after eager slices of the same step
    (so the graph of the other buffer parity starts); a replay past the
    step's last slice is refused; the beam push after the captured slices
    sees their wake."""
    banner(f"scan graph vs sequential, fixed passes, {precision}")
    def make():
        return qpscan.draw(3, grids=[6, 7], ppcs=[2, 3], indz=6, beam_np=(16, 16, 64), seed=4,
                           precision=getattr(qps, precision))
    insts = make()
    try:
        qpscan.run_sequential(insts, 24)
        seq = _fields_and_parts(insts)
        for x in insts:
            x.sim.end_step()
        seq_beam = [x.bm.download() for x in insts]
    finally:
        for x in insts:
            x.close()
    insts = make()   # the same seeds: the same plasma and beams
    try:
        for x in insts:
            x.begin()
        scan = qps.Scan([x.sim for x in insts], iter_max=2, tol=0.0)
        try:
            # three eager slices first (the particles then sit in the other
            # buffer), then the captured slice replayed 2 + 19 times
            for j in (1, 2, 3):
                for x in insts:
                    x.slice(j)
            scan.run(2)
            scan.run(19)
            for i in range(len(insts)):
                assert list(scan.passes(i, 24))[3:] == [2] * 21
            with pytest.raises(qps.QPError):   # past the last slice of the step
                scan.run(insts[0].ctx.nz)
        finally:
            scan.close()
        got = _fields_and_parts(insts)
        for i, (a, b) in enumerate(zip(seq, got)):
            for fa, fb in zip(a, b):
                assert np.array_equal(fa, fb), i
            assert np.abs(a[0]).max() > 0
        # the beam push after the captured slices sees their wake
        for x, want in zip(insts, seq_beam):
            x.sim.end_step()
            assert np.array_equal(x.bm.download(), want)
    finally:
        for x in insts:
            x.close()


def test_graph_scan_to_tolerance_matches_eager():
    """With a residual tolerance the graph enqueues every pass and the
    passes after convergence return at once: the passes per slice and the
    fields equal those of qp_slice_run to the same tolerance, slice for
    slice, with the extrapolated seed counted on the device."""
    banner("scan graph vs sequential, tolerance")
    insts = qpscan.draw(3, grids=[6, 7], ppcs=[2], indz=6, beam_np=(16, 16, 64), seed=8,
                        iter_=20, tol=1e-6, seed_mode=1)
    try:
        qpscan.run_sequential(insts, 32)
        seq = _fields_and_parts(insts)
        seq_passes = [list(x.passes) for x in insts]
        qpscan.run_graph(insts, 32, chunk=7)
        got = _fields_and_parts(insts)
        for i, x in enumerate(insts):
            print(f"instance {i}: passes {x.passes}", flush=True)
            assert x.passes == seq_passes[i], i
            assert max(x.passes) > 2   # the tolerance, not a fixed count, stopped them
            for fa, fb in zip(seq[i], got[i]):
                assert np.array_equal(fa, fb), i
    finally:
        for x in insts:
            x.close()


def test_graph_scan_refuses_diagnostics_and_shared_contexts():
    banner("scan refusals")
    (x,) = qpscan.draw(1, grids=[6], ppcs=[2], indz=5, beam_np=(8, 8, 32), seed=2)
    try:
        x.begin()
        with pytest.raises(qps.QPError):
            qps.Scan([x.sim, x.sim])
        q3 = qps.Field3d(x.ctx, 1)
        x.sim.diag3d(qps.DIAG_PSI, q3)
        with pytest.raises(qps.QPError):
            qps.Scan([x.sim])
        x.sim.diag3d(qps.DIAG_PSI, None)
        scan = qps.Scan([x.sim])
        scan.run(3)
        scan.close()
    finally:
        x.close()
