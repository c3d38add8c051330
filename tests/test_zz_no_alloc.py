"""qp_species_qdp_noalloc (the charge deposit with its scratch in the
library's static device array, no allocation) against qp_species_qdp, both
called through the C ABI on the same particles: equal bit for bit, on two
calls in a row (the second on a smaller grid and particle set, so it reads
the array the first one left behind), in FP32 and FP64; grids whose halo
exceeds the array's 2 MiB are refused with QP_ERR_UNSUPPORTED."""
import numpy as np
import pytest

qps = pytest.importorskip("qpslice")

pytestmark = pytest.mark.gpu

QP_ERR_UNSUPPORTED = -6   # include/qpslice.h


def perturbed_species(ctx, ppc, seed, shift=0.37):
    """The quiet-start plasma with every particle moved by up to `shift`
    cells (and non-uniform charges), uploaded again."""
    sp = qps.Species(ctx, ppc=ppc)
    part = sp.download()
    rng = np.random.default_rng(seed)
    n = part.shape[0]
    part[:, 0] = np.clip(part[:, 0] + rng.uniform(-shift, shift, n), 0.0, ctx.nx - 1e-3)
    part[:, 1] = np.clip(part[:, 1] + rng.uniform(-shift, shift, n), 0.0, ctx.ny - 1e-3)
    part[:, 7] *= rng.uniform(0.5, 1.5, n)
    sp.upload(part)
    return sp


def both(ctx, sp, start):
    out = []
    for fn in (qps.lib().qp_species_qdp, qps.lib().qp_species_qdp_noalloc):
        q = qps.Field2d(ctx, 1)
        q.upload(start, 0)
        rc = fn(sp.h, q.h)
        assert rc == 0, qps.lib().qp_last_error()
        out.append(q.download(0))
    return out


@pytest.mark.parametrize("precision", ["FP32", "FP64"])
def test_noalloc_charge_deposit_equals_qdp(precision):
    prec = getattr(qps, precision)
    # first call: 1024^2 (4096 tiles, the array's limit), 2 x 2 ppc;
    # second call: 128^2, 3 x 3 ppc, a smaller input on the same array
    for call, (ind, ppc) in enumerate(((10, (2, 2)), (7, (3, 3)))):
        ctx = qps.Ctx(ind, ind, 4, precision=prec)
        try:
            sp = perturbed_species(ctx, ppc, seed=11 + call)
            rng = np.random.default_rng(call)
            start = rng.standard_normal((1, ctx.ny + 1, ctx.nx + 1))
            want, got = both(ctx, sp, start)
            print(f"[no-alloc] {precision} call {call + 1}: {ctx.nx}^2, {sp.np} particles, "
                  f"max |q| {np.abs(want).max():.3e}", flush=True)
            assert np.array_equal(got, want)
            # the deposit changed q (the comparison is not of two copies of start)
            assert not np.array_equal(want, start)
        finally:
            ctx.close()


def test_noalloc_at_and_beyond_its_scratch():
    """The array holds 2 MiB of halo: 8192 tiles in FP32 (exactly filled
    by a 2048 x 1024 grid, whose deposit must still equal qdp's) and 4096 in
    FP64 (the same grid needs 4 MiB and is refused with QP_ERR_UNSUPPORTED)."""
    ctx = qps.Ctx(11, 10, 4, precision=qps.FP32)   # 2048 x 1024: 8192 tiles
    try:
        sp = perturbed_species(ctx, (1, 1), seed=5)
        start = np.zeros((1, ctx.ny + 1, ctx.nx + 1))
        want, got = both(ctx, sp, start)
        assert np.array_equal(got, want)
    finally:
        ctx.close()
    ctx = qps.Ctx(11, 10, 4, precision=qps.FP64)
    try:
        sp = qps.Species(ctx, ppc=(1, 1))
        q = qps.Field2d(ctx, 1)
        rc = qps.lib().qp_species_qdp_noalloc(sp.h, q.h)
        assert rc == QP_ERR_UNSUPPORTED, rc
        assert b"scratch" in qps.lib().qp_last_error()
    finally:
        ctx.close()
