"""GPU parity: every hot-path method of libqpslice.so, called through the C ABI,
against the CPU oracle on the same seeded inputs.

Bars (FP64 context): integer results (cell counts, per-tile counts kpic, sort
keys, beam survivor counts) bit-exact; particle records and fields within
1e-10 relative (max-abs error over max-abs value); the beam loader bit-exact.
FP32 context: within 2e-5 relative of the FP64 oracle for single operations.
"""
import numpy as np
import pytest

from oracle import qpo

qps = pytest.importorskip("qpslice")

pytestmark = pytest.mark.gpu

IND = 6                 # 64 x 64 transverse grid
NX = NY = 2 ** IND
DEX = 20.0 / NX
DXI = 10.0 / 64


def rel(a, b):
    s = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(np.asarray(a) - b).max()) / s


def banner(name):
    print(f"[gpu-test] {name}", flush=True)


@pytest.fixture(params=[qps.FP64, qps.FP32], ids=["fp64", "fp32"])
def ctx(request):
    banner(f"ctx precision={request.param}")
    c = qps.Ctx(IND, IND, 6, precision=request.param)
    yield c
    c.close()


def tol(c):
    return 1e-10 if c.precision == qps.FP64 else 2e-5


def smooth(rng, dim, amp=0.05, n=NX):
    y, x = np.mgrid[0:n + 1, 0:n + 1]
    out = np.zeros((dim, n + 1, n + 1))
    for d in range(dim):
        for _ in range(4):
            kx, ky = rng.integers(1, 5, 2)
            out[d] += amp * rng.normal() * np.sin(np.pi * kx * x / n + rng.uniform()) \
                * np.cos(np.pi * ky * y / n + rng.uniform())
    return out


def canon(part):
    """Plasma particles ordered by their charge column, which _jitter makes a
    unique tag; beam particles ordered lexicographically."""
    if part.shape[1] == 8:
        return part[np.argsort(part[:, 7], kind="stable")]
    return part[np.lexsort(part.T[::-1])]


@pytest.mark.parametrize("kind,dim", [(1, 1), (1, 2), (1, 3), (2, 1), (2, 2), (2, 3), (3, 3)])
def test_transforms(ctx, kind, dim):
    banner(f"transforms kind={kind} dim={dim}")
    rng = np.random.default_rng(kind * 10 + dim)
    f = qps.Field2d(ctx, dim)
    r = rng.standard_normal((dim, NY + 1, NX + 1))
    f.upload(r, space=0)
    f.fftrk(kind)
    assert f.state == 1
    assert rel(f.download(1), qpo.r2k(kind, r)) < (1e-12 if ctx.precision == qps.FP64 else 1e-5)
    k = rng.standard_normal((dim, NX + 1, NY + 1))
    f.upload(k, space=1)
    f.fftkr(kind)
    assert f.state == 0
    assert rel(f.download(0), qpo.k2r(kind, k)) < (1e-12 if ctx.precision == qps.FP64 else 1e-5)


@pytest.mark.parametrize("kind,dim", [(1, 1), (2, 3), (3, 3)])
def test_transforms_column_mode_identical(ctx, kind, dim):
    """Every transform variant (QP_TUNE_FFT: the column-pair y pass, the LDS
    paddings) gives the same bits as the default path, and so the oracle's
    result."""
    banner(f"transforms column mode kind={kind} dim={dim}")
    rng = np.random.default_rng(100 + kind * 10 + dim)
    r = rng.standard_normal((dim, NY + 1, NX + 1))
    k = rng.standard_normal((dim, NX + 1, NY + 1))
    out = []
    for mode in range(ctx.tune_get(qps.TUNE_FFT)[1]):
        ctx.tune(qps.TUNE_FFT, mode)
        f = qps.Field2d(ctx, dim)
        f.upload(r, space=0)
        f.fftrk(kind)
        a = f.download(1)
        f.upload(k, space=1)
        f.fftkr(kind)
        out.append((a, f.download(0)))
    ctx.tune(qps.TUNE_FFT, -1)
    for m in range(1, len(out)):
        assert np.array_equal(out[0][0], out[m][0]) and np.array_equal(out[0][1], out[m][1]), m
    assert rel(out[1][0], qpo.r2k(kind, r)) < (1e-12 if ctx.precision == qps.FP64 else 1e-5)


def test_kspace_solvers(ctx):
    banner("kspace solvers")
    rng = np.random.default_rng(3)
    tab = qpo.pois_table(NX, NY)
    ks = lambda d: rng.standard_normal((d, NX + 1, NY + 1))
    q = ks(1)
    fq = qps.Field2d(ctx, 1, state=1)
    fq.upload(q, space=1)
    for name, fn, d in (("pot", qpo.pot, 1), ("smooth", qpo.smooth, 1), ("elf", qpo.elf, 2),
                        ("grad", None, 2)):
        out = qps.Field2d(ctx, d, state=1)
        getattr(fq, name)(out)
        want = qpo.grad(q) if name == "grad" else fn(tab, q)
        assert rel(out.download(1), want) < tol(ctx), name
    cu = ks(3)
    fcu = qps.Field2d(ctx, 3, state=1)
    fcu.upload(cu, space=1)
    out = qps.Field2d(ctx, 3, state=1)
    fcu.bf(out)
    assert rel(out.download(1), qpo.bf(tab, cu)) < tol(ctx)
    dv = qps.Field2d(ctx, 1, state=1)
    fcu.div(dv)
    assert rel(dv.download(1), qpo.div(cu[:2])) < tol(ctx)
    dcu, amu, b = ks(2), ks(3), ks(2)
    fd, fa, fb = qps.Field2d(ctx, 2, 1), qps.Field2d(ctx, 3, 1), qps.Field2d(ctx, 2, 1)
    fd.upload(dcu, space=1)
    fa.upload(amu, space=1)
    fb.upload(b, space=1)
    fcu.bfqp(fd, fa, fb, DEX * DEX, DEX)
    qpo.bfqp(tab, cu, dcu, amu, b, DEX * DEX, DEX)
    assert rel(fb.download(1), b) < tol(ctx)
    assert rel(fd.download(1), dcu) < tol(ctx)


def test_field_ops_and_guards(ctx):
    banner("field ops")
    rng = np.random.default_rng(5)
    a, b = rng.standard_normal((3, NY + 1, NX + 1)), rng.standard_normal((2, NY + 1, NX + 1))
    fa, fb = qps.Field2d(ctx, 3), qps.Field2d(ctx, 2)
    fa.upload(a, space=0)
    fb.upload(b, space=0)
    fa.mult(fb, [-2.0, 3.0], dim=[1, 2], dim1=[2, 1])
    want = a.copy()
    want[0] = b[1] * -2.0
    want[1] = b[0] * 3.0
    assert rel(fa.download(0), want) < tol(ctx)
    fa.sub(fa, fb, dim=[3], dim1=[3], dim2=[2])
    want[2] = want[2] - b[1]
    assert rel(fa.download(0), want) < tol(ctx)
    fa.ag()
    qpo.guard_fold(want)
    assert rel(fa.download(0), want) < tol(ctx)
    fa.cg()
    qpo.copyguard(want)
    assert rel(fa.download(0), want) < tol(ctx)
    f3 = qps.Field3d(ctx, 3)
    f3.as_(0.0)
    fa.cb(f3, 5, [1, 3], [2, 1])
    g = f3.download()
    assert rel(g[1, 4], want[0]) < tol(ctx) and rel(g[0, 4], want[2]) < tol(ctx)
    fq = qps.Field2d(ctx, 1)
    fq.as_(1.0)
    fq.ca(f3, 5, [2], [1])
    assert rel(fq.download(0)[0], 1.0 + want[0]) < tol(ctx)


@pytest.fixture
def sp_pair(ctx):
    sp = qps.Species(ctx, ppc=(2, 2), dt=DXI)
    part = qpo.plasma_init(NX, NY, 2, 2)
    return sp, part


def test_species_load(ctx, sp_pair):
    banner("species load")
    sp, part = sp_pair
    got = sp.download()
    assert got.shape == part.shape
    lex = lambda a: a[np.lexsort(a.T[::-1])]
    want = part if ctx.precision == qps.FP64 else part.astype(np.float32).astype(np.float64)
    assert np.array_equal(lex(got), lex(want))
    cells = np.bincount(qpo.cell_of(NX, NY, part), minlength=NX * NY)
    assert np.array_equal(sp.cell_counts(), cells)
    assert np.array_equal(sp.tile_counts(), np.bincount(qpo.tile_of(NX, NY, part) - 1,
                                                        minlength=(NX // 16) * (NY // 16)))
    # the library keeps its particles sorted by the tile-major cell key
    c = qpo.cell_of(NX, NY, got)
    cx, cy = c % NX, c // NX
    key = ((cy // 16) * (NX // 16) + cx // 16) * 256 + (cy % 16) * 16 + cx % 16
    assert np.all(np.diff(key) >= 0)


def _jitter(rng, part):
    """Moves every particle inside its cell (the loader puts them at 1/4 and
    3/4), gives it momenta, and tags it through a distinct charge
    q = -0.25 - i 2^-18, exact in FP32 and FP64."""
    p = part.copy()
    p[:, 0] += rng.uniform(-0.2, 0.2, len(p))
    p[:, 1] += rng.uniform(-0.2, 0.2, len(p))
    p[:, 2:5] = rng.normal(0.0, 0.3, (len(p), 3))
    p[:, 7] = -0.25 - np.arange(len(p)) * 2.0 ** -18
    return p


def test_species_qdp(ctx, sp_pair):
    banner("species qdp")
    sp, part = sp_pair
    rng = np.random.default_rng(11)
    p = _jitter(rng, part)
    sp.upload(p)
    q = qps.Field2d(ctx, 1)
    q.as_(0.0)
    sp.qdp(q)
    # oracle: folded deposit of the jittered plasma plus qn of the initial one
    mine = np.zeros((1, NY + 1, NX + 1))
    qpo.qdeposit(NX, NY, p, mine)
    qpo.guard_fold(mine)
    qn = np.zeros((1, NY + 1, NX + 1))
    qpo.qdeposit(NX, NY, part, qn)
    qpo.guard_fold(qn)
    want = -qn + mine
    assert rel(q.download(0), want) < tol(ctx) * 10


def test_species_extpsi_amjdp_push(ctx, sp_pair):
    banner("species extpsi/amjdp/push")
    sp, part = sp_pair
    rng = np.random.default_rng(12)
    p = _jitter(rng, part)
    sp.upload(p)
    ref = canon(sp.download())      # the same records in the library's precision
    psi = smooth(rng, 1, 0.3)
    fxy, bxy, psit = smooth(rng, 2), smooth(rng, 3), smooth(rng, 1)
    F = {}
    for name, arr in (("psi", psi), ("fxy", fxy), ("bxy", bxy), ("psit", psit)):
        F[name] = qps.Field2d(ctx, arr.shape[0])
        F[name].upload(arr, space=0)
    # extpsi
    sp.extpsi(F["psi"], DEX)
    qpo.extpsi(ref, psi, -1.0, DEX)
    got = canon(sp.download())
    assert rel(got, ref) < tol(ctx)
    ref = got.copy()                 # continue from the library's state
    # amjdp
    cu, dcu, amu = (qps.Field2d(ctx, d) for d in (3, 2, 3))
    for f in (cu, dcu, amu):
        f.as_(0.0)
    sp.amjdp(F["fxy"], F["bxy"], F["psit"], cu, amu, dcu, DEX)
    rcu, rdcu, ramu = (np.zeros((d, NY + 1, NX + 1)) for d in (3, 2, 3))
    qpo.amjdp(ref, fxy, bxy, psit, rcu, rdcu, ramu, -1.0, DXI, DEX)
    for a in (rcu, rdcu, ramu):
        qpo.guard_fold(a)
    assert rel(cu.download(0), rcu) < tol(ctx) * 10
    assert rel(dcu.download(0), rdcu) < tol(ctx) * 10
    assert rel(amu.download(0), ramu) < tol(ctx) * 10
    got = canon(sp.download())
    assert rel(got, canon(ref)) < tol(ctx)
    # push (with the reorder): per-cell counts are bit-exact in FP64
    ref = got.copy()
    sp.push(F["fxy"], F["bxy"], F["psit"], DEX)
    assert qpo.push(ref, fxy, bxy, psit, -1.0, DXI, DEX) == 0
    got = sp.download()
    if ctx.precision == qps.FP64:
        assert np.array_equal(sp.cell_counts(), np.bincount(qpo.cell_of(NX, NY, ref),
                                                            minlength=NX * NY))
        assert np.array_equal(sp.tile_counts(), qpo.order_by_tile(NX, NY, ref.copy()))
    assert rel(canon(got), canon(ref)) < tol(ctx)


def test_uniform_charge_paths(ctx, sp_pair):
    """Particles that all carry one charge (a uniform plasma, or an upload of
    one) take the kernels' uniform-charge forms, which neither stage nor move
    q: the charge deposit, the current deposit and the push still match the
    oracle, and q downloads intact after the re-sort."""
    banner("uniform-charge kernels")
    sp, part = sp_pair
    rng = np.random.default_rng(13)
    p = _jitter(rng, part)
    p[:, 7] = -0.25
    sp.upload(p)
    q = qps.Field2d(ctx, 1)
    q.as_(0.0)
    sp.qdp(q)
    mine = np.zeros((1, NY + 1, NX + 1))
    qpo.qdeposit(NX, NY, p, mine)
    qpo.guard_fold(mine)
    qn = np.zeros((1, NY + 1, NX + 1))
    qpo.qdeposit(NX, NY, part, qn)
    qpo.guard_fold(qn)
    assert rel(q.download(0), -qn + mine) < tol(ctx) * 10
    ref = canon(sp.download())
    fxy, bxy, psit = smooth(rng, 2), smooth(rng, 3), smooth(rng, 1)
    F = {}
    for name, arr in (("fxy", fxy), ("bxy", bxy), ("psit", psit)):
        F[name] = qps.Field2d(ctx, arr.shape[0])
        F[name].upload(arr, space=0)
    cu, dcu, amu = (qps.Field2d(ctx, d) for d in (3, 2, 3))
    for f in (cu, dcu, amu):
        f.as_(0.0)
    sp.amjdp(F["fxy"], F["bxy"], F["psit"], cu, amu, dcu, DEX)
    rcu, rdcu, ramu = (np.zeros((d, NY + 1, NX + 1)) for d in (3, 2, 3))
    qpo.amjdp(ref, fxy, bxy, psit, rcu, rdcu, ramu, -1.0, DXI, DEX)
    for a in (rcu, rdcu, ramu):
        qpo.guard_fold(a)
    assert rel(cu.download(0), rcu) < tol(ctx) * 10
    assert rel(dcu.download(0), rdcu) < tol(ctx) * 10
    assert rel(amu.download(0), ramu) < tol(ctx) * 10
    # the push does not read q: the same records in the same order, once with
    # the uniform charge and once tagged (the general path, held to the
    # oracle), end bit-identical but for q
    pre = sp.download()
    tagged = pre.copy()
    tagged[:, 7] = -0.25 - np.arange(len(pre)) * 2.0 ** -18
    sp2 = qps.Species(ctx, ppc=(2, 2), dt=DXI)
    sp2.upload(tagged)
    sp.push(F["fxy"], F["bxy"], F["psit"], DEX)
    sp2.push(F["fxy"], F["bxy"], F["psit"], DEX)
    got, got2 = sp.download(), sp2.download()
    assert np.all(got[:, 7] == -0.25)
    assert np.array_equal(got[:, :7], got2[:, :7])
    ref = canon(tagged)
    assert qpo.push(ref, fxy, bxy, psit, -1.0, DXI, DEX) == 0
    assert rel(canon(got2), canon(ref)) < tol(ctx)
    sp2.close()


def test_beam_loader_and_deposit(ctx):
    banner("beam loader/deposit")
    grid = qpo.Grid(IND, IND, 6)
    spec = qpo.BeamSpec(np_=(16, 16, 32), peak=4.0, sigma=(1.0, 1.0, 1.0),
                        center=(0.0, 0.0, 5.0))
    want, qm, _ = qpo.beam_init(grid, spec)
    b = qps.Beam(ctx, center=spec.center, sigma=spec.sigma, sigma_v=spec.sigma_v,
                 np_=spec.np_, gamma=spec.gamma, peak=spec.peak, dt=10.0)
    got = b.download()
    assert np.array_equal(got, want)            # bit-exact loader (ranorm seed 0)
    b.qdp()
    q3 = np.zeros((1, 65, 65, 65))
    qpo.beam_deposit(want, q3)
    assert rel(b.charge(), q3[0]) < 1e-12       # 2^-44 fixed point
    g, qs = b.moments()
    assert abs(qs - want[:, 6].sum()) <= 1e-12 * abs(want[:, 6].sum())
    gam = np.sqrt(1 + (want[:, 3:6] ** 2).sum(1)).sum()
    assert abs(g - gam) <= 1e-12 * gam


def test_twiss_beam_loader(ctx):
    """Profile 2 (fdist3d_002 / PRVDIST32_TWISS): the library's host loader
    and the oracle draw the same ranorm stream; particles bit-exact."""
    banner("twiss beam loader")
    grid = qpo.Grid(IND, IND, 6)
    spec = qpo.TwissSpec(center=(0.3, -0.2, 5.0), alpha=(0.7, -0.4), beta=(10.0, 8.0),
                         emittance=(3.9, 5.0), sigma_z=0.6, sigma_vz=3.0, np_=(16, 16, 32),
                         gamma=100.0, peak=2.5)
    want, qm, _ = qpo.beam_init_twiss(grid, spec)
    b = qps.TwissBeam(ctx, center=spec.center, alpha=spec.alpha, beta=spec.beta,
                      emittance=spec.emittance, sigma_z=spec.sigma_z, sigma_vz=spec.sigma_vz,
                      np_=spec.np_, gamma=spec.gamma, peak=spec.peak, dt=10.0)
    got = b.download()
    assert got.shape == want.shape
    assert np.array_equal(got, want)


def test_beam_push(ctx):
    banner("beam push")
    grid = qpo.Grid(IND, IND, 6)
    spec = qpo.BeamSpec(np_=(8, 8, 16), peak=4.0, sigma=(3.0, 3.0, 1.5),
                        center=(0.0, 0.0, 5.0))
    part, qm, _ = qpo.beam_init(grid, spec)
    b = qps.Beam(ctx, center=spec.center, sigma=spec.sigma, np_=spec.np_, peak=spec.peak,
                 dt=2000.0)
    rng = np.random.default_rng(13)
    e3 = rng.normal(0, 0.2, (3, 65, 65, 65))
    b3 = rng.normal(0, 0.2, (3, 65, 65, 65))
    fe, fb = qps.Field3d(ctx, 3), qps.Field3d(ctx, 3)
    fe.upload(e3)
    fb.upload(b3)
    if ctx.precision == qps.FP32:
        e3 = e3.astype(np.float32).astype(np.float64)
        b3 = b3.astype(np.float32).astype(np.float64)
    b.push(fe, fb, grid.dx, grid.dz)
    want = qpo.beam_push(part.copy(), e3, b3, -1.0, 2000.0, grid.dx, grid.dz)
    got = b.download()
    assert 0 < got.shape[0] < part.shape[0]
    assert got.shape == want.shape              # the same survivors
    assert rel(canon(got), canon(want)) < 1e-12


def test_beam_push_sorts_by_cell_and_deposit_ignores_order(ctx):
    """The push leaves the survivors sorted by cell (z, then y, then x); the
    charge deposit, whose runs of equal cells are summed across the
    wavefront before the atomics, gives the same 3-D charge, bit for bit,
    for the sorted beam and for the same particles reversed or shuffled, and
    agrees with the oracle's particle-by-particle deposit."""
    banner("beam cell sort + deposit order")
    grid = qpo.Grid(IND, IND, 6)
    # a dense core: many particles per cell, so runs of equal cells are long
    spec = qpo.BeamSpec(np_=(32, 32, 64), peak=4.0, sigma=(0.6, 0.6, 0.6),
                        center=(0.0, 0.0, 4.0))
    b = qps.Beam(ctx, center=spec.center, sigma=spec.sigma, np_=spec.np_, peak=spec.peak,
                 dt=200.0)
    rng = np.random.default_rng(29)
    fe, fb = qps.Field3d(ctx, 3), qps.Field3d(ctx, 3)
    fe.upload(rng.normal(0, 0.05, (3, 65, 65, 65)))
    fb.upload(rng.normal(0, 0.05, (3, 65, 65, 65)))
    b.push(fe, fb, grid.dx, grid.dz)
    got = b.download()
    assert got.shape[0] > 0
    cells = (got[:, 2].astype(np.int64) * grid.ny + got[:, 1].astype(np.int64)) * grid.nx \
        + got[:, 0].astype(np.int64)
    assert np.all(np.diff(cells) >= 0)          # sorted by cell
    assert np.unique(cells, return_counts=True)[1].max() > 64   # runs fill whole wavefronts
    b.qdp()
    q_sorted = b.charge()
    want = np.zeros((1, 65, 65, 65))
    qpo.beam_deposit(got, want)
    assert rel(q_sorted, want[0]) < 1e-12
    for perm in (np.arange(got.shape[0])[::-1], rng.permutation(got.shape[0])):
        b2 = qps.Beam(ctx, np_=(0, 0, 0), npmax=got.shape[0] + 2, dt=200.0)
        b2.upload(got[perm])
        b2.qdp()
        assert np.array_equal(b2.charge(), q_sorted)
        b2.close()
    b.close()
