"""The transverse plasma profiles 10, 11 and 12 (qp_species_set_profile,
dist2d_010/011/012 of fdist2d_class.f03:318-950) and the beam profiles 3
and 100 (qp_beam_create_file / _torus, dist3d_003 and dist3d_100 of
fdist3d_class.f03:823-1166) on the GPU against the oracle's loaders and
slice loop: the loaded particles equal the oracle's (positions bit for bit,
charges to 1e-14), and four FP64 slices agree to 1e-10 (with bit-exact
per-cell counts for the plasma profiles); an FP32 run of the
non-uniform-charge plasma profile (11) meets the FP32 bar."""
import numpy as np
import pytest

from oracle import qpo

qps = pytest.importorskip("qpslice")

pytestmark = pytest.mark.gpu

IND = 6
BOX = (-10.0, 10.0, -10.0, 10.0, 0.0, 10.0)
PROFILES = {
    10: dict(center=(0.7, -1.1), inner_radius=(2.1, 1.3), outer_radius=(6.3, 4.9)),
    11: dict(center=(0.4, 0.0), radius=3.0, width=0.8),
    12: dict(center=(0.3, 0.2), radial=((0.5, 2.0, 3.5, 8.0), (0.2, 1.0, 0.4, 0.1))),
}


def rel(a, b):
    return float(np.abs(np.asarray(a) - b).max()) / max(float(np.abs(b).max()), 1e-300)


def key(a):
    return a[np.lexsort(a[:, :2].T[::-1])]


def pair(profile, precision):
    grid = qpo.Grid(IND, IND, IND)
    kw = PROFILES[profile]
    spec = qpo.SpeciesSpec(ppc=(2, 2), profile=profile, **kw)
    bspec = qpo.BeamSpec(center=(0.0, 0.0, 1.5), sigma=(0.8, 0.8, 0.5), np_=(32, 32, 64),
                         peak=4.0)
    osim = qpo.Simulation(grid, [spec], [bspec], dt=10.0, iter_=2)
    ctx = qps.Ctx(IND, IND, IND, box=BOX, precision=precision)
    sp = qps.Species(ctx, ppc=(2, 2))
    sp.set_profile(profile, **kw)
    bm = qps.Beam(ctx, center=bspec.center, sigma=bspec.sigma, sigma_v=bspec.sigma_v,
                  np_=bspec.np_, gamma=bspec.gamma, peak=bspec.peak, dt=10.0)
    return grid, osim, ctx, sp, bm


@pytest.mark.parametrize("profile", [10, 11, 12])
def test_profile_load_and_slices_fp64(profile):
    print(f"[gpu-test] plasma profile {profile}", flush=True)
    grid, osim, ctx, sp, bm = pair(profile, qps.FP64)
    try:
        got, want = key(sp.download()), key(osim.species[0].part)
        assert got.shape == want.shape, (got.shape, want.shape)
        assert np.array_equal(got[:, :7], want[:, :7])
        assert rel(got[:, 7], want[:, 7]) < 1e-14
        gsim = qps.Sim(ctx, [sp], [bm], iter_=2)
        osim.begin_step()
        gsim.begin_step()
        for j in range(1, 5):
            osim.slice(j)
            gsim.slice(j)
            for c in range(3):
                assert rel(gsim.bexyz.plane(c + 1, j + 1), osim.e3[c, j]) < 1e-10, (j, c)
                assert rel(gsim.bbxyz.plane(c + 1, j + 1), osim.b3[c, j]) < 1e-10, (j, c)
            counts = np.bincount(qpo.cell_of(grid.nx, grid.ny, osim.species[0].part),
                                 minlength=grid.nx * grid.ny)
            assert np.array_equal(sp.cell_counts(), counts), j
    finally:
        ctx.close()


def test_gaussian_ring_fp32_against_fp64_oracle():
    grid, osim, ctx, sp, bm = pair(11, qps.FP32)
    try:
        gsim = qps.Sim(ctx, [sp], [bm], iter_=2)
        osim.begin_step()
        gsim.begin_step()
        for j in range(1, 5):
            osim.slice(j)
            gsim.slice(j)
        errs = {f"e{c}": rel(gsim.bexyz.plane(c + 1, 5), osim.e3[c, 4]) for c in range(3)}
        errs.update({f"b{c}": rel(gsim.bbxyz.plane(c + 1, 5), osim.b3[c, 4]) for c in range(3)})
        print("fp32 profile 11, slice 4:", {k: f"{v:.1e}" for k, v in errs.items()}, flush=True)
        for k, v in errs.items():
            assert v < (1e-4 if k == "b2" else 1e-5), (k, v)
    finally:
        ctx.close()


def test_profile_errors():
    ctx = qps.Ctx(IND, IND, 4, box=BOX, precision=qps.FP64)
    try:
        sp = qps.Species(ctx, ppc=(2, 2))
        with pytest.raises(qps.QPError, match="not monotonically increasing"):
            sp.set_profile(12, center=(0.0, 0.0), radial=((0.5, 2.0, 2.0, 8.0), (1, 1, 1, 1)))
        with pytest.raises(qps.QPError, match="profile 7"):
            sp.set_profile(7)
    finally:
        ctx.close()


def _beam_slices(ctx, bm, osim, nsl=4):
    sp = qps.Species(ctx, ppc=(2, 2))
    gsim = qps.Sim(ctx, [sp], [bm], iter_=2)
    osim.begin_step()
    gsim.begin_step()
    for j in range(1, nsl + 1):
        osim.slice(j)
        gsim.slice(j)
    for c in range(3):
        assert rel(gsim.bexyz.plane(c + 1, nsl + 1), osim.e3[c, nsl]) < 1e-10, c
        assert rel(gsim.bbxyz.plane(c + 1, nsl + 1), osim.b3[c, nsl]) < 1e-10, c


def test_file_beam_fp64(tmp_path):
    """Beam profile 3: particles read from a text file (z x y vz vx vy q in
    units of 5.32150254e9/sqrt(n0)), against the oracle's reader, then four
    FP64 slices with it as the driver."""
    print("[gpu-test] beam profile 3", flush=True)
    grid = qpo.Grid(IND, IND, IND)
    n0, center = 1e16, (0.0, 0.0, 0.0)
    cwp = 5.32150254e9 / np.sqrt(n0)
    rng = np.random.default_rng(7)
    n = 3000
    # grid positions around (32, 32, 4) in cells -> file units
    tx = 32 + 3 * rng.standard_normal(n)
    ty = 32 + 3 * rng.standard_normal(n)
    tz = 4 + 1.5 * rng.standard_normal(n)
    x0 = (center[0] - grid.box[0][0]) / grid.dx
    y0 = (center[1] - grid.box[1][0]) / grid.dy
    rec = np.stack([-tz * cwp * grid.dz, (tx - x0) * cwp * grid.dx, (ty - y0) * cwp * grid.dy,
                    2e4 + rng.standard_normal(n), rng.standard_normal(n),
                    rng.standard_normal(n), np.full(n, -0.02)], axis=1)
    path = tmp_path / "beam.txt"
    np.savetxt(path, rec, fmt="%.17e")
    want = qpo.beam_init_file(grid, str(path), n, center, n0)
    ctx = qps.Ctx(IND, IND, IND, box=BOX, precision=qps.FP64)
    try:
        bm = qps.FileBeam(ctx, str(path), n, center=center, n0=n0, dt=10.0)
        got = bm.download()
        assert got.shape == want.shape and 0 < got.shape[0] <= n
        assert np.array_equal(key(got), key(want))
        bspec = qpo.BeamSpec(q=-1.0, m=1.0)
        osim = qpo.Simulation(grid, [qpo.SpeciesSpec(ppc=(2, 2))], [bspec], dt=10.0, iter_=2,
                              beam_parts=[want])
        _beam_slices(ctx, bm, osim)
        with pytest.raises(qps.QPError, match="Read beam file error"):
            qps.FileBeam(ctx, str(path), n + 1, center=center, n0=n0)
    finally:
        ctx.close()


def test_torus_beam_fp64():
    """Beam profile 100: the quiet-start Gaussian torus against the oracle's
    lattice, then four FP64 slices with it as the driver."""
    print("[gpu-test] beam profile 100", flush=True)
    grid = qpo.Grid(IND, IND, IND)
    kw = dict(center=(0.0, 0.0, 1.2), range_x=(-4.0, 4.0), range_y=(-4.0, 4.0),
              range_z=(0.4, 2.0), ppc=(2, 2, 2), q=-1.0, np_=0.5, gamma=2e4, r1=1.5, r2=0.6)
    want = qpo.beam_init_torus(grid, kw["center"], kw["range_x"], kw["range_y"], kw["range_z"],
                               kw["ppc"], kw["q"], kw["np_"], kw["gamma"], kw["r1"], kw["r2"])
    ctx = qps.Ctx(IND, IND, IND, box=BOX, precision=qps.FP64)
    try:
        bm = qps.TorusBeam(ctx, dt=10.0, **kw)
        got = bm.download()
        assert got.shape == want.shape
        a, b = key(got), key(want)
        assert np.array_equal(a[:, :6], b[:, :6])
        assert rel(a[:, 6], b[:, 6]) < 1e-14
        bspec = qpo.BeamSpec(q=-1.0, m=1.0)
        osim = qpo.Simulation(grid, [qpo.SpeciesSpec(ppc=(2, 2))], [bspec], dt=10.0, iter_=2,
                              beam_parts=[want])
        _beam_slices(ctx, bm, osim)
    finally:
        ctx.close()
