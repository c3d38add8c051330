"""Pins the composed parts of the oracle that no reference routine covers
alone (the slice sequence of simulation_class.f03:643-736 and the one-rank
guard glue) with what the mathematics fixes:

* charge conservation of the deposit with the guard fold;
* a single Fourier mode through transform -> pot -> grad -> inverse transform
  equals the closed-form mode answer, including the guard-row copy (Q7);
* the on-axis Ez of a weak bi-Gaussian driver (nb/n0 = 0.01) matches linear
  wake theory: amplitude (nb/n0) sqrt(2 pi) sigma_z exp(-sigma_z^2/2) R(0),
  R(0) = a e^a E1(a), a = sigma_r^2/2 (k_p = 1), with wavelength 2 pi and the
  phase of cos(xi - xi_c) behind the beam;
* the Panofsky-Wenzel relation d(Ex - By)/dxi = -dEz/dx of the same run
  (Faraday's law with d/dt = d/dxi, d/dz = -d/dxi): it ties Ez, which the
  slice forms from the divergence of the plasma current (psit,
  simulation_class.f03:671-674), to the transverse wake, which it forms from
  the charge (psi, :660-667);
* the total transverse magnetic field of the same run (beam plus plasma
  response, simulation_class.f03:677-691,717-724) against its linear-theory
  closed form.  Linear cold-fluid theory with d/dt -> d/dxi gives
  dJz_plasma/dxi = Ez, so curl B_perp = Jz + dEz/dxi = (lap_perp - 1)^-1 rho_b
  + rho_b and div B_perp = dBz/dxi = 0 to first order; hence
  B_perp = z x grad A with A = (lap_perp - 1)^-1 rho_b, local in xi.  The
  plasma part cancels about 92 % of the beam's own field at k_p sigma_r = 2,
  so a wrong sign, a swapped component or a missing dex factor in the plasma
  assembly moves the total by a large factor.  (The transverse wake
  Ex - By alone does not see B_perp: the slice adds the same By to Ex that
  it stores as By.)
* the plasma loader against the quiet-start lattice written out
  independently (fdist2d_class.f03:254-273).
A dropped term, a wrong sign or a wrong factor in the slice composition breaks
one of these.
"""
import numpy as np
import pytest
from scipy.fft import dst, idst
from scipy.special import exp1

from oracle import qpo


def test_charge_conservation_with_guard_fold():
    nx = ny = 32
    part = qpo.plasma_init(nx, ny, 2, 2)
    rng = np.random.default_rng(1)
    part[:, 0] = rng.uniform(0.0, nx - 1e-9, part.shape[0])
    part[:, 1] = rng.uniform(0.0, ny - 1e-9, part.shape[0])
    part[:, 7] = rng.uniform(-1.0, 0.5, part.shape[0])
    q = np.zeros((1, ny + 1, nx + 1))
    qpo.qdeposit(nx, ny, part, q)
    qpo.guard_fold(q)
    assert np.all(q[0, ny] == 0.0)                       # row ny folded into row 0
    assert abs(q.sum() - part[:, 7].sum()) < 1e-11 * np.abs(part[:, 7]).sum()


@pytest.mark.parametrize("a,b", [(1, 1), (3, 2), (7, 5)])
def test_single_mode_psi_and_field(a, b):
    nx = ny = 32
    y, x = np.mgrid[0:ny + 1, 0:nx + 1].astype(float)
    sx, cx = np.sin(np.pi * a * x / nx), np.cos(np.pi * a * x / nx)
    sy, cy = np.sin(np.pi * b * y / ny), np.cos(np.pi * b * y / ny)
    q = (sx * sy)[None]
    g, s = qpo.pois_table(nx, ny)
    qk = qpo.r2k(1, q)
    psi_k = qpo.pot((g, s), qk)
    psi = qpo.k2r(1, psi_k)
    fxy = qpo.k2r(1, qpo.grad(psi_k))
    dkx = 6.28318530717959 / (2 * nx) * a          # the reference's 2 pi literal (Q4)
    dky = 6.28318530717959 / (2 * ny) * b
    # the Green's function with the particle shape factor (fpois2d_lib77.f:1514-1531)
    ax = ay = 0.912871
    gk = np.exp(-0.5 * ((dkx * ax) ** 2 + (dky * ay) ** 2)) / (dkx ** 2 + dky ** 2)
    rows = slice(0, ny)                             # row ny is the guard copy of row 0
    assert np.allclose(psi[0, rows], (gk * sx * sy)[rows], atol=1e-12 * gk, rtol=0)
    assert np.allclose(fxy[0, rows], (dkx * gk * cx * sy)[rows], atol=1e-12 * gk, rtol=0)
    assert np.allclose(fxy[1, rows], (dky * gk * sx * cy)[rows], atol=1e-12 * gk, rtol=0)
    assert np.array_equal(fxy[:, ny], fxy[:, 0])    # Q7: row ny <- row 0


@pytest.fixture(scope="module")
def weak_wake():
    grid = qpo.Grid(6, 6, 8)                        # 64 x 64 x 256 over 20 x 20 x 10
    beam = qpo.BeamSpec(center=(0.0, 0.0, 2.0), sigma=(2.0, 2.0, 0.5), np_=(32, 32, 64),
                        peak=0.01)
    sim = qpo.Simulation(grid, [qpo.SpeciesSpec(ppc=(2, 2))], [beam], dt=10.0, iter_=2)
    sim.begin_step()
    for j in range(1, grid.nz + 1):
        sim.slice(j)
    xi = np.arange(1, grid.nz + 1) * grid.dz
    return xi, sim, beam


def test_linear_wake_on_axis(weak_wake):
    xi, sim, beam = weak_wake
    g = sim.grid
    ez = sim.e3[2, 1:, g.ny // 2, g.nx // 2]
    sr, sz, xc = beam.sigma[0], beam.sigma[2], beam.center[2]
    aa = sr * sr / 2.0
    amp = beam.peak * np.sqrt(2 * np.pi) * sz * np.exp(-sz * sz / 2) * aa * np.exp(aa) * exp1(aa)
    behind = xi > xc + 3.0 * sz
    m = np.vstack([np.cos(xi[behind] - xc), np.sin(xi[behind] - xc)]).T
    (c, s), *_ = np.linalg.lstsq(m, ez[behind], rcond=None)
    fit_amp, phase = np.hypot(c, s), np.arctan2(s, c)
    assert abs(fit_amp / amp - 1.0) < 0.08, (fit_amp, amp)
    assert abs(phase) < 0.1, phase                  # Ez = A cos(xi - xi_c), A > 0
    # wavelength 2 pi c/wp: successive zero crossings pi apart
    zc = xi[behind][np.where(np.diff(np.sign(ez[behind])) != 0)[0]]
    assert len(zc) >= 2
    assert abs(np.diff(zc).mean() / np.pi - 1.0) < 0.02


def test_panofsky_wenzel(weak_wake):
    """d(Ex - By)/dxi = -dEz/dx and d(Ey + Bx)/dxi = -dEz/dy over the wake
    behind the beam, slices j and j+1 (centred differences in x and y, a
    forward difference in xi; Ez averaged over the two slices)."""
    xi, sim, beam = weak_wake
    g = sim.grid
    d, dxi = g.dx, g.dz
    wx = sim.e3[0] - sim.b3[1]
    wy = sim.e3[1] + sim.b3[0]
    ez = sim.e3[2]
    inn = (slice(2, g.ny - 1), slice(2, g.nx - 1))
    worst = 0.0
    for j in range(60, g.nz - 1, 10):
        e = 0.5 * (ez[j] + ez[j + 1])
        for w, axis in ((wx, 1), (wy, 0)):
            lhs = ((w[j + 1] - w[j]) / dxi)[inn]
            rhs = -(np.gradient(e, axis=axis) / d)[inn]
            worst = max(worst, np.abs(lhs - rhs).max() / np.abs(rhs).max())
    assert worst < 0.03, worst        # measured about 1 %


def _screened_field(rho, nx, ny, d):
    """(Bx, By) = (-dA/dy, dA/dx), A = (lap - 1)^-1 rho with conducting walls
    (sine series on the nodes 1..n-1), lap in units of c/wp; derivatives by
    centred differences."""
    inner = rho[1:ny, 1:nx]
    f = dst(dst(inner, type=1, axis=0), type=1, axis=1)
    kx = np.pi * np.arange(1, nx) / nx / d
    ky = np.pi * np.arange(1, ny) / ny / d
    a = np.zeros((ny + 1, nx + 1))
    a[1:ny, 1:nx] = idst(idst(-f / (kx[None, :] ** 2 + ky[:, None] ** 2 + 1.0), type=1, axis=0),
                         type=1, axis=1)
    return -np.gradient(a, axis=0) / d, np.gradient(a, axis=1) / d


def test_linear_transverse_magnetic_field(weak_wake):
    """Stored B_perp inside the driver against z x grad (lap - 1)^-1 rho_b,
    rho_b the beam charge plane the slice used: amplitude within 3 % and the
    shape within 15 % (the rest is the beam's and the plasma's particle
    noise).  Behind the driver B_perp is second order: below 3 % of its
    peak inside the beam."""
    xi, sim, beam = weak_wake
    g = sim.grid
    inn = (slice(2, g.ny - 1), slice(2, g.nx - 1))
    zc, sz = beam.center[2], beam.sigma[2]
    inside = [j for j in range(1, g.nz + 1) if abs(j * g.dz - zc) < 1.0 * sz]
    peak = 0.0
    for j in inside:
        bx, by = _screened_field(sim.q3[0, j], g.nx, g.ny, g.dx)
        for got, want in ((sim.b3[0, j][inn], bx[inn]), (sim.b3[1, j][inn], by[inn])):
            ratio = (got * want).sum() / (want * want).sum()
            shape = np.linalg.norm(got - ratio * want) / np.linalg.norm(got)
            assert abs(ratio - 1.0) < 0.03, (j, ratio)
            assert shape < 0.15, (j, shape)
            peak = max(peak, np.abs(got).max())
    behind = [j for j in range(1, g.nz + 1) if j * g.dz > zc + 5.0 * sz]
    tail = max(np.abs(sim.b3[0:2, j][(slice(None),) + inn]).max() for j in behind)
    assert tail < 0.03 * peak, (tail, peak)


@pytest.mark.parametrize("nx,ny,xppc,yppc,den,q", [(32, 32, 2, 2, 1.0, -1.0),
                                                  (64, 16, 4, 1, 0.5, 1.0),
                                                  (16, 64, 3, 2, 2.0, -2.0)])
def test_plasma_loader_is_the_quiet_start_lattice(nx, ny, xppc, yppc, den, q):
    """dist2d_000 on one rank (noff = 0 < ny, fdist2d_class.f03:254-273):
    cells x = 1..nx-2 and y = 1..ny-1 (0-based), offsets (k + 1/2)/ppc inside
    each cell, charge den sign(q)/(xppc yppc), momenta 0, p6 = p7 = 1.
    Positions are bit-exact: the reference evaluates (k + 0.5)/ppc + i - 1
    left to right with the 1-based cell index i, and so does this lattice."""
    part = qpo.plasma_init(nx, ny, xppc, yppc, den, q)
    ox = (np.arange(xppc) + 0.5) / xppc
    oy = (np.arange(yppc) + 0.5) / yppc
    xs = ((ox[None, :] + np.arange(2, nx)[:, None]) - 1.0).ravel()
    ys = ((oy[None, :] + np.arange(2, ny + 1)[:, None]) - 1.0).ravel()
    assert np.allclose(xs, (np.arange(1, nx - 1)[:, None] + ox[None, :]).ravel(), rtol=0,
                       atol=1e-14)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    want = np.zeros((X.size, 8))
    want[:, 0], want[:, 1] = X.ravel(), Y.ravel()
    want[:, 5] = want[:, 6] = 1.0
    want[:, 7] = den * np.sign(q) / (xppc * yppc)
    assert part.shape == want.shape
    key = lambda a: a[np.lexsort((a[:, 1], a[:, 0]))]   # noqa: E731
    assert np.array_equal(key(part), key(want))


def test_piecewise_density_factor():
    """den_temp of dist2d_000 (fdist2d_class.f03:233-251) at points where the
    linear interpolation is known, and its two error cases."""
    pw = ((0.0, 10.0, 20.0, 20.0, 40.0), (1.0, 3.0, 0.5, 2.0, 0.0))
    for s, want in ((0.0, 1.0), (5.0, 2.0), (10.0, 3.0), (15.0, 1.75), (20.0, 0.5),
                    (30.0, 1.0), (40.0, 0.0)):
        assert abs(qpo.density_factor(pw, s) - want) < 1e-15, s
    assert qpo.density_factor(None, 123.0) == 1.0
    with pytest.raises(ValueError, match="out of the bound"):
        qpo.density_factor(pw, 40.5)
    with pytest.raises(ValueError, match="monotonically"):
        qpo.density_factor(((0.0, 10.0, 5.0, 30.0), (1.0, 1.0, 1.0, 1.0)), 20.0)
    # a decreasing pair after the segment found is not examined (the loop exits)
    assert qpo.density_factor(((0.0, 10.0, 5.0), (1.0, 2.0, 3.0)), 5.0) == 1.5


def test_piecewise_density_scales_the_charge():
    """The plasma at s carries den_temp times the uniform plasma's charge."""
    grid = qpo.Grid(5, 5, 3)
    pw = ((0.0, 100.0), (0.5, 2.5))
    sp = qpo.Species(grid, qpo.SpeciesSpec(ppc=(2, 2), piecewise=pw), 0.1)
    uni = qpo.Species(grid, qpo.SpeciesSpec(ppc=(2, 2)), 0.1)
    assert np.allclose(sp.part[:, 7], 0.5 * uni.part[:, 7], rtol=0, atol=1e-16)
    sp.renew(50.0)
    assert np.allclose(sp.part[:, 7], 1.5 * uni.part[:, 7], rtol=0, atol=1e-16)
    assert np.allclose(sp.qn, 1.5 * uni.qn, rtol=1e-14, atol=0)
