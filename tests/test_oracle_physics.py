"""Pins the composed parts of the oracle that no reference routine covers
alone (the slice sequence of simulation_class.f03:643-736 and the one-rank
guard glue) with what the mathematics fixes:

* charge conservation of the deposit with the guard fold;
* a single Fourier mode through transform -> pot -> grad -> inverse transform
  equals the closed-form mode answer, including the guard-row copy (Q7);
* the on-axis Ez of a weak bi-Gaussian driver (nb/n0 = 0.01) matches linear
  wake theory: amplitude (nb/n0) sqrt(2 pi) sigma_z exp(-sigma_z^2/2) R(0),
  R(0) = a e^a E1(a), a = sigma_r^2/2 (k_p = 1), with wavelength 2 pi and the
  phase of cos(xi - xi_c) behind the beam.  A dropped term, a wrong sign or a
  wrong factor in the slice composition breaks one of these.
"""
import numpy as np
import pytest
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
    ez = []
    for j in range(1, grid.nz + 1):
        sim.slice(j)
        ez.append(sim.e3[2, j][grid.ny // 2, grid.nx // 2])
    xi = np.arange(1, grid.nz + 1) * grid.dz
    return xi, np.array(ez), beam


def test_linear_wake_on_axis(weak_wake):
    xi, ez, beam = weak_wake
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
