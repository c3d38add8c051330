"""Pins of the oracle's transverse plasma profiles (oracle/qpo.py
plasma_init_profile, restating dist2d_010/011/012 of fdist2d_class.f03:
318-950), by what the mathematics fixes rather than by the same formulas:

* 10 (elliptical annulus): the kept particles are exactly the quiet-start
  lattice points whose normalised radius ((x-cx)/a)^2 + ((y-cy)/b)^2 lies in
  [1, 1] of the inner and outer ellipses, computed with the semi-axes
  divided first (a different operation order), on parameters where no
  lattice point lies within 1e-9 of either ellipse;
* 11 (Gaussian ring): the total charge equals the analytic integral of
  den * exp(-(r - r0)^2 / (2 sigma^2)) over the plane (the quiet start is the
  midpoint rule of that integral), to 1e-3;
* 12 (piecewise radial): every kept charge is qm times numpy.interp of the
  radial points at the particle's radius, the dropped particles are exactly
  those outside [r(1), r(n)], and radial points that do not increase raise
  the reference's error.
"""
import numpy as np
import pytest
from scipy.special import erf

from oracle import qpo


def grid(ind=6):
    return qpo.Grid(ind, ind, 4)


def lattice(g, spec):
    return qpo.plasma_init(g.nx, g.ny, spec.ppc[0], spec.ppc[1], spec.density, spec.q)


def test_annulus_is_the_lattice_between_two_ellipses():
    g = grid()
    spec = qpo.SpeciesSpec(ppc=(2, 2), profile=10, center=(0.7, -1.1), inner_radius=(2.1, 1.3),
                           outer_radius=(6.3, 4.9))
    part = qpo.plasma_init_profile(g, spec, spec.density)
    lat = lattice(g, spec)
    cx = (spec.center[0] - g.box[0][0]) / g.dx
    cy = (spec.center[1] - g.box[1][0]) / g.dy
    u = (lat[:, 0] - cx) * g.dx
    v = (lat[:, 1] - cy) * g.dy
    rin = (u / spec.inner_radius[0]) ** 2 + (v / spec.inner_radius[1]) ** 2
    rout = (u / spec.outer_radius[0]) ** 2 + (v / spec.outer_radius[1]) ** 2
    assert np.abs(rin - 1).min() > 1e-9 and np.abs(rout - 1).min() > 1e-9   # no ties
    want = lat[(rin >= 1) & (rout <= 1)]
    key = lambda a: a[np.lexsort(a[:, :2].T[::-1])]   # noqa: E731
    assert part.shape == want.shape and 0 < part.shape[0] < lat.shape[0]
    assert np.array_equal(key(part), key(want))
    assert np.all(part[:, 7] == lat[0, 7])


def test_gaussian_ring_charge_is_the_ring_integral():
    g = grid(7)
    r0, sig = 4.0, 0.9          # c/wp; sigma is 5.8 cells
    spec = qpo.SpeciesSpec(ppc=(2, 2), density=1.3, profile=11, center=(0.0, 0.0), radius=r0,
                           width=sig)
    part = qpo.plasma_init_profile(g, spec, spec.density)
    # charge per unit area den * exp(...): the quiet start puts qm/cell-area
    # at each of xppc*yppc points per cell, so sum(q) * dx*dy is the integral
    total = part[:, 7].sum() * g.dx * g.dy
    a = r0 / (sig * np.sqrt(2.0))
    ring = 2 * np.pi * (sig ** 2 * np.exp(-a * a)
                        + r0 * sig * np.sqrt(np.pi / 2) * (1 + erf(a)))
    want = -spec.density * ring
    assert abs(total / want - 1) < 1e-3, (total, want)


def test_piecewise_radial_density_is_linear_interpolation():
    g = grid()
    rp, fr = (0.5, 2.0, 3.5, 6.0), (0.2, 1.0, 0.4, 0.0)
    spec = qpo.SpeciesSpec(ppc=(2, 2), profile=12, center=(0.3, 0.2), radial=(rp, fr))
    part = qpo.plasma_init_profile(g, spec, spec.density)
    lat = lattice(g, spec)
    cx = (spec.center[0] - g.box[0][0]) / g.dx
    cy = (spec.center[1] - g.box[1][0]) / g.dy

    def radius(a):
        return np.hypot(a[:, 0] - cx, a[:, 1] - cy) * g.dx
    rl = radius(lat)
    inside = (rl >= rp[0]) & (rl <= rp[-1])
    assert part.shape[0] == inside.sum()
    r = radius(part)
    np.testing.assert_allclose(part[:, 7], lat[0, 7] * np.interp(r, rp, fr), rtol=1e-12,
                               atol=1e-15)
    bad = qpo.SpeciesSpec(ppc=(2, 2), profile=12, center=(0.3, 0.2),
                          radial=((0.5, 2.0, 2.0, 6.0), fr))
    with pytest.raises(ValueError, match="not monotonically increasing"):
        qpo.plasma_init_profile(g, bad, 1.0)
