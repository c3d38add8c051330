"""Pins of the oracle's beam profiles 3 and 100 (oracle/qpo.py
beam_init_file, beam_init_torus; dist3d_003 and dist3d_100 of
fdist3d_class.f03:823-1166), by what fixes them independently:

* 3 (particles from a file): records made from chosen grid positions by the
  inverse of the documented unit map come back at those positions, with the
  momenta reordered from (vz, vx, vy) to (vx, vy, vz); the kept set is cut
  exactly at 1 <= x <= nx-1, 0 <= y < ny, 0 <= z < nz; Fortran D exponents
  and commas parse; a short file raises the reference's message.
* 100 (quiet-start torus): the particle count is the clipped range box's
  cells times ppc, and the total charge equals np times the analytic volume
  integral of exp(-(dz^2 + (r - r1)^2)/(2 r2^2)) in cell units (the quiet
  start is its midpoint rule), to 1e-3.
"""
import numpy as np
import pytest
from scipy.special import erf

from oracle import qpo


def test_file_beam_unit_map_and_cuts(tmp_path):
    g = qpo.Grid(6, 6, 6)
    n0, center = 4e16, (0.5, -1.0, 0.0)
    cwp = 5.32150254e9 / np.sqrt(n0)
    x0 = (center[0] - g.box[0][0]) / g.dx
    y0 = (center[1] - g.box[1][0]) / g.dy
    e = 1e-7   # the edges, approached from inside (kept) and outside (dropped)
    want = np.array([[3.25, 10.5, 5.5], [1 + e, e, e], [63 - e, 64 - e, 64 - e],
                     [1 - e, 5.0, 5.0], [63 + e, 5.0, 5.0], [5.0, 64 + e, 5.0],
                     [5.0, 5.0, 64 + e], [5.0, -e, 5.0]])
    keep = np.array([1, 1, 1, 0, 0, 0, 0, 0], bool)
    lines = []
    for k, (tx, ty, tz) in enumerate(want):
        z = -tz * cwp * g.dz
        x = (tx - x0) * cwp * g.dx
        y = (ty - y0) * cwp * g.dy
        sep = "," if k % 2 else " "
        lines.append(sep.join([f"{z:.17e}".replace("e", "D"), f"{x:.17e}", f"{y:.17e}",
                               f"{100.0 + k}", f"{0.1 * k}", f"{-0.2 * k}", f"{-1e-3 * (k + 1)}"]))
    path = tmp_path / "beam.txt"
    path.write_text("\n".join(lines) + "\n")
    got = qpo.beam_init_file(g, str(path), len(want), center, n0)
    assert got.shape == (keep.sum(), 7)
    np.testing.assert_allclose(got[:, :3], want[keep], rtol=0, atol=1e-9)
    k = np.nonzero(keep)[0]
    np.testing.assert_allclose(got[:, 3], 0.1 * k)            # vx
    np.testing.assert_allclose(got[:, 4], -0.2 * k)           # vy
    np.testing.assert_allclose(got[:, 5], 100.0 + k)          # vz
    np.testing.assert_allclose(got[:, 6], -1e-3 * (k + 1))    # q
    with pytest.raises(ValueError, match="Read beam file error"):
        qpo.beam_init_file(g, str(path), len(want) + 1, center, n0)


def test_torus_count_and_charge_integral():
    g = qpo.Grid(6, 6, 6)              # dx = 20/64, dz = 10/64
    ppc, npk, r1, r2 = (2, 2, 2), 1.7, 3.0, 1.0
    part = qpo.beam_init_torus(g, (0.0, 0.0, 5.0), (-8.0, 8.0), (-8.0, 8.0), (1.0, 9.0), ppc,
                               -1.0, npk, 2e4, r1, r2)
    lo = [int((-8 + 10) / g.dx), int((-8 + 10) / g.dy), int(1.0 / g.dz)]
    hi = [int((8 + 10) / g.dx), int((8 + 10) / g.dy), int(9.0 / g.dz)]
    assert part.shape[0] == np.prod([h - l for l, h in zip(lo, hi)]) * 8
    assert np.all(part[:, 3:5] == 0) and np.all(part[:, 5] == 2e4)
    # the integral in cell units: z in z cells, r in x cells (r1, r2 / dx)
    r1g, r2g = r1 / g.dx, r2 / g.dx
    a = r1g / (r2g * np.sqrt(2))
    ring = 2 * np.pi * (r2g ** 2 * np.exp(-a * a) + r1g * r2g * np.sqrt(np.pi / 2) * (1 + erf(a)))
    want = -npk * np.sqrt(2 * np.pi) * r2g * ring
    assert abs(part[:, 6].sum() / want - 1) < 1e-3, (part[:, 6].sum(), want)
