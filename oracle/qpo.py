"""CPU oracle for the quasi-static slice solver -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py import this module, and only as the checker.  The product (the HIP
library in quickpic-opensource_amd/) never calls it.

The heavy loops live in ``libqpo.so`` (oracle/qpo.c, a C restatement of the
reference kernels).  This module binds them with ctypes and restates, on top of
them, the reference's host-side sequencing:

* the species methods qdp / extpsi / amjdp / push / cbq / renew
  (species2d_class.f03:74-418),
* the field state machine and component-mapped arithmetic
  (field2d_class.f03:157-807, ufield2d_class.f03:558-811),
* one xi slice of the 3-D step (simulation_class.f03:643-736) and the 3-D step
  around it (simulation_class.f03:611-765).

Array conventions: real-space fields are ``(dim, ny+1, nx+1)``, k-space fields
``(dim, nx+1, ny+1)`` (the reference's transposed k order,
ufield2d_class.f03:198), plasma particles ``(np, 8)`` and beam particles
``(np, 7)``, 3-D plane stacks ``(dim, nz+1, ny+1, nx+1)``; all float64.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

AX = AY = 0.912871     # particle shape half-width (field2d_class.f03:129)
AFFP = 1.0


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libqpo.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        _declare(L)
        _LIB = L
    return _LIB


_P = C.c_void_p
_I = C.c_int
_L = C.c_long
_D = C.c_double


class Rng(C.Structure):
    _fields_ = [("r1", C.c_int32), ("r2", C.c_int32), ("r4", C.c_int32),
                ("r5", C.c_int32), ("iflg", C.c_int32), ("r0", C.c_double)]


class BeamParams(C.Structure):
    _fields_ = [("nx", _I), ("ny", _I), ("nz", _I),
                ("dx", _D), ("dy", _D), ("dz", _D),
                ("bcx", _D), ("bcy", _D), ("bcz", _D),
                ("sigx", _D), ("sigy", _D), ("sigz", _D),
                ("sigvx", _D), ("sigvy", _D), ("sigvz", _D),
                ("cx", _D * 3), ("cy", _D * 3),
                ("gamma", _D), ("peak", _D), ("q", _D),
                ("npx", _I), ("npy", _I), ("npz", _I), ("quiet", _I)]


def _declare(L):
    sig = {
        "qpo_rng_init": (None, [_P]),
        "qpo_ranorm": (_D, [_P]),
        "qpo_beam_init": (_L, [_P, _P, _L, _P, _P, _P]),
        "qpo_plasma_init": (_L, [_I, _I, _I, _I, _D, _D, _P]),
        "qpo_qdeposit": (None, [_I, _I, _L, _P, _P]),
        "qpo_guard_fold": (None, [_I, _I, _I, _P]),
        "qpo_copyguard": (None, [_I, _I, _I, _P]),
        "qpo_extpsi": (None, [_I, _I, _L, _P, _P, _D, _D]),
        "qpo_amjdp": (None, [_I, _I, _L, _P, _P, _P, _P, _P, _P, _P, _D, _D, _D]),
        "qpo_push": (_I, [_I, _I, _L, _P, _P, _P, _P, _D, _D, _D]),
        "qpo_tile_of": (None, [_I, _I, _L, _P, _P]),
        "qpo_cell_of": (None, [_I, _I, _L, _P, _P]),
        "qpo_order_by_tile": (None, [_I, _I, _L, _P, _P, _P]),
        "qpo_transform_1d": (None, [_I, _I, _I, _P, _P]),
        "qpo_fft_r2k": (None, [_I, _I, _I, _P, _P, _P, _P]),
        "qpo_fft_k2r": (None, [_I, _I, _I, _P, _P, _P, _P]),
        "qpo_kind_types": (None, [_I, _I, _P, _P]),
        "qpo_pois_table": (None, [_I, _I, _D, _D, _D, _P, _P]),
        "qpo_div": (None, [_I, _I, _P, _P]),
        "qpo_grad": (None, [_I, _I, _P, _P]),
        "qpo_pot": (None, [_I, _I, _P, _P, _P, _P]),
        "qpo_smooth": (None, [_I, _I, _P, _P, _P, _P]),
        "qpo_elf": (None, [_I, _I, _P, _P, _P, _P]),
        "qpo_bf": (None, [_I, _I, _P, _P, _P, _P]),
        "qpo_bfqp": (None, [_I, _I, _P, _P, _P, _P, _P, _P, _D, _D]),
        "qpo_beam_deposit": (None, [_I, _I, _I, _L, _P, _P]),
        "qpo_beam_push": (_L, [_I, _I, _I, _L, _P, _P, _P, _D, _D, _D, _D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# random numbers, initial particles
# --------------------------------------------------------------------------
def rng_new():
    s = Rng()
    lib().qpo_rng_init(C.byref(s))
    return s


def ranorm_stream(n, rng=None):
    rng = rng or rng_new()
    return np.array([lib().qpo_ranorm(C.byref(rng)) for _ in range(n)])


@dataclass
class Grid:
    indx: int
    indy: int
    indz: int
    box: tuple = ((-10.0, 10.0), (-10.0, 10.0), (0.0, 10.0))

    @property
    def nx(self):
        return 2 ** self.indx

    @property
    def ny(self):
        return 2 ** self.indy

    @property
    def nz(self):
        return 2 ** self.indz

    @property
    def dx(self):  # simulation_class.f03:176-178
        return (self.box[0][1] - self.box[0][0]) / float(2 ** self.indx)

    @property
    def dy(self):
        return (self.box[1][1] - self.box[1][0]) / float(2 ** self.indy)

    @property
    def dz(self):
        return (self.box[2][1] - self.box[2][0]) / float(2 ** self.indz)


@dataclass
class BeamSpec:
    """Tri-Gaussian beam, profile 0 (fdist3d_class.f03:222-340)."""
    center: tuple = (0.0, 0.0, 2.0)
    sigma: tuple = (0.5, 0.5, 0.5)
    sigma_v: tuple = (1.0, 1.0, 0.0)
    np_: tuple = (64, 64, 256)
    gamma: float = 2.0e4
    peak: float = 4.0
    q: float = -1.0
    m: float = 1.0
    quiet: bool = True
    centroid_x: tuple = (0.0, 0.0, 0.0)
    centroid_y: tuple = (0.0, 0.0, 0.0)


@dataclass
class TwissSpec:
    """Beam profile 2 (fdist3d_class.f03:652-821): Twiss parameters in x, y,
    Gaussian in z.  Lengths in c/wp as in the input deck."""
    center: tuple = (0.0, 0.0, 2.0)
    alpha: tuple = (0.0, 0.0)
    beta: tuple = (1.0, 1.0)
    emittance: tuple = (1e-4, 1e-4)
    sigma_z: float = 0.5
    sigma_vz: float = 0.0
    np_: tuple = (32, 32, 64)
    gamma: float = 2.0e4
    peak: float = 4.0
    q: float = -1.0
    m: float = 1.0
    quiet: bool = True


@dataclass
class SpeciesSpec:
    """A plasma species: uniform (profile 0, fdist2d_class.f03:168-316), an
    elliptical annulus (10, :318-523), a Gaussian ring (11, :525-711) or a
    piecewise-linear radial density (12, :713-950), centred at `center`
    (c/wp, box coordinates); piecewise: an optional longitudinal profile (s
    points, fs factors) of :190-201, for every profile."""
    ppc: tuple = (2, 2)
    density: float = 1.0
    q: float = -1.0
    m: float = 1.0
    piecewise: tuple = None
    profile: int = 0
    center: tuple = (0.0, 0.0)
    inner_radius: tuple = (0.0, 0.0)   # profile 10
    outer_radius: tuple = (0.0, 0.0)   # profile 10
    radius: float = 0.0                # profile 11
    width: float = 1.0                 # profile 11
    radial: tuple = None               # profile 12: (r points, densities)


def density_factor(piecewise, s):
    """den_temp of dist2d_000 (fdist2d_class.f03:233-251): fs interpolated at
    s between the points s(i-1) <= s <= s(i); 1 without a profile."""
    if piecewise is None:
        return 1.0
    ps, fs = piecewise
    if s < ps[0] or s > ps[-1]:
        raise ValueError("The s is out of the bound!")
    for i in range(1, len(ps)):
        if ps[i] < ps[i - 1]:
            raise ValueError("s is not monotonically increasing!")
        if s <= ps[i]:
            return fs[i - 1] + (fs[i] - fs[i - 1]) / (ps[i] - ps[i - 1]) * (s - ps[i - 1])
    return 1.0


def beam_params(grid: Grid, b: BeamSpec) -> BeamParams:
    bp = BeamParams()
    bp.nx, bp.ny, bp.nz = grid.nx, grid.ny, grid.nz
    bp.dx, bp.dy, bp.dz = grid.dx, grid.dy, grid.dz
    bp.bcx = b.center[0] - grid.box[0][0]
    bp.bcy = b.center[1] - grid.box[1][0]
    bp.bcz = b.center[2] - grid.box[2][0]
    bp.sigx, bp.sigy, bp.sigz = b.sigma
    bp.sigvx, bp.sigvy, bp.sigvz = b.sigma_v
    for i in range(3):
        bp.cx[i] = b.centroid_x[i]
        bp.cy[i] = b.centroid_y[i]
    bp.gamma, bp.peak, bp.q = b.gamma, b.peak, b.q
    bp.npx, bp.npy, bp.npz = b.np_
    bp.quiet = 1 if b.quiet else 0
    return bp


def beam_init_twiss(grid: Grid, b: TwissSpec, npmax=None, rng=None):
    """Profile 2: init_fdist3d_002 units and charge (fdist3d_class.f03:720-745)
    and PRVDIST32_TWISS (part3d_lib77.f:2-192) with one partition covering
    the box, drawing from the same ranorm stream.  Returns (particles (np,7),
    qm, ierr).  Pure Python over the C ranorm: for small beams."""
    import math
    rng = rng or rng_new()
    r = lambda: lib().qpo_ranorm(C.byref(rng))  # noqa: E731
    dx, dy, dz = grid.dx, grid.dy, grid.dz
    gamma = b.gamma
    npx, npy, npz = b.np_
    npmax = npmax or npx * npy * npz + 2
    sxp = math.sqrt(b.beta[0] * b.emittance[0] / gamma)
    syp = math.sqrt(b.beta[1] * b.emittance[1] / gamma)
    qm = b.q / abs(b.q) * abs(b.peak) * (2 * 3.1415926535897932) ** 1.5 * sxp * syp * b.sigma_z
    qm = qm / dx / dy / dz
    qm = qm / npx
    qm = qm / npy
    qm = qm / npz
    x0 = (b.center[0] - grid.box[0][0]) / dx
    y0 = (b.center[1] - grid.box[1][0]) / dy
    z0 = (b.center[2] - grid.box[2][0]) / dz
    sigz, vtz, vdz = b.sigma_z / dz, b.sigma_vz, gamma
    ax, ay = b.alpha
    bx, by = b.beta[0] / dx, b.beta[1] / dy
    ex, ey = b.emittance[0] / dx, b.emittance[1] / dy
    nx, ny, nz = grid.nx, grid.ny, grid.nz
    sigx = math.sqrt(bx * ex / gamma)
    sigy = math.sqrt(by * ey / gamma)
    vtx, vty = ex / sigx, ey / sigy
    blx, bly, blz = max(x0 - 3.0 * sigx, 1.0), max(y0 - 3.0 * sigy, 1.0), max(z0 - 3.0 * sigz, 1.0)
    brx, bry = min(x0 + 3.0 * sigx, float(nx - 1)), min(y0 + 3.0 * sigy, float(ny - 1))
    brz = min(z0 + 3.0 * sigz, float(nz - 1))
    out, ierr = [], 0
    j, l = 0, npz
    while j < npx:
        k = 0
        while k < npy:
            l -= npz
            while l < npz:
                while True:
                    tz = z0 + sigz * r()
                    if not (tz >= brz or tz <= blz):
                        break
                while True:                              # label 20
                    while True:
                        tx = r()
                        px_ = x0 + sigx * tx
                        if not (px_ >= brx or px_ <= blx):
                            break
                    while True:                          # label 30
                        ty = r()
                        py_ = y0 + sigy * ty
                        if not (py_ >= bry or py_ <= bly):
                            break
                    if not (tx * tx + ty * ty > 9.0):
                        break
                vx = vtx * r()
                vy = vty * r()
                vz = vtz * r() + vdz
                vx = vx - gamma * ax / bx * (px_ - x0)
                vy = vy - gamma * ay / by * (py_ - y0)
                if 0.0 <= py_ < ny and 0.0 <= tz < nz:
                    for rec in ((px_, py_, tz, vx, vy, vz, qm),
                                (2 * x0 - px_, 2 * y0 - py_, tz, -vx, -vy, vz, qm))[
                                    : 2 if b.quiet else 1]:
                        if len(out) + 1 < npmax:
                            out.append(rec)
                        else:
                            ierr += 1
                l += 1
                if b.quiet:
                    l += 1
            k += 1
        j += 1
    return np.array(out, dtype=np.float64).reshape(-1, 7), qm, ierr


def beam_init_file(grid: Grid, path, np_, center=(0.0, 0.0, 0.0), n0=1e16):
    """dist3d_003 (fdist3d_class.f03:899-974, one partition): np records
    "z x y vz vx vy q" of a text file (blanks or commas, D exponents),
    tx = x/cwp/dx + x0, ty = y/cwp/dy + y0, tz = -z/cwp/dz with cwp =
    5.32150254e9/sqrt(n0) (:846) and x0, y0 the centre in grid units; kept
    when 1 <= tx <= nx-1, 0 <= ty < ny, 0 <= tz < nz.  Returns (n, 7)."""
    cwp = 5.32150254 * 1e9 / np.sqrt(n0)
    x0 = (center[0] - grid.box[0][0]) / grid.dx
    y0 = (center[1] - grid.box[1][0]) / grid.dy
    with open(path) as fh:
        vals = [float(t.replace("D", "e").replace("d", "e"))
                for t in fh.read().replace(",", " ").split()]
    if len(vals) < 7 * np_:
        raise ValueError("Read beam file error")
    rec = np.array(vals[:7 * np_]).reshape(np_, 7)
    tx = rec[:, 1] / cwp / grid.dx + x0
    ty = rec[:, 2] / cwp / grid.dy + y0
    tz = -rec[:, 0] / cwp / grid.dz
    keep = (tx >= 1) & (tx <= grid.nx - 1) & (ty >= 0) & (ty < grid.ny) & (tz >= 0) & \
        (tz < grid.nz)
    out = np.stack([tx, ty, tz, rec[:, 4], rec[:, 5], rec[:, 3], rec[:, 6]], axis=1)
    return np.ascontiguousarray(out[keep])


def beam_init_torus(grid: Grid, center, range_x, range_y, range_z, ppc, q, np_, gamma, r1, r2):
    """dist3d_100 (fdist3d_class.f03:1084-1166, one partition): the lattice
    (ix + 0.5)/xppc + i - 1 over int(lb)..int(ub)-1 of the range box (grid
    units, clipped to the grid), p = (0, 0, gamma), q = qm np exp(-a4/r2^2
    * 0.5) with a4 = (tz - bcz)^2 + (sqrt((tx-bcx)^2 + (ty-bcy)^2) - r1)^2,
    qm = sign(q)/(xppc yppc zppc) and r1, r2 in x cells (:1041-1076).
    Returns (n, 7)."""
    d = (grid.dx, grid.dy, grid.dz)
    mins = (grid.box[0][0], grid.box[1][0], grid.box[2][0])
    nd = (grid.nx, grid.ny, grid.nz)
    bc = [(center[k] - mins[k]) / d[k] for k in range(3)]
    rng_ = (range_x, range_y, range_z)
    lb = [(rng_[k][0] - mins[k]) / d[k] for k in range(3)]
    ub = [(rng_[k][1] - mins[k]) / d[k] for k in range(3)]
    for k in (1, 2):
        if lb[k] > nd[k] or ub[k] < 0:
            return np.zeros((0, 7))
    lb = [max(lb[k], 0.0) for k in range(3)]
    ub = [min(ub[k], float(nd[k])) for k in range(3)]
    qm = q / abs(q) / ppc[0] / ppc[1] / ppc[2]
    r1g, r2g = r1 / grid.dx, r2 / grid.dx
    axes = []
    for k in range(3):
        cells = np.arange(int(lb[k]), int(ub[k]))
        off = (np.arange(ppc[k]) + 0.5) / ppc[k]
        axes.append((off[None, :] + cells[:, None] - 1))
    rows = []
    for i in range(axes[0].shape[0]):
        for j in range(axes[1].shape[0]):
            for k in range(axes[2].shape[0]):
                tx, ty, tz = np.meshgrid(axes[0][i], axes[1][j], axes[2][k], indexing="ij")
                tx, ty, tz = tx.ravel(), ty.ravel(), tz.ravel()
                a1 = np.abs(tz - bc[2])
                a2 = np.sqrt((tx - bc[0]) ** 2 + (ty - bc[1]) ** 2)
                a3 = np.abs(a2 - r1g)
                a4 = a1 ** 2 + a3 ** 2
                qq = qm * np_ * np.exp(-a4 / r2g ** 2 * 0.5)
                z0 = np.zeros_like(tx)
                rows.append(np.stack([tx, ty, tz, z0, z0, z0 + gamma, qq], axis=1))
    return np.concatenate(rows) if rows else np.zeros((0, 7))


def beam_init(grid: Grid, b: BeamSpec, npmax=None, rng=None):
    """Returns (particles (np,7), qm, ierr)."""
    npx, npy, npz = b.np_
    npmax = npmax or npx * npy * npz + 2
    part = np.zeros((npmax, 7))
    rng = rng or rng_new()
    ierr = C.c_int(0)
    qm = C.c_double(0.0)
    bp = beam_params(grid, b)
    n = lib().qpo_beam_init(C.byref(bp), _p(part), npmax, C.byref(rng),
                            C.byref(ierr), C.byref(qm))
    return part[:n].copy(), qm.value, ierr.value


def plasma_init(nx, ny, xppc, yppc, den=1.0, q=-1.0):
    n = (nx - 2) * (ny - 1) * xppc * yppc
    part = np.zeros((n, 8))
    got = lib().qpo_plasma_init(nx, ny, xppc, yppc, den, q, _p(part))
    assert got == n
    return part


def plasma_init_profile(grid, spec, den):
    """dist2d_010 / dist2d_011 / dist2d_012 with one partition (noff = 0):
    the quiet-start lattice of dist2d_000, then the profile's selection and
    charge weight in the reference's operation order (numpy FP64):
      10 keeps a particle unless (x-cx)^2/irx^2 + (y-cy)^2/iry^2 < 1 or
         (x-cx)^2/orx^2 + (y-cy)^2/ory^2 > 1 (fdist2d_class.f03:438-441);
      11 weights q by exp(-(r-r0)^2 * 0.5/sigr^2) (:633-645);
      12 keeps r(1) <= r <= r(n) and weights q by fr linearly interpolated
         at r (:836-856), checking that r increases.
    Radii are in grid units (divided by dx; profile 10's y radii by dy) and
    the centre is (center - box min) / dx (:340-385, 547-588, 740-775)."""
    nx, ny = grid.nx, grid.ny
    part = plasma_init(nx, ny, spec.ppc[0], spec.ppc[1], den, spec.q)
    dx, dy = grid.dx, grid.dy
    cx = (spec.center[0] - grid.box[0][0]) / dx
    cy = (spec.center[1] - grid.box[1][0]) / dy
    x, y = part[:, 0], part[:, 1]
    if spec.profile == 10:
        irx, iry = spec.inner_radius[0] / dx, spec.inner_radius[1] / dy
        orx, ory = spec.outer_radius[0] / dx, spec.outer_radius[1] / dy
        iirx2, iiry2 = 1.0 / irx ** 2, 1.0 / iry ** 2
        iorx2, iory2 = 1.0 / orx ** 2, 1.0 / ory ** 2
        drop = (((x - cx) ** 2 * iirx2 + (y - cy) ** 2 * iiry2) < 1) | \
               (((x - cx) ** 2 * iorx2 + (y - cy) ** 2 * iory2) > 1)
        return np.ascontiguousarray(part[~drop])
    rr = np.sqrt((x - cx) ** 2 + (y - cy) ** 2)
    if spec.profile == 11:
        r0, sigr = spec.radius / dx, spec.width / dx
        isigr2 = 0.5 / sigr ** 2
        part[:, 7] = part[:, 7] * np.exp(-(rr - r0) ** 2 * isigr2)
        return part
    if spec.profile == 12:
        rp = np.asarray(spec.radial[0], dtype=np.float64) / dx
        fr = np.asarray(spec.radial[1], dtype=np.float64)
        if len(rp) != len(fr):
            raise ValueError("The piecewise_radial_density and r array have different sizes!")
        keep = ~((rr < rp[0]) | (rr > rp[-1]))
        w = np.zeros_like(rr)
        for k in np.nonzero(keep)[0]:
            for ii in range(1, len(rp)):
                if rp[ii] <= rp[ii - 1]:
                    raise ValueError("r is not monotonically increasing!")
                if rr[k] <= rp[ii]:
                    w[k] = fr[ii - 1] + (fr[ii] - fr[ii - 1]) / (rp[ii] - rp[ii - 1]) * \
                        (rr[k] - rp[ii - 1])
                    break
        part[:, 7] = part[:, 7] * w
        return np.ascontiguousarray(part[keep])
    raise ValueError(f"plasma profile {spec.profile} is not one of 0, 10, 11, 12")


# --------------------------------------------------------------------------
# particle kernels
# --------------------------------------------------------------------------
def qdeposit(nx, ny, part, q):
    lib().qpo_qdeposit(nx, ny, part.shape[0], _p(part), _p(q))


def guard_fold(f):
    dim, ny1, nx1 = f.shape
    lib().qpo_guard_fold(nx1 - 1, ny1 - 1, dim, _p(f))


def copyguard(f):
    dim, ny1, nx1 = f.shape
    lib().qpo_copyguard(nx1 - 1, ny1 - 1, dim, _p(f))


def extpsi(part, psi, qbm, dex):
    _, ny1, nx1 = psi.shape
    lib().qpo_extpsi(nx1 - 1, ny1 - 1, part.shape[0], _p(part), _p(psi), qbm, dex)


def amjdp(part, fxy, bxy, psit, cu, dcu, amu, qbm, dt, dex):
    _, ny1, nx1 = fxy.shape
    lib().qpo_amjdp(nx1 - 1, ny1 - 1, part.shape[0], _p(part), _p(fxy), _p(bxy),
                    _p(psit), _p(cu), _p(dcu), _p(amu), qbm, dt, dex)


def push(part, fxy, bxy, psit, qbm, dt, dex):
    _, ny1, nx1 = fxy.shape
    return lib().qpo_push(nx1 - 1, ny1 - 1, part.shape[0], _p(part), _p(fxy),
                          _p(bxy), _p(psit), qbm, dt, dex)


def tile_of(nx, ny, part):
    t = np.zeros(part.shape[0], np.int32)
    lib().qpo_tile_of(nx, ny, part.shape[0], _p(part), _p(t))
    return t


def cell_of(nx, ny, part):
    c = np.zeros(part.shape[0], np.int32)
    lib().qpo_cell_of(nx, ny, part.shape[0], _p(part), _p(c))
    return c


def order_by_tile(nx, ny, part):
    """Regroups ``part`` in place by 16x16 tile; returns kpic (per-tile counts)."""
    mx1 = (nx - 1) // 16 + 1
    my1 = (ny - 1) // 16 + 1
    kpic = np.zeros(mx1 * my1, np.int32)
    scratch = np.empty_like(part)
    lib().qpo_order_by_tile(nx, ny, part.shape[0], _p(part), _p(kpic), _p(scratch))
    return kpic


# --------------------------------------------------------------------------
# transforms and k-space operators
# --------------------------------------------------------------------------
def kind_types(kind, dim):
    tx = np.zeros(3, np.int32)
    ty = np.zeros(3, np.int32)
    lib().qpo_kind_types(kind, dim, _p(tx), _p(ty))
    return tx[:dim].copy(), ty[:dim].copy()


def transform_1d(n, typ, vecs):
    vecs = _f64(vecs)
    out = np.zeros_like(vecs)
    lib().qpo_transform_1d(n, typ, vecs.shape[0], _p(vecs), _p(out))
    return out


def r2k(kind, rs):
    rs = _f64(rs)
    dim, ny1, nx1 = rs.shape
    tx, ty = kind_types(kind, dim)
    ks = np.zeros((dim, nx1, ny1))
    lib().qpo_fft_r2k(nx1 - 1, ny1 - 1, dim, _p(tx), _p(ty), _p(rs), _p(ks))
    return ks


def k2r(kind, ks):
    ks = _f64(ks)
    dim, nx1, ny1 = ks.shape
    tx, ty = kind_types(kind, dim)
    rs = np.zeros((dim, ny1, nx1))
    lib().qpo_fft_k2r(nx1 - 1, ny1 - 1, dim, _p(tx), _p(ty), _p(ks), _p(rs))
    return rs


def pois_table(nx, ny, ax=AX, ay=AY, affp=AFFP):
    g = np.zeros((nx, ny))
    s = np.zeros((nx, ny))
    lib().qpo_pois_table(nx, ny, ax, ay, affp, _p(g), _p(s))
    return g, s


def _k_shape(f):
    return f.shape[1] - 1, f.shape[2] - 1


def div(f):
    nx, ny = _k_shape(f)
    out = np.zeros((1, nx + 1, ny + 1))
    lib().qpo_div(nx, ny, _p(_f64(f)), _p(out))
    return out


def grad(f):
    nx, ny = _k_shape(f)
    out = np.zeros((2, nx + 1, ny + 1))
    lib().qpo_grad(nx, ny, _p(_f64(f)), _p(out))
    return out


def pot(tab, q):
    nx, ny = _k_shape(q)
    out = np.zeros((1, nx + 1, ny + 1))
    lib().qpo_pot(nx, ny, _p(tab[0]), _p(tab[1]), _p(_f64(q)), _p(out))
    return out


def smooth(tab, q):
    nx, ny = _k_shape(q)
    out = np.zeros((1, nx + 1, ny + 1))
    lib().qpo_smooth(nx, ny, _p(tab[0]), _p(tab[1]), _p(_f64(q)), _p(out))
    return out


def elf(tab, q):
    nx, ny = _k_shape(q)
    out = np.zeros((2, nx + 1, ny + 1))
    lib().qpo_elf(nx, ny, _p(tab[0]), _p(tab[1]), _p(_f64(q)), _p(out))
    return out


def bf(tab, cu):
    nx, ny = _k_shape(cu)
    out = np.zeros((3, nx + 1, ny + 1))
    lib().qpo_bf(nx, ny, _p(tab[0]), _p(tab[1]), _p(_f64(cu)), _p(out))
    return out


def bfqp(tab, cu, dcu, amu, bxy, aa, dex):
    """Updates dcu and bxy in place (PBPOISD22N_QP isign=1)."""
    nx, ny = _k_shape(cu)
    lib().qpo_bfqp(nx, ny, _p(tab[0]), _p(tab[1]), _p(cu), _p(dcu), _p(amu),
                   _p(bxy), aa, dex)


# --------------------------------------------------------------------------
# beam kernels
# --------------------------------------------------------------------------
def beam_deposit(part, q3):
    _, nz1, ny1, nx1 = q3.shape
    lib().qpo_beam_deposit(nx1 - 1, ny1 - 1, nz1 - 1, part.shape[0], _p(part), _p(q3))


def beam_push(part, e3, b3, qbm, dt, dex, dez):
    """Pushes in place; returns the surviving particles (a view of ``part``)."""
    _, nz1, ny1, nx1 = e3.shape
    n = lib().qpo_beam_push(nx1 - 1, ny1 - 1, nz1 - 1, part.shape[0], _p(part),
                            _p(e3), _p(b3), qbm, dt, dex, dez)
    return part[:n]


# --------------------------------------------------------------------------
# species (species2d_class.f03) and the slice / 3-D step sequencing
# --------------------------------------------------------------------------
class Species:
    """species2d restated over the oracle kernels (species2d_class.f03:74-418)."""

    def __init__(self, grid: Grid, spec: SpeciesSpec, dt: float):
        self.grid, self.spec = grid, spec
        self.qbm = spec.q / spec.m
        self.dt = dt           # species dt is the xi step (simulation_class.f03:578)
        nx, ny = grid.nx, grid.ny
        z = lambda d: np.zeros((d, ny + 1, nx + 1))
        self.q, self.qn = z(1), z(1)
        self.cu, self.dcu, self.amu = z(3), z(2), z(3)
        self.renew()

    def renew(self, s=0.0):
        """init_species2d / renew_species2d: fresh particles at propagation
        distance s, qn = -deposit.  The density factor multiplies the
        density first, as den_temp*den in dist2d_000 (:254)."""
        g = self.grid
        den = density_factor(self.spec.piecewise, s) * self.spec.density
        if self.spec.profile == 0:
            self.part = plasma_init(g.nx, g.ny, self.spec.ppc[0], self.spec.ppc[1], den,
                                    self.spec.q)
        else:
            self.part = plasma_init_profile(g, self.spec, den)
        self.kpic = order_by_tile(g.nx, g.ny, self.part)
        qn = np.zeros((1, g.ny + 1, g.nx + 1))
        qdeposit(g.nx, g.ny, self.part, qn)
        guard_fold(qn)
        self.qn = qn * -1.0

    def qdp(self, q):
        """qdp_species2d :166-185; adds this species' charge and qn into q."""
        g = self.grid
        self.q = np.zeros((1, g.ny + 1, g.nx + 1))
        qdeposit(g.nx, g.ny, self.part, self.q)
        guard_fold(self.q)
        q = self.q + q
        return self.qn + q

    def extpsi(self, psi, dex):
        extpsi(self.part, psi, self.qbm, dex)

    def amjdp(self, fxy, bxy, psit, cu, amu, dcu, dex):
        """amjdp_species2d :187-218; returns the updated (cu, amu, dcu)."""
        g = self.grid
        self.cu = np.zeros((3, g.ny + 1, g.nx + 1))
        self.dcu = np.zeros((2, g.ny + 1, g.nx + 1))
        self.amu = np.zeros((3, g.ny + 1, g.nx + 1))
        amjdp(self.part, fxy, bxy, psit, self.cu, self.dcu, self.amu, self.qbm,
              self.dt, dex)
        guard_fold(self.cu)
        guard_fold(self.dcu)
        guard_fold(self.amu)
        cu = self.cu + cu
        dcu = self.dcu + dcu
        amu = self.amu + amu
        self.cu = self.cu * dex
        return cu, amu, dcu

    def push(self, fxy, bxy, psit, dex):
        """push_species2d :220-237 (push then pmv)."""
        err = push(self.part, fxy, bxy, psit, self.qbm, self.dt, dex)
        if err:
            raise RuntimeError(f"PPPCHECK2L-equivalent failure in tile {err}")
        self.kpic = order_by_tile(self.grid.nx, self.grid.ny, self.part)

    def cbq(self, tab):
        """cbq_species2d :399-418: smoothed q + Jz of this species."""
        q = self.q.copy()
        q[0] = q[0] + self.cu[2]
        qk = r2k(1, q)
        qk = smooth(tab, qk)
        self.q = k2r(1, qk)
        return self.q[0].copy()


class Simulation:
    """The reference's 3-D step on one rank and one stage (Q6: every step starts
    from zero fields and a fresh plasma).  Sequencing restates
    simulation_class.f03:611-765; one slice restates :643-736."""

    def __init__(self, grid: Grid, species: list, beams: list, dt: float,
                 iter_: int = 2, keep_q3=False, beam_parts=None, keep_diag=False):
        """beam_parts: optional list of (np, 7) particle arrays that replace
        the loader's beams (a large beam loaded elsewhere by a loader that is
        itself held bit-exact to beam_init)."""
        self.grid = grid
        self.dex = grid.dx
        self.dxi = grid.dz
        self.dt3d = dt
        self.iter = iter_
        self.tab = pois_table(grid.nx, grid.ny)
        self.species = [Species(grid, s, self.dxi) for s in species]
        self.beam_specs = beams
        self.beams = []
        rng = rng_new()
        for i, b in enumerate(beams):
            if beam_parts is not None:
                part = np.ascontiguousarray(beam_parts[i], dtype=np.float64)
            else:
                part, qm, ierr = beam_init(grid, b, rng=rng)
                assert ierr == 0
            self.beams.append([part, b.q / b.m])
        nx, ny, nz = grid.nx, grid.ny, grid.nz
        self.e3 = np.zeros((3, nz + 1, ny + 1, nx + 1))
        self.b3 = np.zeros((3, nz + 1, ny + 1, nx + 1))
        self.q3 = np.zeros((1, nz + 1, ny + 1, nx + 1))
        self.q3s = [np.zeros((nz + 1, ny + 1, nx + 1)) for _ in species] if keep_q3 else None
        # psi3d and cu3d of the reference's diagnostics (simulation_class.f03:670-671, 696-698)
        self.psi3 = np.zeros((nz + 1, ny + 1, nx + 1)) if keep_diag else None
        self.cu3 = np.zeros((3, nz + 1, ny + 1, nx + 1)) if keep_diag else None
        self.q3_initial()
        self.reset_fields()

    def q3_initial(self):
        """Plane 1 of each species' 3-D charge at a (re)load on the first
        stage (species2d_class.f03:107-113, 151-157): the fresh plasma's
        folded deposit, smoothed."""
        if self.q3s is None:
            return
        for si, s in enumerate(self.species):
            q = s.qn * -1.0
            self.q3s[si][0] = k2r(1, smooth(self.tab, r2k(1, q)))[0]

    def reset_fields(self):
        """pipeline receive on stage 0 zeroes cu, epw, fxy, psit, bxyz
        (ufield2d_class.f03:528-531)."""
        nx, ny = self.grid.nx, self.grid.ny
        self.cu_k = np.zeros((3, nx + 1, ny + 1))
        self.epw_k = np.zeros((2, nx + 1, ny + 1))
        self.fxy = np.zeros((2, ny + 1, nx + 1))
        self.psit = np.zeros((1, ny + 1, nx + 1))
        self.bxyz = np.zeros((3, ny + 1, nx + 1))
        self.dcu_k = np.zeros((2, nx + 1, ny + 1))

    def deposit_beams(self):
        """beam3d%qdp at the start of the step (simulation_class.f03:614-620)."""
        self.q3[:] = 0.0
        for part, _ in self.beams:
            beam_deposit(part, self.q3)

    def begin_step(self):
        self.deposit_beams()
        self.reset_fields()
        # slice-1 boundary copies :639-641
        self.e3[0:2, 0] = self.fxy
        self.e3[2, 0] = self.psit[0]
        self.b3[:, 0] = self.bxyz

    def _bxyz_from_epw(self, epwb):
        d = self.dex
        b = np.zeros_like(self.bxyz_k)
        b[0] = self.epw_k[1] * -d
        b[1] = self.epw_k[0] * d
        b[2] = self.bxyz_k[2] * d
        b[0] = b[0] - epwb[1]
        b[1] = b[1] + epwb[0]
        return b

    def slice(self, j, iters=None, record=None):
        """One xi slice j (1-based, j = 1..nz); simulation_class.f03:643-736."""
        d = self.dex
        tab = self.tab
        iters = self.iter if iters is None else iters
        # beam charge of slice j+1 (copyadd, ufield3d_class.f03:460-490)
        qb = self.q3[:, j].copy()
        qb_k = r2k(1, qb)
        epwb = elf(tab, qb_k)
        qe = np.zeros((1, self.grid.ny + 1, self.grid.nx + 1))
        for s in self.species:
            qe = s.qdp(qe)
        qe_k = r2k(1, qe)
        psi_k = pot(tab, qe_k)
        fxy_k = grad(psi_k)
        self.fxy = k2r(1, fxy_k)
        psi = k2r(1, psi_k)
        for s in self.species:
            s.extpsi(psi, d)
        psi = psi * (d * d)
        self.psi = psi
        if self.psi3 is not None:
            self.psi3[j] = psi[0]
        dv = div(self.cu_k)
        psit_k = pot(tab, dv)
        psit_k = psit_k * -d
        self.psit = k2r(1, psit_k)
        self.bxyz_k = bf(tab, self.cu_k)
        for l in range(1, iters + 1):
            self.bxyz_k = self._bxyz_from_epw(epwb)
            bxyz = k2r(2, self.bxyz_k)
            nx, ny = self.grid.nx, self.grid.ny
            cu = np.zeros((3, ny + 1, nx + 1))
            dcu = np.zeros((2, ny + 1, nx + 1))
            amu = np.zeros((3, ny + 1, nx + 1))
            for s in self.species:
                cu, amu, dcu = s.amjdp(self.fxy, bxyz, self.psit, cu, amu, dcu, d)
            cu = cu * d
            amu = amu * d
            dcu = dcu * d
            if l == iters and self.cu3 is not None:
                self.cu3[:, j] = cu
            if l == iters:
                for si, s in enumerate(self.species):
                    q3 = s.cbq(tab)
                    if self.q3s is not None:
                        self.q3s[si][j] = q3
            self.cu_k = r2k(1, cu)
            self.dcu_k = r2k(1, dcu)
            amu_k = r2k(3, amu)
            epw_prev = self.epw_k.copy()
            bfqp(tab, self.cu_k, self.dcu_k, amu_k, self.epw_k, d * d, d)
            if record is not None:
                record.append(_residual(self.epw_k, epw_prev))
            dv = div(self.cu_k)
            psit_k = pot(tab, dv)
            psit_k = psit_k * -d
            self.psit = k2r(1, psit_k)
            self.bxyz_k = bf(tab, self.cu_k)
        self.bxyz_k = self._bxyz_from_epw(epwb)
        self.bxyz = k2r(2, self.bxyz_k)
        fxy = self.fxy * -1.0
        fxy[0] = fxy[0] + self.bxyz[1]
        fxy[1] = fxy[1] - self.bxyz[0]
        self.fxy = fxy
        self.dcu_k = self.dcu_k * self.dxi
        self.cu_k[0:2] = self.cu_k[0:2] - self.dcu_k[0:2]
        for s in self.species:
            s.push(self.fxy, self.bxyz, self.psit, d)
        self.fxy = self.fxy * d
        self.e3[0:2, j] = self.fxy
        self.e3[2, j] = self.psit[0]
        b = self.bxyz.copy()
        b[0] = b[0] * d
        b[1] = b[1] * d
        self.bxyz = b
        self.b3[:, j] = b

    def push_beams(self):
        """beam3d%push (simulation_class.f03:752-757), PGBPUSH32L_QP."""
        for i, (part, qbm) in enumerate(self.beams):
            self.beams[i][0] = beam_push(part, self.e3, self.b3, qbm, self.dt3d,
                                         self.dex, self.dxi).copy()

    def step(self, nslices=None, s_new=0.0):
        """One 3-D step; the species are renewed at s_new (i*dt after step i,
        simulation_class.f03:761-764)."""
        self.begin_step()
        n = self.grid.nz if nslices is None else nslices
        for j in range(1, n + 1):
            self.slice(j)
        self.push_beams()
        for s in self.species:
            s.renew(s_new)
        self.q3_initial()


def _residual(new, old):
    """||B_new - B_old||_2 / ||B_new||_2 over k space, accumulated in FP64."""
    num = float(np.sqrt(np.sum((new - old) ** 2)))
    den = float(np.sqrt(np.sum(new ** 2)))
    return num / den if den > 0 else (0.0 if num == 0 else float("inf"))
