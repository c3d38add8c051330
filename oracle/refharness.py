"""Harness around the reference's own F77 kernels (oracle/_ref/libqpref.so).

TEST INFRASTRUCTURE ONLY.  ``oracle/Makefile.ref`` compiles the reference's
``part2d_lib77.f``, ``fpois2d_lib77.f``, ``fft2d_lib77.f`` and
``part3d_lib77.f`` from /root/reference in their double-precision
configuration; this module calls them through ctypes with the reference's
Fortran (column-major) layouts.  It is used to pin the C oracle (tests and
tests/golden/make_golden.py) and never by the product.

Where the reference reaches a routine that needs MPI (only in its nvp > 1
branches), the harness performs the one-rank branch itself, in a few lines,
citing it:
  * the transposes PPRTPOSE/PPRNTPOSE for nvp = 1 (fft2d_lib.f03:675-684,
    :794-805) are plain transposes of the (nx+1) x (ny+1) block;
  * PPPMOVE2 for nvp = 1 (part2d_lib.f03:291-312) copies sbufr -> rbufl,
    sbufl -> rbufr and swaps the counts;
  * PPNACGUARD2L / PPNCGUARD2L for nvp = 1 (ufield2d_lib.f03:90-97, :37-41)
    fold row ny into row 0, or copy row 0 into row ny.
Array shapes follow the reference: real fields f(dim, nx+2, 2*ny+1),
k fields g(dim, ny+1, nx+1), tiled particles ppart(idimp, nppmx, mxyp1).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_SO = os.path.join(_HERE, "_ref", "libqpref.so")
MX = MY = 16


def available():
    return os.path.exists(REF_SO)


class Ref:
    def __init__(self, path=REF_SO):
        self.L = C.CDLL(path)
        self.L.ranorm_.restype = C.c_double

    # ---- argument helpers -------------------------------------------------
    @staticmethod
    def i(v):
        return C.byref(C.c_int(int(v)))

    @staticmethod
    def d(v):
        return C.byref(C.c_double(float(v)))

    @staticmethod
    def a(arr):
        assert arr.flags["F_CONTIGUOUS"], "reference arrays must be Fortran-ordered"
        return arr.ctypes.data_as(C.c_void_p)

    # ---- random numbers -----------------------------------------------------
    def ranorm(self):
        return self.L.ranorm_()

    def beam_dist(self, qm, nx, ny, nz, npx, npy, npz, vt, vd, sig, x0, cx, cy,
                  quiet, npmax):
        """PRVDIST32_RANDOM part3d_lib77.f:411-552 with one partition covering
        the box (edges = (0, ny, 0, nz)) and ipbc = 1."""
        idimp = 7
        part = np.zeros((idimp, npmax, 1), order="F")
        edges = np.asfortranarray(np.array([0.0, ny, 0.0, nz]).reshape(4, 1))
        npp = np.zeros(1, np.int32)
        nps = np.ones(1, np.int32)
        cxa = np.asfortranarray(np.array(cx, np.float64))
        cya = np.asfortranarray(np.array(cy, np.float64))
        ierr = C.c_int(0)
        lq = C.c_int(1 if quiet else 0)
        self.L.prvdist32_random_(
            self.a(part), self.d(qm), self.a(edges), self.a(npp), self.a(nps),
            self.d(vt[0]), self.d(vt[1]), self.d(vt[2]), self.d(vd[0]),
            self.d(vd[1]), self.d(vd[2]), self.i(npx), self.i(npy), self.i(npz),
            self.i(nx), self.i(ny), self.i(nz), self.i(1), self.i(idimp),
            self.i(npmax), self.i(1), self.i(1), self.i(4), self.d(sig[0]),
            self.d(sig[1]), self.d(sig[2]), self.d(x0[0]), self.d(x0[1]),
            self.d(x0[2]), self.a(cxa), self.a(cya), C.byref(lq), C.byref(ierr))
        n = int(npp[0])
        return np.ascontiguousarray(part[:, :n, 0].T), ierr.value

    # ---- tiled particles ----------------------------------------------------
    def tiles(self, nx, ny):
        mx1 = (nx - 1) // MX + 1
        myp1 = (ny - 1) // MY + 1
        return mx1, myp1, mx1 * myp1

    def to_tiles(self, part, nx, ny, xtras=0.2):
        """PPDBLKP2L + PPPMOVIN2L (part2d_class.f03:154-177): flat (np, 8) ->
        ppart(8, nppmx0, mxyp1), kpic."""
        mx1, myp1, mxyp1 = self.tiles(nx, ny)
        npp = part.shape[0]
        idimp = part.shape[1]
        pf = np.asfortranarray(part.T)
        kpic = np.zeros(mxyp1, np.int32)
        nppmx = C.c_int(0)
        irc = C.c_int(0)
        self.L.ppdblkp2l_(self.a(pf), self.a(kpic), self.i(npp), self.i(0),
                          C.byref(nppmx), self.i(idimp), self.i(npp), self.i(MX),
                          self.i(MY), self.i(mx1), self.i(mxyp1), C.byref(irc))
        assert irc.value == 0, irc.value
        nppmx0 = int((1.0 + xtras) * nppmx.value)
        ppart = np.zeros((idimp, nppmx0, mxyp1), order="F")
        self.L.pppmovin2l_(self.a(pf), self.a(ppart), self.a(kpic), self.i(npp),
                           self.i(0), self.i(nppmx0), self.i(idimp), self.i(npp),
                           self.i(MX), self.i(MY), self.i(mx1), self.i(mxyp1),
                           C.byref(irc))
        assert irc.value == 0, irc.value
        return ppart, kpic, nppmx.value

    @staticmethod
    def from_tiles(ppart, kpic):
        """Flat (np, idimp) array in tile order (PPPCOPYOUT2 :2286-2332)."""
        rows = [ppart[:, :kpic[k], k].T for k in range(len(kpic))]
        return np.ascontiguousarray(np.concatenate(rows, axis=0))

    def check(self, ppart, kpic, nx, ny):
        mx1, myp1, _ = self.tiles(nx, ny)
        irc = C.c_int(0)
        self.L.pppcheck2l_(self.a(ppart), self.a(kpic), self.i(0), self.i(ny),
                           self.i(ppart.shape[0]), self.i(ppart.shape[1]),
                           self.i(nx), self.i(MX), self.i(MY), self.i(mx1),
                           self.i(myp1), C.byref(irc))
        return irc.value

    # ---- fields in the reference layout ---------------------------------------
    @staticmethod
    def rfield(dim, nx, ny):
        return np.zeros((dim, nx + 2, 2 * ny + 1), order="F")

    @staticmethod
    def kfield(dim, nx, ny):
        return np.zeros((dim, ny + 1, nx + 1), order="F")

    @staticmethod
    def r_from_oracle(f, nx, ny):
        """(dim, ny+1, nx+1) -> reference real layout."""
        dim = f.shape[0]
        g = np.zeros((dim, nx + 2, 2 * ny + 1), order="F")
        g[:, :nx + 1, :ny + 1] = np.transpose(f, (0, 2, 1))
        return g

    @staticmethod
    def r_to_oracle(g, nx, ny):
        return np.ascontiguousarray(np.transpose(g[:, :nx + 1, :ny + 1], (0, 2, 1)))

    @staticmethod
    def k_from_oracle(k):
        """(dim, nx+1, ny+1) -> reference k layout (dim, ny+1, nx+1)."""
        return np.array(np.transpose(k, (0, 2, 1)), order="F", copy=True)

    @staticmethod
    def k_to_oracle(g):
        return np.ascontiguousarray(np.transpose(g, (0, 2, 1)))

    @staticmethod
    def fold(g, ny):
        """PPNACGUARD2L nvp = 1 (ufield2d_lib.f03:90-97), x = 0..nx."""
        nx1 = g.shape[1] - 1
        g[:, :nx1, 0] += g[:, :nx1, ny]
        g[:, :nx1, ny] = 0.0

    @staticmethod
    def copyguard(g, ny):
        """PPNCGUARD2L nvp = 1 (ufield2d_lib.f03:37-41)."""
        g[:, :, ny] = g[:, :, 0]

    # ---- particle kernels ---------------------------------------------------
    def qdeposit(self, ppart, kpic, q, nx, ny):
        """PPGPPOST2L into q(1, nx+2, 2ny+1), then the guard fold."""
        mx1, myp1, mxyp1 = self.tiles(nx, ny)
        q1 = np.asfortranarray(q[0])
        self.L.ppgppost2l_(self.a(ppart), self.a(q1), self.a(kpic), self.i(0),
                           self.i(ppart.shape[0]), self.i(ppart.shape[1]),
                           self.i(MX), self.i(MY), self.i(q1.shape[0]),
                           self.i(q1.shape[1]), self.i(mx1), self.i(mxyp1))
        q[0] = q1

    def extpsi(self, ppart, kpic, psi, qbm, dex, nx, ny):
        mx1, myp1, mxyp1 = self.tiles(nx, ny)
        p1 = np.asfortranarray(psi[0])
        self.L.wpgpsipost2l_qp_(self.a(ppart), self.a(p1), self.a(kpic),
                                self.d(qbm), self.i(0), self.i(ny),
                                self.i(ppart.shape[0]), self.i(ppart.shape[1]),
                                self.i(nx), self.i(MX), self.i(MY),
                                self.i(p1.shape[0]), self.i(p1.shape[1]),
                                self.i(mx1), self.i(mxyp1), self.d(dex))

    def amjdp(self, ppart, kpic, fxy, bxy, psit, cu, dcu, amu, qbm, dt, dex,
              nx, ny):
        mx1, myp1, mxyp1 = self.tiles(nx, ny)
        ps = np.asfortranarray(psit[0])
        self.L.ppgrdcjppost2l_qp_(
            self.a(ppart), self.a(fxy), self.a(bxy), self.a(ps), self.a(cu),
            self.a(dcu), self.a(amu), self.a(kpic), self.i(0), self.i(ny),
            self.d(qbm), self.d(dt), self.d(1.0), self.i(ppart.shape[0]),
            self.i(ppart.shape[1]), self.i(nx), self.i(MX), self.i(MY),
            self.i(fxy.shape[1]), self.i(fxy.shape[2]), self.i(mx1),
            self.i(mxyp1), self.d(dex))

    def push_and_reorder(self, ppart, kpic, fxy, bxy, psit, qbm, dt, dex, nx,
                         ny):
        """PPGRBPPUSHF23L_QP then pmove for one rank (part2d_class.f03:360-594
        without the overflow-recovery branches): PPPORDERF2LA, PPPMOVE2 (nvp=1
        copy), PPPORDER2LB, PPPCHECK2L.  Returns (ppart, kpic, irc)."""
        mx1, myp1, mxyp1 = self.tiles(nx, ny)
        idimp, nppmx, _ = ppart.shape
        ntmax = nppmx  # generous: no ihole overflow
        npbmx = nppmx
        nbmax = nppmx * mx1
        ncl = np.zeros((8, mxyp1), np.int32, order="F")
        ihole = np.zeros((2, ntmax + 1, mxyp1), np.int32, order="F")
        ek = C.c_double(0.0)
        irc = C.c_int(0)
        ps = np.asfortranarray(psit[0])
        self.L.ppgrbppushf23l_qp_(
            self.a(ppart), self.a(fxy), self.a(bxy), self.a(ps), self.a(kpic),
            self.a(ncl), self.a(ihole), self.i(0), self.i(ny), self.d(qbm),
            self.d(dt), self.d(1.0), C.byref(ek), self.i(idimp), self.i(nppmx),
            self.i(nx), self.i(ny), self.i(MX), self.i(MY), self.i(fxy.shape[1]),
            self.i(fxy.shape[2]), self.i(mx1), self.i(mxyp1), self.i(ntmax),
            C.byref(irc), self.d(dex))
        assert irc.value == 0
        ppbuff = np.zeros((idimp, npbmx, mxyp1), order="F")
        sbufl = np.zeros((idimp, nbmax), order="F")
        sbufr = np.zeros((idimp, nbmax), order="F")
        rbufl = np.zeros((idimp, nbmax), order="F")
        rbufr = np.zeros((idimp, nbmax), order="F")
        ncll = np.zeros((3, mx1), np.int32, order="F")
        nclr = np.zeros((3, mx1), np.int32, order="F")
        irc2 = np.zeros(2, np.int32)
        self.L.ppporderf2la_(
            self.a(ppart), self.a(ppbuff), self.a(sbufl), self.a(sbufr),
            self.a(ncl), self.a(ihole), self.a(ncll), self.a(nclr), self.i(idimp),
            self.i(nppmx), self.i(mx1), self.i(myp1), self.i(npbmx),
            self.i(ntmax), self.i(nbmax), self.a(irc2))
        assert irc2[0] == 0, irc2
        # PPPMOVE2, nvp = 1 (part2d_lib.f03:291-312)
        mcll = np.asfortranarray(nclr.copy())
        mclr = np.asfortranarray(ncll.copy())
        rbufl[:, :nclr[2, mx1 - 1]] = sbufr[:, :nclr[2, mx1 - 1]]
        rbufr[:, :ncll[2, mx1 - 1]] = sbufl[:, :ncll[2, mx1 - 1]]
        self.L.ppporder2lb_(
            self.a(ppart), self.a(ppbuff), self.a(rbufl), self.a(rbufr),
            self.a(kpic), self.a(ncl), self.a(ihole), self.a(mcll), self.a(mclr),
            self.i(idimp), self.i(nppmx), self.i(nx), self.i(ny), self.i(mx1),
            self.i(myp1), self.i(npbmx), self.i(ntmax), self.i(nbmax),
            self.a(irc2))
        assert irc2[0] == 0, irc2
        return ppart, kpic, self.check(ppart, kpic, nx, ny)

    # ---- transforms -----------------------------------------------------------
    def fft_tables(self, nx, ny):
        indx, indy = int(np.log2(nx)), int(np.log2(ny))
        nxhyd, nxyd = max(nx // 2, ny), max(nx, ny)
        mixup = np.zeros(nxhyd, np.int32)
        sct = np.zeros(nxyd, np.complex128)
        self.L.wpfst2rinit_(self.a(mixup), self.a(sct), self.i(indx), self.i(indy),
                            self.i(nxhyd), self.i(nxyd))
        return indx, indy, mixup, sct

    # wrapper -> (x routine, y routine, ndim for transpose)
    WRAP = {
        (1, 1): ("ppfst2rmxx_", "ppfst2rmxy_"),
        (1, 2): ("ppfcst2rm2x_", "ppfsct2rm2y_"),
        (1, 3): ("ppfcsst2rm3x_", "ppfscst2rm3y_"),
        (2, 1): ("ppfct2rmxx_", "ppfct2rmxy_"),
        (2, 2): ("ppfsct2rm2x_", "ppfcst2rm2y_"),
        (2, 3): ("ppfscct2rm3x_", "ppfcsct2rm3y_"),
        (3, 3): ("ppfssct2rm23x_", "ppfssct2rm23y_"),
    }

    def transform(self, kind, f, g, isign, nx, ny):
        """The WPPF*RM* wrapper for (kind, dim) on one rank (ntpose = 1):
        isign = -1: x pass of f, transpose f -> g, y pass of g;
        isign = +1: y pass of g, transpose g -> f, x pass of f."""
        dim = f.shape[0]
        xr, yr = self.WRAP[(kind, dim)]
        indx, indy, mixup, sct = self.fft_tables(nx, ny)
        nxvh = f.shape[1] // 2
        kypd = f.shape[2]
        nyv, kxp2d = g.shape[1], g.shape[2]
        nxhyd, nxyd = len(mixup), len(sct)
        scalar = dim == 1
        fa = np.asfortranarray(f[0]) if scalar else f
        ga = np.asfortranarray(g[0]) if scalar else g

        def xpass():
            getattr(self.L, xr)(self.a(fa), self.i(isign), self.a(mixup), self.a(sct),
                                self.i(indx), self.i(indy), self.i(1), self.i(1),
                                self.i(ny + 1), self.i(nxvh), self.i(kypd),
                                self.i(nxhyd), self.i(nxyd))

        def ypass():
            getattr(self.L, yr)(self.a(ga), self.i(isign), self.a(mixup), self.a(sct),
                                self.i(indx), self.i(indy), self.i(1), self.i(1),
                                self.i(nx + 1), self.i(nyv), self.i(kxp2d),
                                self.i(nxhyd), self.i(nxyd))

        if isign < 0:
            xpass()
            if scalar:
                ga[:ny + 1, :nx + 1] = fa[:nx + 1, :ny + 1].T
            else:
                ga[:, :ny + 1, :nx + 1] = np.transpose(fa[:, :nx + 1, :ny + 1], (0, 2, 1))
            ypass()
        else:
            ypass()
            if scalar:
                fa[:nx + 1, :ny + 1] = ga[:ny + 1, :nx + 1].T
            else:
                fa[:, :nx + 1, :ny + 1] = np.transpose(ga[:, :ny + 1, :nx + 1], (0, 2, 1))
            xpass()
        if scalar:
            f[0] = fa
            g[0] = ga

    # ---- k-space --------------------------------------------------------------
    def pois_table(self, nx, ny, ax=0.912871, ay=0.912871, affp=1.0):
        ffc = np.zeros((ny, nx), np.complex128, order="F")
        q = np.zeros((1, 1), order="F")
        fxy = np.zeros((2, 1, 1), order="F")
        we = C.c_double(0.0)
        self.L.mppoisd22_(self.a(q), self.a(fxy), self.i(0), self.a(ffc), self.d(ax),
                          self.d(ay), self.d(affp), C.byref(we), self.i(nx), self.i(ny),
                          self.i(1), self.i(1), self.i(nx), self.i(ny))
        return ffc

    def kcall(self, name, ffc, nx, ny, *arrays):
        we = C.c_double(0.0)
        ptrs = [self.a(a) for a in arrays]
        nyv = ny + 1
        kxp2 = nx
        if name == "pot":
            self.L.mppotpd2_(*ptrs, self.a(ffc), C.byref(we), self.i(nx), self.i(ny),
                             self.i(1), self.i(nyv), self.i(kxp2), self.i(ny))
        elif name == "smooth":
            self.L.mppsmoothd2_(*ptrs, self.a(ffc), self.i(nx), self.i(ny), self.i(1),
                                self.i(nyv), self.i(kxp2), self.i(ny))
        elif name == "elf":
            self.L.mppoisd22_(*ptrs, self.i(-1), self.a(ffc), self.d(0.0), self.d(0.0),
                              self.d(0.0), C.byref(we), self.i(nx), self.i(ny),
                              self.i(1), self.i(nyv), self.i(kxp2), self.i(ny))
        elif name == "bf":
            self.L.mppbbpoisd23_(*ptrs, self.a(ffc), self.d(1.0), C.byref(we),
                                 self.i(nx), self.i(ny), self.i(1), self.i(nyv),
                                 self.i(kxp2), self.i(ny))
        elif name == "div":
            self.L.mppdivfd2_(*ptrs, self.i(nx), self.i(ny), self.i(1),
                              self.i(arrays[0].shape[0]), self.i(nyv), self.i(kxp2))
        elif name == "grad":
            self.L.mppgradfd2_(*ptrs, self.i(nx), self.i(ny), self.i(1),
                               self.i(arrays[1].shape[0]), self.i(nyv), self.i(kxp2))
        else:
            raise KeyError(name)

    def bfqp(self, ffc, cu, dcu, amu, bxy, aa, dex, nx, ny):
        """PBPOISD22N_QP isign = 1 (fpois2d_class.f03:444-495)."""
        dxy = np.zeros((1, 1, 1), order="F")
        wm = C.c_double(0.0)
        self.L.pbpoisd22n_qp_(self.a(cu), self.a(dcu), self.a(amu), self.a(bxy),
                              self.a(dxy), self.i(1), self.a(ffc), self.d(0.0),
                              self.d(0.0), self.d(0.0), self.d(1.0), C.byref(wm),
                              self.i(nx), self.i(ny), self.i(1), self.i(ny + 1),
                              self.i(nx), self.i(1), self.i(ny), self.d(aa),
                              self.d(dex))

    # ---- beam -----------------------------------------------------------------
    def beam_deposit(self, part, nx, ny, nz):
        """PGPOST32L into q(nx+2, ny+1, nz+1) with noff = (0, 0)."""
        n = part.shape[0]
        pf = np.asfortranarray(part.T.reshape(7, n, 1))
        q = np.zeros((nx + 2, ny + 1, nz + 1, 1), order="F")
        npp = np.array([n], np.int32)
        noff = np.zeros((2, 1), np.int32, order="F")
        self.L.pgpost32l_(self.a(pf), self.a(q), self.a(npp), self.a(noff),
                          self.i(7), self.i(n), self.i(1), self.i(nx + 2),
                          self.i(ny + 1), self.i(nz + 1), self.i(2))
        return q[:, :, :, 0]

    def beam_push(self, part, e3, b3, qbm, dt, dex, dez, nx, ny, nz):
        """PGBPUSH32L_QP with ipbc = 1, cofd = 0 (part3d_class.f03:196-199).
        e3, b3: reference layout (3, nx+2, ny+1, nz+1)."""
        n = part.shape[0]
        pf = np.asfortranarray(part.T.reshape(7, n, 1))
        npp = np.array([n], np.int32)
        noff = np.zeros((2, 1), np.int32, order="F")
        ek = C.c_double(0.0)
        e = np.asfortranarray(e3.reshape(e3.shape + (1,)))
        b = np.asfortranarray(b3.reshape(b3.shape + (1,)))
        self.L.pgbpush32l_qp_(self.a(pf), self.a(e), self.a(b), self.a(npp),
                              self.a(noff), self.d(qbm), self.d(dt), self.d(dt),
                              C.byref(ek), self.i(nx), self.i(ny), self.i(nz),
                              self.i(7), self.i(n), self.i(1), self.i(nx + 2),
                              self.i(ny + 1), self.i(nz + 1), self.i(2), self.i(1),
                              self.d(dex), self.d(dez), self.d(0.0))
        return np.ascontiguousarray(pf[:, :npp[0], 0].T)
