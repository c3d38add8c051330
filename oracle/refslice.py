"""The reference's slice loop composed over its own F77 kernels.

TEST INFRASTRUCTURE ONLY (needs oracle/_ref/libqpref.so, built from
/root/reference by oracle/Makefile.ref).  ``RefSimulation`` restates
go_simulation's 3-D step (simulation_class.f03:611-765, one rank, stage 0) on
top of ``refharness.Ref``: every particle kernel, transform, k-space operator
and the beam deposit is the reference's compiled routine, and the glue in
between (the field2d component maps, the species2d additions, the guard fold
and copy, the plasma lattice and the beam normalisation) is restated here from
the reference's class files, in the reference's own array layouts:

* real-space fields f(dim, nx+2, 2*ny+1) (ufield2d_class.f03:129),
* k-space fields g(dim, ny+1, nx+1) (ufield2d_class.f03:198),
* tiled plasma particles ppart(8, nppmx0, mxyp1) (part2d_class.f03:154-177),
* the 3-D beam charge q(nx+2, ny+1, nz+1).

It exists to pin the composed slice of the C/Python oracle
(``qpo.Simulation``): the two are written independently, one over the
oracle's kernels and layouts, this one over the reference's.
"""
from __future__ import annotations

import numpy as np

from oracle import refharness as rh

PI2 = 2 * 3.1415926535897932


class RefSpecies:
    """species2d (species2d_class.f03:74-418) over the reference kernels."""

    def __init__(self, ref: rh.Ref, nx, ny, ppc, den, q, m, dt):
        self.ref, self.nx, self.ny = ref, nx, ny
        self.ppc, self.den, self.qsign = ppc, den, q
        self.qbm = q / m
        self.dt = dt
        self.renew()

    def lattice(self):
        """dist2d_000 on one rank, noff = 0 < ny (fdist2d_class.f03:254-273):
        x cells 2..nx-1, y cells 2..ny (1-based), (i+0.5)/ppc offsets,
        q = den sign/(xppc yppc), p6 = p7 = 1."""
        nx, ny = self.nx, self.ny
        xppc, yppc = self.ppc
        qm = self.den * self.qsign / abs(self.qsign) / float(xppc) / float(yppc)
        # the loop nest i, j, ix, iy (iy fastest), evaluated left to right
        i, j, ix, iy = np.meshgrid(np.arange(2, nx), np.arange(2, ny + 1), np.arange(xppc),
                                   np.arange(yppc), indexing="ij")
        n = i.size
        part = np.zeros((n, 8))
        part[:, 0] = ((ix.ravel() + 0.5) / xppc + i.ravel()) - 1
        part[:, 1] = ((iy.ravel() + 0.5) / yppc + j.ravel()) - 1
        part[:, 5] = part[:, 6] = 1.0
        part[:, 7] = qm
        return part

    def renew(self):
        """renew_species2d :134-164: new lattice, qn = -(deposit, folded)."""
        nx, ny, ref = self.nx, self.ny, self.ref
        self.ppart, self.kpic, _ = ref.to_tiles(self.lattice(), nx, ny)
        qn = ref.rfield(1, nx, ny)
        ref.qdeposit(self.ppart, self.kpic, qn, nx, ny)
        ref.fold(qn, ny)
        self.qn = -qn
        self.cu = ref.rfield(3, nx, ny)

    def qdp(self, q):
        """qdp_species2d :166-185."""
        nx, ny, ref = self.nx, self.ny, self.ref
        self.q = ref.rfield(1, nx, ny)
        ref.qdeposit(self.ppart, self.kpic, self.q, nx, ny)
        ref.fold(self.q, ny)
        q = self.q + q
        return self.qn + q

    def extpsi(self, psi, dex):
        self.ref.extpsi(self.ppart, self.kpic, psi, self.qbm, dex, self.nx, self.ny)

    def amjdp(self, fxy, bxy, psit, cu, amu, dcu, dex):
        """amjdp_species2d :187-218; returns (cu, amu, dcu)."""
        nx, ny, ref = self.nx, self.ny, self.ref
        scu, sdcu, samu = ref.rfield(3, nx, ny), ref.rfield(2, nx, ny), ref.rfield(3, nx, ny)
        ref.amjdp(self.ppart, self.kpic, fxy, bxy, psit, scu, sdcu, samu, self.qbm, self.dt,
                  dex, nx, ny)
        for f in (scu, sdcu, samu):
            ref.fold(f, ny)
        cu = scu + cu
        dcu = sdcu + dcu
        amu = samu + amu
        self.cu = scu * dex
        return cu, amu, dcu

    def push(self, fxy, bxy, psit, dex):
        """push_species2d :220-237 (PPGRBPPUSHF23L_QP, then pmove)."""
        self.ppart, self.kpic, irc = self.ref.push_and_reorder(
            self.ppart, self.kpic, fxy, bxy, psit, self.qbm, self.dt, dex, self.nx, self.ny)
        if irc:
            raise RuntimeError(f"PPPCHECK2L irc = {irc}")

    def particles(self):
        return self.ref.from_tiles(self.ppart, self.kpic)


class RefSimulation:
    """go_simulation (simulation_class.f03:611-765) on one rank, stage 0."""

    def __init__(self, ref: rh.Ref, grid, species, beams, dt, iter_=2):
        self.ref, self.g = ref, grid
        nx, ny, nz = grid.nx, grid.ny, grid.nz
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dex, self.dxi, self.dt3d, self.iter = grid.dx, grid.dz, dt, iter_
        self.ffc = ref.pois_table(nx, ny)
        self.species = [RefSpecies(ref, nx, ny, s.ppc, s.density, s.q, s.m, self.dxi)
                        for s in species]
        self.beams = []
        for b in beams:
            self.beams.append((self._beam_particles(b), b.q / b.m))
        self.e3 = np.zeros((3, nz + 1, ny + 1, nx + 1))
        self.b3 = np.zeros((3, nz + 1, ny + 1, nx + 1))

    def _beam_particles(self, b):
        """init_fdist3d_000 normalisation (fdist3d_class.f03:311-327) and
        PRVDIST32_RANDOM (dist3d_000 :405-415) from the reference's ranorm."""
        g = self.g
        dx, dy, dz = g.dx, g.dy, g.dz
        sx, sy, sz = b.sigma
        qm = b.q / abs(b.q) * abs(b.peak) * PI2 ** 1.5 * sx * sy * sz
        qm = qm / dx / dy / dz
        qm = qm / b.np_[0]
        qm = qm / b.np_[1]
        qm = qm / b.np_[2]
        x0 = ((b.center[0] - g.box[0][0]) / dx, (b.center[1] - g.box[1][0]) / dy,
              (b.center[2] - g.box[2][0]) / dz)
        part, ierr = self.ref.beam_dist(qm, self.nx, self.ny, self.nz, *b.np_, b.sigma_v,
                                        (0.0, 0.0, b.gamma), (sx / dx, sy / dy, sz / dz), x0,
                                        (0.0, 0.0, 0.0), (0.0, 0.0, 0.0), b.quiet,
                                        b.np_[0] * b.np_[1] * b.np_[2] + 2)
        assert ierr == 0
        return part

    # ---- field2d helpers (reference layouts) ---------------------------------
    def rf(self, dim):
        return self.ref.rfield(dim, self.nx, self.ny)

    def kf(self, dim):
        return self.ref.kfield(dim, self.nx, self.ny)

    def fftrk(self, kind, f):
        g = self.kf(f.shape[0])
        self.ref.transform(kind, f.copy(order="F"), g, -1, self.nx, self.ny)
        return g

    def fftkr(self, kind, g):
        """fftkr_field2d :193-226: inverse transform, then the guard copy."""
        f = self.rf(g.shape[0])
        self.ref.transform(kind, f, np.asfortranarray(g.copy()), 1, self.nx, self.ny)
        self.ref.copyguard(f, self.ny)
        return f

    def kop(self, name, src, dout):
        out = self.kf(dout)
        a = np.asfortranarray(src[0]) if src.shape[0] == 1 and name != "div" else \
            np.asfortranarray(src)
        o = np.asfortranarray(out[0]) if dout == 1 else out
        self.ref.kcall(name, self.ffc, self.nx, self.ny, a, o)
        if dout == 1:
            out[0] = o
        return out

    # ---- the step --------------------------------------------------------------
    def begin_step(self):
        nx, ny, nz = self.nx, self.ny, self.nz
        q3 = np.zeros((nx + 2, ny + 1, nz + 1))
        for part, _ in self.beams:
            q3 += self.ref.beam_deposit(part, nx, ny, nz)
        self.q3 = q3
        # pipeline receive at stage 0: zero fields (ufield2d_class.f03:528-531)
        self.cu = self.kf(3)
        self.epw = self.kf(2)
        self.fxy = self.rf(2)
        self.psit = self.rf(1)
        self.bxyz = self.rf(3)
        self._store(0)

    def _store(self, lpos0):
        """cb of fxy, psit, bxyz into the 3-D arrays at 0-based plane lpos0."""
        nx, ny = self.nx, self.ny
        t = lambda f: np.transpose(f[:, :nx + 1, :ny + 1], (0, 2, 1))  # noqa: E731
        self.e3[0:2, lpos0] = t(self.fxy)
        self.e3[2, lpos0] = t(self.psit)[0]
        self.b3[:, lpos0] = t(self.bxyz)

    def _bxyz_k(self, bxyz_k, epwb):
        """bxyz%mult(epw,(1,2),(2,1),(-dex,dex)); bxyz(3) *= dex;
        bxyz(1) -= epwb(2); bxyz(2) += epwb(1) (simulation_class.f03:677-680)."""
        d = self.dex
        b = bxyz_k.copy(order="F")
        b[0] = self.epw[1] * -d
        b[1] = self.epw[0] * d
        b[2] = b[2] * d
        b[0] = b[0] - epwb[1]
        b[1] = b[1] + epwb[0]
        return b

    def slice(self, j):
        """One 2-D step j (simulation_class.f03:643-736)."""
        nx, ny, d = self.nx, self.ny, self.dex
        qb = self.rf(1)
        qb[0, :nx + 1, :ny + 1] += self.q3[:nx + 1, :ny + 1, j]     # qdpcopy, slice j+1
        epwb = self.kop("elf", self.fftrk(1, qb), 2)
        qe = self.rf(1)
        for s in self.species:
            qe = s.qdp(qe)
        psi_k = self.kop("pot", self.fftrk(1, qe), 1)
        fxy_k = self.kop("grad", psi_k, 2)
        self.fxy = self.fftkr(1, fxy_k)
        psi = self.fftkr(1, psi_k)
        for s in self.species:
            s.extpsi(psi, d)
        psit_k = self.kop("pot", self.kop("div", self.cu, 1), 1) * -d
        self.psit = self.fftkr(1, psit_k)
        bxyz_k = self.kop("bf", self.cu, 3)
        for l in range(1, self.iter + 1):
            bxyz_k = self._bxyz_k(bxyz_k, epwb)
            self.bxyz = self.fftkr(2, bxyz_k)
            cu, dcu, amu = self.rf(3), self.rf(2), self.rf(3)
            for s in self.species:
                cu, amu, dcu = s.amjdp(self.fxy, self.bxyz, self.psit, cu, amu, dcu, d)
            cu, amu, dcu = cu * d, amu * d, dcu * d
            if l == self.iter:
                for s in self.species:                           # cbq (diagnostic only)
                    s.q[0] = s.q[0] + s.cu[2]
            self.cu = self.fftrk(1, cu)
            dcu_k = self.fftrk(1, dcu)
            amu_k = self.fftrk(3, amu)
            self.ref.bfqp(self.ffc, self.cu, dcu_k, amu_k, self.epw, d * d, d, nx, ny)
            psit_k = self.kop("pot", self.kop("div", self.cu, 1), 1) * -d
            self.psit = self.fftkr(1, psit_k)
            bxyz_k = self.kop("bf", self.cu, 3)
        bxyz_k = self._bxyz_k(bxyz_k, epwb)
        self.bxyz = self.fftkr(2, bxyz_k)
        fxy = self.fxy * -1.0
        fxy[0] = fxy[0] + self.bxyz[1]
        fxy[1] = fxy[1] - self.bxyz[0]
        self.fxy = np.asfortranarray(fxy)
        dcu_k = dcu_k * self.dxi
        self.cu[0:2] = self.cu[0:2] - dcu_k[0:2]
        for s in self.species:
            s.push(self.fxy, self.bxyz, self.psit, d)
        self.fxy = self.fxy * d
        b = self.bxyz.copy(order="F")
        b[0] = b[0] * d
        b[1] = b[1] * d
        self.bxyz = b
        self._store(j)
