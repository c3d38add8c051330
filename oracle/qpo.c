/*
 * qpo.c -- CPU oracle (FP64, one thread) for the quasi-static slice solver.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  Never part of the product.
 *
 * Every function restates the arithmetic of one reference routine, cited by
 * file:line, in the reference's own operation order wherever rounding could
 * depend on it.  The sine/cosine transforms are an independent restatement of
 * the same mathematics (complex FFT of the symmetric extension) rather than of
 * the reference's Numerical-Recipes recursion, so they agree with it to
 * rounding, not bit for bit.
 */
#define _GNU_SOURCE
#include "qpo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TWOPI_REF 6.28318530717959 /* the reference's literal for 2*pi */

/* ------------------------------------------------------------------------ */
/* ranorm: part3d_lib77.f:1324-1383 (Box-Muller over two 2x31-bit LCGs)      */
/* ------------------------------------------------------------------------ */
void qpo_rng_init(qpo_rng *s)
{
    /* DATA statements, part3d_lib77.f:1335-1337 */
    s->r1 = 885098780;
    s->r2 = 1824280461;
    s->r4 = 1396483093;
    s->r5 = 55318673;
    s->iflg = 0;
    s->r0 = 0.0;
}

double qpo_ranorm(qpo_rng *s)
{
    const double h1l = 65531.0, h1u = 32767.0, h2l = 65525.0;
    if (s->iflg != 0) {
        double out = s->r0;
        s->r0 = 0.0;
        s->iflg = 0;
        return out;
    }
    int32_t isc = 65536, i1;
    double asc = (double)isc, bsc = asc * asc, r3, temp, r0;
    /* first generator pair -> radius */
    i1 = s->r1 - (s->r1 / isc) * isc;
    r3 = h1l * (double)s->r1 + asc * h1u * (double)i1;
    i1 = (int32_t)(r3 / bsc);
    r3 = r3 - (double)i1 * bsc;
    bsc = 0.5 * bsc;
    i1 = s->r2 / isc;
    isc = s->r2 - i1 * isc;
    r0 = h1l * (double)s->r2 + asc * h1u * (double)isc;
    asc = 1.0 / bsc;
    isc = (int32_t)(r0 * asc);
    s->r2 = (int32_t)(r0 - (double)isc * bsc);
    r3 = r3 + ((double)isc + 2.0 * h1u * (double)i1);
    isc = (int32_t)(r3 * asc);
    s->r1 = (int32_t)(r3 - (double)isc * bsc);
    temp = sqrt(-2.0 * log(((double)s->r1 + (double)s->r2 * asc) * asc));
    /* second generator pair -> angle */
    isc = 65536;
    asc = (double)isc;
    bsc = asc * asc;
    i1 = s->r4 - (s->r4 / isc) * isc;
    r3 = h2l * (double)s->r4 + asc * h1u * (double)i1;
    i1 = (int32_t)(r3 / bsc);
    r3 = r3 - (double)i1 * bsc;
    bsc = 0.5 * bsc;
    i1 = s->r5 / isc;
    isc = s->r5 - i1 * isc;
    r0 = h2l * (double)s->r5 + asc * h1u * (double)isc;
    asc = 1.0 / bsc;
    isc = (int32_t)(r0 * asc);
    s->r5 = (int32_t)(r0 - (double)isc * bsc);
    r3 = r3 + ((double)isc + 2.0 * h1u * (double)i1);
    isc = (int32_t)(r3 * asc);
    s->r4 = (int32_t)(r3 - (double)isc * bsc);
    r0 = 6.28318530717959 * (((double)s->r4 + (double)s->r5 * asc) * asc);
    double sn, cs;
    sincos(r0, &sn, &cs);
    s->r0 = temp * cs;
    s->iflg = 1;
    return temp * sn;
}

/* ------------------------------------------------------------------------ */
/* Tri-Gaussian beam: normalisation fdist3d_class.f03:305-335, loading       */
/* PRVDIST32_RANDOM part3d_lib77.f:411-552 (one partition covering the box). */
/* ------------------------------------------------------------------------ */
long qpo_beam_init(const qpo_beam_params *bp, double *part, long npmax,
                   qpo_rng *rng, int *ierr, double *qm_out)
{
    const double dx = bp->dx, dy = bp->dy, dz = bp->dz;
    /* fdist3d_class.f03:311-316 */
    double qm = bp->q / fabs(bp->q) * fabs(bp->peak) *
                pow(2 * 3.1415926535897932, 1.5) * bp->sigx * bp->sigy *
                bp->sigz;
    qm = qm / dx / dy / dz;
    qm = qm / bp->npx;
    qm = qm / bp->npy;
    qm = qm / bp->npz;
    if (qm_out) *qm_out = qm;
    const double x0 = bp->bcx / dx, y0 = bp->bcy / dy, z0 = bp->bcz / dz;
    const double sigx = bp->sigx / dx, sigy = bp->sigy / dy,
                 sigz = bp->sigz / dz;
    const double vtx = bp->sigvx, vty = bp->sigvy, vtz = bp->sigvz;
    const double vdx = 0.0, vdy = 0.0, vdz = bp->gamma;
    double cx[3], cy[3];
    cx[0] = bp->cx[0] * dz * dz / dx;
    cx[1] = bp->cx[1] * dz / dx;
    cx[2] = bp->cx[2] / dx;
    cy[0] = bp->cy[0] * dz * dz / dy;
    cy[1] = bp->cy[1] * dz / dy;
    cy[2] = bp->cy[2] / dy;
    const int nx = bp->nx, ny = bp->ny, nz = bp->nz;
    const double edges[4] = {0.0, (double)ny, 0.0, (double)nz};
    /* part3d_lib77.f:460-468 */
    const double x2 = 2.0 * x0, y2 = 2.0 * y0;
    const double borderlx = fmax(x0 - 5.0 * sigx, 1.0);
    const double borderly = fmax(y0 - 5.0 * sigy, 1.0);
    const double borderlz = fmax(z0 - 5.0 * sigz, 1.0);
    const double borderx = fmin(x0 + 5.0 * sigx, (double)(nx - 1));
    const double bordery = fmin(y0 + 5.0 * sigy, (double)(ny - 1));
    const double borderz = fmin(z0 + 5.0 * sigz, (double)(nz - 1));
    long npp = 0, npt = 1;
    *ierr = 0;
    int j = 0, k, l = bp->npz;
    while (j < bp->npx) {
        k = 0;
        while (k < bp->npy) {
            l = l - bp->npz;
            while (l < bp->npz) {
                double tempz, tempx, tempy;
                do { tempz = z0 + sigz * qpo_ranorm(rng); }
                while (tempz >= borderz || tempz <= borderlz);
                do { tempx = x0 + sigx * qpo_ranorm(rng); }
                while (tempx >= borderx || tempx <= borderlx);
                do { tempy = y0 + sigy * qpo_ranorm(rng); }
                while (tempy >= bordery || tempy <= borderly);
                double tvtx = vtx * qpo_ranorm(rng) + vdx;
                double tvty = vty * qpo_ranorm(rng) + vdy;
                double tvtz = vtz * qpo_ranorm(rng) + vdz;
                tvtz = sqrt(tvtz * tvtz - 1 - tvtx * tvtx - tvty * tvty);
                double tempxx = 0.0, tempyy = 0.0;
                if (tempy >= edges[0] && tempy < edges[1] &&
                    tempz >= edges[2] && tempz < edges[3]) {
                    if (npt < npmax) {
                        npt = npp + 1;
                        double *p = part + 7 * (npt - 1);
                        p[2] = tempz;
                        tempxx = -cx[0] * (p[2] - z0) * (p[2] - z0) -
                                 cx[1] * (p[2] - z0) - cx[2];
                        p[0] = tempx + tempxx;
                        tempyy = -cy[0] * (p[2] - z0) * (p[2] - z0) -
                                 cy[1] * (p[2] - z0) - cy[2];
                        p[1] = tempy + tempyy;
                        p[3] = tvtx;
                        p[4] = tvty;
                        p[5] = tvtz;
                        p[6] = qm;
                        npp = npt;
                    } else {
                        *ierr += 1;
                    }
                    if (bp->quiet) {
                        if (npt < npmax) {
                            npt = npp + 1;
                            double *p = part + 7 * (npt - 1);
                            p[2] = tempz;
                            p[0] = x2 - tempx + tempxx;
                            p[1] = y2 - tempy + tempyy;
                            p[3] = -tvtx;
                            p[4] = -tvty;
                            p[5] = tvtz;
                            p[6] = qm;
                            npp = npt;
                        } else {
                            *ierr += 1;
                        }
                    }
                }
                l = l + 1;
                if (bp->quiet) l = l + 1;
            }
            k = k + 1;
        }
        j = j + 1;
    }
    return npp;
}

/* ------------------------------------------------------------------------ */
/* Uniform plasma, profile 0: fdist2d_class.f03:212-316.  With one rank      */
/* noff = 0 < ny, so the branch at :256-273 applies: x in [1,nx-1),          */
/* y in [1,ny).  Record: x, y, px, py, pz, p6 = 1, p7 = 1, q.                */
/* ------------------------------------------------------------------------ */
long qpo_plasma_init(int nx, int ny, int xppc, int yppc, double den, double q,
                     double *part)
{
    double qm = 1.0 * den * q / fabs(q) / (double)xppc / (double)yppc;
    long n = 0;
    for (int i = 2; i <= nx - 1; ++i)
        for (int j = 2; j <= ny; ++j)
            for (int ix = 0; ix < xppc; ++ix)
                for (int iy = 0; iy < yppc; ++iy) {
                    double *p = part + 8 * n;
                    p[0] = (ix + 0.5) / xppc + i - 1;
                    p[1] = (iy + 0.5) / yppc + j - 1;
                    p[2] = 0.0;
                    p[3] = 0.0;
                    p[4] = 0.0;
                    p[5] = 1.0;
                    p[6] = 1.0;
                    p[7] = qm;
                    ++n;
                }
    return n;
}

/* ------------------------------------------------------------------------ */
/* Charge deposit, PPGPPOST2L part2d_lib77.f:194-309 (CIC, 4 nodes).         */
/* Adds into q; q is (ny+1) x (nx+1).                                        */
/* ------------------------------------------------------------------------ */
void qpo_qdeposit(int nx, int ny, long np, const double *part, double *q)
{
    const int ld = nx + 1;
    (void)ny;
    for (long j = 0; j < np; ++j) {
        const double *p = part + 8 * j;
        double x = p[0], y = p[1], qm = p[7];
        int nn = (int)x, mm = (int)y;
        double dxp = qm * (x - (double)nn);
        double dyp = y - (double)mm;
        double amx = qm - dxp;
        double amy = 1.0 - dyp;
        double *r0 = q + (long)mm * ld + nn, *r1 = r0 + ld;
        r0[0] += amx * amy;
        r0[1] += dxp * amy;
        r1[0] += amx * dyp;
        r1[1] += dxp * dyp;
    }
}

/* Guard add for one rank, PPNACGUARD2L ufield2d_lib.f03:90-97: row ny is   */
/* folded into row 0 and cleared, for x = 0..nx.                             */
void qpo_guard_fold(int nx, int ny, int dim, double *f)
{
    const long plane = (long)(ny + 1) * (nx + 1);
    for (int c = 0; c < dim; ++c) {
        double *g = f + c * plane;
        for (int x = 0; x <= nx; ++x) {
            g[x] = g[x] + g[(long)ny * (nx + 1) + x];
            g[(long)ny * (nx + 1) + x] = 0.0;
        }
    }
}

/* Guard copy for one rank, PPNCGUARD2L ufield2d_lib.f03:37-41: row ny is    */
/* overwritten with row 0 (every component).                                 */
void qpo_copyguard(int nx, int ny, int dim, double *f)
{
    const long plane = (long)(ny + 1) * (nx + 1);
    for (int c = 0; c < dim; ++c) {
        double *g = f + c * plane;
        memcpy(g + (long)ny * (nx + 1), g, sizeof(double) * (nx + 1));
    }
}

/* bilinear gather in the reference's order (e.g. part2d_lib77.f:2265-2268) */
static inline double gather(const double *f, int ld, int nn, int mm, double amx,
                            double amy, double dxp, double dyp)
{
    const double *r0 = f + (long)mm * ld + nn, *r1 = r0 + ld;
    double d = amx * r0[0];
    d = amy * (dxp * r0[1] + d);
    double a = amx * r1[0];
    return d + dyp * (dxp * r1[1] + a);
}

/* ------------------------------------------------------------------------ */
/* psi gather and the quasi-static identities, WPGPSIPOST2L_QP               */
/* part2d_lib77.f:2207-2284.                                                 */
/* ------------------------------------------------------------------------ */
void qpo_extpsi(int nx, int ny, long np, double *part, const double *psi,
                double qbm, double dex)
{
    const int ld = nx + 1;
    const double dx2 = dex * dex;
    (void)ny;
    for (long j = 0; j < np; ++j) {
        double *p = part + 8 * j;
        double x = p[0], y = p[1], vx = p[2], vy = p[3];
        int nn = (int)x, mm = (int)y;
        double dxp = x - (double)nn, dyp = y - (double)mm;
        double amx = 1.0 - dxp, amy = 1.0 - dyp;
        double dx = gather(psi, ld, nn, mm, amx, amy, dxp, dyp);
        dx = -(dx * dx2 * qbm);
        p[6] = 1.0 + dx;
        p[5] = dx + (dx2 * (vx * vx + vy * vy) - dx * dx - 2. * dx) /
                        (2. * (1. + dx)) +
               1.0;
    }
}

/* ------------------------------------------------------------------------ */
/* Current, acceleration and momentum-flux deposit with the half Boris step, */
/* PPGRDCJPPOST2L_QP part2d_lib77.f:311-768.  fxy: 2 comps, bxy: 3, psit: 1; */
/* adds into cu (3), dcu (2), amu (3).  Updates p6 in place (:597).          */
/* ------------------------------------------------------------------------ */
void qpo_amjdp(int nx, int ny, long np, double *part, const double *fxy,
               const double *bxy, const double *psit, double *cu, double *dcu,
               double *amu, double qbm, double dt, double dex)
{
    const int ld = nx + 1;
    const long pl = (long)(ny + 1) * ld;
    const double qtmh = 0.5 * qbm * dt, dti = 1.0 / dt, idex = 1.0 / dex;
    for (long j = 0; j < np; ++j) {
        double *p = part + 8 * j;
        double x = p[0], y = p[1], p6 = p[5], p7 = p[6], qm = p[7];
        int nn = (int)x, mm = (int)y;
        double dxp = x - (double)nn, dyp = y - (double)mm;
        double amx = 1.0 - dxp, amy = 1.0 - dyp;
        double inv_part_7 = 1.0 / p7;
        double qtmh1 = qtmh * inv_part_7;
        double qtmh2 = qtmh1 * p6;
        double dx = gather(fxy, ld, nn, mm, amx, amy, dxp, dyp);
        double dy = gather(fxy + pl, ld, nn, mm, amx, amy, dxp, dyp);
        double dz = gather(psit, ld, nn, mm, amx, amy, dxp, dyp);
        double ox = gather(bxy, ld, nn, mm, amx, amy, dxp, dyp);
        double oy = gather(bxy + pl, ld, nn, mm, amx, amy, dxp, dyp);
        double oz = gather(bxy + 2 * pl, ld, nn, mm, amx, amy, dxp, dyp);
        /* half impulse, :571-572 */
        double ddx = (-1.0) * qtmh2 * dx + qtmh * oy;
        double ddy = (-1.0) * qtmh2 * dy - qtmh * ox;
        double vx = p[2], vy = p[3];
        double acx = vx + ddx, acy = vy + ddy;
        /* rotation about B_z only, :580-585 */
        double omzt = qtmh1 * oz;
        double omt = omzt * omzt;
        double anorm = 2.0 / (1.0 + omt);
        double rot1 = 0.5 * (1.0 - omt);
        double rot2 = omzt;
        double v1 = (rot1 * acx + rot2 * acy) * anorm + ddx;
        double v2 = (rot1 * acy - rot2 * acx) * anorm + ddy;
        /* deposit weights and quantities, :591-615 */
        double qm1 = qm * inv_part_7;
        amx = qm1 * amx;
        dxp = qm1 * dxp;
        ox = 0.5 * (v1 + vx);
        oy = 0.5 * (v2 + vy);
        oz = 0.5 * (1.0 + dex * dex * (ox * ox + oy * oy)) * inv_part_7 -
             0.5 * p7;
        p[5] = p7 + oz;
        oz = oz * idex;
        vx = (v1 - vx) * dti;
        vy = (v2 - vy) * dti;
        dz = qbm * (dz + (dx * ox + dy * oy) * dex * dex * inv_part_7);
        double w00 = amx * amy, w10 = dxp * amy;
        v1 = ox * ox * inv_part_7;
        v2 = oy * oy * inv_part_7;
        double v3 = ox * oy * inv_part_7;
        vx = vx + ox * dz * inv_part_7;
        vy = vy + oy * dz * inv_part_7;
        double w01 = amx * dyp, w11 = dxp * dyp;
        const double wt[4] = {w00, w10, w01, w11};
        const long off[4] = {(long)mm * ld + nn, (long)mm * ld + nn + 1,
                             (long)(mm + 1) * ld + nn,
                             (long)(mm + 1) * ld + nn + 1};
        for (int c = 0; c < 4; ++c) {
            const double w = wt[c];
            const long o = off[c];
            amu[o] += v1 * w;
            amu[pl + o] += v2 * w;
            amu[2 * pl + o] += v3 * w;
            dcu[o] += vx * w;
            dcu[pl + o] += vy * w;
            cu[o] += ox * w;
            cu[pl + o] += oy * w;
            cu[2 * pl + o] += oz * w;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* 3-D Boris push in xi, PPGRBPPUSHF23L_QP part2d_lib77.f:1066-1408, with    */
/* the periodic wrap of the reorder PPPORDER2LB part2d_lib77.f:1936-1977.    */
/* Returns 0, or the 1-based tile index of a particle that moved further     */
/* than one 16x16 tile (the condition PPPCHECK2L :133-192 reports).          */
/* ------------------------------------------------------------------------ */
#define QPO_MX 16
#define QPO_MY 16

int qpo_push(int nx, int ny, long np, double *part, const double *fxy,
             const double *bxy, const double *psit, double qbm, double dt,
             double dex)
{
    const int ld = nx + 1;
    const long pl = (long)(ny + 1) * ld;
    const double idex = 1.0 / dex, qtmh = 0.5 * qbm * dt;
    const double anx = (double)nx, any = (double)ny;
    const int mx1 = (nx - 1) / QPO_MX + 1, my1 = (ny - 1) / QPO_MY + 1;
    int err = 0;
    for (long j = 0; j < np; ++j) {
        double *p = part + 8 * j;
        double x = p[0], y = p[1];
        int nn = (int)x, mm = (int)y;
        double p6 = p[5], p7 = p[6];
        double qtmh1 = qtmh / p7;
        double qtmh2 = qtmh1 * p6;
        double dxp = x - (double)nn, dyp = y - (double)mm;
        double amx = 1.0 - dxp, amy = 1.0 - dyp;
        double dx = gather(fxy, ld, nn, mm, amx, amy, dxp, dyp);
        double dy = gather(fxy + pl, ld, nn, mm, amx, amy, dxp, dyp);
        double dz = gather(psit, ld, nn, mm, amx, amy, dxp, dyp);
        dz = dz * idex;
        double ox = gather(bxy, ld, nn, mm, amx, amy, dxp, dyp);
        double oy = gather(bxy + pl, ld, nn, mm, amx, amy, dxp, dyp);
        double oz = gather(bxy + 2 * pl, ld, nn, mm, amx, amy, dxp, dyp);
        ox = ox * dex;
        oy = oy * dex;
        dx = qtmh2 * dx;
        dy = qtmh2 * dy;
        dz = qtmh2 * dz;
        double acx = p[2] + dx, acy = p[3] + dy, acz = p[4] + dz;
        double omxt = qtmh1 * ox, omyt = qtmh1 * oy, omzt = qtmh1 * oz;
        double omt = omxt * omxt + omyt * omyt + omzt * omzt;
        double anorm = 2.0 / (1.0 + omt);
        omt = 0.5 * (1.0 - omt);
        double rot4 = omxt * omyt, rot7 = omxt * omzt, rot8 = omyt * omzt;
        double rot1 = omt + omxt * omxt;
        double rot5 = omt + omyt * omyt;
        double rot9 = omt + omzt * omzt;
        double rot2 = omzt + rot4;
        rot4 = -omzt + rot4;
        double rot3 = -omyt + rot7;
        rot7 = omyt + rot7;
        double rot6 = omxt + rot8;
        rot8 = -omxt + rot8;
        dx = (rot1 * acx + rot2 * acy + rot3 * acz) * anorm + dx;
        dy = (rot4 * acx + rot5 * acy + rot6 * acz) * anorm + dy;
        dz = (rot7 * acx + rot8 * acy + rot9 * acz) * anorm + dz;
        p[2] = dx;
        p[3] = dy;
        p[4] = dz;
        double dtc1 =
            dt / (sqrt(1 + (dx * dx + dy * dy + dz * dz) * dex * dex) -
                  dz * dex);
        dx = x + dx * dtc1;
        dy = y + dy * dtc1;
        /* leaving direction relative to the old tile, :1350-1389 */
        int tx0 = nn / QPO_MX, ty0 = mm / QPO_MY;
        double edgelx = QPO_MX * tx0, edgely = QPO_MY * ty0;
        double edgerx = edgelx + (nx - edgelx < QPO_MX ? nx - edgelx : QPO_MX);
        double edgery = edgely + (ny - edgely < QPO_MY ? ny - edgely : QPO_MY);
        int dirx = 0, diry = 0;
        if (dx >= edgerx) dirx = 1;
        else if (dx < edgelx) {
            dirx = -1;
            /* roundoff guard, :1358-1366: a tiny negative x that wraps to  */
            /* nx exactly stays in its tile at x = 0                         */
            if (dx < 0.0 && dx + anx >= anx) { dx = 0.0; dirx = 0; }
        }
        if (dy >= edgery) diry = 1;
        else if (dy < edgely) {
            diry = -1;
            if (dy < 0.0 && dy + any >= any) { dy = 0.0; diry = 0; }
        }
        /* periodic wrap of the movers, PPPORDER2LB :1936-1977 */
        if (dx < 0.0) dx = dx + anx;
        if (dx >= anx) dx = dx - anx;
        if (dy < 0.0) dy = dy + any;
        if (dy >= any) dy = dy - any;
        p[0] = dx;
        p[1] = dy;
        int txd = (tx0 + dirx + mx1) % mx1, tyd = (ty0 + diry + my1) % my1;
        int tx1 = (int)dx / QPO_MX, ty1 = (int)dy / QPO_MY;
        if ((tx1 != txd || ty1 != tyd) && err == 0)
            err = 1 + txd + mx1 * tyd;
    }
    return err;
}

/* 1-based tile of every particle, PPDBLKP2L part2d_lib77.f:48-53 */
void qpo_tile_of(int nx, int ny, long np, const double *part, int32_t *tile)
{
    const int mx1 = (nx - 1) / QPO_MX + 1;
    (void)ny;
    for (long j = 0; j < np; ++j) {
        int n = (int)part[8 * j] / QPO_MX + 1;
        int m = (int)part[8 * j + 1] / QPO_MY;
        tile[j] = n + mx1 * m;
    }
}

/* row-major cell index y*nx + x of every particle */
void qpo_cell_of(int nx, int ny, long np, const double *part, int32_t *cell)
{
    (void)ny;
    for (long j = 0; j < np; ++j)
        cell[j] = (int32_t)part[8 * j + 1] * nx + (int32_t)part[8 * j];
}

/* Stable regrouping of the particles by tile with per-tile counts kpic      */
/* (the state PPPMOVIN2L part2d_lib77.f:79-131 and PPPORDER2LB leave; the    */
/* order inside a tile is this oracle's own).  scratch: 8*np doubles.        */
void qpo_order_by_tile(int nx, int ny, long np, double *part, int32_t *kpic,
                       double *scratch)
{
    const int mx1 = (nx - 1) / QPO_MX + 1, my1 = (ny - 1) / QPO_MY + 1;
    const int nt = mx1 * my1;
    long *start = (long *)calloc((size_t)nt + 1, sizeof(long));
    int32_t *tile = (int32_t *)malloc(sizeof(int32_t) * (size_t)(np > 0 ? np : 1));
    qpo_tile_of(nx, ny, np, part, tile);
    for (int k = 0; k < nt; ++k) kpic[k] = 0;
    for (long j = 0; j < np; ++j) kpic[tile[j] - 1]++;
    for (int k = 0; k < nt; ++k) start[k + 1] = start[k] + kpic[k];
    for (long j = 0; j < np; ++j) {
        long d = start[tile[j] - 1]++;
        memcpy(scratch + 8 * d, part + 8 * j, 8 * sizeof(double));
    }
    memcpy(part, scratch, sizeof(double) * 8 * (size_t)np);
    free(tile);
    free(start);
}

/* ------------------------------------------------------------------------ */
/* Sine / cosine transforms (fft2d_lib77.f:4791-4964, :4966-5141 and the     */
/* y-pass and multi-component variants).  One-dimensional definitions, for a */
/* vector f(0..n):                                                           */
/*   S(k) = sum_{j=1}^{n-1} f(j) sin(pi j k/n),  S(0) = S(n) = 0             */
/*   C(k) = f(0)/2 + (-1)^k f(n)/2 + sum_{j=1}^{n-1} f(j) cos(pi j k/n)      */
/* computed by a complex FFT of length 2n of the odd (S) or even (C)         */
/* extension; two vectors ride in the real and imaginary parts.              */
/* ------------------------------------------------------------------------ */
static void fft_inplace(int m, double *re, double *im)
{
    /* radix-2 decimation in time, forward sign e^{-2 pi i jk/m} */
    for (int i = 1, j = 0; i < m; ++i) {
        int bit = m >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= m; len <<= 1) {
        const double ang = -2.0 * M_PI / len;
        for (int k = 0; k < len / 2; ++k) {
            double wr, wi;
            sincos(ang * k, &wi, &wr);
            for (int i = 0; i < m; i += len) {
                int a = i + k, b = i + k + len / 2;
                double tr = re[b] * wr - im[b] * wi;
                double ti = re[b] * wi + im[b] * wr;
                re[b] = re[a] - tr; im[b] = im[a] - ti;
                re[a] += tr; im[a] += ti;
            }
        }
    }
}

/* count vectors of length n+1, contiguous; out receives the un-normalised  */
/* S or C of each.                                                           */
void qpo_transform_1d(int n, int type, int count, const double *in,
                      double *out)
{
    const int m = 2 * n;
    double *re = (double *)malloc(sizeof(double) * m);
    double *im = (double *)malloc(sizeof(double) * m);
    for (int v = 0; v < count; v += 2) {
        const double *a = in + (long)v * (n + 1);
        const double *b = (v + 1 < count) ? a + (n + 1) : NULL;
        for (int j = 0; j <= n; ++j) {
            double fa = a[j], fb = b ? b[j] : 0.0;
            if (type == QPO_SIN && (j == 0 || j == n)) { fa = 0.0; fb = 0.0; }
            re[j] = fa;
            im[j] = fb;
            if (j > 0 && j < n) {
                double s = (type == QPO_SIN) ? -1.0 : 1.0;
                re[m - j] = s * fa;
                im[m - j] = s * fb;
            }
        }
        fft_inplace(m, re, im);
        double *oa = out + (long)v * (n + 1);
        double *ob = b ? oa + (n + 1) : NULL;
        for (int k = 0; k <= n; ++k) {
            if (type == QPO_COS) {
                oa[k] = 0.5 * re[k];
                if (ob) ob[k] = 0.5 * im[k];
            } else {
                oa[k] = (k == 0 || k == n) ? 0.0 : -0.5 * im[k];
                if (ob) ob[k] = (k == 0 || k == n) ? 0.0 : 0.5 * re[k];
            }
        }
    }
    free(re);
    free(im);
}

/* Component transform types of the wrappers used by field2d%fftrk/fftkr:    */
/*  kind 1: dim1 WPPFSST2RM (S,S); dim2 WPPFCST2RM2 (C,S),(S,C);             */
/*          dim3 WPPFCST2RM3 (C,S),(S,C),(S,S)                               */
/*  kind 2: dim1 WPPFCCT2RM (C,C); dim2 WPPFSCT2RM2 (S,C),(C,S);             */
/*          dim3 WPPFSCT2RM3 (S,C),(C,S),(C,C)                               */
/*  kind 3: dim3 WPPFSST2RM23 (S,S),(S,S),(C,C)                              */
/* (x type, y type); fft2d_class.f03:121-413, fft2d_lib77.f:4479, :4713,     */
/* :5495, :5577, :6495, :6577, :8640 and the X/Y passes they call.           */
void qpo_kind_types(int kind, int dim, int *tx, int *ty)
{
    static const int k1[3][3][2] = {
        {{0, 0}}, {{1, 0}, {0, 1}}, {{1, 0}, {0, 1}, {0, 0}}};
    static const int k2[3][3][2] = {
        {{1, 1}}, {{0, 1}, {1, 0}}, {{0, 1}, {1, 0}, {1, 1}}};
    static const int k3[3][2] = {{0, 0}, {0, 0}, {1, 1}};
    for (int c = 0; c < dim; ++c) {
        if (kind == 1) { tx[c] = k1[dim - 1][c][0]; ty[c] = k1[dim - 1][c][1]; }
        else if (kind == 2) { tx[c] = k2[dim - 1][c][0]; ty[c] = k2[dim - 1][c][1]; }
        else { tx[c] = k3[c][0]; ty[c] = k3[c][1]; }
    }
}

/* real -> k (isign = -1): x pass scaled by 1/(nx*ny), then y pass.          */
void qpo_fft_r2k(int nx, int ny, int dim, const int *tx, const int *ty,
                 const double *rs, double *ks)
{
    const long rpl = (long)(ny + 1) * (nx + 1);
    double *t = (double *)malloc(sizeof(double) * rpl);
    double *u = (double *)malloc(sizeof(double) * rpl);
    const double ani = 1.0 / ((double)nx * (double)ny);
    for (int c = 0; c < dim; ++c) {
        qpo_transform_1d(nx, tx[c], ny + 1, rs + c * rpl, t); /* t[y][kx] */
        for (int y = 0; y <= ny; ++y)
            for (int kx = 0; kx <= nx; ++kx)
                u[(long)kx * (ny + 1) + y] = ani * t[(long)y * (nx + 1) + kx];
        qpo_transform_1d(ny, ty[c], nx + 1, u, ks + c * rpl); /* [kx][ky] */
    }
    free(t);
    free(u);
}

/* k -> real (isign = +1): y pass and x pass, each with a factor 2, then the */
/* guard copy of fftkr (field2d_class.f03:223).                              */
void qpo_fft_k2r(int nx, int ny, int dim, const int *tx, const int *ty,
                 const double *ks, double *rs)
{
    const long rpl = (long)(ny + 1) * (nx + 1);
    double *t = (double *)malloc(sizeof(double) * rpl);
    double *u = (double *)malloc(sizeof(double) * rpl);
    for (int c = 0; c < dim; ++c) {
        qpo_transform_1d(ny, ty[c], nx + 1, ks + c * rpl, t); /* t[kx][y] */
        for (int kx = 0; kx <= nx; ++kx)
            for (int y = 0; y <= ny; ++y)
                u[(long)y * (nx + 1) + kx] = 2.0 * t[(long)kx * (ny + 1) + y];
        qpo_transform_1d(nx, tx[c], ny + 1, u, rs + c * rpl);
        for (long i = 0; i < rpl; ++i) rs[c * rpl + i] *= 2.0;
    }
    qpo_copyguard(nx, ny, dim, rs);
    free(t);
    free(u);
}

/* ------------------------------------------------------------------------ */
/* k-space operators.  ffg = real(ffc) (Green's function with one shape     */
/* factor), ffs = aimag(ffc) (shape factor), both [kx][ky] for kx < nx,      */
/* ky < ny, stride ny.                                                       */
/* ------------------------------------------------------------------------ */

/* MPPOISD22 isign = 0, fpois2d_lib77.f:1514-1531 */
void qpo_pois_table(int nx, int ny, double ax, double ay, double affp,
                    double *ffg, double *ffs)
{
    const double dnx = TWOPI_REF / (double)(nx + nx);
    const double dny = TWOPI_REF / (double)(ny + ny);
    for (int j = 1; j <= nx; ++j) {
        double dkx = dnx * (double)(j - 1);
        double at1 = dkx * dkx;
        double at2 = (dkx * ax) * (dkx * ax);
        for (int k = 1; k <= ny; ++k) {
            double dky = dny * (double)(k - 1);
            double at3 = dky * dky + at1;
            double at4 = exp(-0.5 * ((dky * ay) * (dky * ay) + at2));
            long o = (long)(j - 1) * ny + (k - 1);
            if (at3 == 0.0) { ffg[o] = affp; ffs[o] = 1.0; }
            else { ffg[o] = affp * at4 / at3; ffs[o] = at4; }
        }
    }
}

#define KIDX(kx, ky) ((long)(kx) * (ny + 1) + (ky))
#define FIDX(kx, ky) ((long)(kx) * ny + (ky))

/* MPPDIVFD2 fft2d_lib77.f:9170-9229: df = -(kx f1 + ky f2) */
void qpo_div(int nx, int ny, const double *f, double *df)
{
    const long pl = (long)(nx + 1) * (ny + 1);
    const double dnx = TWOPI_REF / (double)(nx + nx);
    const double dny = TWOPI_REF / (double)(ny + ny);
    for (int kx = 0; kx < nx; ++kx) {
        double dkx = dnx * (double)kx;
        if (kx > 0)
            for (int ky = 1; ky < ny; ++ky) {
                double dky = dny * (double)ky;
                df[KIDX(kx, ky)] =
                    -(dkx * f[KIDX(kx, ky)] + dky * f[pl + KIDX(kx, ky)]);
            }
        df[KIDX(kx, 0)] = 0.0;
        df[KIDX(kx, ny)] = 0.0;
    }
    for (int ky = 1; ky < ny; ++ky) df[KIDX(0, ky)] = 0.0;
    for (int ky = 0; ky <= ny; ++ky) df[KIDX(nx, ky)] = 0.0;
}

/* MPPGRADFD2 fft2d_lib77.f:9231-9295: f = (kx df, ky df) */
void qpo_grad(int nx, int ny, const double *df, double *f)
{
    const long pl = (long)(nx + 1) * (ny + 1);
    const double dnx = TWOPI_REF / (double)(nx + nx);
    const double dny = TWOPI_REF / (double)(ny + ny);
    for (int kx = 0; kx < nx; ++kx) {
        double dkx = dnx * (double)kx;
        if (kx > 0)
            for (int ky = 1; ky < ny; ++ky) {
                double dky = dny * (double)ky;
                f[KIDX(kx, ky)] = dkx * df[KIDX(kx, ky)];
                f[pl + KIDX(kx, ky)] = dky * df[KIDX(kx, ky)];
            }
        f[KIDX(kx, 0)] = 0.0;
        f[pl + KIDX(kx, 0)] = 0.0;
        f[KIDX(kx, ny)] = 0.0;
        f[pl + KIDX(kx, ny)] = 0.0;
    }
    for (int ky = 1; ky < ny; ++ky) {
        f[KIDX(0, ky)] = 0.0;
        f[pl + KIDX(0, ky)] = 0.0;
    }
    for (int ky = 0; ky <= ny; ++ky) {
        f[KIDX(nx, ky)] = 0.0;
        f[pl + KIDX(nx, ky)] = 0.0;
    }
}

/* MPPOTPD2 fpois2d_lib77.f:1701-1779: pot = g q */
void qpo_pot(int nx, int ny, const double *ffg, const double *ffs,
             const double *q, double *pot)
{
    (void)ffs;
    for (int kx = 0; kx < nx; ++kx) {
        if (kx > 0)
            for (int ky = 1; ky < ny; ++ky)
                pot[KIDX(kx, ky)] = ffg[FIDX(kx, ky)] * q[KIDX(kx, ky)];
        pot[KIDX(kx, 0)] = 0.0;
        pot[KIDX(kx, ny)] = 0.0;
    }
    for (int ky = 1; ky < ny; ++ky) pot[KIDX(0, ky)] = 0.0;
    for (int ky = 0; ky <= ny; ++ky) pot[KIDX(nx, ky)] = 0.0;
}

/* MPPSMOOTHD2 fpois2d_lib77.f:1781-1843: qs = s q */
void qpo_smooth(int nx, int ny, const double *ffg, const double *ffs,
                const double *q, double *qs)
{
    (void)ffg;
    for (int kx = 0; kx < nx; ++kx) {
        if (kx > 0)
            for (int ky = 1; ky < ny; ++ky)
                qs[KIDX(kx, ky)] = ffs[FIDX(kx, ky)] * q[KIDX(kx, ky)];
        qs[KIDX(kx, 0)] = 0.0;
        qs[KIDX(kx, ny)] = 0.0;
    }
    for (int ky = 1; ky < ny; ++ky) qs[KIDX(0, ky)] = 0.0;
    for (int ky = 0; ky <= ny; ++ky) qs[KIDX(nx, ky)] = 0.0;
}

/* MPPOISD22 isign = -1, fpois2d_lib77.f:1533-1574: E = -k g s q */
void qpo_elf(int nx, int ny, const double *ffg, const double *ffs,
             const double *q, double *fxy)
{
    const long pl = (long)(nx + 1) * (ny + 1);
    const double dnx = TWOPI_REF / (double)(nx + nx);
    const double dny = TWOPI_REF / (double)(ny + ny);
    for (int kx = 0; kx < nx; ++kx) {
        double dkx = dnx * (double)kx;
        if (kx > 0)
            for (int ky = 1; ky < ny; ++ky) {
                double at1 = ffg[FIDX(kx, ky)] * ffs[FIDX(kx, ky)];
                double at3 = -at1 * q[KIDX(kx, ky)];
                double at2 = dkx * at3;
                at3 = dny * (double)ky * at3;
                fxy[KIDX(kx, ky)] = at2;
                fxy[pl + KIDX(kx, ky)] = at3;
            }
        fxy[KIDX(kx, 0)] = 0.0;
        fxy[pl + KIDX(kx, 0)] = 0.0;
        fxy[KIDX(kx, ny)] = 0.0;
        fxy[pl + KIDX(kx, ny)] = 0.0;
    }
    for (int ky = 1; ky < ny; ++ky) {
        fxy[KIDX(0, ky)] = 0.0;
        fxy[pl + KIDX(0, ky)] = 0.0;
    }
    for (int ky = 0; ky <= ny; ++ky) {
        fxy[KIDX(nx, ky)] = 0.0;
        fxy[pl + KIDX(nx, ky)] = 0.0;
    }
}

/* MPPBBPOISD23 fpois2d_lib77.f:1845-1956 (ci = 1): B from J */
void qpo_bf(int nx, int ny, const double *ffg, const double *ffs,
            const double *cu, double *bxy)
{
    const long pl = (long)(nx + 1) * (ny + 1);
    const double dnx = TWOPI_REF / (double)(nx + nx);
    const double dny = TWOPI_REF / (double)(ny + ny);
    const double ci2 = 1.0;
    for (int kx = 0; kx < nx; ++kx) {
        double dkx = dnx * (double)kx;
        if (kx > 0) {
            for (int ky = 1; ky < ny; ++ky) {
                double dky = dny * (double)ky;
                double at1 = ci2 * ffg[FIDX(kx, ky)] * ffs[FIDX(kx, ky)];
                double at2 = dky * at1, at3 = dkx * at1;
                long o = KIDX(kx, ky);
                bxy[o] = at2 * cu[2 * pl + o];
                bxy[pl + o] = -at3 * cu[2 * pl + o];
                bxy[2 * pl + o] = at3 * cu[pl + o] - at2 * cu[o];
            }
            double at1 = ci2 * ffg[FIDX(kx, 0)] * ffs[FIDX(kx, 0)];
            double at2 = dkx * at1;
            long o = KIDX(kx, 0);
            bxy[o] = 0.0;
            bxy[pl + o] = 0.0;
            bxy[2 * pl + o] = at2 * cu[pl + o];
        }
        long o = KIDX(kx, ny);
        bxy[o] = 0.0;
        bxy[pl + o] = 0.0;
        bxy[2 * pl + o] = 0.0;
    }
    for (int ky = 1; ky < ny; ++ky) {
        double dky = dny * (double)ky;
        double at1 = ci2 * ffg[FIDX(0, ky)] * ffs[FIDX(0, ky)];
        double at2 = dky * at1;
        long o = KIDX(0, ky);
        bxy[o] = 0.0;
        bxy[pl + o] = 0.0;
        bxy[2 * pl + o] = -at2 * cu[o];
    }
    bxy[KIDX(0, 0)] = 0.0;
    bxy[pl + KIDX(0, 0)] = 0.0;
    bxy[2 * pl + KIDX(0, 0)] = 0.0;
    for (int ky = 0; ky <= ny; ++ky) {
        long o = KIDX(nx, ky);
        bxy[o] = 0.0;
        bxy[pl + o] = 0.0;
        bxy[2 * pl + o] = 0.0;
    }
}

/* PBPOISD22N_QP isign = 1, fpois2d_lib77.f:1322-1421 (ci = 1): the         */
/* quasi-static B_perp update.  dcu is overwritten with -dcu + k x amu.      */
void qpo_bfqp(int nx, int ny, const double *ffg, const double *ffs,
              const double *cu, double *dcu, const double *amu, double *bxy,
              double aa, double dex)
{
    const long pl = (long)(nx + 1) * (ny + 1);
    const double dnx = TWOPI_REF / (double)(nx + nx);
    const double dny = TWOPI_REF / (double)(ny + ny);
    const double idex = 1.0 / dex, ci2 = 1.0;
    for (int kx = 0; kx < nx; ++kx) {
        double dkx = dnx * (double)kx;
        if (kx > 0) {
            for (int ky = 1; ky < ny; ++ky) {
                double dky = dny * (double)ky;
                long o = KIDX(kx, ky);
                dcu[o] = -dcu[o] + dkx * amu[o] - dky * amu[2 * pl + o];
                double aat1 = dcu[o] - dkx * cu[2 * pl + o] * idex;
                dcu[pl + o] = -dcu[pl + o] - dkx * amu[2 * pl + o] +
                              dky * amu[pl + o];
                double aat2 = dcu[pl + o] - dky * cu[2 * pl + o] * idex;
                double at2 = ffg[FIDX(kx, ky)];
                at2 = ci2 * at2 / (1 + aa * at2);
                double at1 = at2 * ffs[FIDX(kx, ky)];
                bxy[o] = at1 * aat1 + aa * at2 * bxy[o];
                bxy[pl + o] = at1 * aat2 + aa * at2 * bxy[pl + o];
            }
            long o = KIDX(kx, 0);
            dcu[o] = -dcu[o] + dkx * amu[o];
            dcu[pl + o] = -dcu[pl + o] - dkx * amu[2 * pl + o];
            double at2 = ffg[FIDX(kx, 0)];
            at2 = ci2 * at2 / (1 + aa * at2);
            double at1 = at2 * ffs[FIDX(kx, 0)];
            bxy[o] = 0.0;
            bxy[pl + o] = at1 * dcu[pl + o] + aa * at2 * bxy[pl + o];
        }
        bxy[KIDX(kx, ny)] = 0.0;
        bxy[pl + KIDX(kx, ny)] = 0.0;
    }
    for (int ky = 1; ky < ny; ++ky) {
        double dky = dny * (double)ky;
        long o = KIDX(0, ky);
        dcu[o] = -dcu[o] - dky * amu[2 * pl + o];
        dcu[pl + o] = -dcu[pl + o] + dky * amu[pl + o];
        double at2 = ffg[FIDX(0, ky)];
        at2 = ci2 * at2 / (1 + aa * at2);
        double at1 = at2 * ffs[FIDX(0, ky)];
        bxy[o] = at1 * dcu[o] + aa * at2 * bxy[o];
        bxy[pl + o] = 0.0;
    }
    bxy[KIDX(0, 0)] = 0.0;
    bxy[pl + KIDX(0, 0)] = 0.0;
    for (int ky = 0; ky <= ny; ++ky) {
        bxy[KIDX(nx, ky)] = 0.0;
        bxy[pl + KIDX(nx, ky)] = 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* Beam charge deposit, PGPOST32L part3d_lib77.f:554-628 (8 nodes).          */
/* With one rank and one stage noff = (0,0), so particle z in [l, l+1)       */
/* deposits to z planes l and l+1.  q3 is [nz+1][ny+1][nx+1].                */
/* ------------------------------------------------------------------------ */
void qpo_beam_deposit(int nx, int ny, int nz, long np, const double *part,
                      double *q3)
{
    const long ld = nx + 1, pl = (long)(ny + 1) * ld;
    (void)nz;
    for (long j = 0; j < np; ++j) {
        const double *p = part + 7 * j;
        int nn = (int)p[0], mm = (int)p[1], ll = (int)p[2];
        double qm = p[6];
        double dxp = qm * (p[0] - (double)nn);
        double dyp = p[1] - (double)mm;
        double dzp = p[2] - (double)ll;
        double amx = qm - dxp;
        double amy = 1. - dyp;
        double dx1 = dxp * dyp;
        dyp = amx * dyp;
        amx = amx * amy;
        double amz = 1. - dzp;
        amy = dxp * amy;
        double *b = q3 + (long)ll * pl + (long)mm * ld + nn;
        b[0] += amx * amz;
        b[1] += amy * amz;
        b[ld] += dyp * amz;
        b[ld + 1] += dx1 * amz;
        b[pl] += amx * dzp;
        b[pl + 1] += amy * dzp;
        b[pl + ld] += dyp * dzp;
        b[pl + ld + 1] += dx1 * dzp;
    }
}

/* trilinear gather in the reference's order, part3d_lib77.f:689-709 */
static inline double gather3(const double *f, long ld, long pl, double amx,
                             double amy, double dyp, double dx1, double amz,
                             double dzp)
{
    return amz * (amx * f[0] + amy * f[1] + dyp * f[ld] + dx1 * f[ld + 1]) +
           dzp * (amx * f[pl] + amy * f[pl + 1] + dyp * f[pl + ld] +
                  dx1 * f[pl + ld + 1]);
}

/* ------------------------------------------------------------------------ */
/* Beam push, PGBPUSH32L_QP part3d_lib77.f:630-786 with ipbc = 1 (drop at    */
/* [1, n-1)) and cofd = 0 (part3d_class.f03:196-199).  Dropped particles are */
/* replaced by the last live one and that one is pushed in their place,      */
/* exactly as :748-776.  e3: 3 comps, b3: 3 comps (only 1..2 read), each     */
/* [nz+1][ny+1][nx+1].  Returns the new particle count.                      */
/* ------------------------------------------------------------------------ */
long qpo_beam_push(int nx, int ny, int nz, long np, double *part,
                   const double *e3, const double *b3, double qbm, double dt,
                   double dex, double dez)
{
    const long ld = nx + 1, pl = (long)(ny + 1) * ld;
    const long vol = pl * (nz + 1);
    const double qtmh = qbm * dt;
    const double dtcx = dt / dex, dtcz = dt / dez;
    const double cofd1 = 0.0 / dt;
    const double edgelx = 1., edgely = 1., edgelz = 1.;
    const double edgerx = (double)(nx - 1), edgery = (double)(ny - 1),
                 edgerz = (double)(nz - 1);
    long inpp = np;
    for (long j = 0; j < np; ++j) {
        if (j >= inpp) break;
    again:;
        double *p = part + 7 * j;
        int nn = (int)p[0], mm = (int)p[1], ll = (int)p[2];
        double dxp = p[0] - (double)nn;
        double dyp = p[1] - (double)mm;
        double dzp = p[2] - (double)ll;
        double amx = 1. - dxp, amy = 1. - dyp;
        double dx1 = dxp * dyp;
        dyp = amx * dyp;
        amx = amx * amy;
        double amz = 1. - dzp;
        amy = dxp * amy;
        long o = (long)ll * pl + (long)mm * ld + nn;
        double dx = gather3(e3 + o, ld, pl, amx, amy, dyp, dx1, amz, dzp);
        double dy = gather3(e3 + vol + o, ld, pl, amx, amy, dyp, dx1, amz, dzp);
        double dz = gather3(e3 + 2 * vol + o, ld, pl, amx, amy, dyp, dx1, amz, dzp);
        double ox = gather3(b3 + o, ld, pl, amx, amy, dyp, dx1, amz, dzp);
        double oy = gather3(b3 + vol + o, ld, pl, amx, amy, dyp, dx1, amz, dzp);
        dx = dx - oy;
        dy = dy + ox;
        double delx = qtmh * dx, dely = qtmh * dy, delz = qtmh * dz;
        dx = p[3] + delx;
        dy = p[4] + dely;
        dz = p[5] + delz;
        p[3] = dx;
        p[4] = dy;
        double p2t = dx * dx + dy * dy;
        double p2l = dz * dz;
        double p2 = p2t + p2l;
        double ngamma = sqrt(1.0 + p2);
        double dtgx = dtcx / ngamma, dtgy = dtcx / ngamma;
        double one_minus_vz = (1.0 + p2t) / (dz * (dz + ngamma));
        dz = dz - cofd1 * ngamma *
                      ((delx * delx + dely * dely) * dz -
                       delz * (delx * dx + dely * dy));
        p[5] = dz;
        dx = p[0] + dx * dtgx;
        dy = p[1] + dy * dtgy;
        dz = p[2] + one_minus_vz * dtcz;
        if ((dx < edgelx) || (dx >= edgerx) || (dy < edgely) ||
            (dy >= edgery) || (dz < edgelz) || (dz >= edgerz)) {
            if (j == inpp - 1) { inpp -= 1; break; }
            memcpy(p, part + 7 * (inpp - 1), 7 * sizeof(double));
            inpp -= 1;
            goto again;
        }
        p[0] = dx;
        p[1] = dy;
        p[2] = dz;
    }
    return inpp;
}
