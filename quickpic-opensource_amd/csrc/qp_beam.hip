// Beam (beam3d_class.f03): tri-Gaussian loader, 3-D charge deposit and the
// relativistic push with drop-at-boundary.
//
// Beam particles stay in FP64 in both precision modes (positions over a long
// propagation).  The 3-D charge is accumulated in int64 fixed point (2^-44
// resolution) with 64-bit atomics, so it is independent of the order in which
// particles land and therefore identical run to run; it is converted to Real
// when a slice is copied out (qdpcopy).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "qp_common.h"

struct qp_beam {
    qp_ctx *ctx;
    qp_beam_params p;
    double qbm, dt;
    int64_t np, npmax;
    qp::DevBuf a[7], b[7];       // x, y, z, px, py, pz, q (double)
    qp::DevBuf alive, pos, scan_tmp, moments;
    qp::DevBuf key[2], perm[2];  // push: cell keys and particle indices, before and after the sort
    size_t scan_bytes;           // temporary bytes of the scan and the sort
    int key_shift, key_bits;     // cell key = linear cell index >> key_shift, below 2^key_bits
    unsigned key_drop;           // key of a dropped particle (sorts last)
    qp::DevBuf q3;               // int64 [nz+1][ny+1][ldr]
};

namespace qp {

// ---- ranorm (part3d_lib77.f:1324-1383): Box-Muller over two LCG pairs ----
struct Ranorm {
    int32_t r1, r2, r4, r5, iflg;
    double r0;
    explicit Ranorm(uint64_t seed)
    {
        iflg = 0;
        r0 = 0.0;
        if (seed == 0) {   // the reference's DATA state (:1335)
            r1 = 885098780; r2 = 1824280461; r4 = 1396483093; r5 = 55318673;
            return;
        }
        uint64_t z = seed;
        auto mix = [&z]() {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
            return t ^ (t >> 31);
        };
        r1 = (int32_t)(mix() & 0x7FFFFFFF);
        r2 = (int32_t)((mix() & 0x7FFFFFFF) | 1);
        r4 = (int32_t)(mix() & 0x7FFFFFFF);
        r5 = (int32_t)((mix() & 0x7FFFFFFF) | 1);
    }
    // the stream of a context (seed 0 continues it)
    static Ranorm of(const qp_ctx *c)
    {
        Ranorm r(0);
        r.r1 = c->rn_state[0]; r.r2 = c->rn_state[1]; r.r4 = c->rn_state[2];
        r.r5 = c->rn_state[3]; r.iflg = c->rn_state[4]; r.r0 = c->rn_r0;
        return r;
    }
    void save(qp_ctx *c) const
    {
        c->rn_state[0] = r1; c->rn_state[1] = r2; c->rn_state[2] = r4;
        c->rn_state[3] = r5; c->rn_state[4] = iflg; c->rn_r0 = r0;
    }
    double next()
    {
        const double h1l = 65531.0, h1u = 32767.0, h2l = 65525.0;
        if (iflg != 0) {
            double out = r0;
            r0 = 0.0;
            iflg = 0;
            return out;
        }
        int32_t isc = 65536, i1;
        double asc = (double)isc, bsc = asc * asc, r3, temp, rr;
        i1 = r1 - (r1 / isc) * isc;
        r3 = h1l * (double)r1 + asc * h1u * (double)i1;
        i1 = (int32_t)(r3 / bsc);
        r3 = r3 - (double)i1 * bsc;
        bsc = 0.5 * bsc;
        i1 = r2 / isc;
        isc = r2 - i1 * isc;
        rr = h1l * (double)r2 + asc * h1u * (double)isc;
        asc = 1.0 / bsc;
        isc = (int32_t)(rr * asc);
        r2 = (int32_t)(rr - (double)isc * bsc);
        r3 = r3 + ((double)isc + 2.0 * h1u * (double)i1);
        isc = (int32_t)(r3 * asc);
        r1 = (int32_t)(r3 - (double)isc * bsc);
        temp = std::sqrt(-2.0 * std::log(((double)r1 + (double)r2 * asc) * asc));
        isc = 65536;
        asc = (double)isc;
        bsc = asc * asc;
        i1 = r4 - (r4 / isc) * isc;
        r3 = h2l * (double)r4 + asc * h1u * (double)i1;
        i1 = (int32_t)(r3 / bsc);
        r3 = r3 - (double)i1 * bsc;
        bsc = 0.5 * bsc;
        i1 = r5 / isc;
        isc = r5 - i1 * isc;
        rr = h2l * (double)r5 + asc * h1u * (double)isc;
        asc = 1.0 / bsc;
        isc = (int32_t)(rr * asc);
        r5 = (int32_t)(rr - (double)isc * bsc);
        r3 = r3 + ((double)isc + 2.0 * h1u * (double)i1);
        isc = (int32_t)(r3 * asc);
        r4 = (int32_t)(r3 - (double)isc * bsc);
        rr = 6.28318530717959 * (((double)r4 + (double)r5 * asc) * asc);
        double sn, cs;
        sincos(rr, &sn, &cs);
        r0 = temp * cs;
        iflg = 1;
        return temp * sn;
    }
};

// fdist3d_000 normalisation (fdist3d_class.f03:305-335) and
// PRVDIST32_RANDOM (part3d_lib77.f:411-552) with one partition covering the
// box.  Fills h[7][n]; returns the particle count or -1 on overflow.
static int64_t load_beam(qp_ctx *c, const qp_beam_params &bp, uint64_t seed,
                         std::vector<double> (&h)[7], int64_t npmax, double *qm_out)
{
    const double dx = c->dex, dy = c->dey, dz = c->dxi;
    double qm = bp.q / std::fabs(bp.q) * std::fabs(bp.peak) *
                std::pow(2 * 3.1415926535897932, 1.5) * bp.sigma[0] * bp.sigma[1] * bp.sigma[2];
    qm = qm / dx / dy / dz;
    qm = qm / bp.np[0];
    qm = qm / bp.np[1];
    qm = qm / bp.np[2];
    *qm_out = qm;
    const double x0 = (bp.center[0] - c->box[0]) / dx;
    const double y0 = (bp.center[1] - c->box[2]) / dy;
    const double z0 = (bp.center[2] - c->box[4]) / dz;
    const double sigx = bp.sigma[0] / dx, sigy = bp.sigma[1] / dy, sigz = bp.sigma[2] / dz;
    const double vtx = bp.sigma_v[0], vty = bp.sigma_v[1], vtz = bp.sigma_v[2];
    const double vdx = 0.0, vdy = 0.0, vdz = bp.gamma;
    const double cx[3] = {bp.centroid_x[0] * dz * dz / dx, bp.centroid_x[1] * dz / dx,
                          bp.centroid_x[2] / dx};
    const double cy[3] = {bp.centroid_y[0] * dz * dz / dy, bp.centroid_y[1] * dz / dy,
                          bp.centroid_y[2] / dy};
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const double x2 = 2.0 * x0, y2 = 2.0 * y0;
    const double blx = std::fmax(x0 - 5.0 * sigx, 1.0), bly = std::fmax(y0 - 5.0 * sigy, 1.0);
    const double blz = std::fmax(z0 - 5.0 * sigz, 1.0);
    const double brx = std::fmin(x0 + 5.0 * sigx, (double)(nx - 1));
    const double bry = std::fmin(y0 + 5.0 * sigy, (double)(ny - 1));
    const double brz = std::fmin(z0 + 5.0 * sigz, (double)(nz - 1));
    Ranorm rng = seed == 0 ? Ranorm::of(c) : Ranorm(seed);
    int64_t npp = 0, npt = 1;
    int j = 0, k, l = bp.np[2];
    auto put = [&](double x, double y, double z, double px, double py, double pz) {
        npt = npp + 1;
        size_t i = (size_t)(npt - 1);
        h[0][i] = x; h[1][i] = y; h[2][i] = z;
        h[3][i] = px; h[4][i] = py; h[5][i] = pz; h[6][i] = qm;
        npp = npt;
    };
    while (j < bp.np[0]) {
        k = 0;
        while (k < bp.np[1]) {
            l = l - bp.np[2];
            while (l < bp.np[2]) {
                double tz, tx, ty;
                do { tz = z0 + sigz * rng.next(); } while (tz >= brz || tz <= blz);
                do { tx = x0 + sigx * rng.next(); } while (tx >= brx || tx <= blx);
                do { ty = y0 + sigy * rng.next(); } while (ty >= bry || ty <= bly);
                double tvx = vtx * rng.next() + vdx;
                double tvy = vty * rng.next() + vdy;
                double tvz = vtz * rng.next() + vdz;
                tvz = std::sqrt(tvz * tvz - 1 - tvx * tvx - tvy * tvy);
                double txx = 0.0, tyy = 0.0;
                if (ty >= 0.0 && ty < (double)ny && tz >= 0.0 && tz < (double)nz) {
                    if (npt < npmax) {
                        txx = -cx[0] * (tz - z0) * (tz - z0) - cx[1] * (tz - z0) - cx[2];
                        tyy = -cy[0] * (tz - z0) * (tz - z0) - cy[1] * (tz - z0) - cy[2];
                        put(tx + txx, ty + tyy, tz, tvx, tvy, tvz);
                    } else {
                        return -1;
                    }
                    if (bp.quiet) {
                        if (npt < npmax) put(x2 - tx + txx, y2 - ty + tyy, tz, -tvx, -tvy, tvz);
                        else return -1;
                    }
                }
                l = l + 1;
                if (bp.quiet) l = l + 1;
            }
            k = k + 1;
        }
        j = j + 1;
    }
    if (seed == 0) rng.save(c);
    return npp;
}

// fdist3d_002 normalisation (fdist3d_class.f03:720-745) and PRVDIST32_TWISS
// (part3d_lib77.f:2-192) with one partition covering the box.  Fills
// h[7][n]; returns the particle count or -1 on overflow.
static int64_t load_beam_twiss(qp_ctx *c, const qp_beam_twiss_params &bp, uint64_t seed,
                               std::vector<double> (&h)[7], int64_t npmax, double *qm_out)
{
    const double dx = c->dex, dy = c->dey, dz = c->dxi;
    const double gamma = bp.gamma;
    // charge per particle from the physical sizes (:729-736)
    const double sx_phys = std::sqrt(bp.beta[0] * bp.emittance[0] / gamma);
    const double sy_phys = std::sqrt(bp.beta[1] * bp.emittance[1] / gamma);
    double qm = bp.q / std::fabs(bp.q) * std::fabs(bp.peak) *
                std::pow(2 * 3.1415926535897932, 1.5) * sx_phys * sy_phys * bp.sigma_z;
    qm = qm / dx / dy / dz;
    qm = qm / bp.np[0];
    qm = qm / bp.np[1];
    qm = qm / bp.np[2];
    *qm_out = qm;
    // grid units (:737-745)
    const double x0 = (bp.center[0] - c->box[0]) / dx;
    const double y0 = (bp.center[1] - c->box[2]) / dy;
    const double z0 = (bp.center[2] - c->box[4]) / dz;
    const double sigz = bp.sigma_z / dz, vtz = bp.sigma_vz, vdz = gamma;
    const double alpha_x = bp.alpha[0], alpha_y = bp.alpha[1];
    const double beta_x = bp.beta[0] / dx, beta_y = bp.beta[1] / dy;
    const double emt_x = bp.emittance[0] / dx, emt_y = bp.emittance[1] / dy;
    // PRVDIST32_TWISS (:62-110)
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const double x2 = 2.0 * x0, y2 = 2.0 * y0;
    const double sigx = std::sqrt(beta_x * emt_x / gamma);
    const double sigy = std::sqrt(beta_y * emt_y / gamma);
    const double vtx = emt_x / sigx, vty = emt_y / sigy;
    const double blx = std::fmax(x0 - 3.0 * sigx, 1.0), bly = std::fmax(y0 - 3.0 * sigy, 1.0);
    const double blz = std::fmax(z0 - 3.0 * sigz, 1.0);
    const double brx = std::fmin(x0 + 3.0 * sigx, (double)(nx - 1));
    const double bry = std::fmin(y0 + 3.0 * sigy, (double)(ny - 1));
    const double brz = std::fmin(z0 + 3.0 * sigz, (double)(nz - 1));
    Ranorm rng = seed == 0 ? Ranorm::of(c) : Ranorm(seed);
    int64_t npp = 0, npt = 1;
    int j = 0, k, l = bp.np[2];
    auto put = [&](double x, double y, double z, double px, double py, double pz) {
        npt = npp + 1;
        size_t i = (size_t)(npt - 1);
        h[0][i] = x; h[1][i] = y; h[2][i] = z;
        h[3][i] = px; h[4][i] = py; h[5][i] = pz; h[6][i] = qm;
        npp = npt;
    };
    while (j < bp.np[0]) {
        k = 0;
        while (k < bp.np[1]) {
            l = l - bp.np[2];
            while (l < bp.np[2]) {
                double tz, tx, ty, rx, ry;
                do { tz = z0 + sigz * rng.next(); } while (tz >= brz || tz <= blz);
                // x then y, each within its 3-sigma window, redrawn together
                // while the pair lies outside radius 3 (:98-106)
                for (;;) {
                    do { rx = rng.next(); tx = x0 + sigx * rx; } while (tx >= brx || tx <= blx);
                    do { ry = rng.next(); ty = y0 + sigy * ry; } while (ty >= bry || ty <= bly);
                    if (!(rx * rx + ry * ry > 9.0)) break;
                }
                double tvx = vtx * rng.next();
                double tvy = vty * rng.next();
                const double tvz = vtz * rng.next() + vdz;
                tvx = tvx - gamma * alpha_x / beta_x * (tx - x0);
                tvy = tvy - gamma * alpha_y / beta_y * (ty - y0);
                if (ty >= 0.0 && ty < (double)ny && tz >= 0.0 && tz < (double)nz) {
                    if (npt < npmax) put(tx, ty, tz, tvx, tvy, tvz);
                    else return -1;
                    if (bp.quiet) {
                        if (npt < npmax) put(x2 - tx, y2 - ty, tz, -tvx, -tvy, tvz);
                        else return -1;
                    }
                }
                l = l + 1;
                if (bp.quiet) l = l + 1;
            }
            k = k + 1;
        }
        j = j + 1;
    }
    if (seed == 0) rng.save(c);
    return npp;
}

// PGPOST32L (part3d_lib77.f:554-628), int64 fixed-point accumulation.
// Particles of one wavefront that sit in the same cell on neighbouring lanes
// (the push leaves the beam sorted by cell) are summed first, in int64 across
// the wavefront, and the first lane of each such run adds the sums with one
// atomic per node. Integer sums make the charge the same, bit for bit, as one
// atomic per particle and node, whatever the particle order.
__global__ __launch_bounds__(256) void k_beam_deposit(
    int64_t np, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ q, unsigned long long *q3, int ld,
    size_t pl, int nx, int ny, int nz)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    constexpr unsigned long long NONE = ~0ull;
    unsigned long long cell = NONE;
    long long w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (j < np) {
        const double px = x[j], py = y[j], pz = z[j], qm = q[j];
        // a particle outside the grid has no nodes to deposit to
        if (px >= 0.0 && px < (double)nx && py >= 0.0 && py < (double)ny && pz >= 0.0 &&
            pz < (double)nz) {
            int nn = (int)px, mm = (int)py, ll = (int)pz;
            double dxp = qm * (px - (double)nn);
            double dyp = py - (double)mm;
            double dzp = pz - (double)ll;
            double amx = qm - dxp;
            double amy = 1. - dyp;
            double dx1 = dxp * dyp;
            dyp = amx * dyp;
            amx = amx * amy;
            double amz = 1. - dzp;
            amy = dxp * amy;
            cell = (size_t)ll * pl + (size_t)mm * ld + nn;
            const double f[8] = {amx * amz, amy * amz, dyp * amz, dx1 * amz,
                                 amx * dzp, amy * dzp, dyp * dzp, dx1 * dzp};
#pragma unroll
            for (int k = 0; k < 8; ++k) w[k] = __double2ll_rn(f[k] * BEAM_FIX);
        }
    }
    // a run of equal cells starts where the lane before holds another cell
    const unsigned long long prev = __shfl_up(cell, 1, 64);
    const bool head = lane == 0 || prev != cell;
    const unsigned long long heads = __ballot(head);
    if (heads != ~0ull) {
        // run number of each lane, then suffix sums inside each run: after
        // the step of stride d a lane holds its run's sum over [lane, lane + 2d)
        const int run = __popcll(heads & ((2ull << lane) - 1ull));
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int orun = __shfl_down(run, d, 64);
            const bool take = lane + d < 64 && orun == run;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const long long o = __shfl_down(w[k], d, 64);
                if (take) w[k] += o;
            }
        }
    }
    if (!head || cell == NONE) return;
    unsigned long long *b = q3 + cell;
    const size_t o[8] = {0, 1, (size_t)ld, (size_t)ld + 1, pl, pl + 1, pl + ld, pl + ld + 1};
#pragma unroll
    for (int k = 0; k < 8; ++k) atomicAdd(b + o[k], (unsigned long long)w[k]);
}

// qdpcopy / copyadd of one xi plane (ufield3d_class.f03:460-490) from the
// fixed-point charge
// (assign: q2 is taken as zero, as after asc(0), so it is not read)
template <typename R>
__global__ void k_beam_slice_add(R *q2, const long long *q3, size_t n, int assign,
                                 const int *jdev, int nz)
{
    // jdev: plane (slice - 1) = j read from device memory (captured slices),
    // 1 <= j <= nz (the host checks it before a replay)
    if (jdev) {
        const int j = *jdev;
        if (j < 1 || j > nz) return;
        q3 += (size_t)j * n;
    }
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const R v = (R)((double)q3[i] * (1.0 / BEAM_FIX));
        q2[i] = assign ? R(0) + v : q2[i] + v;
    }
}

// Sort key of a beam particle: its cell's linear index (z, then y, then x)
// shifted right by `shift` bits so that it fits 32 bits; dropped particles
// carry `drop`, above every cell's key.
struct CellKey {
    int nx, ny, shift;
    unsigned drop;
    __device__ unsigned of(double X, double Y, double Z) const
    {
        const unsigned long long c =
            ((unsigned long long)(int)Z * (unsigned)ny + (unsigned)(int)Y) * (unsigned)nx +
            (unsigned)(int)X;
        return (unsigned)(c >> shift);
    }
};

// PGBPUSH32L_QP (part3d_lib77.f:630-786) with ipbc = 1 and cofd = 0; flags
// the particles that leave [1, n-1) in any direction, and writes each
// particle's sort key (its new cell) with its index beside it.
template <typename R>
__global__ void k_beam_push(int64_t np, double *x, double *y, double *z, double *px, double *py,
                            double *pz, const R *__restrict__ e3, const R *__restrict__ b3,
                            size_t vol, int ld, size_t pl, double qtmh, double dtcx, double dtcz,
                            double rx, double ry, double rz, int *alive, CellKey ck,
                            unsigned *key, unsigned *idx)
{
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= np) return;
    idx[j] = (unsigned)j;
    double X = x[j], Y = y[j], Z = z[j];
    // outside the grid nothing can be gathered: dropped without touching the
    // fields (the loader and the push keep every particle in [1, n-1))
    if (!(X >= 0.0 && X < rx + 1.0 && Y >= 0.0 && Y < ry + 1.0 && Z >= 0.0 && Z < rz + 1.0)) {
        alive[j] = 0;
        key[j] = ck.drop;
        return;
    }
    int nn = (int)X, mm = (int)Y, ll = (int)Z;
    double dxp = X - (double)nn, dyp = Y - (double)mm, dzp = Z - (double)ll;
    double amx = 1. - dxp, amy = 1. - dyp;
    double dx1 = dxp * dyp;
    dyp = amx * dyp;
    amx = amx * amy;
    double amz = 1. - dzp;
    amy = dxp * amy;
    const size_t o = (size_t)ll * pl + (size_t)mm * ld + nn;
    auto g3 = [&](const R *f) -> double {
        const R *p = f + o;
        return amz * (amx * (double)p[0] + amy * (double)p[1] + dyp * (double)p[ld] +
                      dx1 * (double)p[ld + 1]) +
               dzp * (amx * (double)p[pl] + amy * (double)p[pl + 1] + dyp * (double)p[pl + ld] +
                      dx1 * (double)p[pl + ld + 1]);
    };
    double dx = g3(e3), dy = g3(e3 + vol), dz = g3(e3 + 2 * vol);
    double ox = g3(b3), oy = g3(b3 + vol);
    dx = dx - oy;
    dy = dy + ox;
    double delx = qtmh * dx, dely = qtmh * dy, delz = qtmh * dz;
    dx = px[j] + delx;
    dy = py[j] + dely;
    dz = pz[j] + delz;
    double p2t = dx * dx + dy * dy;
    double p2 = p2t + dz * dz;
    double ngamma = sqrt(1.0 + p2);
    double dtg = dtcx / ngamma;
    double omvz = (1.0 + p2t) / (dz * (dz + ngamma));
    px[j] = dx;
    py[j] = dy;
    pz[j] = dz;
    double nxp = X + dx * dtg, nyp = Y + dy * dtg, nzp = Z + omvz * dtcz;
    bool live = !((nxp < 1.0) || (nxp >= rx) || (nyp < 1.0) || (nyp >= ry) || (nzp < 1.0) ||
                  (nzp >= rz));
    if (live) { x[j] = nxp; y[j] = nyp; z[j] = nzp; }
    alive[j] = live ? 1 : 0;
    key[j] = live ? ck.of(nxp, nyp, nzp) : ck.drop;
}

struct Ptr7 {
    double *p[7];
};

// survivor i of the cell-sorted beam is particle perm[i] (i < *nlive)
__global__ void k_permute(int64_t np, const int *nlive, const unsigned *perm, Ptr7 src, Ptr7 dst)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= np || i >= *nlive) return;
    const unsigned s = perm[i];
#pragma unroll
    for (int k = 0; k < 7; ++k) dst.p[k][i] = src.p[k][s];
}

__global__ void k_beam_moments(int64_t np, const double *px, const double *py, const double *pz,
                               const double *q, double *partial)
{
    __shared__ double s[2][256];
    double g = 0.0, qq = 0.0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < np;
         j += (int64_t)gridDim.x * blockDim.x) {
        g += sqrt(1.0 + px[j] * px[j] + py[j] * py[j] + pz[j] * pz[j]);
        qq += q[j];
    }
    s[0][threadIdx.x] = g;
    s[1][threadIdx.x] = qq;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            s[0][threadIdx.x] += s[0][threadIdx.x + st];
            s[1][threadIdx.x] += s[1][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s[0][0];
        partial[2 * blockIdx.x + 1] = s[1][0];
    }
}

}  // namespace qp

using namespace qp;

#define BCHECK(b) QP_REQUIRE((b) && (b)->ctx, QP_ERR_ARG, "%s: null beam", __func__)

static int beam_alloc(qp_beam *b, int64_t n)
{
    qp_ctx *c = b->ctx;
    for (int k = 0; k < 7; ++k) {
        int rc = b->a[k].alloc((size_t)n * 8);
        if (!rc) rc = b->b[k].alloc((size_t)n * 8);
        if (rc) return rc;
    }
    int rc = b->alive.alloc((size_t)(n + 1) * 4);
    if (!rc) rc = b->pos.alloc((size_t)(n + 1) * 4);
    for (int k = 0; k < 2 && !rc; ++k) {
        rc = b->key[k].alloc((size_t)(n + 1) * 4);
        if (!rc) rc = b->perm[k].alloc((size_t)(n + 1) * 4);
    }
    if (!rc) rc = b->moments.alloc(sizeof(double) * 2 * 1024 + 64);
    if (!rc) rc = b->q3.alloc((size_t)(c->nz + 1) * c->rplane * 8);
    if (rc) return rc;
    // the smallest shift that puts every cell's key below the drop key < 2^32
    const unsigned long long ncell = (unsigned long long)c->nx * c->ny * c->nz;
    b->key_shift = 0;
    while (((ncell - 1) >> b->key_shift) >= 0xFFFFFFFFull) ++b->key_shift;
    b->key_drop = (unsigned)(((ncell - 1) >> b->key_shift) + 1);
    b->key_bits = 1;
    while (b->key_bits < 32 && (b->key_drop >> b->key_bits) != 0) ++b->key_bits;
    size_t scan = 0, sort = 0;
    QP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan, b->alive.as<int>(), b->pos.as<int>(),
                                            (int)(n + 1), c->stream));
    QP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort, b->key[0].as<unsigned>(),
                                              b->key[1].as<unsigned>(), b->perm[0].as<unsigned>(),
                                              b->perm[1].as<unsigned>(), (int)n, 0, b->key_bits,
                                              c->stream));
    b->scan_bytes = std::max(scan, sort);
    return b->scan_tmp.alloc(b->scan_bytes + 256);
}

static int beam_set(qp_beam *b, std::vector<double> (&h)[7], int64_t n)
{
    qp_ctx *c = b->ctx;
    for (int k = 0; k < 7; ++k)
        QP_HIP(hipMemcpyAsync(b->a[k].p, h[k].data(), (size_t)n * 8, hipMemcpyHostToDevice,
                              c->stream));
    QP_HIP(hipStreamSynchronize(c->stream));
    b->np = n;
    return QP_OK;
}

extern "C" {

int qp_beam_create(qp_ctx *c, const qp_beam_params *p, double dt, uint64_t seed, qp_beam **out)
{
    QP_REQUIRE(c && p && out, QP_ERR_ARG, "qp_beam_create: null argument");
    const bool empty = p->np[0] == 0 && p->np[1] == 0 && p->np[2] == 0;
    QP_REQUIRE(p->q != 0.0 && p->m != 0.0 &&
                   ((p->np[0] > 0 && p->np[1] > 0 && p->np[2] > 0) || (empty && p->npmax > 0)),
               QP_ERR_ARG, "qp_beam_create: bad beam parameters");
    int64_t want = (int64_t)p->np[0] * p->np[1] * p->np[2];
    int64_t npmax = p->npmax > 0 ? p->npmax : want + 2;
    QP_REQUIRE(npmax < ((int64_t)1 << 31), QP_ERR_UNSUPPORTED, "qp_beam_create: npmax too large");
    qp_beam *b = new qp_beam();
    b->ctx = c;
    b->p = *p;
    b->qbm = p->q / p->m;
    b->dt = dt;
    b->npmax = npmax;
    std::vector<double> h[7];
    for (auto &v : h) v.resize((size_t)npmax);
    double qm = 0.0;
    int64_t n = empty ? 0 : load_beam(c, *p, seed, h, npmax, &qm);
    int rc = n < 0 ? (set_error("qp_beam_create: npmax %lld too small", (long long)npmax),
                      QP_ERR_ARG)
                   : beam_alloc(b, npmax);
    if (!rc) rc = beam_set(b, h, n);
    if (rc) { qp_beam_destroy(b); return rc; }
    *out = b;
    return QP_OK;
}

// A beam whose particles the host has built (profiles 3 and 100): the
// record of profile 0's parameter block that the push and diagnostics use
static int beam_from_host(qp_ctx *c, double q, double m, int evolution, int64_t npmax,
                          std::vector<double> (&h)[7], int64_t n, double dt, qp_beam **out)
{
    QP_REQUIRE(npmax < ((int64_t)1 << 31), QP_ERR_UNSUPPORTED, "beam: npmax too large");
    QP_REQUIRE(n <= npmax, QP_ERR_ARG, "beam: %lld particles exceed npmax %lld", (long long)n,
               (long long)npmax);
    qp_beam_params bp{};
    bp.q = q;
    bp.m = m;
    bp.evolution = evolution;
    bp.npmax = npmax;
    qp_beam *b = new qp_beam();
    b->ctx = c;
    b->p = bp;
    b->qbm = q / m;
    b->dt = dt;
    b->npmax = npmax;
    int rc = beam_alloc(b, npmax);
    if (!rc) rc = beam_set(b, h, n);
    if (rc) { qp_beam_destroy(b); return rc; }
    *out = b;
    return QP_OK;
}

// init_fdist3d_003 + dist3d_003 (fdist3d_class.f03:823-974): np records
// "z x y vz vx vy q" of a text file (list-directed: blanks or commas between
// values, Fortran D exponents), positions in units of cwp =
// 5.32150254e9/sqrt(n0) converted to grid units, x and y about the centre,
// z = -z/cwp/dz; kept if 1 <= x <= nx-1, 0 <= y < ny and 0 <= z < nz (one
// partition); stored as x, y, z, vx, vy, vz, q.  A file with fewer than np
// records fails with the reference's message.
int qp_beam_create_file(qp_ctx *c, const qp_beam_file_params *p, double dt, qp_beam **out)
{
    QP_REQUIRE(c && p && out && p->file_name, QP_ERR_ARG, "qp_beam_create_file: null argument");
    QP_REQUIRE(p->np >= 0 && p->n0 > 0.0 && p->q != 0.0 && p->m != 0.0, QP_ERR_ARG,
               "qp_beam_create_file: bad beam parameters");
    FILE *f = std::fopen(p->file_name, "r");
    QP_REQUIRE(f, QP_ERR_ARG, "dist3d_003: cannot open %s", p->file_name);
    const double cwp = 5.32150254 * 1e9 / std::sqrt(p->n0);
    const double dx = c->dex, dy = c->dey, dz = c->dxi;
    const double x0 = (p->center[0] - c->box[0]) / dx, y0 = (p->center[1] - c->box[2]) / dy;
    const int64_t npmax = p->npmax > 0 ? p->npmax : (int64_t)p->np + 2;
    std::vector<double> h[7];
    for (auto &v : h) v.resize((size_t)npmax);
    int64_t n = 0;
    bool short_file = false;
    std::string tok;
    auto next = [&](double &v) -> bool {
        tok.clear();
        int ch;
        while ((ch = std::fgetc(f)) != EOF && (std::isspace(ch) || ch == ',')) {}
        if (ch == EOF) return false;
        do {
            tok.push_back((char)(ch == 'd' || ch == 'D' ? 'e' : ch));
        } while ((ch = std::fgetc(f)) != EOF && !std::isspace(ch) && ch != ',');
        char *end = nullptr;
        v = std::strtod(tok.c_str(), &end);
        return end && *end == '\0';
    };
    for (int i = 0; i < p->np && !short_file; ++i) {
        double t[7];
        for (int k = 0; k < 7; ++k)
            if (!next(t[k])) { short_file = true; break; }
        if (short_file) break;
        const double tx = t[1] / cwp / dx + x0, ty = t[2] / cwp / dy + y0, tz = -t[0] / cwp / dz;
        if (tx >= 1 && tx <= (double)(c->nx - 1) && ty >= 0.0 && ty < (double)c->ny && tz >= 0.0 &&
            tz < (double)c->nz) {
            if (n >= npmax) { std::fclose(f); QP_REQUIRE(false, QP_ERR_ARG, "dist3d_003: npmax %lld too small", (long long)npmax); }
            h[0][n] = tx; h[1][n] = ty; h[2][n] = tz;
            h[3][n] = t[4]; h[4][n] = t[5]; h[5][n] = t[3]; h[6][n] = t[6];
            ++n;
        }
    }
    std::fclose(f);
    QP_REQUIRE(!short_file, QP_ERR_ARG, "dist3d_003: Read beam file error");
    return beam_from_host(c, p->q, p->m, p->evolution, npmax, h, n, dt, out);
}

// init_fdist3d_100 + dist3d_100 (fdist3d_class.f03:976-1166): a quiet-start
// lattice over the range box (clipped to the grid), tx = (ix + 0.5)/xppc +
// i - 1 for i = int(lb)..int(ub)-1 (likewise y, z), momentum (0, 0, gamma),
// charge qm np exp(-((tz - bcz)^2 + (|r - bc| - r1)^2) / r2^2 / 2) with
// qm = sign(q)/(xppc yppc zppc), r1 and r2 in x-cells
int qp_beam_create_torus(qp_ctx *c, const qp_beam_torus_params *p, double dt, qp_beam **out)
{
    QP_REQUIRE(c && p && out, QP_ERR_ARG, "qp_beam_create_torus: null argument");
    QP_REQUIRE(p->ppc[0] > 0 && p->ppc[1] > 0 && p->ppc[2] > 0 && p->q != 0.0 && p->m != 0.0 &&
                   p->r2 != 0.0,
               QP_ERR_ARG, "qp_beam_create_torus: bad beam parameters");
    const double dx = c->dex, dy = c->dey, dz = c->dxi;
    const double bcx = (p->center[0] - c->box[0]) / dx, bcy = (p->center[1] - c->box[2]) / dy;
    const double bcz = (p->center[2] - c->box[4]) / dz;
    double lb[3] = {(p->range_x[0] - c->box[0]) / dx, (p->range_y[0] - c->box[2]) / dy,
                    (p->range_z[0] - c->box[4]) / dz};
    double ub[3] = {(p->range_x[1] - c->box[0]) / dx, (p->range_y[1] - c->box[2]) / dy,
                    (p->range_z[1] - c->box[4]) / dz};
    const double r1 = p->r1 / dx, r2 = p->r2 / dx;
    const double qm = p->q / std::fabs(p->q) / p->ppc[0] / p->ppc[1] / p->ppc[2];
    const int nd[3] = {c->nx, c->ny, c->nz};
    std::vector<double> h[7];
    int64_t n = 0;
    bool outside = false;
    for (int d = 0; d < 3; ++d) {
        if (d > 0 && (lb[d] > (double)nd[d] || ub[d] < 0.0)) outside = true;   // :1108-1124
        lb[d] = std::fmax(lb[d], 0.0);
        ub[d] = std::fmin(ub[d], (double)nd[d]);
    }
    const int64_t cap = outside ? 0 : (int64_t)std::max(0, (int)ub[0] - (int)lb[0]) *
                                          std::max(0, (int)ub[1] - (int)lb[1]) *
                                          std::max(0, (int)ub[2] - (int)lb[2]) * p->ppc[0] *
                                          p->ppc[1] * p->ppc[2];
    const int64_t npmax = p->npmax > 0 ? p->npmax : cap + 2;
    QP_REQUIRE(cap <= npmax, QP_ERR_ARG, "dist3d_100: %lld particles exceed npmax %lld",
               (long long)cap, (long long)npmax);
    for (auto &v : h) v.resize((size_t)npmax);
    if (!outside)
        for (int i = (int)lb[0]; i <= (int)ub[0] - 1; ++i)
            for (int j = (int)lb[1]; j <= (int)ub[1] - 1; ++j)
                for (int k = (int)lb[2]; k <= (int)ub[2] - 1; ++k)
                    for (int ix = 0; ix < p->ppc[0]; ++ix)
                        for (int iy = 0; iy < p->ppc[1]; ++iy)
                            for (int iz = 0; iz < p->ppc[2]; ++iz) {
                                const double tx = (ix + 0.5) / p->ppc[0] + i - 1;
                                const double ty = (iy + 0.5) / p->ppc[1] + j - 1;
                                const double tz = (iz + 0.5) / p->ppc[2] + k - 1;
                                const double a1 = std::fabs(tz - bcz);
                                const double a2 = std::sqrt((tx - bcx) * (tx - bcx) +
                                                            (ty - bcy) * (ty - bcy));
                                const double a3 = std::fabs(a2 - r1);
                                const double a4 = a1 * a1 + a3 * a3;
                                h[0][n] = tx; h[1][n] = ty; h[2][n] = tz;
                                h[3][n] = 0.0; h[4][n] = 0.0; h[5][n] = p->gamma;
                                h[6][n] = qm * p->np * std::exp(-a4 / (r2 * r2) * 0.5);
                                ++n;
                            }
    return beam_from_host(c, p->q, p->m, p->evolution, npmax, h, n, dt, out);
}

int qp_beam_create_twiss(qp_ctx *c, const qp_beam_twiss_params *p, double dt, uint64_t seed,
                         qp_beam **out)
{
    QP_REQUIRE(c && p && out, QP_ERR_ARG, "qp_beam_create_twiss: null argument");
    QP_REQUIRE(p->q != 0.0 && p->m != 0.0 && p->gamma > 0.0 && p->beta[0] > 0.0 &&
                   p->beta[1] > 0.0 && p->emittance[0] > 0.0 && p->emittance[1] > 0.0 &&
                   p->sigma_z > 0.0 && p->np[0] > 0 && p->np[1] > 0 && p->np[2] > 0,
               QP_ERR_ARG, "qp_beam_create_twiss: bad beam parameters");
    int64_t want = (int64_t)p->np[0] * p->np[1] * p->np[2];
    int64_t npmax = p->npmax > 0 ? p->npmax : want + 2;
    QP_REQUIRE(npmax < ((int64_t)1 << 31), QP_ERR_UNSUPPORTED,
               "qp_beam_create_twiss: npmax too large");
    // the beam record keeps the profile-0 parameter block for push and diagnostics
    qp_beam_params bp{};
    for (int d = 0; d < 3; ++d) { bp.center[d] = p->center[d]; bp.np[d] = p->np[d]; }
    bp.sigma[0] = std::sqrt(p->beta[0] * p->emittance[0] / p->gamma);
    bp.sigma[1] = std::sqrt(p->beta[1] * p->emittance[1] / p->gamma);
    bp.sigma[2] = p->sigma_z;
    bp.sigma_v[2] = p->sigma_vz;
    bp.gamma = p->gamma; bp.peak = p->peak; bp.q = p->q; bp.m = p->m;
    bp.quiet = p->quiet; bp.evolution = p->evolution; bp.npmax = p->npmax;
    qp_beam *b = new qp_beam();
    b->ctx = c;
    b->p = bp;
    b->qbm = p->q / p->m;
    b->dt = dt;
    b->npmax = npmax;
    std::vector<double> h[7];
    for (auto &v : h) v.resize((size_t)npmax);
    double qm = 0.0;
    int64_t n = load_beam_twiss(c, *p, seed, h, npmax, &qm);
    int rc = n < 0 ? (set_error("qp_beam_create_twiss: npmax %lld too small", (long long)npmax),
                      QP_ERR_ARG)
                   : beam_alloc(b, npmax);
    if (!rc) rc = beam_set(b, h, n);
    if (rc) { qp_beam_destroy(b); return rc; }
    *out = b;
    return QP_OK;
}

int qp_beam_destroy(qp_beam *b)
{
    if (!b) return QP_OK;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    for (int k = 0; k < 7; ++k) { b->a[k].release(); b->b[k].release(); }
    b->alive.release(); b->pos.release(); b->scan_tmp.release(); b->moments.release();
    for (int k = 0; k < 2; ++k) { b->key[k].release(); b->perm[k].release(); }
    b->q3.release();
    delete b;
    return QP_OK;
}

// qdeposit_beam3d (beam3d_class.f03:111-134): q = 0, deposit; the one-rank
// guard operations are no-ops (ufield3d_lib77.f:230-232, 280-282)
int qp_beam_qdp(qp_beam *b)
{
    BCHECK(b);
    qp_ctx *c = b->ctx;
    QP_HIP(hipMemsetAsync(b->q3.p, 0, b->q3.bytes, c->stream));
    if (b->np > 0) {
        {
        ProfScope _ps(c, QP_PROF_BEAM_DEPOSIT, 4.0 * 8.0 * (double)b->np);
        hipLaunchKernelGGL(k_beam_deposit, dim3((unsigned)((b->np + 255) / 256)), dim3(256), 0,
                           c->stream, b->np, b->a[0].as<double>(), b->a[1].as<double>(),
                           b->a[2].as<double>(), b->a[6].as<double>(),
                           b->q3.as<unsigned long long>(), c->ldr, c->rplane, c->nx, c->ny,
                           c->nz);
        }
        QP_LAUNCH_CHECK();
    }
    return QP_OK;
}

// qdpcopy_beam3d (beam3d_class.f03:136-153): q += beam q at plane `slice`,
// or with jdev at plane j + 1, j read from device memory when the copy runs
static int slice_copy(qp_beam *b, qp_field2d *q, int slice, int assign, const int *jdev = nullptr)
{
    BCHECK(b);
    QP_REQUIRE(q && q->dim >= 1, QP_ERR_ARG, "qdpcopy: bad field");
    qp_ctx *c = b->ctx;
    QP_REQUIRE(jdev || (slice >= 1 && slice <= c->nz + 1), QP_ERR_ARG, "qdpcopy: slice %d", slice);
    const long long *src = b->q3.as<long long>() + (jdev ? 0 : (size_t)(slice - 1) * c->rplane);
    size_t n = c->rplane;
    dim3 grid((unsigned)((n + 255) / 256));
    {
    ProfScope _ps(c, QP_PROF_ELEMENTWISE, (assign ? 12.0 : 16.0) * n);
    if (c->prec == QP_FP64)
        hipLaunchKernelGGL(k_beam_slice_add<double>, grid, dim3(256), 0, c->stream,
                           q->rs.as<double>(), src, n, assign, jdev, c->nz);
    else
        hipLaunchKernelGGL(k_beam_slice_add<float>, grid, dim3(256), 0, c->stream,
                           q->rs.as<float>(), src, n, assign, jdev, c->nz);
    }
    QP_LAUNCH_CHECK();
    q->state = 0;
    q->gcells = 1;
    return QP_OK;
}

int qp_beam_qdpcopy(qp_beam *b, qp_field2d *q, int slice) { return slice_copy(b, q, slice, 0); }

}  // extern "C"

namespace qp {
// the slice loop's first beam charge copy: qb = 0 + beam slice (asc then copyadd)
int beam_slice_assign(qp_beam *b, qp_field2d *q, int slice, const int *jdev)
{
    return slice_copy(b, q, slice, 1, jdev);
}
// the copies of the other beams (qdpcopy), with the slice index on the device
int beam_slice_add_dev(qp_beam *b, qp_field2d *q, const int *jdev)
{
    return slice_copy(b, q, 0, 0, jdev);
}
}  // namespace qp

extern "C" {

int qp_beam_push(qp_beam *b, const qp_field3d *ef, const qp_field3d *bf, double dex, double dez)
{
    BCHECK(b);
    QP_REQUIRE(ef && bf && ef->dim == 3 && bf->dim >= 2, QP_ERR_ARG, "beam push: bad fields");
    if (!b->p.evolution) return QP_OK;       // beam3d_class.f03:170-173
    qp_ctx *c = b->ctx;
    if (b->np == 0) return QP_OK;
    const int64_t np = b->np;
    const size_t vol = (size_t)(c->nz + 1) * c->rplane;
    dim3 grid((unsigned)((np + 255) / 256));
    const double qtmh = b->qbm * b->dt, dtcx = b->dt / dex, dtcz = b->dt / dez;
    const CellKey ck{c->nx, c->ny, b->key_shift, b->key_drop};
    {
    ProfScope _ps(c, QP_PROF_BEAM_PUSH, (12.0 * 8.0 + 12.0) * (double)np);   // 6 read, 6 written, flag, key, index
    if (c->prec == QP_FP64)
        hipLaunchKernelGGL(k_beam_push<double>, grid, dim3(256), 0, c->stream, np,
                           b->a[0].as<double>(), b->a[1].as<double>(), b->a[2].as<double>(),
                           b->a[3].as<double>(), b->a[4].as<double>(), b->a[5].as<double>(),
                           ef->buf.as<double>(), bf->buf.as<double>(), vol, c->ldr, c->rplane,
                           qtmh, dtcx, dtcz, (double)(c->nx - 1), (double)(c->ny - 1),
                           (double)(c->nz - 1), b->alive.as<int>(), ck, b->key[0].as<unsigned>(),
                           b->perm[0].as<unsigned>());
    else
        hipLaunchKernelGGL(k_beam_push<float>, grid, dim3(256), 0, c->stream, np,
                           b->a[0].as<double>(), b->a[1].as<double>(), b->a[2].as<double>(),
                           b->a[3].as<double>(), b->a[4].as<double>(), b->a[5].as<double>(),
                           ef->buf.as<float>(), bf->buf.as<float>(), vol, c->ldr, c->rplane,
                           qtmh, dtcx, dtcz, (double)(c->nx - 1), (double)(c->ny - 1),
                           (double)(c->nz - 1), b->alive.as<int>(), ck, b->key[0].as<unsigned>(),
                           b->perm[0].as<unsigned>());
    }
    QP_LAUNCH_CHECK();
    // The survivors, sorted by cell (a stable sort: equal cells keep their
    // order; the dropped particles' key sorts them last). The reference
    // instead moves the last particle into each hole (part3d_lib77.f:748-776);
    // the survivor set is the same, and the order is the one the gathers of
    // the next push and the atomics of the next charge deposit want.
    QP_HIP(hipMemsetAsync(b->alive.as<int>() + np, 0, sizeof(int), c->stream));
    size_t tb = b->scan_bytes;
    QP_HIP(hipcub::DeviceScan::ExclusiveSum(b->scan_tmp.p, tb, b->alive.as<int>(),
                                            b->pos.as<int>(), (int)(np + 1), c->stream));
    int live = 0;
    QP_HIP(hipMemcpyAsync(&live, b->pos.as<int>() + np, sizeof(int), hipMemcpyDeviceToHost,
                          c->stream));
    {
    ProfScope _ps(c, QP_PROF_BEAM_PUSH, 16.0 * (double)np);   // keys and indices read and written once
    size_t need = 0;
    QP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, need, b->key[0].as<unsigned>(),
                                              b->key[1].as<unsigned>(), b->perm[0].as<unsigned>(),
                                              b->perm[1].as<unsigned>(), (int)np, 0, b->key_bits,
                                              c->stream));
    QP_REQUIRE(need <= b->scan_bytes, QP_ERR_ALLOC, "beam push: sort needs %zu bytes, %zu held",
               need, b->scan_bytes);
    QP_HIP(hipcub::DeviceRadixSort::SortPairs(b->scan_tmp.p, need, b->key[0].as<unsigned>(),
                                              b->key[1].as<unsigned>(), b->perm[0].as<unsigned>(),
                                              b->perm[1].as<unsigned>(), (int)np, 0, b->key_bits,
                                              c->stream));
    }
    Ptr7 src, dst;
    for (int k = 0; k < 7; ++k) { src.p[k] = b->a[k].as<double>(); dst.p[k] = b->b[k].as<double>(); }
    {
    ProfScope _ps(c, QP_PROF_BEAM_PUSH, (7.0 * 8.0 * 2.0 + 4.0) * (double)np);
    hipLaunchKernelGGL(k_permute, grid, dim3(256), 0, c->stream, np, b->pos.as<int>() + np,
                       b->perm[1].as<unsigned>(), src, dst);
    }
    QP_LAUNCH_CHECK();
    QP_HIP(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 7; ++k) std::swap(b->a[k], b->b[k]);
    b->np = live;
    return QP_OK;
}

int qp_beam_count(const qp_beam *b, int64_t *np)
{
    BCHECK(b);
    *np = b->np;
    return QP_OK;
}

int qp_beam_download(const qp_beam *b, double *part, int64_t ld)
{
    BCHECK(b);
    QP_REQUIRE(part && ld >= b->np, QP_ERR_ARG, "qp_beam_download: bad buffer");
    qp_ctx *c = b->ctx;
    std::vector<double> h((size_t)b->np);
    for (int k = 0; k < 7; ++k) {
        QP_HIP(hipMemcpyAsync(h.data(), b->a[k].p, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
        QP_HIP(hipStreamSynchronize(c->stream));
        for (int64_t i = 0; i < b->np; ++i) part[k + 7 * i] = h[i];
    }
    return QP_OK;
}

int qp_beam_upload(qp_beam *b, const double *part, int64_t np, int64_t ld)
{
    BCHECK(b);
    QP_REQUIRE(part && np >= 0 && np <= b->npmax && ld >= np, QP_ERR_ARG,
               "qp_beam_upload: bad buffer");
    std::vector<double> h[7];
    for (int k = 0; k < 7; ++k) {
        h[k].resize((size_t)(np > 0 ? np : 1));
        for (int64_t i = 0; i < np; ++i) h[k][i] = part[k + 7 * i];
    }
    return beam_set(b, h, np);
}

int qp_beam_charge_download(const qp_beam *b, double *host, int ld1, int ld2, int ld3)
{
    BCHECK(b);
    qp_ctx *c = b->ctx;
    QP_REQUIRE(host && ld1 >= c->nx + 1 && ld2 >= c->ny + 1 && ld3 >= c->nz + 1, QP_ERR_ARG,
               "qp_beam_charge_download: bad buffer");
    std::vector<long long> h(b->q3.bytes / 8);
    QP_HIP(hipMemcpyAsync(h.data(), b->q3.p, b->q3.bytes, hipMemcpyDeviceToHost, c->stream));
    QP_HIP(hipStreamSynchronize(c->stream));
    for (int z = 0; z <= c->nz; ++z)
        for (int y = 0; y <= c->ny; ++y)
            for (int x = 0; x <= c->nx; ++x)
                host[x + (size_t)ld1 * (y + (size_t)ld2 * z)] =
                    (double)h[z * c->rplane + (size_t)y * c->ldr + x] * (1.0 / BEAM_FIX);
    return QP_OK;
}

int qp_beam_moments(const qp_beam *b, double *sum_gamma, double *sum_q)
{
    BCHECK(b);
    qp_ctx *c = b->ctx;
    double *part = b->moments.as<double>();
    hipLaunchKernelGGL(k_beam_moments, dim3(1024), dim3(256), 0, c->stream, b->np,
                       b->a[3].as<double>(), b->a[4].as<double>(), b->a[5].as<double>(),
                       b->a[6].as<double>(), part);
    QP_LAUNCH_CHECK();
    std::vector<double> h(2048);
    QP_HIP(hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    QP_HIP(hipStreamSynchronize(c->stream));
    double g = 0.0, q = 0.0;
    for (int i = 0; i < 1024; ++i) { g += h[2 * i]; q += h[2 * i + 1]; }
    if (sum_gamma) *sum_gamma = g;
    if (sum_q) *sum_q = q;
    return QP_OK;
}

}  // extern "C"
