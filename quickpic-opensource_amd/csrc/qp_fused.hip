// Fused field kernels of the xi-slice loop.  Each one restates a fixed
// sequence of field2d calls of go_simulation (simulation_class.f03:643-736)
// mode by mode or node by node, with the same operations in the same order,
// so its results are those of the separate calls; it only saves the launches
// and the intermediate planes.
#include "qp_common.h"

namespace qp {

// pot then grad (simulation_class.f03:659-660): psi = g q (MPPOTPD2,
// fpois2d_lib77.f:1701-1779), fxy = (dkx psi, dky psi) (MPPGRADFD2,
// fft2d_lib77.f:9231-9295); both zero off 0 < kx < nx, 0 < ky < ny.
template <typename R>
__global__ void k_psi_fxy(KGrid g, const R *__restrict__ ffg, const R *__restrict__ q, R *psi,
                          R *fxy)
{
    int kx, ky;
    if (!kmode(g, kx, ky)) return;
    const size_t o = (size_t)kx * g.ldk + ky;
    R p = R(0), a = R(0), b = R(0);
    if (kx > 0 && kx < g.nx && ky > 0 && ky < g.ny) {
        p = ffg[(size_t)kx * g.ny + ky] * q[o];
        a = dk<R>(g.nx, kx) * p;
        b = dk<R>(g.ny, ky) * p;
    }
    psi[o] = p;
    fxy[o] = a;
    fxy[g.cs + o] = b;
}

// One corrector pass in k space (simulation_class.f03:697-709 then the
// B assembly of :677-680 that opens the next pass or follows the loop):
//   QP:  bfqp(cu, dcu, amu, epw)   PBPOISD22N_QP isign = 1 (epw, dcu updated)
//   psit = pot(div(cu)) * (-dex)   MPPDIVFD2, MPPOTPD2
//   b    = bf(cu)                  MPPBBPOISD23
//   bxyz = (-dex epw_y - epwb_y, dex epw_x + epwb_x, dex b_z)
// Without QP (the predictor) epw is read as it stands.  With QP the FP64
// squared change and squared new value of epw are summed per block for the
// corrector's stopping test.  With MIX the change f = epw_new - epw_old is
// also kept in fprev, and <f, f - f_prev> and |f - f_prev|^2 (f_prev = 0 on
// the first pass, hist = 0) are summed per block into mixp for k_mix.
// The predictor (no QP) also applies the corrector seed by extrapolation in
// xi to each mode first: seed = 1 keeps epw (B(j-1)) in ehist, seed = 2 also
// replaces epw by 2 epw - ehist; with ehist2 (quadratic), ehist2 keeps the
// slice before and seed = 3 replaces epw by 3 epw - 3 B(j-2) + B(j-3).  Only the iteration count to a given residual
// depends on the seed.
template <typename R, bool QP, bool MIX>
__global__ void __launch_bounds__(256)
k_pass(KGrid g, const R *__restrict__ ffg, const R *__restrict__ ffs, const R *__restrict__ cu,
       R *dcu, const R *__restrict__ amu, R *epw, const R *__restrict__ epwb, R *psit, R *bxyz,
       R aa, R idex, R dex, double *partial, int *flags, const int *skip, R *fprev, int hist,
       double *mixp, R *ehist, int seed, const int *seed_hist, R *ehist2)
{
    __shared__ double s_num[256], s_den[256], s_ma[MIX ? 256 : 1], s_mb[MIX ? 256 : 1];
    if (skip && *skip) return;
    // captured slices: the slices recorded in ehist so far come from device
    // memory (2 or more: extrapolate; fewer: only record)
    if (seed && seed_hist) {
        const int n = *seed_hist;
        seed = n >= 3 ? 3 : (n >= 2 ? 2 : 1);
    }
    if (seed == 3 && !ehist2) seed = 2;
    if (!QP && blockIdx.x == 0 && threadIdx.x < FLAG_WORDS) flags[threadIdx.x] = 0;
    double num = 0.0, den = 0.0, ma = 0.0, mb = 0.0;
    const size_t cs = g.cs;
    const int total = (g.nx + 1) * (g.ny + 1);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int kx = t / (g.ny + 1), ky = t - kx * (g.ny + 1);
        const size_t o = (size_t)kx * g.ldk + ky;
        const bool tab = kx < g.nx && ky < g.ny;
        const size_t tt = (size_t)kx * g.ny + ky;
        const size_t ti = tab ? tt : 0;
        const R fg = tab ? ffg[ti] : R(0), fs = tab ? ffs[ti] : R(0);
        const R dkx = dk<R>(g.nx, kx), dky = dk<R>(g.ny, ky);
        const R c1 = cu[o], c2 = cu[cs + o], c3 = cu[2 * cs + o];
        R e1 = epw[o], e2 = epw[cs + o];
        if (!QP && seed) {
            const R h1 = ehist[o], h2 = ehist[cs + o];
            R g1 = R(0), g2 = R(0);
            if (ehist2) {
                g1 = ehist2[o];
                g2 = ehist2[cs + o];
                ehist2[o] = h1;
                ehist2[cs + o] = h2;
            }
            ehist[o] = e1;
            ehist[cs + o] = e2;
            if (seed == 2) {
                e1 = R(2) * e1 - h1;
                e2 = R(2) * e2 - h2;
                epw[o] = e1;
                epw[cs + o] = e2;
            } else if (seed == 3) {
                e1 = (R(3) * e1 - R(3) * h1) + g1;
                e2 = (R(3) * e2 - R(3) * h2) + g2;
                epw[o] = e1;
                epw[cs + o] = e2;
            }
        }
        if (QP) {
            const R old1 = e1, old2 = e2;
            R n1 = R(0), n2 = R(0);
            if (tab) {
                R at2 = R(1) * fg / (R(1) + aa * fg);
                R at1 = at2 * fs;
                if (kx > 0 && ky > 0) {
                    R d1 = -dcu[o] + dkx * amu[o] - dky * amu[2 * cs + o];
                    R aat1 = d1 - dkx * c3 * idex;
                    R d2 = -dcu[cs + o] - dkx * amu[2 * cs + o] + dky * amu[cs + o];
                    R aat2 = d2 - dky * c3 * idex;
                    dcu[o] = d1;
                    dcu[cs + o] = d2;
                    n1 = at1 * aat1 + aa * at2 * old1;
                    n2 = at1 * aat2 + aa * at2 * old2;
                } else if (kx > 0) {
                    R d1 = -dcu[o] + dkx * amu[o];
                    R d2 = -dcu[cs + o] - dkx * amu[2 * cs + o];
                    dcu[o] = d1;
                    dcu[cs + o] = d2;
                    n2 = at1 * d2 + aa * at2 * old2;
                } else if (ky > 0) {
                    R d1 = -dcu[o] - dky * amu[2 * cs + o];
                    R d2 = -dcu[cs + o] + dky * amu[cs + o];
                    dcu[o] = d1;
                    dcu[cs + o] = d2;
                    n1 = at1 * d1 + aa * at2 * old1;
                }
            }
            epw[o] = n1;
            epw[cs + o] = n2;
            const double q1 = (double)n1 - (double)old1, q2 = (double)n2 - (double)old2;
            num += q1 * q1 + q2 * q2;
            den += (double)n1 * n1 + (double)n2 * n2;
            if (MIX) {
                const R f1 = n1 - old1, f2 = n2 - old2;
                const R p1 = hist ? fprev[o] : R(0), p2 = hist ? fprev[cs + o] : R(0);
                const double d1 = (double)f1 - (double)p1, d2 = (double)f2 - (double)p2;
                ma += (double)f1 * d1 + (double)f2 * d2;
                mb += d1 * d1 + d2 * d2;
                fprev[o] = f1;
                fprev[cs + o] = f2;
            }
            e1 = n1;
            e2 = n2;
        }
        // psit: div (MPPDIVFD2) -> pot (MPPOTPD2) -> * (-dex)
        R ps = R(0);
        if (kx > 0 && kx < g.nx && ky > 0 && ky < g.ny) ps = fg * (-(dkx * c1 + dky * c2));
        psit[o] = ps * (-dex);
        // bf (MPPBBPOISD23, ci = 1): only its z component survives the assembly
        R b3 = R(0);
        if (tab) {
            const R at1 = R(1) * fg * fs;
            if (kx > 0 && ky > 0) b3 = (dkx * at1) * c2 - (dky * at1) * c1;
            else if (kx > 0) b3 = dkx * at1 * c2;
            else if (ky > 0) b3 = -(dky * at1) * c1;
        }
        bxyz[o] = e2 * (-dex) - epwb[cs + o];
        bxyz[cs + o] = e1 * dex + epwb[o];
        bxyz[2 * cs + o] = b3 * dex;
    }
    if (!QP) return;
    s_num[threadIdx.x] = num;
    s_den[threadIdx.x] = den;
    if (MIX) {
        s_ma[threadIdx.x] = ma;
        s_mb[threadIdx.x] = mb;
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_num[threadIdx.x] += s_num[threadIdx.x + s];
            s_den[threadIdx.x] += s_den[threadIdx.x + s];
            if (MIX) {
                s_ma[threadIdx.x] += s_ma[threadIdx.x + s];
                s_mb[threadIdx.x] += s_mb[threadIdx.x + s];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s_num[0];
        partial[2 * blockIdx.x + 1] = s_den[0];
        if (MIX) {
            mixp[2 * blockIdx.x] = s_ma[0];
            mixp[2 * blockIdx.x + 1] = s_mb[0];
        }
    }
}

// Anderson mixing of depth 1 after corrector pass l.  With g = epw (the
// pass's result), g' = gprev (the previous pass's) and gamma = <f, f - f'> /
// |f - f'|^2 from k_pass's sums, the next iterate is g - gamma (g - g') and
// B_perp = (-dex e_y - epwb_y, dex e_x + epwb_x) is assembled from it.  Every
// block sums the partials in the residual's own order, so all blocks, and the
// host, agree on whether pass l converged: a converged pass is left as it is
// (the plain iteration's result).  mix = 0 only keeps g for the next pass.
template <typename R>
__global__ void __launch_bounds__(256)
k_mix(KGrid g, R *epw, R *gprev, const R *__restrict__ epwb, R *bxyz, R dex,
      const double *partial, const double *mixp, int nblk, double tol, int mix, const int *skip)
{
    __shared__ double s_gam;
    __shared__ int s_act;
    if (skip && *skip) return;
    if (threadIdx.x < 64) {
        double num, den, a, b;
        wave_pair_sum(partial, nblk, num, den);
        wave_pair_sum(mixp, nblk, a, b);
        if (threadIdx.x == 0) {
            const bool conv = tol > 0.0 && residual_of(num, den) < tol;
            s_act = conv ? 0 : (mix ? 2 : 1);   // 0: nothing, 1: keep g, 2: mix
            s_gam = b > 0.0 ? a / b : 0.0;
        }
    }
    __syncthreads();
    const int act = s_act;
    if (act == 0) return;
    const R gam = (R)s_gam;
    const size_t cs = g.cs;
    const int total = (g.nx + 1) * (g.ny + 1);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        const int kx = t / (g.ny + 1), ky = t - kx * (g.ny + 1);
        const size_t o = (size_t)kx * g.ldk + ky;
        const R e1 = epw[o], e2 = epw[cs + o];
        if (act == 2) {
            const R x1 = e1 - gam * (e1 - gprev[o]), x2 = e2 - gam * (e2 - gprev[cs + o]);
            epw[o] = x1;
            epw[cs + o] = x2;
            bxyz[o] = x2 * (-dex) - epwb[cs + o];
            bxyz[cs + o] = x1 * dex + epwb[o];
        }
        gprev[o] = e1;
        gprev[cs + o] = e2;
    }
}

// after the loop (simulation_class.f03:725-726): dcu *= dxi; cu(1:2) -= dcu(1:2)
template <typename R>
__global__ void k_predict_cu(size_t n, size_t cs, R *cu, R *dcu, R dxi)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const R d1 = dcu[e] * dxi, d2 = dcu[cs + e] * dxi;
        dcu[e] = d1;
        dcu[cs + e] = d2;
        cu[e] = cu[e] - d1;
        cu[cs + e] = cu[cs + e] - d2;
    }
}

// (simulation_class.f03:721-723) fxy = -fxy; fxy(1) += bxyz(2); fxy(2) -= bxyz(1)
template <typename R>
__global__ void k_efield(size_t n, size_t cs, R *fxy, const R *__restrict__ bxyz)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const R f1 = fxy[e] * R(-1), f2 = fxy[cs + e] * R(-1);
        fxy[e] = f1 + bxyz[cs + e];
        fxy[cs + e] = f2 - bxyz[e];
    }
}

// slice writes (simulation_class.f03:731-735): fxy *= dex, copied to
// bexyz(1:2); psit to bexyz(3); bxyz(1:2) *= dex; bxyz to bbxyz(1:3); and
// the psi *= dex^2 of :670 (psi has no reader after extpsi in the slice)
template <typename R>
__global__ void k_slice_write(size_t n, size_t cs, R *fxy, const R *__restrict__ psit, R *bxyz,
                              R *psi, R *be1, R *be2, R *be3, R *bb1, R *bb2, R *bb3, R dex,
                              R dd, const int *jdev, int nz, int *ctl, int jnext, int hist)
{
    // eager slices keep the device slice word of captured ones in step
    if (ctl && blockIdx.x == 0 && threadIdx.x == 0) {
        ctl[0] = jnext;
        ctl[1] = hist;
    }
    if (jdev) {   // plane lpos - 1 = j, read from device memory (captured slices)
        const int j = *jdev;
        if (j < 1 || j > nz) return;   // the host checks it before a replay
        const size_t z = (size_t)j * n;
        be1 += z; be2 += z; be3 += z; bb1 += z; bb2 += z; bb3 += z;
    }
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        psi[e] = psi[e] * dd;
        const R f1 = fxy[e] * dex, f2 = fxy[cs + e] * dex;
        const R b1 = bxyz[e] * dex, b2 = bxyz[cs + e] * dex;
        fxy[e] = f1;
        fxy[cs + e] = f2;
        bxyz[e] = b1;
        bxyz[cs + e] = b2;
        be1[e] = f1;
        be2[e] = f2;
        be3[e] = psit[e];
        bb1[e] = b1;
        bb2[e] = b2;
        bb3[e] = bxyz[2 * cs + e];
    }
}

static dim3 flat_grid(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return dim3((unsigned)(b ? b : 1));
}

template <typename R>
static int psi_fxy_t(qp_field2d *q, qp_field2d *psi, qp_field2d *fxy)
{
    qp_ctx *c = q->ctx;
    {
    ProfScope _ps(c, QP_PROF_KSPACE, 5.0 * (c->nx + 1.0) * (c->ny + 1.0) * sizeof(R));
    hipLaunchKernelGGL(k_psi_fxy<R>, kgrid_blocks(c), dim3(256), 0, c->stream, kgrid(c),
                       c->ffg.as<R>(), q->ks.as<R>(), psi->ks.as<R>(), fxy->ks.as<R>());
    }
    QP_LAUNCH_CHECK();
    psi->state = fxy->state = 1;
    return QP_OK;
}

int fused_psi_fxy(qp_field2d *q, qp_field2d *psi, qp_field2d *fxy)
{
    return q->ctx->prec == QP_FP64 ? psi_fxy_t<double>(q, psi, fxy)
                                   : psi_fxy_t<float>(q, psi, fxy);
}

template <typename R>
static int pass_t(const PassFields &f, double dex, bool qp, int l, double tol, int seed,
                  qp_field2d *ehist, const int *seed_hist, qp_field2d *ehist2)
{
    qp_ctx *c = f.cu->ctx;
    double *part = c->red.as<double>();
    // cu 3, dcu 2+2, amu 3, epw 2+2, epwb 2, tables 2, psit 1, bxyz 3
    const double planes = qp ? 22.0 : 12.0;
    {
    ProfScope _ps(c, QP_PROF_KSPACE, planes * (c->nx + 1.0) * (c->ny + 1.0) * sizeof(R));
    const bool mix = qp && f.fprev;
    auto kern = mix ? k_pass<R, true, true> : qp ? k_pass<R, true, false> : k_pass<R, false, false>;
    hipLaunchKernelGGL(kern, dim3(BFQP_BLOCKS), dim3(256), 0, c->stream, kgrid(c),
                       c->ffg.as<R>(), c->ffs.as<R>(), f.cu->ks.as<R>(), f.dcu->ks.as<R>(),
                       f.amu->ks.as<R>(), f.epw->ks.as<R>(), f.epwb->ks.as<R>(),
                       f.psit->ks.as<R>(), f.bxyz->ks.as<R>(), (R)(dex * dex), (R)(1.0 / dex),
                       (R)dex, part, c->flags.as<int>(), qp ? c->skip : nullptr,
                       mix ? f.fprev->ks.as<R>() : nullptr, l >= 2 ? 1 : 0,
                       part + MIX_PARTIALS, seed ? ehist->ks.as<R>() : nullptr, qp ? 0 : seed,
                       qp ? nullptr : seed_hist,
                       seed && ehist2 ? ehist2->ks.as<R>() : nullptr);
    }
    QP_LAUNCH_CHECK();
    if (qp) {
        // the next transform launch sums the partials (see ResJob)
        int *fl = c->flags.as<int>();
        c->res_job = ResJob{part, BFQP_BLOCKS, l > 0 && l < FLAG_WORDS ? fl + l - 1 : nullptr,
                            l > 0 && l < FLAG_WORDS ? fl + l : nullptr,
                            c->h_res + (l % RES_SLOTS), part + 2 * BFQP_BLOCKS, tol};
        f.epw->state = f.dcu->state = 1;
    }
    f.psit->state = f.bxyz->state = 1;
    return QP_OK;
}

int fused_pass(const PassFields &f, double dex, bool qp, int l, double tol, int seed,
               qp_field2d *ehist, const int *seed_hist, qp_field2d *ehist2)
{
    QP_REQUIRE(!seed || (ehist && !qp), QP_ERR_ARG, "fused_pass: seed without history");

    return f.cu->ctx->prec == QP_FP64
               ? pass_t<double>(f, dex, qp, l, tol, seed, ehist, seed_hist, ehist2)
               : pass_t<float>(f, dex, qp, l, tol, seed, ehist, seed_hist, ehist2);
}

template <typename R>
static int mix_t(const PassFields &f, double dex, int l, double tol, bool mix)
{
    qp_ctx *c = f.epw->ctx;
    const double *part = c->red.as<double>();
    {
    // epw 2+2, gprev 2+2, epwb 2, bxyz 2
    ProfScope _ps(c, QP_PROF_KSPACE, 12.0 * (c->nx + 1.0) * (c->ny + 1.0) * sizeof(R));
    hipLaunchKernelGGL(k_mix<R>, dim3(BFQP_BLOCKS), dim3(256), 0, c->stream, kgrid(c),
                       f.epw->ks.as<R>(), f.gprev->ks.as<R>(), f.epwb->ks.as<R>(),
                       f.bxyz->ks.as<R>(), (R)dex, part, part + MIX_PARTIALS, BFQP_BLOCKS, tol,
                       mix ? 1 : 0, c->skip);
    }
    QP_LAUNCH_CHECK();
    (void)l;
    return QP_OK;
}

int fused_mix(const PassFields &f, double dex, int l, double tol, bool mix)
{
    QP_REQUIRE(f.gprev && f.fprev, QP_ERR_ARG, "fused_mix: no corrector history");
    return f.epw->ctx->prec == QP_FP64 ? mix_t<double>(f, dex, l, tol, mix)
                                       : mix_t<float>(f, dex, l, tol, mix);
}

template <typename R>
static int predict_cu_t(qp_field2d *cu, qp_field2d *dcu, double dxi)
{
    qp_ctx *c = cu->ctx;
    const size_t n = c->kplane;
    {
    ProfScope _ps(c, QP_PROF_ELEMENTWISE, 8.0 * n * sizeof(R));
    hipLaunchKernelGGL(k_predict_cu<R>, flat_grid(n), dim3(256), 0, c->stream, n, n,
                       cu->ks.as<R>(), dcu->ks.as<R>(), (R)dxi);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int fused_predict_cu(qp_field2d *cu, qp_field2d *dcu, double dxi)
{
    return cu->ctx->prec == QP_FP64 ? predict_cu_t<double>(cu, dcu, dxi)
                                    : predict_cu_t<float>(cu, dcu, dxi);
}

template <typename R>
static int efield_t(qp_field2d *fxy, const qp_field2d *bxyz)
{
    qp_ctx *c = fxy->ctx;
    const size_t n = c->rplane;
    {
    ProfScope _ps(c, QP_PROF_ELEMENTWISE, 6.0 * n * sizeof(R));
    hipLaunchKernelGGL(k_efield<R>, flat_grid(n), dim3(256), 0, c->stream, n, n,
                       fxy->rs.as<R>(), bxyz->rs.as<R>());
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int fused_efield(qp_field2d *fxy, const qp_field2d *bxyz)
{
    return fxy->ctx->prec == QP_FP64 ? efield_t<double>(fxy, bxyz) : efield_t<float>(fxy, bxyz);
}

template <typename R>
static int slice_write_t(qp_field2d *fxy, const qp_field2d *psit, qp_field2d *bxyz,
                         qp_field2d *psi, qp_field3d *bexyz, qp_field3d *bbxyz, int lpos,
                         double dex, double dd, const int *jdev, int *ctl, int hist)
{
    qp_ctx *c = fxy->ctx;
    const size_t n = c->rplane, st = (size_t)(c->nz + 1) * n;
    const size_t z = jdev ? 0 : (size_t)(lpos - 1) * n;
    R *be = bexyz->buf.as<R>() + z, *bb = bbxyz->buf.as<R>() + z;
    {
    ProfScope _ps(c, QP_PROF_ELEMENTWISE, 18.0 * n * sizeof(R));
    hipLaunchKernelGGL(k_slice_write<R>, flat_grid(n), dim3(256), 0, c->stream, n, n,
                       fxy->rs.as<R>(), psit->rs.as<R>(), bxyz->rs.as<R>(), psi->rs.as<R>(), be,
                       be + st, be + 2 * st, bb, bb + st, bb + 2 * st, (R)dex, (R)dd, jdev,
                       c->nz, ctl, lpos, hist);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int fused_slice_write(qp_field2d *fxy, const qp_field2d *psit, qp_field2d *bxyz, qp_field2d *psi,
                      qp_field3d *bexyz, qp_field3d *bbxyz, int lpos, double dex, double dd,
                      const int *jdev, int *ctl, int hist)
{
    return fxy->ctx->prec == QP_FP64
               ? slice_write_t<double>(fxy, psit, bxyz, psi, bexyz, bbxyz, lpos, dex, dd, jdev,
                                       ctl, hist)
               : slice_write_t<float>(fxy, psit, bxyz, psi, bexyz, bbxyz, lpos, dex, dd, jdev,
                                      ctl, hist);
}

}  // namespace qp
