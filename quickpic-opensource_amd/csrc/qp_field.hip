// field2d / field3d objects: state machine, component-mapped arithmetic,
// guard operations, slice copies and the k-space solvers.
#include <cmath>
#include <cstring>

#include "qp_common.h"

namespace qp {

// ------------------------------------------------------------------------
// elementwise kernels
// ------------------------------------------------------------------------
template <typename R>
__global__ void k_fill(R *p, size_t n, R v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

struct MapArgs {
    int n, op;            // op: 0 add, 1 sub, 2 mult
    int d[3], d1[3], d2[3];
    double v[3];
};

// ufield2d sum2/minus2/multiply2 (ufield2d_class.f03:611-618 etc.): per
// element, the components are processed in order i = 1..n, so aliasing of
// this/a1/a2 behaves as in the reference.
template <typename R>
__global__ void k_mapop(R *dst, const R *a1, const R *a2, size_t cs, size_t nel, MapArgs m)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < nel;
         e += (size_t)gridDim.x * blockDim.x) {
        for (int i = 0; i < m.n; ++i) {
            R x = a1[m.d1[i] * cs + e];
            R r;
            if (m.op == 0) r = x + a2[m.d2[i] * cs + e];
            else if (m.op == 1) r = x - a2[m.d2[i] * cs + e];
            else r = x * (R)m.v[i];
            dst[m.d[i] * cs + e] = r;
        }
    }
}

// PPNACGUARD2L nvp = 1 (ufield2d_lib.f03:90-97): row ny added into row 0,
// then cleared, for x = 0..nx
template <typename R>
__global__ void k_guard_fold(R *f, size_t cs, int ld, int nx, int ny)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x > nx) return;
    R *g = f + blockIdx.y * cs;
    g[x] = g[x] + g[(size_t)ny * ld + x];
    g[(size_t)ny * ld + x] = R(0);
}

// PPNCGUARD2L nvp = 1 (ufield2d_lib.f03:37-41): row ny <- row 0
template <typename R>
__global__ void k_copyguard(R *f, size_t cs, int ld, int ny)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ld) return;
    R *g = f + blockIdx.y * cs;
    g[(size_t)ny * ld + x] = g[x];
}

template <typename R>
__global__ void k_plane_copy(R *dst, const R *src, size_t n, int add)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = add ? dst[i] + src[i] : src[i];
}

static dim3 grid_for(size_t n)
{
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    return dim3((unsigned)(b ? b : 1));
}

template <typename R>
static int fill_t(qp_ctx *c, void *p, size_t n, double v)
{
    {
    ProfScope _ps(c, QP_PROF_ELEMENTWISE, (double)n * sizeof(R));
    hipLaunchKernelGGL(k_fill<R>, grid_for(n), dim3(256), 0, c->stream, (R *)p, n, (R)v);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int launch_fill(qp_ctx *c, void *p, size_t n, double v)
{
    return c->prec == QP_FP64 ? fill_t<double>(c, p, n, v) : fill_t<float>(c, p, n, v);
}

template <typename R>
static int fold_t(qp_ctx *c, void *planes, int dim)
{
    {
    ProfScope _ps(c, QP_PROF_ELEMENTWISE, 3.0 * dim * (c->nx + 1) * sizeof(R));
    hipLaunchKernelGGL(k_guard_fold<R>, dim3((c->nx + 256) / 256, dim), dim3(256), 0, c->stream,
                       (R *)planes, c->rplane, c->ldr, c->nx, c->ny);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int guard_fold(qp_ctx *c, void *planes, int dim)
{
    return c->prec == QP_FP64 ? fold_t<double>(c, planes, dim) : fold_t<float>(c, planes, dim);
}

template <typename R>
static int copyguard_t(qp_field2d *f)
{
    qp_ctx *c = f->ctx;
    {
    ProfScope _ps(c, QP_PROF_ELEMENTWISE, 2.0 * f->dim * (c->nx + 1) * sizeof(R));
    hipLaunchKernelGGL(k_copyguard<R>, dim3((c->ldr + 255) / 256, f->dim), dim3(256), 0,
                       c->stream, f->rs.as<R>(), c->rplane, c->ldr, c->ny);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

int copyguard(qp_field2d *f)
{
    return f->ctx->prec == QP_FP64 ? copyguard_t<double>(f) : copyguard_t<float>(f);
}

// ------------------------------------------------------------------------
// k-space solvers.  Arrays are [kx][ky] with pitch ldk; tables [kx][ky]
// with pitch ny (kx < nx, ky < ny).  Each kernel restates the index ranges
// of its reference routine; one thread per mode (kx, ky), 0 <= kx <= nx,
// 0 <= ky <= ny.
// ------------------------------------------------------------------------
// MPPDIVFD2 fft2d_lib77.f:9170-9229
template <typename R>
__global__ void k_div(KGrid g, const R *f, R *df)
{
    int kx, ky;
    if (!kmode(g, kx, ky)) return;
    size_t o = (size_t)kx * g.ldk + ky;
    R v = R(0);
    if (kx > 0 && kx < g.nx && ky > 0 && ky < g.ny)
        v = -(dk<R>(g.nx, kx) * f[o] + dk<R>(g.ny, ky) * f[g.cs + o]);
    df[o] = v;
}

// MPPGRADFD2 fft2d_lib77.f:9231-9295
template <typename R>
__global__ void k_grad(KGrid g, const R *df, R *f)
{
    int kx, ky;
    if (!kmode(g, kx, ky)) return;
    size_t o = (size_t)kx * g.ldk + ky;
    R a = R(0), b = R(0);
    if (kx > 0 && kx < g.nx && ky > 0 && ky < g.ny) {
        a = dk<R>(g.nx, kx) * df[o];
        b = dk<R>(g.ny, ky) * df[o];
    }
    f[o] = a;
    f[g.cs + o] = b;
}

// MPPOTPD2 (mode 0, fpois2d_lib77.f:1701-1779) and MPPSMOOTHD2 (mode 1,
// :1781-1843): pot = g q, qs = s q on 0 < kx < nx, 0 < ky < ny
template <typename R>
__global__ void k_scale_tab(KGrid g, const R *tab, const R *q, R *out)
{
    int kx, ky;
    if (!kmode(g, kx, ky)) return;
    size_t o = (size_t)kx * g.ldk + ky;
    R v = R(0);
    if (kx > 0 && kx < g.nx && ky > 0 && ky < g.ny) v = tab[(size_t)kx * g.ny + ky] * q[o];
    out[o] = v;
}

// MPPOISD22 isign = -1 (fpois2d_lib77.f:1533-1574): E = -k g s q
template <typename R>
__global__ void k_elf(KGrid g, const R *ffg, const R *ffs, const R *q, R *fxy)
{
    int kx, ky;
    if (!kmode(g, kx, ky)) return;
    size_t o = (size_t)kx * g.ldk + ky;
    R a = R(0), b = R(0);
    if (kx > 0 && kx < g.nx && ky > 0 && ky < g.ny) {
        size_t t = (size_t)kx * g.ny + ky;
        R at1 = ffg[t] * ffs[t];
        R at3 = -at1 * q[o];
        a = dk<R>(g.nx, kx) * at3;
        b = dk<R>(g.ny, ky) * at3;
    }
    fxy[o] = a;
    fxy[g.cs + o] = b;
}

// MPPBBPOISD23 (fpois2d_lib77.f:1845-1956) with ci = 1
template <typename R>
__global__ void k_bf(KGrid g, const R *ffg, const R *ffs, const R *cu, R *bxy)
{
    int kx, ky;
    if (!kmode(g, kx, ky)) return;
    size_t o = (size_t)kx * g.ldk + ky;
    R b1 = R(0), b2 = R(0), b3 = R(0);
    if (kx < g.nx && ky < g.ny) {
        size_t t = (size_t)kx * g.ny + ky;
        R at1 = R(1) * ffg[t] * ffs[t];
        if (kx > 0 && ky > 0) {
            R at2 = dk<R>(g.ny, ky) * at1, at3 = dk<R>(g.nx, kx) * at1;
            b1 = at2 * cu[2 * g.cs + o];
            b2 = -at3 * cu[2 * g.cs + o];
            b3 = at3 * cu[g.cs + o] - at2 * cu[o];
        } else if (kx > 0) {           // ky = 0
            b3 = dk<R>(g.nx, kx) * at1 * cu[g.cs + o];
        } else if (ky > 0) {           // kx = 0
            b3 = -(dk<R>(g.ny, ky) * at1) * cu[o];
        }
    }
    bxy[o] = b1;
    bxy[g.cs + o] = b2;
    bxy[2 * g.cs + o] = b3;
}

// PBPOISD22N_QP isign = 1 (fpois2d_lib77.f:1322-1421) with ci = 1.  dcu is
// overwritten with -dcu + k x amu.  Also accumulates, in FP64, the squared
// change and the squared new value of B_perp for the corrector's stopping
// test (per-block partials, summed in a fixed order afterwards).
template <typename R>
__global__ void __launch_bounds__(256)
k_bfqp(KGrid g, const R *ffg, const R *ffs, const R *cu, R *dcu, const R *amu, R *bxy,
       R aa, R idex, double *partial)
{
    __shared__ double s_num[256], s_den[256];
    double num = 0.0, den = 0.0;
    const int total = (g.nx + 1) * (g.ny + 1);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
        int kx = t / (g.ny + 1), ky = t - kx * (g.ny + 1);
        size_t o = (size_t)kx * g.ldk + ky;
        R old1 = bxy[o], old2 = bxy[g.cs + o];
        R n1 = R(0), n2 = R(0);
        if (kx < g.nx && ky < g.ny) {
            size_t tt = (size_t)kx * g.ny + ky;
            R at2 = ffg[tt];
            at2 = R(1) * at2 / (R(1) + aa * at2);
            R at1 = at2 * ffs[tt];
            R dkx = dk<R>(g.nx, kx), dky = dk<R>(g.ny, ky);
            if (kx > 0 && ky > 0) {
                R d1 = -dcu[o] + dkx * amu[o] - dky * amu[2 * g.cs + o];
                R aat1 = d1 - dkx * cu[2 * g.cs + o] * idex;
                R d2 = -dcu[g.cs + o] - dkx * amu[2 * g.cs + o] + dky * amu[g.cs + o];
                R aat2 = d2 - dky * cu[2 * g.cs + o] * idex;
                dcu[o] = d1;
                dcu[g.cs + o] = d2;
                n1 = at1 * aat1 + aa * at2 * old1;
                n2 = at1 * aat2 + aa * at2 * old2;
            } else if (kx > 0) {            // ky = 0
                R d1 = -dcu[o] + dkx * amu[o];
                R d2 = -dcu[g.cs + o] - dkx * amu[2 * g.cs + o];
                dcu[o] = d1;
                dcu[g.cs + o] = d2;
                n2 = at1 * d2 + aa * at2 * old2;
            } else if (ky > 0) {            // kx = 0
                R d1 = -dcu[o] - dky * amu[2 * g.cs + o];
                R d2 = -dcu[g.cs + o] + dky * amu[g.cs + o];
                dcu[o] = d1;
                dcu[g.cs + o] = d2;
                n1 = at1 * d1 + aa * at2 * old1;
            }
        }
        bxy[o] = n1;
        bxy[g.cs + o] = n2;
        double e1 = (double)n1 - (double)old1, e2 = (double)n2 - (double)old2;
        num += e1 * e1 + e2 * e2;
        den += (double)n1 * n1 + (double)n2 * n2;
    }
    s_num[threadIdx.x] = num;
    s_den[threadIdx.x] = den;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_num[threadIdx.x] += s_num[threadIdx.x + s];
            s_den[threadIdx.x] += s_den[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s_num[0];
        partial[2 * blockIdx.x + 1] = s_den[0];
    }
}

// fixed-order sum of the per-block partials
__global__ void k_sum_partials(const double *partial, int nblk, double *out)
{
    __shared__ double s[2][256];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x) {
        a += partial[2 * i];
        b += partial[2 * i + 1];
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            s[0][threadIdx.x] += s[0][threadIdx.x + st];
            s[1][threadIdx.x] += s[1][threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = s[0][0]; out[1] = s[1][0]; }
}


template <typename R>
int bfqp_resid(qp_field2d *cu, qp_field2d *dcu, qp_field2d *amu, qp_field2d *bxy, double cc,
               double dex, double *num_den_dev)
{
    qp_ctx *c = cu->ctx;
    double *part = c->red.as<double>();
    {
    ProfScope _ps(c, QP_PROF_KSPACE, 16.0 * (c->nx + 1.0) * (c->ny + 1.0) * sizeof(R));
    hipLaunchKernelGGL(k_bfqp<R>, dim3(BFQP_BLOCKS), dim3(256), 0, c->stream, kgrid(c),
                       c->ffg.as<R>(), c->ffs.as<R>(), cu->ks.as<R>(), dcu->ks.as<R>(),
                       amu->ks.as<R>(), bxy->ks.as<R>(), (R)cc, (R)(1.0 / dex), part);
    }
    QP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, c->stream, part, BFQP_BLOCKS,
                       num_den_dev);
    QP_LAUNCH_CHECK();
    return QP_OK;
}
template int bfqp_resid<float>(qp_field2d *, qp_field2d *, qp_field2d *, qp_field2d *, double,
                               double, double *);
template int bfqp_resid<double>(qp_field2d *, qp_field2d *, qp_field2d *, qp_field2d *, double,
                                double, double *);

int bfqp_impl(qp_field2d *cu, qp_field2d *dcu, qp_field2d *amu, qp_field2d *bxy, double cc,
              double dex, double *resid_out)
{
    qp_ctx *c = cu->ctx;
    double *nd = c->red.as<double>() + 2 * BFQP_BLOCKS;
    int rc = c->prec == QP_FP64 ? bfqp_resid<double>(cu, dcu, amu, bxy, cc, dex, nd)
                                : bfqp_resid<float>(cu, dcu, amu, bxy, cc, dex, nd);
    if (rc) return rc;
    bxy->state = 1;
    if (resid_out) {
        double h[2];
        QP_HIP(hipMemcpyAsync(h, nd, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        QP_HIP(hipStreamSynchronize(c->stream));
        *resid_out = h[1] > 0 ? std::sqrt(h[0] / h[1]) : (h[0] > 0 ? INFINITY : 0.0);
    }
    return QP_OK;
}

// ------------------------------------------------------------------------
// host helpers for the field objects
// ------------------------------------------------------------------------
template <typename R>
static int upload_planes(qp_ctx *c, void *dev, const double *host, int dim, int rows, int cols,
                         int ld, size_t cs, int hld1, int hld2, bool kspace)
{
    // host f(dim, hld1, hld2) column-major; real: hld1 indexes x (cols),
    // hld2 indexes y (rows); k: hld1 indexes ky (cols), hld2 indexes kx (rows)
    std::vector<R> h(cs * dim, R(0));
    for (int d = 0; d < dim; ++d)
        for (int r = 0; r < rows; ++r)
            for (int col = 0; col < cols; ++col)
                h[d * cs + (size_t)r * ld + col] =
                    (R)host[d + (size_t)dim * (col + (size_t)hld1 * r)];
    (void)hld2;
    (void)kspace;
    QP_HIP(hipMemcpyAsync(dev, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice, c->stream));
    QP_HIP(hipStreamSynchronize(c->stream));
    return QP_OK;
}

template <typename R>
static int download_planes(qp_ctx *c, const void *dev, double *host, int dim, int rows, int cols,
                           int ld, size_t cs, int hld1)
{
    std::vector<R> h(cs * dim);
    QP_HIP(hipMemcpyAsync(h.data(), dev, h.size() * sizeof(R), hipMemcpyDeviceToHost, c->stream));
    QP_HIP(hipStreamSynchronize(c->stream));
    for (int d = 0; d < dim; ++d)
        for (int r = 0; r < rows; ++r)
            for (int col = 0; col < cols; ++col)
                host[d + (size_t)dim * (col + (size_t)hld1 * r)] =
                    (double)h[d * cs + (size_t)r * ld + col];
    return QP_OK;
}

template <typename R>
static int mapop_t(qp_ctx *c, void *dst, const void *a1, const void *a2, size_t cs,
                   const MapArgs &m)
{
    {
    ProfScope _ps(c, QP_PROF_ELEMENTWISE, (m.op == 2 ? 2.0 : 3.0) * m.n * (double)cs * sizeof(R));
    hipLaunchKernelGGL(k_mapop<R>, grid_for(cs), dim3(256), 0, c->stream, (R *)dst,
                       (const R *)a1, (const R *)a2, cs, cs, m);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

static int mapop(qp_ctx *c, void *dst, const void *a1, const void *a2, size_t cs,
                 const MapArgs &m)
{
    return c->prec == QP_FP64 ? mapop_t<double>(c, dst, a1, a2, cs, m)
                              : mapop_t<float>(c, dst, a1, a2, cs, m);
}

static void *space_ptr(const qp_field2d *f, int state)
{
    return state == 0 ? f->rs.p : f->ks.p;
}

static size_t space_cs(const qp_field2d *f, int state)
{
    return state == 0 ? f->ctx->rplane : f->ctx->kplane;
}

static int build_map(MapArgs &m, int op, int n, const int *dim, const int *dim1,
                     const int *dim2, const double *value, int fdim, int d1max, int d2max)
{
    m.op = op;
    if (n == 0) {
        m.n = fdim;
        for (int i = 0; i < fdim; ++i) {
            m.d[i] = m.d1[i] = m.d2[i] = i;
            m.v[i] = value ? value[0] : 0.0;
        }
    } else {
        QP_REQUIRE(n > 0 && n <= 3 && dim && dim1, QP_ERR_ARG, "bad component map (n=%d)", n);
        m.n = n;
        for (int i = 0; i < n; ++i) {
            m.d[i] = dim[i] - 1;
            m.d1[i] = dim1[i] - 1;
            m.d2[i] = dim2 ? dim2[i] - 1 : 0;
            m.v[i] = value ? value[i] : 0.0;
            QP_REQUIRE(m.d[i] >= 0 && m.d[i] < fdim && m.d1[i] >= 0 && m.d1[i] < d1max &&
                           m.d2[i] >= 0 && m.d2[i] < d2max,
                       QP_ERR_ARG, "component index out of range");
        }
    }
    for (int i = m.n; i < 3; ++i) { m.d[i] = m.d1[i] = m.d2[i] = 0; m.v[i] = 0.0; }
    return QP_OK;
}

template <typename R>
static int slice_copy_t(qp_ctx *c, R *dst, const R *src, int add)
{
    {
    ProfScope _ps(c, QP_PROF_ELEMENTWISE, (add ? 3.0 : 2.0) * (double)c->rplane * sizeof(R));
    hipLaunchKernelGGL(k_plane_copy<R>, grid_for(c->rplane), dim3(256), 0, c->stream, dst, src,
                       c->rplane, add);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

}  // namespace qp

using namespace qp;

#define FCHECK(f) QP_REQUIRE((f) && (f)->ctx, QP_ERR_ARG, "%s: null field", __func__)

extern "C" {

int qp_field2d_create(qp_ctx *c, int dim, int state, qp_field2d **out)
{
    QP_REQUIRE(c && out, QP_ERR_ARG, "qp_field2d_create: null argument");
    QP_REQUIRE(dim >= 1 && dim <= 3, QP_ERR_ARG, "qp_field2d_create: dim %d", dim);
    QP_REQUIRE(state == 0 || state == 1, QP_ERR_ARG, "qp_field2d_create: state %d", state);
    qp_field2d *f = new qp_field2d();
    f->ctx = c;
    f->dim = dim;
    f->state = state;
    f->gcells = 0;
    int rc = f->rs.alloc(dim * c->rplane * c->esize);
    if (!rc) rc = f->ks.alloc(dim * c->kplane * c->esize);
    if (rc) { qp_field2d_destroy(f); return rc; }
    *out = f;
    return QP_OK;
}

int qp_field2d_destroy(qp_field2d *f)
{
    if (!f) return QP_OK;
    if (f->ctx) (void)hipStreamSynchronize(f->ctx->stream);
    f->rs.release();
    f->ks.release();
    delete f;
    return QP_OK;
}

int qp_field2d_state(const qp_field2d *f, int *state, int *dim)
{
    FCHECK(f);
    if (state) *state = f->state;
    if (dim) *dim = f->dim;
    return QP_OK;
}

int qp_field2d_upload(qp_field2d *f, int space, const double *host, int ld1, int ld2)
{
    FCHECK(f);
    qp_ctx *c = f->ctx;
    QP_REQUIRE(host, QP_ERR_ARG, "qp_field2d_upload: null host array");
    int rows = space == 0 ? c->ny + 1 : c->nx + 1;
    int cols = space == 0 ? c->nx + 1 : c->ny + 1;
    QP_REQUIRE(ld1 >= cols && ld2 >= rows, QP_ERR_ARG, "qp_field2d_upload: leading dims too small");
    int ld = space == 0 ? c->ldr : c->ldk;
    size_t cs = space == 0 ? c->rplane : c->kplane;
    void *dev = space == 0 ? f->rs.p : f->ks.p;
    int rc = c->prec == QP_FP64
                 ? upload_planes<double>(c, dev, host, f->dim, rows, cols, ld, cs, ld1, ld2, space)
                 : upload_planes<float>(c, dev, host, f->dim, rows, cols, ld, cs, ld1, ld2, space);
    if (!rc) f->state = space;
    return rc;
}

int qp_field2d_download(const qp_field2d *f, int space, double *host, int ld1, int ld2)
{
    FCHECK(f);
    qp_ctx *c = f->ctx;
    QP_REQUIRE(host, QP_ERR_ARG, "qp_field2d_download: null host array");
    int rows = space == 0 ? c->ny + 1 : c->nx + 1;
    int cols = space == 0 ? c->nx + 1 : c->ny + 1;
    QP_REQUIRE(ld1 >= cols && ld2 >= rows, QP_ERR_ARG, "qp_field2d_download: leading dims too small");
    int ld = space == 0 ? c->ldr : c->ldk;
    size_t cs = space == 0 ? c->rplane : c->kplane;
    const void *dev = space == 0 ? f->rs.p : f->ks.p;
    return c->prec == QP_FP64 ? download_planes<double>(c, dev, host, f->dim, rows, cols, ld, cs, ld1)
                              : download_planes<float>(c, dev, host, f->dim, rows, cols, ld, cs, ld1);
}

int qp_field2d_device(const qp_field2d *f, int space, void **ptr, int *ld, size_t *comp_stride)
{
    FCHECK(f);
    QP_REQUIRE(space == 0 || space == 1, QP_ERR_ARG, "qp_field2d_device: space %d", space);
    if (ptr) *ptr = space == 0 ? f->rs.p : f->ks.p;
    if (ld) *ld = space == 0 ? f->ctx->ldr : f->ctx->ldk;
    if (comp_stride) *comp_stride = space == 0 ? f->ctx->rplane : f->ctx->kplane;
    return QP_OK;
}

// asc (field2d_class.f03:493-510): real space set to value, state 0
int qp_field2d_as(qp_field2d *f, double value)
{
    FCHECK(f);
    int rc = launch_fill(f->ctx, f->rs.p, f->dim * f->ctx->rplane, value);
    if (!rc) { f->state = 0; f->gcells = 1; }
    return rc;
}

// asa (field2d_class.f03:624-642)
int qp_field2d_asa(qp_field2d *f, const qp_field2d *that)
{
    FCHECK(f);
    FCHECK(that);
    QP_REQUIRE(f->dim == that->dim, QP_ERR_ARG, "qp_field2d_asa: dim mismatch");
    qp_ctx *c = f->ctx;
    size_t bytes = f->dim * space_cs(f, that->state) * c->esize;
    QP_HIP(hipMemcpyAsync(space_ptr(f, that->state), space_ptr(that, that->state), bytes,
                          hipMemcpyDeviceToDevice, c->stream));
    f->state = that->state;
    f->gcells = that->gcells;
    return QP_OK;
}

int qp_field2d_fftrk(qp_field2d *f, int kind)
{
    FCHECK(f);
    QP_REQUIRE(f->state == 0, QP_ERR_STATE, "fftrk: data is in ks");
    QP_REQUIRE(kind >= 1 && kind <= 3, QP_ERR_ARG, "fftrk: kind %d", kind);
    QP_REQUIRE(kind != 3 || f->dim == 3, QP_ERR_ARG, "fftrk: wrong fft kind");
    int rc = fft_r2k(f, kind);
    if (!rc) f->state = 1;
    return rc;
}

int qp_field2d_fftkr(qp_field2d *f, int kind)
{
    FCHECK(f);
    QP_REQUIRE(f->state == 1, QP_ERR_STATE, "fftkr: data is in rs");
    QP_REQUIRE(kind >= 1 && kind <= 3, QP_ERR_ARG, "fftkr: kind %d", kind);
    QP_REQUIRE(kind != 3 || f->dim == 3, QP_ERR_ARG, "fftkr: wrong fft kind");
    int rc = fft_k2r(f, kind);
    if (!rc) f->state = 0;
    return rc;
}

#define KSTATE(f, name) QP_REQUIRE((f)->state == 1, QP_ERR_STATE, name ": data is in rs")

}  // extern "C"

template <typename R>
static int kop_launch(int op, qp_field2d *in, qp_field2d *out)
{
    qp_ctx *c = in->ctx;
    KGrid g = kgrid(c);
    const R *a = in->ks.as<R>();
    R *o = out->ks.as<R>();
    static const double planes[6] = {3, 3, 3, 3, 5, 8};   // read + written + tables
    ProfScope _ps(c, QP_PROF_KSPACE, planes[op] * (c->nx + 1.0) * (c->ny + 1.0) * sizeof(R));
    switch (op) {
    case 0: hipLaunchKernelGGL(k_div<R>, kgrid_blocks(c), dim3(256), 0, c->stream, g, a, o); break;
    case 1: hipLaunchKernelGGL(k_grad<R>, kgrid_blocks(c), dim3(256), 0, c->stream, g, a, o); break;
    case 2: hipLaunchKernelGGL(k_scale_tab<R>, kgrid_blocks(c), dim3(256), 0, c->stream, g,
                               c->ffg.as<R>(), a, o); break;
    case 3: hipLaunchKernelGGL(k_scale_tab<R>, kgrid_blocks(c), dim3(256), 0, c->stream, g,
                               c->ffs.as<R>(), a, o); break;
    case 4: hipLaunchKernelGGL(k_elf<R>, kgrid_blocks(c), dim3(256), 0, c->stream, g,
                               c->ffg.as<R>(), c->ffs.as<R>(), a, o); break;
    case 5: hipLaunchKernelGGL(k_bf<R>, kgrid_blocks(c), dim3(256), 0, c->stream, g,
                               c->ffg.as<R>(), c->ffs.as<R>(), a, o); break;
    }
    QP_LAUNCH_CHECK();
    out->state = 1;
    return QP_OK;
}

extern "C" {

static int kop(int op, qp_field2d *in, qp_field2d *out, int din_min, int dout)
{
    FCHECK(in);
    FCHECK(out);
    QP_REQUIRE(in->state == 1, QP_ERR_STATE, "k-space operator: data is in rs");
    QP_REQUIRE(in->dim >= din_min && out->dim >= dout, QP_ERR_ARG,
               "k-space operator: component count (in %d, out %d)", in->dim, out->dim);
    return in->ctx->prec == QP_FP64 ? kop_launch<double>(op, in, out)
                                    : kop_launch<float>(op, in, out);
}

int qp_field2d_div(qp_field2d *f, qp_field2d *out) { return kop(0, f, out, 2, 1); }
int qp_field2d_grad(qp_field2d *f, qp_field2d *out) { return kop(1, f, out, 1, 2); }
int qp_field2d_pot(qp_field2d *q, qp_field2d *out) { return kop(2, q, out, 1, 1); }
int qp_field2d_smooth(qp_field2d *q, qp_field2d *out) { return kop(3, q, out, 1, 1); }
int qp_field2d_elf(qp_field2d *q, qp_field2d *out) { return kop(4, q, out, 1, 2); }
int qp_field2d_bf(qp_field2d *cu, qp_field2d *out) { return kop(5, cu, out, 3, 3); }

int qp_field2d_bfqp(qp_field2d *cu, qp_field2d *dcu, qp_field2d *amu, qp_field2d *bxy,
                    double cc, double dex)
{
    FCHECK(cu); FCHECK(dcu); FCHECK(amu); FCHECK(bxy);
    QP_REQUIRE(cu->state == 1 && dcu->state == 1 && amu->state == 1, QP_ERR_STATE,
               "bfield_qp: data is in rs");
    QP_REQUIRE(cu->dim == 3 && dcu->dim >= 2 && amu->dim == 3 && bxy->dim >= 2, QP_ERR_ARG,
               "bfield_qp: component counts");
    return bfqp_impl(cu, dcu, amu, bxy, cc, dex, nullptr);
}

static int field_op(qp_field2d *f, const qp_field2d *a1, const qp_field2d *a2, int op, int n,
                    const int *dim, const int *dim1, const int *dim2, const double *value)
{
    FCHECK(f);
    FCHECK(a1);
    if (op != 2) {
        FCHECK(a2);
        QP_REQUIRE(a1->state == a2->state, QP_ERR_STATE, "field op: states are different");
    }
    MapArgs m;
    int rc = build_map(m, op, n, dim, dim1, dim2, value, f->dim, a1->dim, a2 ? a2->dim : 1);
    if (rc) return rc;
    if (n == 0) {
        QP_REQUIRE(a1->dim == f->dim && (op == 2 || a2->dim == f->dim), QP_ERR_ARG,
                   "field op: dim mismatch");
    }
    int st = a1->state;
    rc = mapop(f->ctx, space_ptr(f, st), space_ptr(a1, st), a2 ? space_ptr(a2, st) : nullptr,
               space_cs(f, st), m);
    if (!rc) {
        f->state = st;
        if (st == 0) f->gcells = a1->gcells;
    }
    return rc;
}

int qp_field2d_add(qp_field2d *f, const qp_field2d *a1, const qp_field2d *a2, int n,
                   const int *dim, const int *dim1, const int *dim2)
{
    return field_op(f, a1, a2, 0, n, dim, dim1, dim2, nullptr);
}

int qp_field2d_sub(qp_field2d *f, const qp_field2d *a1, const qp_field2d *a2, int n,
                   const int *dim, const int *dim1, const int *dim2)
{
    return field_op(f, a1, a2, 1, n, dim, dim1, dim2, nullptr);
}

int qp_field2d_mult(qp_field2d *f, const qp_field2d *a1, int n, const int *dim, const int *dim1,
                    const double *value)
{
    QP_REQUIRE(value, QP_ERR_ARG, "qp_field2d_mult: null value");
    return field_op(f, a1, nullptr, 2, n, dim, dim1, nullptr, value);
}

int qp_field2d_ag(qp_field2d *f)
{
    FCHECK(f);
    QP_REQUIRE(f->state == 0, QP_ERR_STATE, "acopyguard: data is in ks");
    int rc = guard_fold(f->ctx, f->rs.p, f->dim);
    if (!rc) f->gcells = 1;
    return rc;
}

int qp_field2d_cg(qp_field2d *f)
{
    FCHECK(f);
    QP_REQUIRE(f->state == 0, QP_ERR_STATE, "copyguard: data is in ks");
    int rc = copyguard(f);
    if (!rc) f->gcells = 1;
    return rc;
}

static int slice_args(const qp_field2d *f, const qp_field3d *f3, int lpos, int n, const int *s,
                      const int *d, bool to3d)
{
    QP_REQUIRE(f && f3 && f->ctx == f3->ctx, QP_ERR_ARG, "slice copy: bad objects");
    QP_REQUIRE(lpos >= 1 && lpos <= f->ctx->nz + 1, QP_ERR_ARG, "slice copy: lpos %d", lpos);
    QP_REQUIRE(n >= 1 && n <= 3 && s && d, QP_ERR_ARG, "slice copy: bad component map");
    for (int i = 0; i < n; ++i) {
        int sd = s[i] - 1, dd = d[i] - 1;
        int smax = to3d ? f->dim : f3->dim, dmax = to3d ? f3->dim : f->dim;
        QP_REQUIRE(sd >= 0 && sd < smax && dd >= 0 && dd < dmax, QP_ERR_ARG,
                   "slice copy: component out of range");
    }
    return QP_OK;
}

int qp_field2d_cb(const qp_field2d *f, qp_field3d *f3, int lpos, int n, const int *sdim,
                  const int *ddim)
{
    int rc = slice_args(f, f3, lpos, n, sdim, ddim, true);
    if (rc) return rc;
    QP_REQUIRE(f->state == 0, QP_ERR_STATE, "copyto: data is in ks");
    qp_ctx *c = f->ctx;
    size_t e = c->esize, pl = c->rplane, st = (size_t)(c->nz + 1) * pl;
    for (int i = 0; i < n; ++i) {
        char *dst = (char *)f3->buf.p + ((ddim[i] - 1) * st + (size_t)(lpos - 1) * pl) * e;
        const char *src = (const char *)f->rs.p + (size_t)(sdim[i] - 1) * pl * e;
        QP_HIP(hipMemcpyAsync(dst, src, pl * e, hipMemcpyDeviceToDevice, c->stream));
    }
    return QP_OK;
}

int qp_field2d_ca(qp_field2d *f, const qp_field3d *f3, int lpos, int n, const int *sdim,
                  const int *ddim)
{
    int rc = slice_args(f, f3, lpos, n, sdim, ddim, false);
    if (rc) return rc;
    qp_ctx *c = f->ctx;
    size_t pl = c->rplane, st = (size_t)(c->nz + 1) * pl;
    for (int i = 0; i < n; ++i) {
        size_t so = (sdim[i] - 1) * st + (size_t)(lpos - 1) * pl;
        size_t dof = (size_t)(ddim[i] - 1) * pl;
        rc = c->prec == QP_FP64
                 ? slice_copy_t<double>(c, f->rs.as<double>() + dof, f3->buf.as<double>() + so, 1)
                 : slice_copy_t<float>(c, f->rs.as<float>() + dof, f3->buf.as<float>() + so, 1);
        if (rc) return rc;
    }
    f->state = 0;
    f->gcells = 1;
    return QP_OK;
}

// ---- field3d ----
int qp_field3d_create(qp_ctx *c, int dim, qp_field3d **out)
{
    QP_REQUIRE(c && out && dim >= 1 && dim <= 3, QP_ERR_ARG, "qp_field3d_create: bad argument");
    qp_field3d *f = new qp_field3d();
    f->ctx = c;
    f->dim = dim;
    int rc = f->buf.alloc((size_t)dim * (c->nz + 1) * c->rplane * c->esize);
    if (rc) { delete f; return rc; }
    *out = f;
    return QP_OK;
}

int qp_field3d_destroy(qp_field3d *f)
{
    if (!f) return QP_OK;
    if (f->ctx) (void)hipStreamSynchronize(f->ctx->stream);
    f->buf.release();
    delete f;
    return QP_OK;
}

int qp_field3d_as(qp_field3d *f, double v)
{
    QP_REQUIRE(f && f->ctx, QP_ERR_ARG, "qp_field3d_as: null field");
    return launch_fill(f->ctx, f->buf.p, (size_t)f->dim * (f->ctx->nz + 1) * f->ctx->rplane, v);
}

int qp_field3d_upload(qp_field3d *f, const double *host, int ld1, int ld2, int ld3)
{
    QP_REQUIRE(f && f->ctx && host, QP_ERR_ARG, "qp_field3d_upload: bad argument");
    qp_ctx *c = f->ctx;
    QP_REQUIRE(ld1 >= c->nx + 1 && ld2 >= c->ny + 1 && ld3 >= c->nz + 1, QP_ERR_ARG,
               "qp_field3d_upload: leading dims too small");
    int dim = f->dim;
    size_t pl = c->rplane, st = (size_t)(c->nz + 1) * pl;
    std::vector<double> h(dim * st, 0.0);
    for (int d = 0; d < dim; ++d)
        for (int z = 0; z <= c->nz; ++z)
            for (int y = 0; y <= c->ny; ++y)
                for (int x = 0; x <= c->nx; ++x)
                    h[d * st + z * pl + (size_t)y * c->ldr + x] =
                        host[d + (size_t)dim * (x + (size_t)ld1 * (y + (size_t)ld2 * z))];
    if (c->prec == QP_FP64) {
        QP_HIP(hipMemcpyAsync(f->buf.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, c->stream));
        QP_HIP(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> hf(h.begin(), h.end());
        QP_HIP(hipMemcpyAsync(f->buf.p, hf.data(), hf.size() * 4, hipMemcpyHostToDevice, c->stream));
        QP_HIP(hipStreamSynchronize(c->stream));
    }
    return QP_OK;
}

int qp_field3d_download(const qp_field3d *f, double *host, int ld1, int ld2, int ld3)
{
    QP_REQUIRE(f && f->ctx && host, QP_ERR_ARG, "qp_field3d_download: bad argument");
    qp_ctx *c = f->ctx;
    QP_REQUIRE(ld1 >= c->nx + 1 && ld2 >= c->ny + 1 && ld3 >= c->nz + 1, QP_ERR_ARG,
               "qp_field3d_download: leading dims too small");
    int dim = f->dim;
    size_t pl = c->rplane, st = (size_t)(c->nz + 1) * pl;
    std::vector<double> h(dim * st);
    if (c->prec == QP_FP64) {
        QP_HIP(hipMemcpyAsync(h.data(), f->buf.p, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
        QP_HIP(hipStreamSynchronize(c->stream));
    } else {
        std::vector<float> hf(h.size());
        QP_HIP(hipMemcpyAsync(hf.data(), f->buf.p, hf.size() * 4, hipMemcpyDeviceToHost, c->stream));
        QP_HIP(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < h.size(); ++i) h[i] = hf[i];
    }
    for (int d = 0; d < dim; ++d)
        for (int z = 0; z <= c->nz; ++z)
            for (int y = 0; y <= c->ny; ++y)
                for (int x = 0; x <= c->nx; ++x)
                    host[d + (size_t)dim * (x + (size_t)ld1 * (y + (size_t)ld2 * z))] =
                        h[d * st + z * pl + (size_t)y * c->ldr + x];
    return QP_OK;
}

int qp_field3d_download_plane(const qp_field3d *f, int comp, int lpos, double *host, int ld1,
                              int ld2)
{
    QP_REQUIRE(f && f->ctx && host, QP_ERR_ARG, "qp_field3d_download_plane: bad argument");
    qp_ctx *c = f->ctx;
    QP_REQUIRE(comp >= 1 && comp <= f->dim && lpos >= 1 && lpos <= c->nz + 1, QP_ERR_ARG,
               "qp_field3d_download_plane: comp %d lpos %d", comp, lpos);
    QP_REQUIRE(ld1 >= c->nx + 1 && ld2 >= c->ny + 1, QP_ERR_ARG,
               "qp_field3d_download_plane: leading dims too small");
    size_t pl = c->rplane, st = (size_t)(c->nz + 1) * pl;
    size_t off = (comp - 1) * st + (size_t)(lpos - 1) * pl;
    return c->prec == QP_FP64
               ? download_planes<double>(c, f->buf.as<double>() + off, host, 1, c->ny + 1,
                                         c->nx + 1, c->ldr, pl, ld1)
               : download_planes<float>(c, f->buf.as<float>() + off, host, 1, c->ny + 1,
                                        c->nx + 1, c->ldr, pl, ld1);
}

}  // extern "C"
