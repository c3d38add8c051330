// Sine / cosine transforms of field2d%fftrk / fftkr on the GPU.
//
// One-dimensional definitions on nodes f(0..n) (the reference's x and y
// passes, fft2d_lib77.f:4791-5495 and the multi-component variants):
//   S(k) = sum_{j=1}^{n-1} f(j) sin(pi j k / n),            S(0) = S(n) = 0
//   C(k) = f(0)/2 + (-1)^k f(n)/2 + sum_{j=1}^{n-1} f(j) cos(pi j k / n)
// real -> k (isign = -1): x pass times 1/(nx*ny), then y pass;
// k -> real (isign = +1): y pass times 2, x pass times 2, then the guard copy
// of fftkr (field2d_class.f03:223): row ny <- row 0.
//
// GPU method (not the reference's recursion): two rows of the same type are
// packed as the real and imaginary part of the 2n-point odd (S) or even (C)
// extension and transformed by one complex FFT held in LDS (Stockham autosort,
// radix 4 with one radix-2 pass when log2(2n) is odd).  For the even extension
// X(k) = 2 C(k) is real; for the odd one X(k) = -2i S(k); so
//   C: Re Z = 2 C_a, Im Z = 2 C_b;   S: Re Z = 2 S_b, Im Z = -2 S_a.
// A pass reads its rows coalesced, transforms them in LDS and writes them
// back; an LDS-tiled transpose sits between the x and the y pass.
#include "qp_common.h"

namespace qp {

template <typename R>
__device__ __forceinline__ cplx<R> cmul(cplx<R> a, cplx<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ cplx<R> cadd(cplx<R> a, cplx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R>
__device__ __forceinline__ cplx<R> csub(cplx<R> a, cplx<R> b) { return {a.x - b.x, a.y - b.y}; }

// Transform rows [0, nrows) of every component along their contiguous axis.
// grid.x = row pairs, grid.y = component.  src may equal dst.  If guard_row
// >= 0 the result of row 0 is also written to row guard_row.
template <typename R>
__global__ void __launch_bounds__(256)
k_rowfft(const R *__restrict__ src, size_t src_cs, int ld_src, R *dst, size_t dst_cs,
         int ld_dst, int nrows, int n, int log2m, int types, R scale,
         const cplx<R> *__restrict__ W, int guard_row)
{
    extern __shared__ unsigned char smem_raw[];
    const int M = 2 * n;
    cplx<R> *A = reinterpret_cast<cplx<R> *>(smem_raw);
    cplx<R> *B = A + M;
    const int comp = blockIdx.y;
    const int type = (types >> comp) & 1;   // 0 sine, 1 cosine
    const int ra = 2 * blockIdx.x;
    const bool hb = ra + 1 < nrows;
    const R *sa = src + comp * src_cs + (size_t)ra * ld_src;
    const R *sb = sa + ld_src;
    for (int j = threadIdx.x; j < M; j += blockDim.x) {
        int idx = j <= n ? j : M - j;
        R va = sa[idx];
        R vb = hb ? sb[idx] : R(0);
        if (type == 0) {
            if (idx == 0 || idx == n) { va = R(0); vb = R(0); }
            if (j > n) { va = -va; vb = -vb; }
        }
        A[j] = {va, vb};
    }
    __syncthreads();
    cplx<R> *in = A, *out = B;
    int Ns = 1;
    if (log2m & 1) {
        for (int jj = threadIdx.x; jj < M / 2; jj += blockDim.x) {
            cplx<R> a0 = in[jj], a1 = in[jj + M / 2];
            out[2 * jj] = cadd(a0, a1);
            out[2 * jj + 1] = csub(a0, a1);
        }
        __syncthreads();
        cplx<R> *t = in; in = out; out = t;
        Ns = 2;
    }
    const int q = M / 4;
    while (Ns < M) {
        const int stride = M / (4 * Ns);
        for (int jj = threadIdx.x; jj < q; jj += blockDim.x) {
            const int k = jj & (Ns - 1);
            cplx<R> a0 = in[jj], a1 = in[jj + q], a2 = in[jj + 2 * q], a3 = in[jj + 3 * q];
            if (k) {
                a1 = cmul(a1, W[k * stride]);
                a2 = cmul(a2, W[2 * k * stride]);
                a3 = cmul(a3, W[3 * k * stride]);
            }
            cplx<R> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3);
            cplx<R> d = csub(a1, a3);
            cplx<R> t3 = {d.y, -d.x};   // -i * (a1 - a3)
            const int base = (jj - k) * 4 + k;
            out[base] = cadd(t0, t2);
            out[base + Ns] = cadd(t1, t3);
            out[base + 2 * Ns] = csub(t0, t2);
            out[base + 3 * Ns] = csub(t1, t3);
        }
        __syncthreads();
        cplx<R> *t = in; in = out; out = t;
        Ns *= 4;
    }
    const R half = scale * R(0.5);
    R *da = dst + comp * dst_cs + (size_t)ra * ld_dst;
    R *db = da + ld_dst;
    R *dg = guard_row >= 0 && ra == 0 ? dst + comp * dst_cs + (size_t)guard_row * ld_dst : nullptr;
    for (int k = threadIdx.x; k <= n; k += blockDim.x) {
        cplx<R> z = in[k];
        R oa, ob;
        if (type == 1) {
            oa = z.x * half;
            ob = z.y * half;
        } else if (k == 0 || k == n) {
            oa = R(0);
            ob = R(0);
        } else {
            oa = -z.y * half;
            ob = z.x * half;
        }
        da[k] = oa;
        if (hb) db[k] = ob;
        if (dg) dg[k] = oa;
    }
}

// dst(c)[col][row] = src(c)[row][col] for row < rows, col < cols
template <typename R>
__global__ void __launch_bounds__(256)
k_transpose(const R *__restrict__ src, size_t src_cs, int ld_src, R *__restrict__ dst,
            size_t dst_cs, int ld_dst, int rows, int cols)
{
    __shared__ R tile[32][33];
    const int comp = blockIdx.z;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const R *s = src + comp * src_cs;
    R *d = dst + comp * dst_cs;
    for (int i = ty; i < 32; i += 8) {
        int r = r0 + i, c = c0 + tx;
        if (r < rows && c < cols) tile[i][tx] = s[(size_t)r * ld_src + c];
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        int c = c0 + i, r = r0 + tx;
        if (r < rows && c < cols) d[(size_t)c * ld_dst + r] = tile[tx][i];
    }
}

void kind_types(int kind, int dim, int *tx, int *ty)
{
    // fft2d_class.f03:121-413 and the X/Y passes of the WPPF*RM* wrappers:
    // kind 1: (S,S) | (C,S),(S,C) | (C,S),(S,C),(S,S)
    // kind 2: (C,C) | (S,C),(C,S) | (S,C),(C,S),(C,C)
    // kind 3: (S,S),(S,S),(C,C)
    static const int k1[3][3][2] = {{{0, 0}}, {{1, 0}, {0, 1}}, {{1, 0}, {0, 1}, {0, 0}}};
    static const int k2[3][3][2] = {{{1, 1}}, {{0, 1}, {1, 0}}, {{0, 1}, {1, 0}, {1, 1}}};
    static const int k3[3][2] = {{0, 0}, {0, 0}, {1, 1}};
    for (int c = 0; c < dim; ++c) {
        if (kind == 1) { tx[c] = k1[dim - 1][c][0]; ty[c] = k1[dim - 1][c][1]; }
        else if (kind == 2) { tx[c] = k2[dim - 1][c][0]; ty[c] = k2[dim - 1][c][1]; }
        else { tx[c] = k3[c][0]; ty[c] = k3[c][1]; }
    }
}

static int ilog2(int v) { int l = 0; while ((1 << l) < v) ++l; return l; }

template <typename R>
static int rowfft(qp_ctx *c, const R *src, size_t src_cs, int ld_src, R *dst, size_t dst_cs,
                  int ld_dst, int dim, int nrows, int n, const int *types, double scale,
                  const cplx<R> *W, int guard_row)
{
    const int M = 2 * n;
    const size_t lds = 2 * (size_t)M * sizeof(cplx<R>);
    QP_REQUIRE(lds <= c->max_lds, QP_ERR_UNSUPPORTED,
               "transform of length %d needs %zu B of LDS (limit %zu)", n, lds, c->max_lds);
    int tmask = 0;
    for (int i = 0; i < dim; ++i) tmask |= (types[i] & 1) << i;
    dim3 grid((nrows + 1) / 2, dim);
    {
    ProfScope _ps(c, QP_PROF_FFT, 2.0 * dim * nrows * (n + 1.0) * sizeof(R));
    hipLaunchKernelGGL(k_rowfft<R>, grid, dim3(256), lds, c->stream, src, src_cs, ld_src, dst,
                       dst_cs, ld_dst, nrows, n, ilog2(M), tmask, (R)scale, W, guard_row);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

template <typename R>
static int transpose(qp_ctx *c, const R *src, size_t src_cs, int ld_src, R *dst, size_t dst_cs,
                     int ld_dst, int dim, int rows, int cols)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32, dim);
    {
    ProfScope _ps(c, QP_PROF_TRANSPOSE, 2.0 * dim * (double)rows * cols * sizeof(R));
    hipLaunchKernelGGL(k_transpose<R>, grid, dim3(256), 0, c->stream, src, src_cs, ld_src, dst,
                       dst_cs, ld_dst, rows, cols);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

template <typename R>
static int r2k_t(qp_field2d *f, int kind)
{
    qp_ctx *c = f->ctx;
    int tx[3], ty[3];
    kind_types(kind, f->dim, tx, ty);
    R *rs = f->rs.as<R>(), *ks = f->ks.as<R>(), *sr = c->scr_r.as<R>();
    const double ani = 1.0 / ((double)c->nx * (double)c->ny);
    int rc = rowfft<R>(c, rs, c->rplane, c->ldr, sr, c->rplane, c->ldr, f->dim, c->ny + 1,
                       c->nx, tx, ani, c->twx.as<cplx<R>>(), -1);
    if (!rc) rc = transpose<R>(c, sr, c->rplane, c->ldr, ks, c->kplane, c->ldk, f->dim,
                               c->ny + 1, c->nx + 1);
    if (!rc) rc = rowfft<R>(c, ks, c->kplane, c->ldk, ks, c->kplane, c->ldk, f->dim, c->nx + 1,
                            c->ny, ty, 1.0, c->twy.as<cplx<R>>(), -1);
    return rc;
}

template <typename R>
static int k2r_t(qp_field2d *f, int kind)
{
    qp_ctx *c = f->ctx;
    int tx[3], ty[3];
    kind_types(kind, f->dim, tx, ty);
    R *rs = f->rs.as<R>(), *ks = f->ks.as<R>(), *sk = c->scr_k.as<R>();
    int rc = rowfft<R>(c, ks, c->kplane, c->ldk, sk, c->kplane, c->ldk, f->dim, c->nx + 1,
                       c->ny, ty, 2.0, c->twy.as<cplx<R>>(), -1);
    if (!rc) rc = transpose<R>(c, sk, c->kplane, c->ldk, rs, c->rplane, c->ldr, f->dim,
                               c->nx + 1, c->ny + 1);
    // rows 0..ny-1; row ny receives row 0 (guard copy, PPNCGUARD2L nvp = 1)
    if (!rc) rc = rowfft<R>(c, rs, c->rplane, c->ldr, rs, c->rplane, c->ldr, f->dim, c->ny,
                            c->nx, tx, 2.0, c->twx.as<cplx<R>>(), c->ny);
    return rc;
}

int fft_r2k(qp_field2d *f, int kind)
{
    return f->ctx->prec == QP_FP64 ? r2k_t<double>(f, kind) : r2k_t<float>(f, kind);
}

int fft_k2r(qp_field2d *f, int kind)
{
    return f->ctx->prec == QP_FP64 ? k2r_t<double>(f, kind) : k2r_t<float>(f, kind);
}

}  // namespace qp
