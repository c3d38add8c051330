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
// packed as the real and imaginary part of the M = 2n point odd (S) or even
// (C) extension and transformed by one complex FFT.  For the even extension
// X(k) = 2 C(k) is real; for the odd one X(k) = -2i S(k); so
//   C: Re Z = 2 C_a, Im Z = 2 C_b;   S: Re Z = 2 S_b, Im Z = -2 S_a.
// The FFT is a Stockham autosort with radix-8 butterflies held in registers
// (M/8 threads per transform, 8 points each) and one closing radix-2 or
// radix-4 pass when log2(M) is not a multiple of 3.  The first pass reads the
// extension straight from memory, the last writes the n+1 outputs straight
// to memory, and the passes between exchange through LDS; the twiddle table
// entries are fetched before the first butterfly (three per radix-8 pass,
// the other powers formed by complex products).  A batch holds up to 8 components of
// any fields (grid.y); an LDS-tiled transpose sits between the x and the y
// pass.
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
// -i z
template <typename R>
__device__ __forceinline__ cplx<R> cmi(cplx<R> a) { return {a.y, -a.x}; }

// forward DFT of 4 points in place (natural order)
template <typename R>
__device__ __forceinline__ void dft4(cplx<R> &a0, cplx<R> &a1, cplx<R> &a2, cplx<R> &a3)
{
    cplx<R> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = cmi(csub(a1, a3));
    a0 = cadd(t0, t2);
    a1 = cadd(t1, t3);
    a2 = csub(t0, t2);
    a3 = csub(t1, t3);
}

// forward DFT of 8 points in place (natural order): decimation in frequency,
// X[2m] = DFT4(v[n] + v[n+4]), X[2m+1] = DFT4((v[n] - v[n+4]) w^n), w = e^{-i pi/4}
template <typename R>
__device__ __forceinline__ void dft8(cplx<R> *v)
{
    const R h = (R)0.70710678118654752440;
    cplx<R> a0 = cadd(v[0], v[4]), a1 = cadd(v[1], v[5]), a2 = cadd(v[2], v[6]),
            a3 = cadd(v[3], v[7]);
    cplx<R> c0 = csub(v[0], v[4]), c1 = csub(v[1], v[5]), c2 = csub(v[2], v[6]),
            c3 = csub(v[3], v[7]);
    c1 = {(c1.x + c1.y) * h, (c1.y - c1.x) * h};     // * e^{-i pi/4}
    c2 = cmi(c2);                                      // * e^{-i pi/2}
    c3 = {(c3.y - c3.x) * h, -(c3.x + c3.y) * h};    // * e^{-3i pi/4}
    dft4(a0, a1, a2, a3);
    dft4(c0, c1, c2, c3);
    v[0] = a0; v[2] = a1; v[4] = a2; v[6] = a3;
    v[1] = c0; v[3] = c1; v[5] = c2; v[7] = c3;
}

// LDS index of FFT point i: one pad point after every 2^PS points
template <int PS> __device__ __forceinline__ int padk(int i) { return i + (i >> PS); }

template <int L, int PS = 3> struct FftShape {
    static constexpr int M = 1 << L;                // complex points
    static constexpr int N = M / 2;                 // transform length n
    static constexpr int TF = M / 8;                // threads per transform
    static constexpr int BS = TF > 256 ? TF : 256;  // threads per block
    static constexpr int FB = BS / TF;              // transforms per block
    static constexpr int NR8 = L / 3;               // radix-8 passes
    static constexpr int REM = L % 3;               // closing radix-2 (1) / radix-4 (2)
    static constexpr int LDSM = M + (M >> PS) + 1;  // padded LDS points per transform
};

template <typename R> struct CompSet {
    const R *src[8];
    R *dst[8];
};

// Transform rows [0, nrows) of every component of the batch along their
// contiguous axis; row pair p of component blockIdx.y is transform p.  src
// may equal dst (a block reads its rows completely before writing).  If
// guard_row >= 0 the result of row 0 is also written to row guard_row.
// COL = 1: the transform rows are the column pairs of a row-major source
// (element e of row r at src[e * ld_src + r]); COL = 2: the results are
// written as column pairs of a row-major destination (dst[k * ld_dst + r]).
// Either way a lane moves both rows of its pair with one 2-element access,
// and block b takes pair (b mod 8) * (gridDim.x / 8) + b / 8, so the blocks
// on one XCD (dispatched round-robin by block id) share cache lines.
template <typename R, int L, int COL, int PS>
__global__ void __launch_bounds__(FftShape<L>::BS)
k_fft_rows(CompSet<R> cs, int nrows, int ld_src, int ld_dst, int types, R scale,
           const cplx<R> *__restrict__ W, int guard_row, const int *skip, ResJob rj)
{
    using S = FftShape<L, PS>;
    if (skip && *skip) {
        // a pass skipped after convergence (skip is the residual job's
        // flag_in) still passes the converged flag on to the next pass
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && rj.nblk > 0 && rj.flag_out)
            *rj.flag_out = 1;
        return;
    }
    if (blockIdx.x == 0 && blockIdx.y == 0) run_res_job(rj);
    constexpr int M = S::M, n = S::N, TF = S::TF;
    extern __shared__ __attribute__((aligned(16))) unsigned char smem_raw[];
    const int comp = blockIdx.y;
    const int type = (types >> comp) & 1;   // 0 sine, 1 cosine
    const int f = threadIdx.x / TF, t = threadIdx.x % TF;
    const int bx = COL ? (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3) : blockIdx.x;
    const int ra = 2 * (bx * S::FB + f);
    const bool live = ra < nrows, hb = ra + 1 < nrows;
    cplx<R> *buf = reinterpret_cast<cplx<R> *>(smem_raw) + f * S::LDSM;

    // first pass: radix 8 on the extension, straight from memory.  The 16
    // data loads are issued first and unconditionally (a missing second row
    // re-reads the first, a block past the last row reads row 0; the values
    // are then discarded), so one memory latency covers them and the twiddle
    // loads issued after them.
    R la[8], lb[8];
    {
        const R *base = cs.src[comp];
        const R *sa = live ? base + (COL == 1 ? (size_t)ra : (size_t)ra * ld_src) : base;
        // rows are read as rows except by the column-pair reader (COL = 1)
        const R *sb = (COL != 1 && live && hb) ? sa + ld_src : sa;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int idx = t + r * TF;
            const int e = idx <= n ? idx : M - idx;
            if (COL == 1) {
                // the pair of columns ra, ra+1 of row e (ra even, ld_src even)
                const cplx<R> p2 = *reinterpret_cast<const cplx<R> *>(sa + (size_t)e * ld_src);
                la[r] = p2.x;
                lb[r] = p2.y;
            } else {
                la[r] = sa[e];
                lb[r] = sb[e];
            }
        }
    }

    // twiddles of every pass after the first: pass p (Ns = 8^p) needs
    // W^(r e), e = (t mod Ns) M / (8 Ns), r = 1..7: W^e, W^2e and W^4e come
    // from the table, the other four are their products; the closing radix-4
    // pass needs W^j, W^2j (table) and W^3j, the closing radix-2 pass W^j
    constexpr int NTW8 = S::NR8 > 1 ? 3 * (S::NR8 - 1) : 1;
    cplx<R> tw8[NTW8];
#pragma unroll
    for (int p = 1; p < S::NR8; ++p) {
        const int Ns = 1 << (3 * p);
        const int e = (t & (Ns - 1)) * (M / (8 * Ns));
        tw8[3 * (p - 1)] = W[e];
        tw8[3 * (p - 1) + 1] = W[2 * e];
        tw8[3 * (p - 1) + 2] = W[4 * e];
    }
    constexpr int NTWR = S::REM == 2 ? 4 : (S::REM == 1 ? 4 : 1);
    cplx<R> twr[NTWR];
    if (S::REM == 2) {
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            twr[2 * s] = W[t + s * TF];
            twr[2 * s + 1] = W[2 * (t + s * TF)];
        }
    } else if (S::REM == 1) {
#pragma unroll
        for (int s = 0; s < 4; ++s) twr[s] = W[t + s * TF];
    }

    cplx<R> v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int idx = t + r * TF;
        const int e = idx <= n ? idx : M - idx;
        R va = live ? la[r] : R(0);
        R vb = (live && hb) ? lb[r] : R(0);
        if (type == 0) {
            const bool edge = e == 0 || e == n;
            const R sg = idx > n ? R(-1) : R(1);
            va = edge ? R(0) : sg * va;
            vb = edge ? R(0) : sg * vb;
        }
        v[r] = {va, vb};
    }
    dft8(v);

    const R half = scale * R(0.5);
    R *da = cs.dst[comp] + (COL == 2 ? (size_t)ra : (size_t)ra * ld_dst);
    R *db = da + ld_dst;
    R *dg = (guard_row >= 0 && ra == 0) ? cs.dst[comp] + (size_t)guard_row * ld_dst : nullptr;
    auto emit = [&](int k, cplx<R> z) {
        if (!live || k > n) return;
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
        if (COL == 2) {
            R *d = da + (size_t)k * ld_dst;
            if (hb) *reinterpret_cast<cplx<R> *>(d) = cplx<R>{oa, ob};
            else d[0] = oa;
        } else {
            da[k] = oa;
            if (hb) db[k] = ob;
            if (dg) dg[k] = oa;
        }
    };

    if (S::NR8 == 1 && S::REM == 0) {   // M = 8 is below the supported sizes
#pragma unroll
        for (int r = 0; r < 8; ++r) emit(t + r * TF, v[r]);
        return;
    }
    // first pass output (Ns = 1): index 8 t + r
#pragma unroll
    for (int r = 0; r < 8; ++r) buf[padk<PS>(8 * t + r)] = v[r];

    // middle radix-8 passes
#pragma unroll
    for (int p = 1; p < S::NR8; ++p) {
        const int Ns = 1 << (3 * p);
        __syncthreads();
        {
            const cplx<R> w1 = tw8[3 * (p - 1)], w2 = tw8[3 * (p - 1) + 1],
                          w4 = tw8[3 * (p - 1) + 2];
            const cplx<R> w3 = cmul(w1, w2);
            const cplx<R> w[8] = {w1, w1, w2, w3, w4, cmul(w1, w4), cmul(w2, w4), cmul(w3, w4)};
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                v[r] = buf[padk<PS>(t + r * TF)];
                if (r) v[r] = cmul(v[r], w[r]);
            }
        }
        dft8(v);
        if (p == S::NR8 - 1 && S::REM == 0) {
            // last pass (Ns = M/8): output index t + r M/8 goes to memory
#pragma unroll
            for (int r = 0; r < 8; ++r) emit(t + r * TF, v[r]);
            return;
        }
        __syncthreads();
        const int base = (t / Ns) * Ns * 8 + (t & (Ns - 1));
#pragma unroll
        for (int r = 0; r < 8; ++r) buf[padk<PS>(base + r * Ns)] = v[r];
    }
    __syncthreads();
    if (S::REM == 2) {
        // closing radix 4 (Ns = M/4): butterflies t and t + TF
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int j = t + s * TF;
            const cplx<R> w1 = twr[2 * s], w2 = twr[2 * s + 1];
            cplx<R> a0 = buf[padk<PS>(j)];
            cplx<R> a1 = cmul(buf[padk<PS>(j + M / 4)], w1);
            cplx<R> a2 = cmul(buf[padk<PS>(j + M / 2)], w2);
            cplx<R> a3 = cmul(buf[padk<PS>(j + 3 * M / 4)], cmul(w1, w2));
            dft4(a0, a1, a2, a3);
            emit(j, a0);
            emit(j + M / 4, a1);
            emit(j + M / 2, a2);
        }
    } else {
        // closing radix 2 (Ns = M/2): butterflies t + s TF
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const int j = t + s * TF;
            cplx<R> a0 = buf[padk<PS>(j)];
            cplx<R> a1 = cmul(buf[padk<PS>(j + M / 2)], twr[s]);
            emit(j, cadd(a0, a1));
            emit(j + M / 2, csub(a0, a1));
        }
    }
}

// dst(c)[col][row] = src(c)[row][col] for row < rows, col < cols, by TS x TS
// tiles through LDS; each of the 256 threads moves TS*TS/256 elements, all
// of its loads issued before the first LDS store
template <typename R, int TS>
__global__ void __launch_bounds__(256)
k_transpose(CompSet<R> cs, int ld_src, int ld_dst, int rows, int cols, const int *skip)
{
    constexpr int RS = 256 / TS, K = TS / RS;   // rows per sweep, sweeps
    __shared__ R tile[TS][TS + 1];
    if (skip && *skip) return;
    const int comp = blockIdx.z;
    const int c0 = blockIdx.x * TS, r0 = blockIdx.y * TS;
    const int tx = threadIdx.x % TS, ty = threadIdx.x / TS;
    const R *s = cs.src[comp];
    R *d = cs.dst[comp];
    R v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = ty + k * RS, r = r0 + i, c = c0 + tx;
        v[k] = (r < rows && c < cols) ? s[(size_t)r * ld_src + c] : R(0);
    }
#pragma unroll
    for (int k = 0; k < K; ++k) tile[ty + k * RS][tx] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int i = ty + k * RS, c = c0 + i, r = r0 + tx;
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

template <typename R, int L, int COL, int PS>
static int rows_launch(qp_ctx *c, const CompSet<R> &cs, int ncomp, int nrows, int ld_src,
                       int ld_dst, int types, double scale, const cplx<R> *W, int guard_row)
{
    using S = FftShape<L, PS>;
    const size_t lds = (size_t)S::FB * S::LDSM * sizeof(cplx<R>);
    QP_REQUIRE(lds <= c->max_lds, QP_ERR_UNSUPPORTED,
               "transform of length %d needs %zu B of LDS (limit %zu)", S::N, lds, c->max_lds);
    QP_REQUIRE(!COL || guard_row < 0, QP_ERR_ARG, "column transforms take no guard row");
    auto kern = k_fft_rows<R, L, COL, PS>;
    if (lds > 65536)
        QP_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
    const int pairs = (nrows + 1) / 2;
    int gx = (pairs + S::FB - 1) / S::FB;
    if (COL) gx = (gx + 7) & ~7;   // whole rounds of the 8 XCDs
    dim3 grid(gx, ncomp);
    const ResJob rj = c->res_job;
    c->res_job = ResJob{};
    {
    ProfScope _ps(c, QP_PROF_FFT, 2.0 * ncomp * nrows * (S::N + 1.0) * sizeof(R));
    hipLaunchKernelGGL(kern, grid, dim3(S::BS), lds, c->stream, cs, nrows, ld_src, ld_dst, types,
                       (R)scale, W, guard_row, c->skip, rj);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

template <typename R, int COL, int PS>
static int rowfft(qp_ctx *c, const CompSet<R> &cs, int ncomp, int nrows, int n, int ld_src,
                  int ld_dst, int types, double scale, const cplx<R> *W, int guard_row)
{
    switch (ilog2(2 * n)) {
    case 5: return rows_launch<R, 5, COL, PS>(c, cs, ncomp, nrows, ld_src, ld_dst, types, scale, W, guard_row);
    case 6: return rows_launch<R, 6, COL, PS>(c, cs, ncomp, nrows, ld_src, ld_dst, types, scale, W, guard_row);
    case 7: return rows_launch<R, 7, COL, PS>(c, cs, ncomp, nrows, ld_src, ld_dst, types, scale, W, guard_row);
    case 8: return rows_launch<R, 8, COL, PS>(c, cs, ncomp, nrows, ld_src, ld_dst, types, scale, W, guard_row);
    case 9: return rows_launch<R, 9, COL, PS>(c, cs, ncomp, nrows, ld_src, ld_dst, types, scale, W, guard_row);
    case 10: return rows_launch<R, 10, COL, PS>(c, cs, ncomp, nrows, ld_src, ld_dst, types, scale, W, guard_row);
    case 11: return rows_launch<R, 11, COL, PS>(c, cs, ncomp, nrows, ld_src, ld_dst, types, scale, W, guard_row);
    case 12: return rows_launch<R, 12, COL, PS>(c, cs, ncomp, nrows, ld_src, ld_dst, types, scale, W, guard_row);
    case 13: return rows_launch<R, 13, COL, PS>(c, cs, ncomp, nrows, ld_src, ld_dst, types, scale, W, guard_row);
    }
    set_error("transform length %d unsupported", n);
    return QP_ERR_UNSUPPORTED;
}

template <typename R>
static int transpose(qp_ctx *c, const CompSet<R> &cs, int ncomp, int ld_src, int ld_dst,
                     int rows, int cols)
{
    // 64 x 64 tiles in FP32 (16 elements per thread), 32 x 32 in FP64
    constexpr int TS = sizeof(R) == 4 ? 64 : 32;
    dim3 grid((cols + TS - 1) / TS, (rows + TS - 1) / TS, ncomp);
    {
    ProfScope _ps(c, QP_PROF_TRANSPOSE, 2.0 * ncomp * (double)rows * cols * sizeof(R));
    auto kern = k_transpose<R, TS>;
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, c->stream, cs, ld_src, ld_dst, rows, cols,
                       c->skip);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

// LDS padding of the transforms by default: one pad point per 16 points in
// FP32 (a third fewer LDS bank conflicts than per 8 at 2048 points, measured
// on MI355X), per 8 in FP64
template <typename R> constexpr int pad_shift() { return sizeof(R) == 4 ? 4 : 3; }

// the row transforms of the transpose path; QP_TUNE_FFT values 2-4 select
// another LDS padding (one pad point per 8, 32 or 4 points)
template <typename R>
static int rowfft0(qp_ctx *c, const CompSet<R> &cs, int ncomp, int nrows, int n, int ld_src,
                   int ld_dst, int types, double scale, const cplx<R> *W, int guard_row)
{
    constexpr int PD = pad_shift<R>();
    switch (c->tune[QP_TUNE_FFT]) {
    case 2: return rowfft<R, 0, 3>(c, cs, ncomp, nrows, n, ld_src, ld_dst, types, scale, W, guard_row);
    case 3: return rowfft<R, 0, 5>(c, cs, ncomp, nrows, n, ld_src, ld_dst, types, scale, W, guard_row);
    case 4: return rowfft<R, 0, 2>(c, cs, ncomp, nrows, n, ld_src, ld_dst, types, scale, W, guard_row);
    default: return rowfft<R, 0, PD>(c, cs, ncomp, nrows, n, ld_src, ld_dst, types, scale, W, guard_row);
    }
}

// components of a batch of fields: plane pointers and type masks
template <typename R>
static int gather_comps(qp_field2d *const *fs, const int *kinds, int nf, R **rs, R **ks,
                        int *mx, int *my)
{
    int nc = 0;
    *mx = *my = 0;
    for (int i = 0; i < nf; ++i) {
        qp_field2d *f = fs[i];
        int tx[3], ty[3];
        kind_types(kinds[i], f->dim, tx, ty);
        for (int d = 0; d < f->dim; ++d) {
            QP_REQUIRE(nc < MAX_BATCH, QP_ERR_ARG, "transform batch exceeds %d components",
                       MAX_BATCH);
            rs[nc] = f->rs.as<R>() + d * f->ctx->rplane;
            ks[nc] = f->ks.as<R>() + d * f->ctx->kplane;
            *mx |= tx[d] << nc;
            *my |= ty[d] << nc;
            ++nc;
        }
    }
    return nc;
}

// real -> k of nc components: rs[i] (real planes, pitch ldr) to ks[i]
// (k planes, pitch ldk); sr: nc scratch planes of (ny+1) x ldr for the x
// pass; mx, my: per-component sine/cosine bit masks; W: twiddles of 2nx and
// 2ny points
template <typename R>
static int r2k_core(qp_ctx *c, R *const *rs, R *const *ks, int nc, int mx, int my, R *sr,
                    int ldr, int ldk, const cplx<R> *twx, const cplx<R> *twy)
{
    const size_t pl = (size_t)(c->ny + 1) * ldr;
    CompSet<R> a{}, b{}, y{};
    for (int i = 0; i < nc; ++i) {
        a.src[i] = rs[i];
        a.dst[i] = sr + i * pl;
        b.src[i] = a.dst[i];
        b.dst[i] = ks[i];
        y.src[i] = ks[i];
        y.dst[i] = ks[i];
    }
    const double ani = 1.0 / ((double)c->nx * (double)c->ny);
    int rc = rowfft0<R>(c, a, nc, c->ny + 1, c->nx, ldr, ldr, mx, ani, twx, -1);
    if (rc) return rc;
    // y pass straight from the column pairs of the x pass (QP_TUNE_FFT 1, or 6:
    // this direction only)
    if (c->tune[QP_TUNE_FFT] == 1 || c->tune[QP_TUNE_FFT] == 6)
        return rowfft<R, 1, pad_shift<R>()>(c, b, nc, c->nx + 1, c->ny, ldr, ldk, my, 1.0,
                                            twy, -1);
    rc = transpose<R>(c, b, nc, ldr, ldk, c->ny + 1, c->nx + 1);
    if (!rc) rc = rowfft0<R>(c, y, nc, c->nx + 1, c->ny, ldk, ldk, my, 1.0, twy, -1);
    return rc;
}

// k -> real of nc components, ks[i] to rs[i]; sk: nc scratch planes of
// (nx+1) x ldk for the y pass (unused by the column-pair form)
template <typename R>
static int k2r_core(qp_ctx *c, R *const *rs, R *const *ks, int nc, int mx, int my, R *sk,
                    int ldr, int ldk, const cplx<R> *twx, const cplx<R> *twy)
{
    const size_t pl = (size_t)(c->nx + 1) * ldk;
    CompSet<R> y{}, b{}, a{}, yc{};
    for (int i = 0; i < nc; ++i) {
        y.src[i] = ks[i];
        y.dst[i] = sk + i * pl;
        b.src[i] = y.dst[i];
        b.dst[i] = rs[i];
        a.src[i] = rs[i];
        a.dst[i] = rs[i];
        yc.src[i] = ks[i];
        yc.dst[i] = rs[i];
    }
    int rc;
    // y pass written as column pairs of rs (QP_TUNE_FFT 1, or 5: this
    // direction only)
    if (c->tune[QP_TUNE_FFT] == 1 || c->tune[QP_TUNE_FFT] == 5) {
        rc = rowfft<R, 2, pad_shift<R>()>(c, yc, nc, c->nx + 1, c->ny, ldk, ldr, my, 2.0, twy,
                                          -1);
    } else {
        rc = rowfft0<R>(c, y, nc, c->nx + 1, c->ny, ldk, ldk, my, 2.0, twy, -1);
        if (!rc) rc = transpose<R>(c, b, nc, ldk, ldr, c->nx + 1, c->ny + 1);
    }
    // rows 0..ny-1; row ny receives row 0 (guard copy, PPNCGUARD2L nvp = 1)
    if (!rc) rc = rowfft0<R>(c, a, nc, c->ny, c->nx, ldr, ldr, mx, 2.0, twx, c->ny);
    return rc;
}

template <typename R>
static int r2k_t(qp_ctx *c, qp_field2d *const *fs, const int *kinds, int nf)
{
    R *rs[MAX_BATCH], *ks[MAX_BATCH];
    int mx, my;
    const int nc = gather_comps<R>(fs, kinds, nf, rs, ks, &mx, &my);
    if (nc <= 0) return nc;
    return r2k_core<R>(c, rs, ks, nc, mx, my, c->scr_r.as<R>(), c->ldr, c->ldk,
                       c->twx.as<cplx<R>>(), c->twy.as<cplx<R>>());
}

template <typename R>
static int k2r_t(qp_ctx *c, qp_field2d *const *fs, const int *kinds, int nf)
{
    R *rs[MAX_BATCH], *ks[MAX_BATCH];
    int mx, my;
    const int nc = gather_comps<R>(fs, kinds, nf, rs, ks, &mx, &my);
    if (nc <= 0) return nc;
    return k2r_core<R>(c, rs, ks, nc, mx, my, c->scr_k.as<R>(), c->ldr, c->ldk,
                       c->twx.as<cplx<R>>(), c->twy.as<cplx<R>>());
}

int fft_r2k_batch(qp_ctx *c, qp_field2d *const *fs, const int *kinds, int nf)
{
    return c->prec == QP_FP64 ? r2k_t<double>(c, fs, kinds, nf) : r2k_t<float>(c, fs, kinds, nf);
}

int fft_k2r_batch(qp_ctx *c, qp_field2d *const *fs, const int *kinds, int nf)
{
    return c->prec == QP_FP64 ? k2r_t<double>(c, fs, kinds, nf) : k2r_t<float>(c, fs, kinds, nf);
}

int fft_r2k(qp_field2d *f, int kind) { return fft_r2k_batch(f->ctx, &f, &kind, 1); }

int fft_k2r(qp_field2d *f, int kind) { return fft_k2r_batch(f->ctx, &f, &kind, 1); }

}  // namespace qp

