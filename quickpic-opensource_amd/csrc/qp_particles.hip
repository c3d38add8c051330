// Plasma species on the GPU: initial load, sort by cell, charge deposit,
// psi gather, current/acceleration/momentum-flux deposit and the push.
//
// Particles are a structure of arrays sorted by a tile-major cell key (see
// qp_common.h).  Every particle kernel is cell-centric: one workgroup per
// 16x16-cell tile and one thread per cell.  A thread reads the four corner
// values of the fields its cell needs once, then walks the cell's particles
// (a contiguous run of the sorted arrays), so the per-particle gathers need no
// memory traffic and the per-particle deposits accumulate in registers.  The
// four corner sums of every cell are combined into the tile's 17x17 nodes in
// LDS in a fixed order; nodes inside the tile are written directly and the 64
// border nodes go to a halo buffer that a second kernel sums, again in a fixed
// order, with the one-rank guard fold of PPNACGUARD2L
// (ufield2d_lib.f03:90-97).  No atomics touch the deposits.
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstring>
#include <type_traits>

#include "qp_common.h"

struct qp_species {
    qp_ctx *ctx;
    qp_species_params p;
    double qbm, dt;
    int64_t np;
    qp::DevBuf a[8], b[8];     // x, y, px, py, pz, p6, p7, q (current, alternate)
    qp::DevBuf key, slot;      // per particle: destination bin and rank (push), key and index (load)
    qp::DevBuf bins;           // per tile: counts, then starts, of the 9 x 256 destination bins
    qp::DevBuf count, off;     // per cell (ncells + 1)
    qp::DevBuf off_new;        // the push's new cell offsets, swapped with off
    // every particle has charge q_value (a uniform plasma, or an upload that
    // says so): q is then filled in both buffers and not moved, staged or read
    bool q_uniform;
    double q_value;
    qp::DevBuf scan_tmp, err;
    size_t scan_bytes;
    qp_field2d *q, *qn, *cu, *dcu, *amu;
    // deferred move check of the slice loop: pinned copy of err and its event
    int *h_err;
    hipEvent_t err_ev;
    bool err_pending;
    // piecewise longitudinal density (fdist2d_class.f03:233-251): points
    // s and factors fs; empty for a uniform plasma
    std::vector<double> lp_s, lp_fs;
    double den_factor = 1.0;   // the factor at the last load
    // transverse profile (0 uniform, 10, 11, 12; fdist2d_class.f03:318-950),
    // lengths in grid units
    int profile = 0;
    double pcx = 0, pcy = 0, irx = 0, iry = 0, orx = 0, ory = 0, r0 = 0, sigr = 1;
    std::vector<double> pr, pfr;
    std::string profile_err;   // a failed check of the last load
};

namespace qp {

template <typename R> struct Soa {
    R *x, *y, *px, *py, *pz, *p6, *p7, *q;
};

template <typename R> static Soa<R> soa(DevBuf *a)
{
    return Soa<R>{a[0].as<R>(), a[1].as<R>(), a[2].as<R>(), a[3].as<R>(),
                  a[4].as<R>(), a[5].as<R>(), a[6].as<R>(), a[7].as<R>()};
}

struct TileGeom {
    int nx, ny, mx1, my1, ldr;
};

// Particle positions.  FP64 records hold x and y in grid units, as the
// reference does.  FP32 records hold them relative to the origin of the
// particle's 16 x 16-cell tile (the tile is implied by the sorted order):
// a float near 1000 resolves 6e-5 of a cell, one in [0, 16) 1e-6, and the
// plasma's response ahead of a driver is a displacement of that order per
// slice.  Downloads and uploads convert at the boundary (in FP64), and the
// push re-bases a particle that changes tile.
template <typename R> struct TileRel {
    static constexpr bool on = std::is_same<R, float>::value;
};

// v mod m for -m <= v < 2m (a tile index plus a direction in {-1, 0, 1}):
// the periodic neighbour without an integer division
__device__ __forceinline__ int wrap1(int v, int m) { return v < 0 ? v + m : (v >= m ? v - m : v); }

__host__ __device__ __forceinline__ int cell_key(int cx, int cy, int mx1)
{
    return ((cy >> 4) * mx1 + (cx >> 4)) * TCELLS + ((cy & 15) << 4) + (cx & 15);
}
static inline int host_cell_key(int cx, int cy, int mx1) { return cell_key(cx, cy, mx1); }

// this thread's cell: key = blockIdx.x*256 + threadIdx.x
__device__ __forceinline__ void my_cell(const TileGeom &g, int &cx, int &cy)
{
    int t = blockIdx.x, l = threadIdx.x;
    cx = (t % g.mx1) * TILE + (l & 15);
    cy = (t / g.mx1) * TILE + (l >> 4);
}

template <typename R> struct Corners {
    R f00, f10, f01, f11;
};

template <typename R>
__device__ __forceinline__ Corners<R> load_corners(const R *plane, int ld, int cx, int cy)
{
    const R *r0 = plane + (size_t)cy * ld + cx, *r1 = r0 + ld;
    return Corners<R>{r0[0], r0[1], r1[0], r1[1]};
}

// bilinear gather in the reference's order (e.g. part2d_lib77.f:2265-2268)
template <typename R>
__device__ __forceinline__ R gather(const Corners<R> &f, R amx, R amy, R dxp, R dyp)
{
    R d = amx * f.f00;
    d = amy * (dxp * f.f10 + d);
    R a = amx * f.f01;
    return d + dyp * (dxp * f.f11 + a);
}

__device__ __forceinline__ int halo_index(int i, int j)
{
    if (j == 0) return i;
    if (j == TILE) return 17 + i;
    if (i == 0) return 34 + (j - 1);
    return 49 + (j - 1);
}

// Where a deposit's node sums go.  For component c and node value v:
//   t = v (+ inner[c][node]);  t *= scale;  p[c][node] = (outer[c][node] +) t
//   p2[c][node] = v * scale2                                (when p2[c] is set)
// With null inner/outer and unit scales this is a plain overwrite; the other
// settings restate the field additions and scalings that species2d and the
// slice loop apply right after a deposit (species2d_class.f03:166-218,
// simulation_class.f03:682-685), in the same order.
template <typename R> struct DepOut {
    R *p[8];
    const R *inner[8];
    const R *outer[8];
    R *p2[8];
    R scale, scale2;
};

template <typename R>
__device__ __forceinline__ void dep_store(const DepOut<R> &o, int c, size_t idx, R v)
{
    R t = v;
    if (o.inner[c]) t = t + o.inner[c][idx];
    t = t * o.scale;
    if (o.p2[c]) o.p2[c][idx] = v * o.scale2;
    o.p[c][idx] = o.outer[c] ? o.outer[c][idx] + t : t;
}

template <typename R> static DepOut<R> dep_plain(R *const *planes, int n)
{
    DepOut<R> o{};
    for (int c = 0; c < n; ++c) o.p[c] = planes[c];
    o.scale = o.scale2 = R(1);
    return o;
}

// Sum the tile-border contributions of every border node in a fixed order;
// with fold, row ny is added into row 0 and cleared (PPNACGUARD2L nvp = 1).
// One thread per border node and component (one thread taking every
// component of a node measured 1.5x slower: too few waves per CU).
template <typename R>
__global__ void k_deposit_border(TileGeom g, const R *__restrict__ halo, size_t halo_cs,
                                 DepOut<R> out, int fold, const int *skip)
{
    if (skip && *skip) return;
    const int comp = blockIdx.y;
    const int rows_n = (g.my1 + 1) * (g.nx + 1);
    const int cols_n = (g.mx1 + 1) * g.my1 * (TILE - 1);
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    int X, Y;
    if (id < rows_n) {
        Y = (id / (g.nx + 1)) * TILE;
        X = id % (g.nx + 1);
    } else if (id < rows_n + cols_n) {
        int k = id - rows_n;
        int per_col = g.my1 * (TILE - 1);
        X = (k / per_col) * TILE;
        int r = k % per_col;
        Y = (r / (TILE - 1)) * TILE + 1 + r % (TILE - 1);
    } else {
        return;
    }
    const R *h = halo + comp * halo_cs;
    auto node_sum = [&](int XX, int YY) -> R {
        R s = R(0);
        for (int tyc = (YY >> 4) - 1; tyc <= (YY >> 4); ++tyc) {
            if (tyc < 0 || tyc >= g.my1) continue;
            int ly = YY - TILE * tyc;
            if (ly < 0 || ly > TILE) continue;
            for (int txc = (XX >> 4) - 1; txc <= (XX >> 4); ++txc) {
                if (txc < 0 || txc >= g.mx1) continue;
                int lx = XX - TILE * txc;
                if (lx < 0 || lx > TILE) continue;
                s += h[(size_t)(tyc * g.mx1 + txc) * HALO + halo_index(lx, ly)];
            }
        }
        return s;
    };
    R v;
    if (fold && Y == g.ny) v = R(0);
    else if (fold && Y == 0) { v = node_sum(X, 0); v = v + node_sum(X, g.ny); }
    else v = node_sum(X, Y);
    dep_store(out, comp, (size_t)Y * g.ldr + X, v);
}

// LDS slot of staged particle p: one pad word every 32 entries spreads the
// per-cell strided reads over the banks
__device__ __forceinline__ int lpad(int p) { return p + (p >> 5); }

// 1/x: IEEE division in FP64 (the reference's arithmetic); the hardware
// reciprocal (1 ulp) in FP32, whose results are compared with FP64 at 1e-5
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }
// a / b and sqrt: IEEE in FP64 (the reference's arithmetic); the hardware
// reciprocal and square root (1 ulp) in FP32
__device__ __forceinline__ double divq(double a, double b) { return a / b; }
__device__ __forceinline__ float divq(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double sqrtq(double x) { return sqrt(x); }
__device__ __forceinline__ float sqrtq(float x) { return __builtin_amdgcn_sqrtf(x); }

template <int CH> struct Stage {
    static constexpr int N = CH + CH / 32 + 1;
};

// Stages particles [c0, c0 + n) of NA arrays into LDS (padded index lpad):
// thread tid takes p = tid + k*step, k < K.  Every load of the thread is
// issued before the first LDS store, so the chunk costs one memory latency.
template <typename R, int NA, int K>
__device__ __forceinline__ void stage_in(const R *const (&src)[NA], R *const (&dst)[NA], int c0,
                                         int n, int tid, int step)
{
    R t[K][NA];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = tid + k * step;
        if (p < n) {
#pragma unroll
            for (int a = 0; a < NA; ++a) t[k][a] = src[a][c0 + p];
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int p = tid + k * step;
        if (p < n) {
            const int l = lpad(p);
#pragma unroll
            for (int a = 0; a < NA; ++a) dst[a][l] = t[k][a];
        }
    }
}

// Interior nodes of the tile's 17x17 block to the output plane, border nodes
// to the halo buffer (summed over tiles by k_deposit_border).
template <typename R>
__device__ __forceinline__ void emit_nodes(const R *nacc, const DepOut<R> &out, int c, int ldr,
                                           R *halo, int tile, int tx, int ty)
{
    for (int n = threadIdx.x; n < 17 * 17; n += TCELLS) {
        const int i = n % 17, j = n / 17;
        const R v = nacc[n];
        if (i >= 1 && i < TILE && j >= 1 && j < TILE)
            dep_store(out, c, (size_t)(ty * TILE + j) * ldr + tx * TILE + i, v);
        else
            halo[(size_t)tile * HALO + halo_index(i, j)] = v;
    }
}

// Deposit skeleton.  The tile's cells are taken RG rows at a time; inside a
// group every cell owns T = 16/RG consecutive lanes (lane cl*T + sub), lane
// sub taking particles sub, sub+T, ... of its cell.  The group's particles
// (a contiguous run of the sorted arrays) are staged through LDS in chunks of
// at most CH with coalesced loads.  Per-lane register sums are reduced over
// the T lanes with xor shuffles and added to the tile's node accumulators in
// four phases (one per cell corner), so every sum has a fixed order and no
// atomic touches the deposit.

// PPGPPOST2L (part2d_lib77.f:194-309): q(n,m) += qm (1-dx)(1-dy) etc.
// UQ: every particle's charge is quni (not staged or read).
template <typename R, int RG, int CH, int WPE, bool UQ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_deposit_q(TileGeom g, Soa<R> s, const int *__restrict__ off, DepOut<R> out, R *halo, R quni)
{
    constexpr int T = TILE / RG, NP = Stage<CH>::N;
    extern __shared__ unsigned char smem_raw[];
    R *sx = reinterpret_cast<R *>(smem_raw), *sy = sx + NP, *sq = sy + NP, *nacc = sq + NP;
    const int tid = threadIdx.x, tile = blockIdx.x;
    const int tx = tile % g.mx1, ty = tile / g.mx1, kb = tile * TCELLS;
    for (int n = tid; n < 289; n += TCELLS) nacc[n] = R(0);
    const int cl = tid / T, sub = tid % T;
    for (int row = 0; row < TILE; row += RG) {
        const int lrow = row + cl / TILE, lcol = cl % TILE;
        const int key = kb + lrow * TILE + lcol;
        const int i0 = off[key], i1 = off[key + 1];
        const int G0 = off[kb + row * TILE], G1 = off[kb + (row + RG) * TILE];
        const R fx = (R)(TileRel<R>::on ? lcol : tx * TILE + lcol);
        const R fy = (R)(TileRel<R>::on ? lrow : ty * TILE + lrow);
        R a00 = 0, a10 = 0, a01 = 0, a11 = 0;
        for (int c0 = G0; c0 < G1; c0 += CH) {
            const int n = min(CH, G1 - c0);
            __syncthreads();
            if (UQ) {
                const R *src[2] = {s.x, s.y};
                R *dst[2] = {sx, sy};
                stage_in<R, 2, CH / TCELLS>(src, dst, c0, n, tid, TCELLS);
            } else {
                const R *src[3] = {s.x, s.y, s.q};
                R *dst[3] = {sx, sy, sq};
                stage_in<R, 3, CH / TCELLS>(src, dst, c0, n, tid, TCELLS);
            }
            __syncthreads();
            int a = i0 + sub;
            if (a < c0) a += ((c0 - a + T - 1) / T) * T;
            const int b = min(i1, c0 + n);
            for (int i = a; i < b; i += T) {
                const int l = lpad(i - c0);
                R qm = UQ ? quni : sq[l];
                R dxp = qm * (sx[l] - fx);
                R dyp = sy[l] - fy;
                R amx = qm - dxp;
                R amy = R(1) - dyp;
                a00 += amx * amy;
                a10 += dxp * amy;
                a01 += amx * dyp;
                a11 += dxp * dyp;
            }
        }
#pragma unroll
        for (int o = T / 2; o > 0; o >>= 1) {
            a00 += __shfl_xor(a00, o);
            a10 += __shfl_xor(a10, o);
            a01 += __shfl_xor(a01, o);
            a11 += __shfl_xor(a11, o);
        }
        const R av[4] = {a00, a10, a01, a11};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            __syncthreads();
            if (sub == 0) nacc[(lrow + (k >> 1)) * 17 + lcol + (k & 1)] += av[k];
        }
    }
    __syncthreads();
    emit_nodes(nacc, out, 0, g.ldr, halo, tile, tx, ty);
}

// Record words 6 and 7 (gamma and gamma - p_z, both 1 for a particle at
// rest).  FP64 stores them as the reference does.  FP32 stores p6 - 1 and
// p7 - 1: a float holding 1 + e keeps e only to 6e-8 absolute, so psi's
// imprint on p7 (e = -qbm psi dex^2) and the p_z formed from p6 and p7 (a
// difference of two numbers near 1/2, part2d_lib77.f:597-598) would carry
// errors of 6e-8/|e| relative, 1e-4 where psi is 1e-3.  The deviations are
// formed without that cancellation and are mathematically the same
// quantities; downloads and uploads convert at the boundary.
template <typename R> struct P67Dev {
    static constexpr bool on = std::is_same<R, float>::value;
};

// WPGPSIPOST2L_QP (part2d_lib77.f:2207-2284): p7 = 1 + dx and
// p6 = dx + (dx2 |v|^2 - dx^2 - 2 dx) / (2 (1 + dx)) + 1 for dx = -qbm psi dex^2;
// in the deviation form p7 - 1 = dx and p6 - 1 = (dx2 |v|^2 + dx^2) / (2 (1 + dx))
template <typename R>
__device__ __forceinline__ void p67_from_psi(R dx, R dx2, R vx, R vy, R &p6, R &p7)
{
    if constexpr (P67Dev<R>::on) {
        p7 = dx;
        p6 = (dx2 * (vx * vx + vy * vy) + dx * dx) / (R(2) * (R(1) + dx));
    } else {
        p7 = R(1) + dx;
        p6 = dx + (dx2 * (vx * vx + vy * vy) - dx * dx - R(2) * dx) / (R(2) * (R(1) + dx)) + R(1);
    }
}

template <typename R>
__global__ void __launch_bounds__(256)
k_extpsi(int64_t np, int nx, int ny, int ldr, Soa<R> s, const R *psi, R qbm, R dex,
         const int *__restrict__ off, int mx1, int ntiles)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= np) return;
    const R x = s.x[i], y = s.y[i];
    int cx, cy;
    R dxp, dyp;
    if constexpr (TileRel<R>::on) {
        // the particle's tile: the last one starting at or before i
        int lo = 0, hi = ntiles - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid * TCELLS] <= i) lo = mid;
            else hi = mid - 1;
        }
        const int lx = min(max((int)x, 0), TILE - 1), ly = min(max((int)y, 0), TILE - 1);
        cx = (lo % mx1) * TILE + lx;
        cy = (lo / mx1) * TILE + ly;
        dxp = x - (R)lx;
        dyp = y - (R)ly;
    } else {
        cx = min(max((int)x, 0), nx - 1);
        cy = min(max((int)y, 0), ny - 1);
        dxp = x - (R)cx;
        dyp = y - (R)cy;
    }
    const Corners<R> f = load_corners(psi, ldr, cx, cy);
    const R dx2 = dex * dex;
    R amx = R(1) - dxp, amy = R(1) - dyp;
    R vx = s.px[i], vy = s.py[i];
    R dx = gather(f, amx, amy, dxp, dyp);
    dx = -(dx * dx2 * qbm);
    R p6, p7;
    p67_from_psi(dx, dx2, vx, vy, p6, p7);
    s.p7[i] = p7;
    s.p6[i] = p6;
}

// WPGPSIPOST2L_QP for one particle of the cell at (fx, fy): p7 and p6 from
// the psi gather, in k_extpsi's operation order (the first corrector pass
// of a slice applies it just before amjdp_one)
template <typename R>
__device__ __forceinline__ void extpsi_one(const Corners<R> &f, R fx, R fy, R x, R y, R vx, R vy,
                                           R qbm, R dex, R &p6, R &p7)
{
    const R dx2 = dex * dex;
    R dxp = x - fx, dyp = y - fy;
    R amx = R(1) - dxp, amy = R(1) - dyp;
    R dx = gather(f, amx, amy, dxp, dyp);
    dx = -(dx * dx2 * qbm);
    p67_from_psi(dx, dx2, vx, vy, p6, p7);
}

// the six fields amjdp and push gather, at the four corners of one cell
template <typename R> struct Fields6 {
    Corners<R> ex, ey, ez, bx, by, bz;
};

template <typename R>
__device__ __forceinline__ Fields6<R> load_fields6(const R *fxy, const R *bxy, const R *psit,
                                                   size_t cs, int ld, int cx, int cy)
{
    return Fields6<R>{load_corners(fxy, ld, cx, cy),      load_corners(fxy + cs, ld, cx, cy),
                      load_corners(psit, ld, cx, cy),     load_corners(bxy, ld, cx, cy),
                      load_corners(bxy + cs, ld, cx, cy), load_corners(bxy + 2 * cs, ld, cx, cy)};
}

// FP32: a cell's corner values rewritten once as bilinear coefficients,
// f(dxp, dyp) = c0 + dxp c1 + dyp (c2 + dxp c3), so each gather of the cell's
// particles is three FMAs; FP64 keeps the corner values and the reference's
// operation order (gather above)
template <typename R>
__device__ __forceinline__ void prep_fields(Fields6<R> &F)
{
    if constexpr (std::is_same<R, float>::value) {
        Corners<R> *c[6] = {&F.ex, &F.ey, &F.ez, &F.bx, &F.by, &F.bz};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const R f00 = c[k]->f00, f10 = c[k]->f10, f01 = c[k]->f01, f11 = c[k]->f11;
            *c[k] = Corners<R>{f00, f10 - f00, f01 - f00, (f11 - f10) - (f01 - f00)};
        }
    }
}

template <typename R>
__device__ __forceinline__ R cell_gather(const Corners<R> &f, R amx, R amy, R dxp, R dyp)
{
    if constexpr (std::is_same<R, float>::value)
        return fmaf(dyp, fmaf(dxp, f.f11, f.f01), fmaf(dxp, f.f10, f.f00));
    else
        return gather(f, amx, amy, dxp, dyp);
}

// One particle of PPGRDCJPPOST2L_QP (part2d_lib77.f:311-768), in the
// reference's operation order: gather, half Boris step with the B_z
// rotation, the 8 deposited quantities times the 4 corner weights added to
// acc; returns the new p6 (part2d_lib77.f:597).
template <typename R>
__device__ __forceinline__ R amjdp_one(const Fields6<R> &F, R fx, R fy, R x, R y, R px, R py,
                                       R p6, R p7, R qm, R qtmh, R qbm, R dti, R dex, R idex,
                                       R (&acc)[8][4])
{
    constexpr bool DEV = P67Dev<R>::on;
    R dxp = x - fx, dyp = y - fy;
    R amx = R(1) - dxp, amy = R(1) - dyp;
    R inv_part_7 = recip(DEV ? R(1) + p7 : p7);
    R qtmh1 = qtmh * inv_part_7;
    R qtmh2 = DEV ? qtmh1 + qtmh1 * p6 : qtmh1 * p6;
    R dx = cell_gather(F.ex, amx, amy, dxp, dyp);
    R dy = cell_gather(F.ey, amx, amy, dxp, dyp);
    R dz = cell_gather(F.ez, amx, amy, dxp, dyp);
    R ox = cell_gather(F.bx, amx, amy, dxp, dyp);
    R oy = cell_gather(F.by, amx, amy, dxp, dyp);
    R oz = cell_gather(F.bz, amx, amy, dxp, dyp);
    R ddx = (R(-1)) * qtmh2 * dx + qtmh * oy;
    R ddy = (R(-1)) * qtmh2 * dy - qtmh * ox;
    R vx = px, vy = py;
    R acx = vx + ddx, acy = vy + ddy;
    R omzt = qtmh1 * oz;
    R omt = omzt * omzt;
    R anorm = R(2) * recip(R(1) + omt);
    R rot1 = R(0.5) * (R(1) - omt);
    R rot2 = omzt;
    R v1 = (rot1 * acx + rot2 * acy) * anorm + ddx;
    R v2 = (rot1 * acy - rot2 * acx) * anorm + ddy;
    R qm1 = qm * inv_part_7;
    amx = qm1 * amx;
    dxp = qm1 * dxp;
    ox = R(0.5) * (v1 + vx);
    oy = R(0.5) * (v2 + vy);
    // p_z = (1 + |p_perp|^2)/(2 p7) - p7/2; with p7 = 1 + e7 it is
    // (|p_perp|^2 - e7 (2 + e7)) / (2 p7), formed without the cancellation
    if constexpr (DEV)
        oz = R(0.5) * (dex * dex * (ox * ox + oy * oy) - p7 * (R(2) + p7)) * inv_part_7;
    else
        oz = R(0.5) * (R(1) + dex * dex * (ox * ox + oy * oy)) * inv_part_7 - R(0.5) * p7;
    const R p6n = p7 + oz;   // gamma = (gamma - p_z) + p_z (both forms)
    oz = oz * idex;
    vx = (v1 - vx) * dti;
    vy = (v2 - vy) * dti;
    dz = qbm * (dz + (dx * ox + dy * oy) * dex * dex * inv_part_7);
    const R w[4] = {amx * amy, dxp * amy, amx * dyp, dxp * dyp};
    v1 = ox * ox * inv_part_7;
    v2 = oy * oy * inv_part_7;
    R v3 = ox * oy * inv_part_7;
    vx = vx + ox * dz * inv_part_7;
    vy = vy + oy * dz * inv_part_7;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        acc[0][k] += ox * w[k];
        acc[1][k] += oy * w[k];
        acc[2][k] += oz * w[k];
        acc[3][k] += vx * w[k];
        acc[4][k] += vy * w[k];
        acc[5][k] += v1 * w[k];
        acc[6][k] += v2 * w[k];
        acc[7][k] += v3 * w[k];
    }
    return p6n;
}

// PPGRDCJPPOST2L_QP (part2d_lib77.f:311-768): half Boris step with the B_z
// rotation, deposit of J (3), dJ/dxi (2) and momentum flux (3); p6 updated.
// PSI = false: a pass without the psi gather (psi is ignored), compiled
// without its corner registers and branches.  (Measured slower: spreading a
// chunk's cells over all four waves at half their lanes, 2 %; each lane
// storing its particles' new p6 itself instead of the copy-back, 1.6 %.)
// UQ: every particle's charge is quni (not staged).
// RG = 0: the rows per group are chosen per tile, as 16 / T for the smallest
// T (lanes per cell: 1, 2, 4, 8 or 16) with which the tile's particles,
// spread evenly over its T groups, stay within THR/4 staging chunks per
// group; a chunk then holds more of its group's cells than with a fixed RG
// in dense tiles (where the lanes of the cells a chunk does not hold wait),
// and a sparse tile takes fewer, larger groups.  (Compiling each T as its
// own path was measured slower: 443-444 against 454.6 slices/s at THR = 4.)
template <typename R, int RG, int CH, int WPE, bool PSI, bool UQ, int THR = 4>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, 8)))
k_amjdp(TileGeom g, Soa<R> s, const int *__restrict__ off, const R *fxy, const R *bxy,
        const R *psit, size_t cs, DepOut<R> out, R *halo, size_t halo_cs, R qbm, R dt, R dex,
        const int *skip, const R *psi_arg, R quni)
{
    const R *psi = PSI ? psi_arg : nullptr;
    if (skip && *skip) return;
    constexpr int NP = Stage<CH>::N;
    extern __shared__ unsigned char smem_raw[];
    R *sx = reinterpret_cast<R *>(smem_raw), *sy = sx + NP, *spx = sy + NP, *spy = spx + NP;
    R *sp6 = spy + NP, *sp7 = sp6 + NP, *sq = sp7 + NP, *nacc = sq + NP;   // nacc: 8 x 289
    const int tid = threadIdx.x, tile = blockIdx.x;
    const int tx = tile % g.mx1, ty = tile / g.mx1, kb = tile * TCELLS;
    for (int n = tid; n < 8 * 289; n += TCELLS) nacc[n] = R(0);
    int T = RG > 0 ? TILE / (RG > 0 ? RG : 1) : 1;
    if (RG == 0) {
        const int ntile = off[kb + TCELLS] - off[kb];
        while (T < TILE && 4 * ntile > THR * CH * T) T *= 2;
    }
    const int rg = TILE / T;   // rows per group
    const int cl = tid / T, sub = tid % T;
    const R qtmh = R(0.5) * qbm * dt, dti = R(1) / dt, idex = R(1) / dex;
    for (int row = 0; row < TILE; row += rg) {
        const int lrow = row + cl / TILE, lcol = cl % TILE;
        const int key = kb + lrow * TILE + lcol;
        const int i0 = off[key], i1 = off[key + 1];
        const int G0 = off[kb + row * TILE], G1 = off[kb + (row + rg) * TILE];
        const int cx = tx * TILE + lcol, cy = ty * TILE + lrow;
        Fields6<R> F = load_fields6(fxy, bxy, psit, cs, g.ldr, cx, cy);
        prep_fields(F);
        const Corners<R> FP = psi ? load_corners(psi, g.ldr, cx, cy) : Corners<R>{};
        const R fx = (R)(TileRel<R>::on ? lcol : cx), fy = (R)(TileRel<R>::on ? lrow : cy);
        R acc[8][4];
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[c][k] = R(0);
        for (int c0 = G0; c0 < G1; c0 += CH) {
            const int n = min(CH, G1 - c0);
            __syncthreads();
            if (UQ) {
                const R *src[6] = {s.x, s.y, s.px, s.py, s.p6, s.p7};
                R *dst[6] = {sx, sy, spx, spy, sp6, sp7};
                stage_in<R, 6, CH / TCELLS>(src, dst, c0, n, tid, TCELLS);
            } else {
                const R *src[7] = {s.x, s.y, s.px, s.py, s.p6, s.p7, s.q};
                R *dst[7] = {sx, sy, spx, spy, sp6, sp7, sq};
                stage_in<R, 7, CH / TCELLS>(src, dst, c0, n, tid, TCELLS);
            }
            __syncthreads();
            int a = i0 + sub;
            if (a < c0) a += ((c0 - a + T - 1) / T) * T;
            const int b = min(i1, c0 + n);
            for (int i = a; i < b; i += T) {
                const int l = lpad(i - c0);
                R p6 = sp6[l], p7 = sp7[l];
                if (psi) {
                    extpsi_one(FP, fx, fy, sx[l], sy[l], spx[l], spy[l], qbm, dex, p6, p7);
                    sp7[l] = p7;
                }
                sp6[l] = amjdp_one(F, fx, fy, sx[l], sy[l], spx[l], spy[l], p6, p7,
                                   UQ ? quni : sq[l], qtmh, qbm, dti, dex, idex, acc);
            }
            __syncthreads();
            for (int p = tid; p < n; p += TCELLS) s.p6[c0 + p] = sp6[lpad(p)];
            if (psi)
                for (int p = tid; p < n; p += TCELLS) s.p7[c0 + p] = sp7[lpad(p)];
        }
        for (int o = T / 2; o > 0; o >>= 1)
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int k = 0; k < 4; ++k) acc[c][k] += __shfl_xor(acc[c][k], o);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            __syncthreads();
            if (sub == 0) {
                const int node = (lrow + (k >> 1)) * 17 + lcol + (k & 1);
#pragma unroll
                for (int c = 0; c < 8; ++c) nacc[c * 289 + node] += acc[c][k];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 8; ++c)
        emit_nodes(nacc + c * 289, out, c, g.ldr, halo + c * halo_cs, tile, tx, ty);
}

// The same deposit with the next chunk's particles loaded into registers
// while the current chunk is computed, so a workgroup's global loads overlap
// its own arithmetic (the staging chunk of k_amjdp waits for its loads).
// The tile's particles are one contiguous run, row group after row group, so
// the chunk after the last of a group is the first of the next; a chunk that
// was not prefetched (after an empty group) is loaded when it is reached.
// Cells, lanes, sums and their order are those of k_amjdp with the same RG.
template <typename R, int RG, int CH>
__global__ void __launch_bounds__(256)
k_amjdp_pf(TileGeom g, Soa<R> s, const int *__restrict__ off, const R *fxy, const R *bxy,
           const R *psit, size_t cs, DepOut<R> out, R *halo, size_t halo_cs, R qbm, R dt, R dex,
           const int *skip, const R *psi, R)
{
    if (skip && *skip) return;
    constexpr int T = TILE / RG, NP = Stage<CH>::N, PF = CH / TCELLS;
    extern __shared__ unsigned char smem_raw[];
    R *sx = reinterpret_cast<R *>(smem_raw), *sy = sx + NP, *spx = sy + NP, *spy = spx + NP;
    R *sp6 = spy + NP, *sp7 = sp6 + NP, *sq = sp7 + NP, *nacc = sq + NP;   // nacc: 8 x 289
    const int tid = threadIdx.x, tile = blockIdx.x;
    const int tx = tile % g.mx1, ty = tile / g.mx1, kb = tile * TCELLS;
    for (int n = tid; n < 8 * 289; n += TCELLS) nacc[n] = R(0);
    const int cl = tid / T, sub = tid % T;
    const R qtmh = R(0.5) * qbm * dt, dti = R(1) / dt, idex = R(1) / dex;
    R rx[PF], ry[PF], rpx[PF], rpy[PF], rp6[PF], rp7[PF], rq[PF];
    int pc0 = -1, pn = 0;   // the chunk held in the registers
    auto fetch = [&](int b0, int nn) {
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const int p = tid + k * TCELLS;
            if (p < nn) {
                const int q = b0 + p;
                rx[k] = s.x[q];
                ry[k] = s.y[q];
                rpx[k] = s.px[q];
                rpy[k] = s.py[q];
                rp6[k] = s.p6[q];
                rp7[k] = s.p7[q];
                rq[k] = s.q[q];
            }
        }
        pc0 = b0;
        pn = nn;
    };
    for (int row = 0; row < TILE; row += RG) {
        const int lrow = row + cl / TILE, lcol = cl % TILE;
        const int key = kb + lrow * TILE + lcol;
        const int i0 = off[key], i1 = off[key + 1];
        const int G0 = off[kb + row * TILE], G1 = off[kb + (row + RG) * TILE];
        const int G2 = row + RG < TILE ? off[kb + (row + 2 * RG) * TILE] : G1;
        const int cx = tx * TILE + lcol, cy = ty * TILE + lrow;
        Fields6<R> F = load_fields6(fxy, bxy, psit, cs, g.ldr, cx, cy);
        prep_fields(F);
        const Corners<R> FP = psi ? load_corners(psi, g.ldr, cx, cy) : Corners<R>{};
        const R fx = (R)(TileRel<R>::on ? lcol : cx), fy = (R)(TileRel<R>::on ? lrow : cy);
        R acc[8][4];
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[c][k] = R(0);
        for (int c0 = G0; c0 < G1; c0 += CH) {
            const int n = min(CH, G1 - c0);
            if (pc0 != c0 || pn != n) fetch(c0, n);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const int p = tid + k * TCELLS;
                if (p < n) {
                    const int l = lpad(p);
                    sx[l] = rx[k];
                    sy[l] = ry[k];
                    spx[l] = rpx[k];
                    spy[l] = rpy[k];
                    sp6[l] = rp6[k];
                    sp7[l] = rp7[k];
                    sq[l] = rq[k];
                }
            }
            __syncthreads();
            // the next chunk: the rest of this group, or the start of the next
            const int nc0 = c0 + n;
            const int nn = nc0 < G1 ? min(CH, G1 - nc0) : min(CH, G2 - nc0);
            if (nn > 0) fetch(nc0, nn);
            else pc0 = -1;
            int a = i0 + sub;
            if (a < c0) a += ((c0 - a + T - 1) / T) * T;
            const int b = min(i1, c0 + n);
            for (int i = a; i < b; i += T) {
                const int l = lpad(i - c0);
                R p6 = sp6[l], p7 = sp7[l];
                if (psi) {
                    extpsi_one(FP, fx, fy, sx[l], sy[l], spx[l], spy[l], qbm, dex, p6, p7);
                    sp7[l] = p7;
                }
                sp6[l] = amjdp_one(F, fx, fy, sx[l], sy[l], spx[l], spy[l], p6, p7, sq[l], qtmh,
                                   qbm, dti, dex, idex, acc);
            }
            __syncthreads();
            for (int p = tid; p < n; p += TCELLS) s.p6[c0 + p] = sp6[lpad(p)];
            if (psi)
                for (int p = tid; p < n; p += TCELLS) s.p7[c0 + p] = sp7[lpad(p)];
        }
#pragma unroll
        for (int o = T / 2; o > 0; o >>= 1)
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int k = 0; k < 4; ++k) acc[c][k] += __shfl_xor(acc[c][k], o);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            __syncthreads();
            if (sub == 0) {
                const int node = (lrow + (k >> 1)) * 17 + lcol + (k & 1);
#pragma unroll
                for (int c = 0; c < 8; ++c) nacc[c * 289 + node] += acc[c][k];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 8; ++c)
        emit_nodes(nacc + c * 289, out, c, g.ldr, halo + c * halo_cs, tile, tx, ty);
}

// FP32 deposit with the particle chunks copied from memory into LDS by the
// LDS-DMA loads (global_load_lds_dword: each lane's dword lands at the wave's
// LDS base + 4 lane) into two buffers: the next chunk is in flight while the
// current one is computed, and the copy uses no registers.  Chunks hold CH
// particles; LDS index of particle p of a chunk: p + p/64 (one pad per
// wave block).  Cells, lanes, sums and their order are those of k_amjdp.
__device__ __forceinline__ int lpad64(int p) { return p + (p >> 6); }

template <int RG, int CH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 8)))
k_amjdp_dma(TileGeom g, Soa<float> s, const int *__restrict__ off, const float *fxy,
            const float *bxy, const float *psit, size_t cs, DepOut<float> out, float *halo,
            size_t halo_cs, float qbm, float dt, float dex, const int *skip, const float *psi,
            float)
{
    using R = float;
    if (skip && *skip) return;
    constexpr int T = TILE / RG, NB = CH + CH / 64, K = CH / TCELLS;
    extern __shared__ unsigned char smem_raw[];
    R *base = reinterpret_cast<R *>(smem_raw);          // [2 buffers][7 arrays][NB]
    R *nacc = base + 2 * 7 * NB;                         // 8 x 289
    const int tid = threadIdx.x, tile = blockIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tx = tile % g.mx1, ty = tile / g.mx1, kb = tile * TCELLS;
    for (int n = tid; n < 8 * 289; n += TCELLS) nacc[n] = R(0);
    const int cl = tid / T, sub = tid % T;
    const R qtmh = R(0.5) * qbm * dt, dti = R(1) / dt, idex = R(1) / dex;
    const R *src[7] = {s.x, s.y, s.px, s.py, s.p6, s.p7, s.q};
    // copy chunk [b0, b0 + nn) into buffer bf: wave w takes blocks k*256 + 64 w
    auto issue = [&](int b0, int nn, int bf) {
#pragma unroll
        for (int a = 0; a < 7; ++a) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int blk = k * TCELLS + wave * 64;
                if (blk + lane < nn)
                    __builtin_amdgcn_global_load_lds(
                        src[a] + b0 + blk + lane,
                        (__attribute__((address_space(3))) void *)(base + (bf * 7 + a) * NB +
                                                                   lpad64(blk)),
                        4, 0, 0);
            }
        }
    };
    int pc0 = -1, pn = 0, cur = 0;
    for (int row = 0; row < TILE; row += RG) {
        const int lrow = row + cl / TILE, lcol = cl % TILE;
        const int key = kb + lrow * TILE + lcol;
        const int i0 = off[key], i1 = off[key + 1];
        const int G0 = off[kb + row * TILE], G1 = off[kb + (row + RG) * TILE];
        const int G2 = row + RG < TILE ? off[kb + (row + 2 * RG) * TILE] : G1;
        const int cx = tx * TILE + lcol, cy = ty * TILE + lrow;
        Fields6<R> F = load_fields6(fxy, bxy, psit, cs, g.ldr, cx, cy);
        prep_fields(F);
        const Corners<R> FP = psi ? load_corners(psi, g.ldr, cx, cy) : Corners<R>{};
        const R fx = (R)(TileRel<R>::on ? lcol : cx), fy = (R)(TileRel<R>::on ? lrow : cy);
        R acc[8][4];
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[c][k] = R(0);
        for (int c0 = G0; c0 < G1; c0 += CH) {
            const int n = min(CH, G1 - c0);
            if (pc0 != c0 || pn != n) issue(c0, n, cur);
            __builtin_amdgcn_s_waitcnt(0);
            __syncthreads();
            // the next chunk: the rest of this group, or the start of the next
            const int nc0 = c0 + n;
            const int nn = nc0 < G1 ? min(CH, G1 - nc0) : min(CH, G2 - nc0);
            if (nn > 0) {
                issue(nc0, nn, cur ^ 1);
                pc0 = nc0;
                pn = nn;
            } else {
                pc0 = -1;
            }
            R *sx = base + (cur * 7 + 0) * NB, *sy = sx + NB, *spx = sy + NB, *spy = spx + NB;
            R *sp6 = spy + NB, *sp7 = sp6 + NB, *sq = sp7 + NB;
            int a = i0 + sub;
            if (a < c0) a += ((c0 - a + T - 1) / T) * T;
            const int b = min(i1, c0 + n);
            for (int i = a; i < b; i += T) {
                const int l = lpad64(i - c0);
                R p6 = sp6[l], p7 = sp7[l];
                if (psi) {
                    extpsi_one(FP, fx, fy, sx[l], sy[l], spx[l], spy[l], qbm, dex, p6, p7);
                    sp7[l] = p7;
                }
                sp6[l] = amjdp_one(F, fx, fy, sx[l], sy[l], spx[l], spy[l], p6, p7, sq[l], qtmh,
                                   qbm, dti, dex, idex, acc);
            }
            __syncthreads();
            for (int p = tid; p < n; p += TCELLS) s.p6[c0 + p] = sp6[lpad64(p)];
            if (psi)
                for (int p = tid; p < n; p += TCELLS) s.p7[c0 + p] = sp7[lpad64(p)];
            cur ^= 1;
        }
#pragma unroll
        for (int o = T / 2; o > 0; o >>= 1)
#pragma unroll
            for (int c = 0; c < 8; ++c)
#pragma unroll
                for (int k = 0; k < 4; ++k) acc[c][k] += __shfl_xor(acc[c][k], o);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            __syncthreads();
            if (sub == 0) {
                const int node = (lrow + (k >> 1)) * 17 + lcol + (k & 1);
#pragma unroll
                for (int c = 0; c < 8; ++c) nacc[c * 289 + node] += acc[c][k];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 8; ++c)
        emit_nodes(nacc + c * 289, out, c, g.ldr, halo + c * halo_cs, tile, tx, ty);
}

// Orders LDS accesses among the lanes of one wave (no workgroup barrier).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// The same deposit with the waves of a tile working independently: wave w
// takes the tile's row pairs w and w+4 (32 cells, two lanes per cell),
// stages that pair's particles through its own LDS area in chunks of CHW,
// and adds its cell sums into its own 3 x 17 node rows in four corner phases
// (wave-level ordering only).  One workgroup barrier at the end, then each
// tile node is the sum, in row-pair order, of the (one or two) row-pair node
// rows that hold it.  A heavy row pair delays only its own wave.
template <typename R, int CHW>
__global__ void __launch_bounds__(256)
k_amjdp_w(TileGeom g, Soa<R> s, const int *__restrict__ off, const R *fxy, const R *bxy,
          const R *psit, size_t cs, DepOut<R> out, R *halo, size_t halo_cs, R qbm, R dt, R dex,
          const int *skip, const R *psi, R)
{
    if (skip && *skip) return;
    constexpr int NP = Stage<CHW>::N, NR = 3 * 17;   // node rows of a row pair
    extern __shared__ unsigned char smem_raw[];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    R *base = reinterpret_cast<R *>(smem_raw);
    R *sx = base + wave * 7 * NP, *sy = sx + NP, *spx = sy + NP, *spy = spx + NP;
    R *sp6 = spy + NP, *sp7 = sp6 + NP, *sq = sp7 + NP;
    R *rnode = base + 4 * 7 * NP;                     // [row pair 8][comp 8][3][17]
    const int tile = blockIdx.x, tx = tile % g.mx1, ty = tile / g.mx1, kb = tile * TCELLS;
    for (int n = tid; n < 8 * 8 * NR; n += TCELLS) rnode[n] = R(0);
    __syncthreads();
    const int cl = lane >> 1, sub = lane & 1;
    const R qtmh = R(0.5) * qbm * dt, dti = R(1) / dt, idex = R(1) / dex;
    for (int it = 0; it < 2; ++it) {
        const int rp = wave + 4 * it;
        const int lrow = 2 * rp + (cl >> 4), lcol = cl & 15;
        const int key = kb + lrow * TILE + lcol;
        const int i0 = off[key], i1 = off[key + 1];
        const int G0 = off[kb + 2 * rp * TILE], G1 = off[kb + (2 * rp + 2) * TILE];
        const int cx = tx * TILE + lcol, cy = ty * TILE + lrow;
        Fields6<R> F = load_fields6(fxy, bxy, psit, cs, g.ldr, cx, cy);
        prep_fields(F);
        const Corners<R> FP = psi ? load_corners(psi, g.ldr, cx, cy) : Corners<R>{};
        const R fx = (R)(TileRel<R>::on ? lcol : cx), fy = (R)(TileRel<R>::on ? lrow : cy);
        R acc[8][4];
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[c][k] = R(0);
        for (int c0 = G0; c0 < G1; c0 += CHW) {
            const int n = min(CHW, G1 - c0);
            wave_sync();
            {
                const R *src[7] = {s.x, s.y, s.px, s.py, s.p6, s.p7, s.q};
                R *dst[7] = {sx, sy, spx, spy, sp6, sp7, sq};
                stage_in<R, 7, CHW / 64>(src, dst, c0, n, lane, 64);
            }
            wave_sync();
            int a = i0 + sub;
            if (a < c0) a += ((c0 - a + 1) / 2) * 2;
            const int b = min(i1, c0 + n);
            for (int i = a; i < b; i += 2) {
                const int l = lpad(i - c0);
                R p6 = sp6[l], p7 = sp7[l];
                if (psi) {
                    extpsi_one(FP, fx, fy, sx[l], sy[l], spx[l], spy[l], qbm, dex, p6, p7);
                    sp7[l] = p7;
                }
                sp6[l] = amjdp_one(F, fx, fy, sx[l], sy[l], spx[l], spy[l], p6, p7, sq[l], qtmh,
                                   qbm, dti, dex, idex, acc);
            }
            wave_sync();
            for (int p = lane; p < n; p += 64) s.p6[c0 + p] = sp6[lpad(p)];
            if (psi)
                for (int p = lane; p < n; p += 64) s.p7[c0 + p] = sp7[lpad(p)];
        }
#pragma unroll
        for (int c = 0; c < 8; ++c)
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[c][k] += __shfl_xor(acc[c][k], 1);
        R *rn = rnode + rp * 8 * NR;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            wave_sync();
            if (sub == 0) {
                const int node = ((cl >> 4) + (k >> 1)) * 17 + lcol + (k & 1);
#pragma unroll
                for (int c = 0; c < 8; ++c) rn[c * NR + node] += acc[c][k];
            }
        }
    }
    __syncthreads();
    for (int n = tid; n < 17 * 17; n += TCELLS) {
        const int i = n % 17, j = n / 17;
        const int rp = j >> 1;                 // row pair whose node row j%2 is node row j
        const bool top = (j & 1) == 0 && j > 0;   // also row 2 of row pair rp-1
        for (int c = 0; c < 8; ++c) {
            R v;
            if (j == 16) v = rnode[(7 * 8 + c) * NR + 2 * 17 + i];
            else {
                v = rnode[(rp * 8 + c) * NR + (j & 1) * 17 + i];
                if (top) v = rnode[((rp - 1) * 8 + c) * NR + 2 * 17 + i] + v;
            }
            if (i >= 1 && i < TILE && j >= 1 && j < TILE)
                dep_store(out, c, (size_t)(ty * TILE + j) * g.ldr + tx * TILE + i, v);
            else
                halo[c * halo_cs + (size_t)tile * HALO + halo_index(i, j)] = v;
        }
    }
}

// PPGRBPPUSHF23L_QP (part2d_lib77.f:1066-1408): 3-D Boris push in xi, then the
// periodic wrap of the reorder (PPPORDER2LB :1936-1977).  One workgroup per
// source tile walks the tile's particles (a contiguous run of the sorted
// arrays) 256 at a time, one thread per particle.  A particle leaves for the
// neighbour tile in its direction of motion (dir in {-1,0,1}^2, bin
// nb = 3 (diry+1) + dirx + 1) and lands in cell cl of it; its rank among the
// tile's particles bound for the same (nb, cl) follows their order in the
// tile: lanes of a wave rank by lane order, waves add their counts to the
// tile's LDS histogram in turn.  So the new order is a function of the old
// one, with no atomics: the sort, and with it every later sum, is the same
// run to run.  The histogram goes to cnt[tile][nb][cl].  A particle whose new
// tile is not its old tile plus its direction (the condition PPPCHECK2L
// reports) sets *err and is kept in its own tile.
//
// MODE 0 stores the pushed particle in place with its bin and rank (k_scatter
// then moves it).  MODE 1 stores nothing per particle: only the histogram,
// counted with LDS integer atomics (the same counts in any order; no ranks).
// MODE 2 runs after MODE 1 and k_bin_starts (cnt then holds each bin's start
// within its cell) and the scan of the new cell counts into off_new: it
// repeats the push on the unchanged input, so bins and ranks are those of
// MODE 1, and stores each particle at off_new[cell] + start + rank of the
// alternate arrays o.  The two passes read the particle record twice and
// write it once, where push and scatter read it twice and write it twice.
constexpr int NBINS = 9 * TCELLS;

template <typename R, int MODE>
__global__ void __launch_bounds__(256)
k_push(TileGeom g, Soa<R> s, Soa<R> o, const int *__restrict__ off,
       const int *__restrict__ off_new, const R *fxy, const R *bxy, const R *psit, size_t cs,
       R qbm, R dt, R dex, int *bin, int *rank, int *cnt, int *err, int64_t np, int uq)
{
    __shared__ int hist[NBINS];
    __shared__ R fnode[6][289];   // the six fields at the tile's 17 x 17 nodes
    const int S = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = tid; k < NBINS; k += TCELLS) hist[k] = 0;
    const int tx0 = S % g.mx1, ty0 = S / g.mx1;
    const int p0 = off[S * TCELLS], p1 = off[(S + 1) * TCELLS];
    // the particle record of chunk c0 is loaded while chunk c0 - 256 is ranked
    R nx_ = R(0), ny_ = R(0), npx = R(0), npy = R(0), npz = R(0), np6 = R(1), np7 = R(1);
    R nq = R(0);
    auto fetch = [&](int c) {
        const int i = c + tid;
        if (i < p1) {
            nx_ = s.x[i];
            ny_ = s.y[i];
            npx = s.px[i];
            npy = s.py[i];
            npz = s.pz[i];
            np6 = s.p6[i];
            np7 = s.p7[i];
            if (MODE == 2 && !uq) nq = s.q[i];
        }
    };
    fetch(p0);
    // every particle of the tile lies in one of its cells (the sort put it
    // there), so the corners it gathers are tile nodes: read them once
    {
        constexpr int NF = (6 * 289 + TCELLS - 1) / TCELLS;
        R tv[NF];
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const int n = tid + k * TCELLS;
            if (n < 6 * 289) {
                const int f = n / 289, m = n % 289;
                const R *pl = f == 0 ? fxy : f == 1 ? fxy + cs : f == 2 ? psit
                            : f == 3 ? bxy : f == 4 ? bxy + cs : bxy + 2 * cs;
                tv[k] = pl[(size_t)(ty0 * TILE + m / 17) * g.ldr + tx0 * TILE + m % 17];
            }
        }
#pragma unroll
        for (int k = 0; k < NF; ++k) {
            const int n = tid + k * TCELLS;
            if (n < 6 * 289) fnode[n / 289][n % 289] = tv[k];
        }
    }
    __syncthreads();
    const R idex = R(1) / dex, qtmh = R(0.5) * qbm * dt;
    const R anx = (R)g.nx, any = (R)g.ny;
    const R edgelx = (R)(tx0 * TILE), edgely = (R)(ty0 * TILE);
    const R edgerx = edgelx + (R)min(TILE, g.nx - tx0 * TILE);
    const R edgery = edgely + (R)min(TILE, g.ny - ty0 * TILE);
    for (int c0 = p0; c0 < p1; c0 += TCELLS) {
        const int i = c0 + tid;
        const bool valid = i < p1;
        const R x = nx_, y = ny_, ppx = npx, ppy = npy, ppz = npz, p6 = np6, p7 = np7, q = nq;
        if (c0 + TCELLS < p1) fetch(c0 + TCELLS);
        int b = 0, dcell = 0;
        R ox_ = R(0), oy_ = R(0), opx = R(0), opy = R(0), opz = R(0);
        if (valid) {
            // absolute cell (FP64) or cell within the tile (FP32 tile-relative)
            const int cx = TileRel<R>::on ? min(max((int)x, 0), TILE - 1)
                                          : min(max((int)x, 0), g.nx - 1);
            const int cy = TileRel<R>::on ? min(max((int)y, 0), TILE - 1)
                                          : min(max((int)y, 0), g.ny - 1);
            // local node of the cell, clamped into the tile (a particle flagged
            // by the move check may sit outside it; its result is reported)
            const int li = TileRel<R>::on ? cx : min(max(cx - tx0 * TILE, 0), TILE - 1);
            const int lj = TileRel<R>::on ? cy : min(max(cy - ty0 * TILE, 0), TILE - 1);
            const int m = lj * 17 + li;
            auto lc = [&](int f) {
                return Corners<R>{fnode[f][m], fnode[f][m + 1], fnode[f][m + 17], fnode[f][m + 18]};
            };
            const Corners<R> ex = lc(0), ey = lc(1), ez = lc(2), bx = lc(3), by = lc(4), bz = lc(5);
            R qtmh1 = divq(qtmh, P67Dev<R>::on ? R(1) + p7 : p7);
            R qtmh2 = P67Dev<R>::on ? qtmh1 + qtmh1 * p6 : qtmh1 * p6;
            R dxp = x - (R)cx, dyp = y - (R)cy;
            R amx = R(1) - dxp, amy = R(1) - dyp;
            R dx = gather(ex, amx, amy, dxp, dyp);
            R dy = gather(ey, amx, amy, dxp, dyp);
            R dz = gather(ez, amx, amy, dxp, dyp);
            dz = dz * idex;
            R ox = gather(bx, amx, amy, dxp, dyp);
            R oy = gather(by, amx, amy, dxp, dyp);
            R oz = gather(bz, amx, amy, dxp, dyp);
            ox = ox * dex;
            oy = oy * dex;
            dx = qtmh2 * dx;
            dy = qtmh2 * dy;
            dz = qtmh2 * dz;
            R acx = ppx + dx, acy = ppy + dy, acz = ppz + dz;
            R omxt = qtmh1 * ox, omyt = qtmh1 * oy, omzt = qtmh1 * oz;
            R omt = omxt * omxt + omyt * omyt + omzt * omzt;
            R anorm = divq(R(2), R(1) + omt);
            omt = R(0.5) * (R(1) - omt);
            R rot4 = omxt * omyt, rot7 = omxt * omzt, rot8 = omyt * omzt;
            R rot1 = omt + omxt * omxt;
            R rot5 = omt + omyt * omyt;
            R rot9 = omt + omzt * omzt;
            R rot2 = omzt + rot4;
            rot4 = -omzt + rot4;
            R rot3 = -omyt + rot7;
            rot7 = omyt + rot7;
            R rot6 = omxt + rot8;
            rot8 = -omxt + rot8;
            dx = (rot1 * acx + rot2 * acy + rot3 * acz) * anorm + dx;
            dy = (rot4 * acx + rot5 * acy + rot6 * acz) * anorm + dy;
            dz = (rot7 * acx + rot8 * acy + rot9 * acz) * anorm + dz;
            if (MODE == 0) {
                s.px[i] = dx;
                s.py[i] = dy;
                s.pz[i] = dz;
            }
            opx = dx;
            opy = dy;
            opz = dz;
            R dtc1 = divq(dt, sqrtq(R(1) + (dx * dx + dy * dy + dz * dz) * dex * dex) - dz * dex);
            dx = x + dx * dtc1;
            dy = y + dy * dtc1;
            int dirx = 0, diry = 0, ncx, ncy;
            bool far = false;   // beyond the neighbouring tiles (PPPCHECK2L)
            if constexpr (TileRel<R>::on) {
                // A position that leaves [0, 16) moves to the neighbouring
                // tile and is re-based by 16 (exact for -16 <= x < 32).  One
                // just below 0 whose re-based value rounds to 16 stays at 0,
                // as the absolute form keeps a particle whose wrap rounds to
                // the box edge.  Tile indices wrap periodically below.
                const R T = (R)TILE;
                if (dx >= T) { dirx = 1; dx = dx - T; }
                else if (dx < R(0)) {
                    if (dx + T >= T) dx = R(0);
                    else { dirx = -1; dx = dx + T; }
                }
                if (dy >= T) { diry = 1; dy = dy - T; }
                else if (dy < R(0)) {
                    if (dy + T >= T) dy = R(0);
                    else { diry = -1; dy = dy + T; }
                }
                far = !(dx >= R(0) && dx < T && dy >= R(0) && dy < T);
                if (far && !(isfinite(dx) && isfinite(dy)) && MODE != 2) atomicMax(err, 1 << 30);
                ncx = far ? min(max((int)dx, 0), TILE - 1) : (int)dx;
                ncy = far ? min(max((int)dy, 0), TILE - 1) : (int)dy;
            } else {
                if (dx >= edgerx) dirx = 1;
                else if (dx < edgelx) {
                    dirx = -1;
                    if (dx < R(0) && dx + anx >= anx) { dx = R(0); dirx = 0; }
                }
                if (dy >= edgery) diry = 1;
                else if (dy < edgely) {
                    diry = -1;
                    if (dy < R(0) && dy + any >= any) { dy = R(0); diry = 0; }
                }
                if (dx < R(0)) dx = dx + anx;
                if (dx >= anx) dx = dx - anx;
                if (dy < R(0)) dy = dy + any;
                if (dy >= any) dy = dy - any;
                ncx = (int)dx;
                ncy = (int)dy;
                // a non-finite or out-of-range position must not become an index
                if (!(ncx >= 0 && ncx < g.nx && ncy >= 0 && ncy < g.ny)) {
                    if (MODE != 2) atomicMax(err, 1 << 30);
                    ncx = min(max(ncx, 0), g.nx - 1);
                    ncy = min(max(ncy, 0), g.ny - 1);
                }
            }
            if (MODE == 0) {
                s.x[i] = dx;
                s.y[i] = dy;
            }
            ox_ = dx;
            oy_ = dy;
            const int txd = wrap1(tx0 + dirx, g.mx1), tyd = wrap1(ty0 + diry, g.my1);
            // destination tile: the neighbour in the direction of motion, or,
            // for a particle the move check flags, its own tile
            int dt = tyd * g.mx1 + txd;
            if (!TileRel<R>::on) far = (ncx >> 4) != txd || (ncy >> 4) != tyd;
            if (far) {
                if (MODE != 2) atomicMax(err, 1 + txd + g.mx1 * tyd);
                dirx = diry = 0;
                dt = S;
            }
            b = (3 * (diry + 1) + dirx + 1) * TCELLS + ((ncy & 15) << 4) + (ncx & 15);
            dcell = dt * TCELLS + (b & 255);
        }
        if (MODE == 1) {
            // counting only (integer adds give the same counts in any order):
            // one LDS atomic per run of equal bins along the lanes, of the
            // run's length (valid lanes are a prefix of the wave)
            const int bprev = __shfl_up(b, 1);
            const bool head = valid && (lane == 0 || bprev != b);
            const unsigned long long heads = __ballot(head);
            const int nvalid = __popcll(__ballot(valid));
            const unsigned long long above = heads & ~((2ull << lane) - 1ull);
            const int end = above ? __ffsll((long long)above) - 1 : nvalid;
            if (head) atomicAdd(&hist[b], end - lane);
            continue;
        }
        // rank among this chunk's earlier particles of the same bin, in lane order
        unsigned long long todo = __ballot(valid);
        int rk = 0, nlead = 0;
        while (todo) {
            const int leader = __ffsll((long long)todo) - 1;
            const int lb = __shfl(b, leader);
            const unsigned long long m = __ballot(valid && b == lb);
            if (valid && b == lb) rk = __popcll(m & ((1ull << lane) - 1ull));
            if (lane == leader) nlead = __popcll(m);
            todo &= ~m;
        }
        // waves add to the tile histogram in order (the first barrier also
        // orders the last wave's update of the chunk before)
        for (int w = 0; w < TCELLS / 64; ++w) {
            __syncthreads();
            if (wave == w) {
                if (valid) rk += hist[b];
                if (nlead) hist[b] += nlead;
            }
        }
        if (MODE == 0 && valid) {
            bin[i] = S * NBINS + b;
            rank[i] = rk;
        }
        if (MODE == 2 && valid) {
            const int64_t d = (int64_t)off_new[dcell] + cnt[(size_t)S * NBINS + b] + rk;
            if (d >= 0 && d < np) {
                o.x[d] = ox_;
                o.y[d] = oy_;
                o.px[d] = opx;
                o.py[d] = opy;
                o.pz[d] = opz;
                o.p6[d] = p6;
                o.p7[d] = p7;
                if (!uq) o.q[d] = q;   // a uniform q is already in place
            } else {
                atomicMax(err, 1 << 29);   // the two passes disagree: never expected
            }
        }
    }
    if (MODE == 2) return;
    __syncthreads();
    for (int k = tid; k < NBINS; k += TCELLS) cnt[(size_t)S * NBINS + k] = hist[k];
}

// Per destination cell d (in tile T, local cell cl): its particles come from
// bins (T - dir(nb), nb, cl), nb = 0..8.  Taken in nb order, the bin counts
// are turned into each bin's start within the cell, and the cell count is
// written for the scan; count[ncells] = 0.
__global__ void k_bin_starts(TileGeom g, int *cnt, int *count)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncells = g.mx1 * g.my1 * TCELLS;
    if (d == ncells) count[d] = 0;
    if (d >= ncells) return;
    const int T = d >> 8, cl = d & 255;
    const int tx = T % g.mx1, ty = T / g.mx1;
    int run = 0;
    for (int nb = 0; nb < 9; ++nb) {
        const int dirx = nb % 3 - 1, diry = nb / 3 - 1;
        const int sx = wrap1(tx - dirx, g.mx1), sy = wrap1(ty - diry, g.my1);
        const size_t k = (size_t)(sy * g.mx1 + sx) * NBINS + nb * TCELLS + cl;
        const int c = cnt[k];
        cnt[k] = run;
        run += c;
    }
    count[d] = run;
}

// destination of particle i: off[cell] + start of its bin in the cell + rank
template <typename R>
__global__ void k_scatter(int64_t np, Soa<R> a, Soa<R> b, const int *bin, const int *rank,
                          const int *cnt, const int *off, int mx1, int my1)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= np) return;
    // the record is read completely, and first, before any store (the
    // compiler may not move a load of a past a store to b, which it cannot
    // prove disjoint); its loads overlap those of the destination
    const R v0 = a.x[i], v1 = a.y[i], v2 = a.px[i], v3 = a.py[i], v4 = a.pz[i], v5 = a.p6[i],
            v6 = a.p7[i], v7 = a.q[i];
    __builtin_amdgcn_sched_barrier(0);   // keep these loads ahead of the lookups
    const int bi = bin[i], rk = rank[i];
    const int S = bi / NBINS, bb = bi - S * NBINS, nb = bb >> 8, cl = bb & 255;
    const int tx = wrap1(S % mx1 + nb % 3 - 1, mx1), ty = wrap1(S / mx1 + nb / 3 - 1, my1);
    const int d = off[(ty * mx1 + tx) * TCELLS + cl] + cnt[bi] + rk;
    b.x[d] = v0;
    b.y[d] = v1;
    b.px[d] = v2;
    b.py[d] = v3;
    b.pz[d] = v4;
    b.p6[d] = v5;
    b.p7[d] = v6;
    b.q[d] = v7;
}

// cell keys of unsorted particles (initial load and uploads), with their
// index for a stable sort and a per-cell count
template <typename R>
__global__ void k_keys(int64_t np, const R *x, const R *y, int nx, int ny, int mx1, int *key,
                       int *idx, int *count, int *err)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= np) return;
    R xv = x[i], yv = y[i];
    int cx = (int)xv, cy = (int)yv;
    if (!(xv >= R(0) && yv >= R(0) && cx < nx && cy < ny)) {
        atomicMax(err, 1 << 30);
        cx = min(max(cx, 0), nx - 1);
        cy = min(max(cy, 0), ny - 1);
    }
    int k = cell_key(cx, cy, mx1);
    key[i] = k;
    idx[i] = (int)i;
    atomicAdd(&count[k], 1);
}

// the same for keys computed on the host (FP32 tile-relative records)
__global__ void k_count_keys(int64_t np, const int *key, int *idx, int *count)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= np) return;
    idx[i] = (int)i;
    atomicAdd(&count[key[i]], 1);
}

// b[p] = a[src[p]]
template <typename R>
__global__ void k_gather(int64_t np, Soa<R> a, Soa<R> b, const int *src)
{
    int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int i = src[p];
    const R v0 = a.x[i], v1 = a.y[i], v2 = a.px[i], v3 = a.py[i], v4 = a.pz[i], v5 = a.p6[i],
            v6 = a.p7[i], v7 = a.q[i];
    b.x[p] = v0;
    b.y[p] = v1;
    b.px[p] = v2;
    b.py[p] = v3;
    b.pz[p] = v4;
    b.p6[p] = v5;
    b.p7[p] = v6;
    b.q[p] = v7;
}

// the move-check word into pinned host memory, in stream order (a kernel
// store instead of a copy-engine transfer between the slice's kernels):
// h_err[0] = the word; h_err[1] = the slice of the first post that saw it
// set (the word stays set until the host clears it, so that is the slice
// whose push set it)
__global__ void k_post_err(const int *err, int *h_err, int slice)
{
    if (threadIdx.x == 0) {
        const int e = *err;
        h_err[0] = e;
        if (e && h_err[1] == 0) h_err[1] = slice;
        __threadfence_system();
    }
}

__global__ void k_zero_int(int *p, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

static TileGeom geom(qp_ctx *c) { return TileGeom{c->nx, c->ny, c->mx1, c->my1, c->ldr}; }

static dim3 pgrid(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

// cell offsets from the per-cell counts, into off (or the buffer given)
static int scan_counts(qp_species *sp, DevBuf *dst = nullptr)
{
    qp_ctx *c = sp->ctx;
    size_t tb = sp->scan_bytes;
    QP_HIP(hipcub::DeviceScan::ExclusiveSum(sp->scan_tmp.p, tb, sp->count.as<int>(),
                                            (dst ? dst : &sp->off)->as<int>(), c->ncells + 1,
                                            c->stream));
    return QP_OK;
}

static int zero_counts(qp_species *sp)
{
    qp_ctx *c = sp->ctx;
    hipLaunchKernelGGL(k_zero_int, dim3(1024), dim3(256), 0, c->stream, sp->count.as<int>(),
                       c->ncells + 1);
    QP_LAUNCH_CHECK();
    return QP_OK;
}

static int check_err(qp_species *sp, const char *what)
{
    qp_ctx *c = sp->ctx;
    int h = 0;
    QP_HIP(hipMemcpyAsync(&h, sp->err.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    QP_HIP(hipStreamSynchronize(c->stream));
    if (h) {
        QP_HIP(hipMemsetAsync(sp->err.p, 0, sizeof(int), c->stream));
        if (h >= (1 << 30)) set_error("%s: particle position outside the grid", what);
        else if (h >= (1 << 29)) set_error("%s: the push's two passes disagree", what);
        else set_error("%s: particle left its tile by more than one tile (tile %d)", what, h);
        return QP_ERR_CHECK;
    }
    return QP_OK;
}

// sort an arbitrary particle order by cell: keys, a stable radix sort of
// (key, index), offsets, then a gather into the alternate arrays.  hkeys:
// the cell keys, computed on the host (FP32 records hold tile-relative
// positions, from which the device cannot tell the tile)
template <typename R>
static int rebin_unsorted(qp_species *sp, const std::vector<int> *hkeys = nullptr)
{
    qp_ctx *c = sp->ctx;
    int rc = zero_counts(sp);
    if (rc) return rc;
    const int64_t np = sp->np;
    DevBuf keys2, idx2, tmp;
    rc = keys2.alloc((size_t)np * 4 + 4);
    if (!rc) rc = idx2.alloc((size_t)np * 4 + 4);
    if (rc) return rc;
    size_t tb = 0;
    int bits = 1;
    while ((1 << bits) < c->ncells) ++bits;
    QP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, sp->key.as<int>(), keys2.as<int>(),
                                              sp->slot.as<int>(), idx2.as<int>(), (int)np, 0,
                                              bits, c->stream));
    rc = tmp.alloc(tb + 256);
    if (rc) return rc;
    if (hkeys) {
        QP_REQUIRE((int64_t)hkeys->size() == np, QP_ERR_ARG, "particle load: %zu keys for %lld",
                   hkeys->size(), (long long)np);
        QP_HIP(hipMemcpyAsync(sp->key.p, hkeys->data(), (size_t)np * sizeof(int),
                              hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(k_count_keys, pgrid(np), dim3(256), 0, c->stream, np,
                           sp->key.as<const int>(), sp->slot.as<int>(), sp->count.as<int>());
    } else {
        ProfScope _ps(c, QP_PROF_SORT, 2.0 * sizeof(R) * (double)np);
        hipLaunchKernelGGL(k_keys<R>, pgrid(np), dim3(256), 0, c->stream, np, sp->a[0].as<R>(),
                           sp->a[1].as<R>(), c->nx, c->ny, c->mx1, sp->key.as<int>(),
                           sp->slot.as<int>(), sp->count.as<int>(), sp->err.as<int>());
    }
    QP_LAUNCH_CHECK();
    QP_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, sp->key.as<int>(), keys2.as<int>(),
                                              sp->slot.as<int>(), idx2.as<int>(), (int)np, 0,
                                              bits, c->stream));
    rc = scan_counts(sp);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gather<R>, pgrid(np), dim3(256), 0, c->stream, np, soa<R>(sp->a),
                       soa<R>(sp->b), idx2.as<const int>());
    QP_LAUNCH_CHECK();
    for (int k = 0; k < 8; ++k) std::swap(sp->a[k], sp->b[k]);
    QP_HIP(hipStreamSynchronize(c->stream));   // the temporaries are freed on return
    return check_err(sp, "particle load");
}

template <typename R>
static int deposit_border(qp_ctx *c, const DepOut<R> &out, int ncomp, int fold,
                          R *halo = nullptr)
{
    QP_REQUIRE(ncomp >= 1 && ncomp <= 8, QP_ERR_ARG, "deposit border: %d components", ncomp);
    TileGeom g = geom(c);
    int n = (c->my1 + 1) * (c->nx + 1) + (c->mx1 + 1) * c->my1 * (TILE - 1);
    {
    ProfScope _ps(c, QP_PROF_BORDER, 2.0 * ncomp * (double)n * sizeof(R));
    hipLaunchKernelGGL(k_deposit_border<R>, dim3((n + 255) / 256, ncomp), dim3(256), 0, c->stream,
                       g, halo ? halo : c->halo.as<R>(), (size_t)c->ntiles * HALO, out, fold,
                       c->skip);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

// species charge into a real-space plane (overwritten), guard-folded
// Deposit launch shapes: rows per group RG (16/RG lanes per cell; 0: chosen
// per tile), staging chunk CH and a waves-per-SIMD floor WPE (4 caps the
// kernel at 128 VGPRs).  Variant 0 of each table is the default; qp_ctx_tune
// selects another.  The FP32 current deposit's default chooses its rows per
// group per tile, keeping a group within two staging chunks (THR = 8):
// 472.4-472.5 slices/s on whole 1024^2, 16 ppc steps at nb/n0 = 10, against
// 464.0 within 1.5 chunks (variant 12), 455.0 within one (variant 14) and
// 434.8-437.2 with a fixed RG = 8 (variant 11); FP64 likewise, 242.9-243.4
// against 233.9-234.3 with a fixed RG = 8 (variant 11).
struct DepShape {
    int rg, ch, wpe;
    int wave;          // current deposit: 1 k_amjdp_w (independent waves, staging
                       // chunk ch per wave); 2 k_amjdp_pf (next chunk prefetched
                       // into registers); 3 k_amjdp_dma (FP32: chunks copied into
                       // LDS by LDS-DMA, double-buffered)
    int thr = 4;       // rg = 0: a group's share of the tile's particles is kept
                       // at or below thr/4 of the chunk
};
template <typename R> struct DepTab;
template <> struct DepTab<float> {
    static constexpr DepShape q[] = {{16, 4096, 1, 0}, {16, 1024, 1, 0}, {8, 2048, 1, 0},
                                     {16, 2048, 1, 0}};
    static constexpr DepShape j[] = {{0, 1024, 4, 0, 8}, {8, 1024, 1, 0}, {4, 1024, 4, 0},
                                     {16, 1024, 4, 0}, {4, 512, 4, 0},  {8, 512, 1, 0},
                                     {2, 256, 1, 1},   {2, 128, 1, 1},  {8, 1024, 1, 2},
                                     {8, 512, 1, 2},   {8, 512, 1, 3},  {8, 1024, 4, 0},
                                     {0, 1024, 4, 0, 6}, {0, 1024, 4, 0, 10}, {0, 1024, 4, 0, 4}};
};
template <> struct DepTab<double> {
    static constexpr DepShape q[] = {{8, 1024, 1, 0}, {8, 2048, 1, 0}, {16, 1024, 1, 0},
                                     {8, 512, 1, 0}};
    static constexpr DepShape j[] = {{0, 1024, 1, 0, 8}, {4, 1024, 1, 0}, {4, 512, 1, 0},
                                     {8, 512, 1, 0},  {2, 512, 1, 0},  {16, 512, 1, 0},
                                     {2, 128, 1, 1},  {2, 64, 1, 1},   {8, 512, 1, 2},
                                     {8, 256, 1, 2},   {8, 256, 1, 0},  {8, 1024, 1, 0},
                                     {0, 1024, 1, 0, 6}, {0, 1024, 1, 0, 10}, {0, 1024, 1, 0, 4}};
};

// kernels using more than 64 KB of dynamic LDS ask for it once
template <typename K>
static int allow_lds(K kernel, size_t bytes, qp_ctx *c)
{
    QP_REQUIRE(bytes <= c->max_lds, QP_ERR_UNSUPPORTED, "kernel needs %zu B of LDS (limit %zu)",
               bytes, c->max_lds);
    if (bytes > 65536)
        QP_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)bytes));
    return QP_OK;
}

template <typename R, int V>
static int deposit_q_v(qp_species *sp, const DepOut<R> &out, R *halo)
{
    qp_ctx *c = sp->ctx;
    constexpr DepShape d = DepTab<R>::q[V];
    const size_t lds = (3 * (size_t)Stage<d.ch>::N + 289) * sizeof(R);
    auto kern = sp->q_uniform ? k_deposit_q<R, d.rg, d.ch, d.wpe, true>
                              : k_deposit_q<R, d.rg, d.ch, d.wpe, false>;
    int rc = allow_lds(kern, lds, c);
    if (rc) return rc;
    {
    // x, y (and q unless uniform) per particle, one plane
    const double words = sp->q_uniform ? 2.0 : 3.0;
    ProfScope _ps(c, QP_PROF_DEPOSIT_Q, (words * sp->np + (c->nx + 1.0) * (c->ny + 1.0)) * sizeof(R));
    hipLaunchKernelGGL(kern, dim3(c->ntiles), dim3(256), lds, c->stream, geom(c),
                       soa<R>(sp->a), sp->off.as<int>(), out, halo, (R)sp->q_value);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

// halo: the tile-border buffer (null: the context's)
template <typename R>
static int deposit_q_out(qp_species *sp, const DepOut<R> &out, R *halo = nullptr)
{
    int rc;
    if (!halo) halo = sp->ctx->halo.as<R>();
    switch (sp->ctx->tune[QP_TUNE_DEPOSIT_Q]) {
    case 1: rc = deposit_q_v<R, 1>(sp, out, halo); break;
    case 2: rc = deposit_q_v<R, 2>(sp, out, halo); break;
    case 3: rc = deposit_q_v<R, 3>(sp, out, halo); break;
    default: rc = deposit_q_v<R, 0>(sp, out, halo); break;
    }
    if (rc) return rc;
    return deposit_border<R>(sp->ctx, out, 1, 1, halo);
}

template <typename R>
static int deposit_q_t(qp_species *sp, R *plane, R *halo = nullptr)
{
    return deposit_q_out<R>(sp, dep_plain<R>(&plane, 1), halo);
}

// dist2d_000 (fdist2d_class.f03:212-316), one rank: x in [1,nx-1),
// y in [1,ny); record x, y, 0, 0, 0, 1, 1, qm in the reference's order.
// Marks the species' charge uniform (or not); a uniform q is written into
// both particle buffers, so no permutation has to move it
static int set_q_uniform(qp_species *sp, bool uniform, double v)
{
    sp->q_uniform = uniform;
    sp->q_value = v;
    if (!uniform || sp->np == 0) return QP_OK;
    int rc = launch_fill(sp->ctx, sp->a[7].p, (size_t)sp->np, v);
    if (!rc) rc = launch_fill(sp->ctx, sp->b[7].p, (size_t)sp->np, v);
    return rc;
}

// The transverse profile at a lattice point (x, y) in grid units: false
// drops it; w is its charge weight.  In the reference's operation order:
//   10 (dist2d_010, fdist2d_class.f03:438-441): dropped inside the inner or
//      outside the outer ellipse;
//   11 (dist2d_011, :633-645): w = exp(-(r - r0)^2 * 0.5 / sigr^2);
//   12 (dist2d_012, :836-856): dropped outside [r(1), r(n)], else w = fr
//      interpolated linearly at r, with the check that r increases (its
//      failure is left in sp->profile_err).
static bool profile_weight(qp_species *sp, double x, double y, double &w)
{
    w = 1.0;
    if (sp->profile == 0) return true;
    const double cx = sp->pcx, cy = sp->pcy;
    if (sp->profile == 10) {
        const double iirx2 = 1.0 / (sp->irx * sp->irx), iiry2 = 1.0 / (sp->iry * sp->iry);
        const double iorx2 = 1.0 / (sp->orx * sp->orx), iory2 = 1.0 / (sp->ory * sp->ory);
        const double ddx = (x - cx) * (x - cx), ddy = (y - cy) * (y - cy);
        return !((ddx * iirx2 + ddy * iiry2 < 1.0) || (ddx * iorx2 + ddy * iory2 > 1.0));
    }
    const double rr = std::sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy));
    if (sp->profile == 11) {
        const double isigr2 = 0.5 / (sp->sigr * sp->sigr);
        w = std::exp(-((rr - sp->r0) * (rr - sp->r0)) * isigr2);
        return true;
    }
    const std::vector<double> &r = sp->pr, &fr = sp->pfr;
    const size_t m = r.size();
    if (rr < r[0] || rr > r[m - 1]) return false;
    for (size_t ii = 1; ii < m; ++ii) {
        if (r[ii] <= r[ii - 1]) {
            sp->profile_err = "dist2d_012: r is not monotonically increasing!";
            return false;
        }
        if (rr <= r[ii]) {
            w = fr[ii - 1] + (fr[ii] - fr[ii - 1]) / (r[ii] - r[ii - 1]) * (rr - r[ii - 1]);
            break;
        }
    }
    return true;
}

template <typename R>
static int distribute(qp_species *sp)
{
    qp_ctx *c = sp->ctx;
    const int nx = c->nx, ny = c->ny, xppc = sp->p.ppc[0], yppc = sp->p.ppc[1];
    const int64_t np = (int64_t)(nx - 2) * (ny - 1) * xppc * yppc;
    const double qm = sp->den_factor * sp->p.density * sp->p.q / std::fabs(sp->p.q) / (double)xppc /
                      (double)yppc;
    std::vector<R> hx((size_t)np), hy((size_t)np), hq;
    sp->profile_err.clear();
    std::vector<int> hkeys;
    if (TileRel<R>::on) hkeys.resize((size_t)np);
    const bool weighted = sp->profile == 11 || sp->profile == 12;
    if (weighted) hq.resize((size_t)np);
    int64_t n = 0;
    for (int i = 2; i <= nx - 1; ++i)
        for (int j = 2; j <= ny; ++j)
            for (int ix = 0; ix < xppc; ++ix)
                for (int iy = 0; iy < yppc; ++iy) {
                    const double x = (ix + 0.5) / xppc + i - 1, y = (iy + 0.5) / yppc + j - 1;
                    double w = 1.0;
                    if (!profile_weight(sp, x, y, w)) continue;
                    if (weighted) hq[n] = (R)(qm * w);
                    if (TileRel<R>::on) {
                        const int cx = (int)x, cy = (int)y;
                        hkeys[n] = host_cell_key(cx, cy, c->mx1);
                        hx[n] = (R)(x - (double)(cx & ~(TILE - 1)));
                        hy[n] = (R)(y - (double)(cy & ~(TILE - 1)));
                    } else {
                        hx[n] = (R)x;
                        hy[n] = (R)y;
                    }
                    ++n;
                }
    QP_REQUIRE(sp->profile_err.empty(), QP_ERR_ARG, "%s", sp->profile_err.c_str());
    sp->np = n;
    if (TileRel<R>::on) hkeys.resize((size_t)n);
    const size_t bytes = (size_t)n * sizeof(R);
    QP_HIP(hipMemcpyAsync(sp->a[0].p, hx.data(), bytes, hipMemcpyHostToDevice, c->stream));
    QP_HIP(hipMemcpyAsync(sp->a[1].p, hy.data(), bytes, hipMemcpyHostToDevice, c->stream));
    if (weighted)
        QP_HIP(hipMemcpyAsync(sp->a[7].p, hq.data(), bytes, hipMemcpyHostToDevice, c->stream));
    QP_HIP(hipStreamSynchronize(c->stream));
    // p6 = p7 = 1 at rest (0 in the FP32 deviation form)
    const double p67 = P67Dev<R>::on ? 0.0 : 1.0;
    const double fillv[6] = {0.0, 0.0, 0.0, p67, p67, qm};
    for (int k = 2; k < (weighted ? 7 : 8); ++k) {
        int rc = launch_fill(c, sp->a[k].p, (size_t)n, fillv[k - 2]);
        if (rc) return rc;
    }
    int rc = rebin_unsorted<R>(sp, TileRel<R>::on ? &hkeys : nullptr);
    if (rc) return rc;
    return set_q_uniform(sp, !weighted, qm);
}

// The density factor at propagation distance s of dist2d_000
// (fdist2d_class.f03:233-251): 1 for a uniform plasma; else the linear
// interpolation of fs between the points s(i-1) <= s <= s(i), with the
// reference's checks (s inside the points; the points increasing up to the
// segment found)
static int den_factor_at(qp_species *sp, double s)
{
    sp->den_factor = 1.0;
    const size_t n = sp->lp_s.size();
    if (n == 0) return QP_OK;
    const std::vector<double> &ps = sp->lp_s, &fs = sp->lp_fs;
    QP_REQUIRE(!(s < ps[0] || s > ps[n - 1]), QP_ERR_ARG,
               "dist2d_000: The s is out of the bound! (s = %g)", s);
    for (size_t i = 1; i < n; ++i) {
        QP_REQUIRE(!(ps[i] < ps[i - 1]), QP_ERR_ARG,
                   "dist2d_000: s is not monotonically increasing!");
        if (s <= ps[i]) {
            sp->den_factor = fs[i - 1] + (fs[i] - fs[i - 1]) / (ps[i] - ps[i - 1]) * (s - ps[i - 1]);
            break;
        }
    }
    return QP_OK;
}

// renew: particles and qn = -(folded deposit of the fresh plasma)
template <typename R>
static int renew_t(qp_species *sp)
{
    int rc = distribute<R>(sp);
    if (rc) return rc;
    rc = deposit_q_t<R>(sp, sp->qn->rs.as<R>());
    if (rc) return rc;
    sp->qn->state = 0;
    double m1 = -1.0;
    return qp_field2d_mult(sp->qn, sp->qn, 0, nullptr, nullptr, &m1);
}

template <typename R>
static int alloc_particles(qp_species *sp, int64_t np)
{
    qp_ctx *c = sp->ctx;
    for (int k = 0; k < 8; ++k) {
        int rc = sp->a[k].alloc((size_t)np * sizeof(R));
        if (!rc) rc = sp->b[k].alloc((size_t)np * sizeof(R));
        if (rc) return rc;
    }
    int rc = sp->key.alloc((size_t)np * 4);
    if (!rc) rc = sp->slot.alloc((size_t)np * 4);
    if (!rc) rc = sp->count.alloc((size_t)(c->ncells + 1) * 4);
    if (!rc) rc = sp->bins.alloc((size_t)c->ntiles * NBINS * 4);
    if (!rc) rc = sp->off.alloc((size_t)(c->ncells + 1) * 4);
    if (!rc) rc = sp->off_new.alloc((size_t)(c->ncells + 1) * 4);
    if (!rc) rc = sp->err.alloc(sizeof(int));
    if (rc) return rc;
    QP_HIP(hipHostMalloc((void **)&sp->h_err, 2 * sizeof(int), hipHostMallocDefault));
    sp->h_err[0] = sp->h_err[1] = 0;
    QP_HIP(hipEventCreateWithFlags(&sp->err_ev, hipEventDisableTiming));
    sp->scan_bytes = 0;
    QP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, sp->scan_bytes, sp->count.as<int>(),
                                            sp->off.as<int>(), c->ncells + 1, c->stream));
    return sp->scan_tmp.alloc(sp->scan_bytes + 256);
}

template <typename R>
static int extpsi_t(qp_species *sp, const qp_field2d *psi, double dex)
{
    qp_ctx *c = sp->ctx;
    if (sp->np == 0) return QP_OK;
    {
    ProfScope _ps(c, QP_PROF_EXTPSI, (6.0 * sp->np + (c->nx + 1.0) * (c->ny + 1.0)) * sizeof(R));
    hipLaunchKernelGGL(k_extpsi<R>, pgrid(sp->np), dim3(256), 0, c->stream, sp->np, c->nx, c->ny,
                       c->ldr, soa<R>(sp->a), psi->rs.as<R>(), (R)sp->qbm, (R)dex,
                       sp->off.as<int>(), c->mx1, c->ntiles);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

template <typename R, int V>
static int amjdp_v(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                   const qp_field2d *psit, double dex, const DepOut<R> &out,
                   const qp_field2d *psi)
{
    qp_ctx *c = sp->ctx;
    const size_t pl = c->rplane;
    const size_t hcs = (size_t)c->ntiles * HALO;
    constexpr DepShape d = DepTab<R>::j[V];
    size_t lds;
    void (*kern)(TileGeom, Soa<R>, const int *, const R *, const R *, const R *, size_t,
                 DepOut<R>, R *, size_t, R, R, R, const int *, const R *, R);
    if constexpr (d.wave == 1) {
        lds = (4 * 7 * (size_t)Stage<d.ch>::N + 8 * 8 * 51) * sizeof(R);
        kern = k_amjdp_w<R, d.ch>;
    } else if constexpr (d.wave == 2) {
        lds = (7 * (size_t)Stage<d.ch>::N + 8 * 289) * sizeof(R);
        kern = k_amjdp_pf<R, d.rg, d.ch>;
    } else if constexpr (d.wave == 3) {
        static_assert(std::is_same<R, float>::value, "LDS-DMA staging is FP32 only");
        lds = (2 * 7 * (size_t)(d.ch + d.ch / 64) + 8 * 289) * sizeof(R);
        kern = k_amjdp_dma<d.rg, d.ch>;
    } else {
        lds = (7 * (size_t)Stage<d.ch>::N + 8 * 289) * sizeof(R);
        if (sp->q_uniform)
            kern = psi ? k_amjdp<R, d.rg, d.ch, d.wpe, true, true, d.thr>
                       : k_amjdp<R, d.rg, d.ch, d.wpe, false, true, d.thr>;
        else
            kern = psi ? k_amjdp<R, d.rg, d.ch, d.wpe, true, false, d.thr>
                       : k_amjdp<R, d.rg, d.ch, d.wpe, false, false, d.thr>;
    }
    int rc = allow_lds(kern, lds, c);
    if (rc) return rc;
    {
    // with psi: one more word per particle (p7 written) and one more plane;
    // a uniform q is not read (by the kernels with a uniform-charge form)
    const double words = (psi ? 9.0 : 8.0) - (sp->q_uniform && d.wave == 0 ? 1.0 : 0.0);
    const double planes = psi ? 15.0 : 14.0;
    ProfScope _ps(c, QP_PROF_AMJDP, (words * sp->np + planes * (c->nx + 1.0) * (c->ny + 1.0)) * sizeof(R));
    hipLaunchKernelGGL(kern, dim3(c->ntiles), dim3(256), lds, c->stream, geom(c),
                       soa<R>(sp->a), sp->off.as<int>(), ef->rs.as<R>(), bf->rs.as<R>(),
                       psit->rs.as<R>(), pl, out, c->halo.as<R>(), hcs, (R)sp->qbm, (R)sp->dt,
                       (R)dex, c->skip, psi ? psi->rs.as<R>() : nullptr, (R)sp->q_value);
    }
    QP_LAUNCH_CHECK();
    return QP_OK;
}

template <typename R>
static int amjdp_out(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                     const qp_field2d *psit, double dex, const DepOut<R> &out,
                     const qp_field2d *psi = nullptr)
{
    int rc;
    switch (sp->ctx->tune[QP_TUNE_AMJDP]) {
    case 1: rc = amjdp_v<R, 1>(sp, ef, bf, psit, dex, out, psi); break;
    case 2: rc = amjdp_v<R, 2>(sp, ef, bf, psit, dex, out, psi); break;
    case 3: rc = amjdp_v<R, 3>(sp, ef, bf, psit, dex, out, psi); break;
    case 4: rc = amjdp_v<R, 4>(sp, ef, bf, psit, dex, out, psi); break;
    case 5: rc = amjdp_v<R, 5>(sp, ef, bf, psit, dex, out, psi); break;
    case 6: rc = amjdp_v<R, 6>(sp, ef, bf, psit, dex, out, psi); break;
    case 7: rc = amjdp_v<R, 7>(sp, ef, bf, psit, dex, out, psi); break;
    case 8: rc = amjdp_v<R, 8>(sp, ef, bf, psit, dex, out, psi); break;
    case 9: rc = amjdp_v<R, 9>(sp, ef, bf, psit, dex, out, psi); break;
    case 10: rc = amjdp_v<R, 10>(sp, ef, bf, psit, dex, out, psi); break;
    case 11: rc = amjdp_v<R, 11>(sp, ef, bf, psit, dex, out, psi); break;
    case 12: rc = amjdp_v<R, 12>(sp, ef, bf, psit, dex, out, psi); break;
    case 13: rc = amjdp_v<R, 13>(sp, ef, bf, psit, dex, out, psi); break;
    case 14: rc = amjdp_v<R, 14>(sp, ef, bf, psit, dex, out, psi); break;
    default: rc = amjdp_v<R, 0>(sp, ef, bf, psit, dex, out, psi); break;
    }
    if (rc) return rc;
    return deposit_border<R>(sp->ctx, out, 8, 1);
}

// the species' own cu, dcu, amu (overwritten)
template <typename R>
static int amjdp_t(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                   const qp_field2d *psit, double dex)
{
    const size_t pl = sp->ctx->rplane;
    R *cu = sp->cu->rs.as<R>(), *dcu = sp->dcu->rs.as<R>(), *amu = sp->amu->rs.as<R>();
    R *planes[8] = {cu, cu + pl, cu + 2 * pl, dcu, dcu + pl, amu, amu + pl, amu + 2 * pl};
    int rc = amjdp_out<R>(sp, ef, bf, psit, dex, dep_plain<R>(planes, 8));
    sp->cu->state = sp->dcu->state = sp->amu->state = 0;
    return rc;
}

template <typename R>
static int push_t(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                  const qp_field2d *psit, double dex, bool check)
{
    qp_ctx *c = sp->ctx;
    const TileGeom g = geom(c);
    const bool scatter = c->tune[QP_TUNE_PUSH] == 1;
    const double fields = 6.0 * (c->nx + 1.0) * (c->ny + 1.0) * sizeof(R);
    const Soa<R> a = soa<R>(sp->a), b = soa<R>(sp->b);
    int *off = sp->off.as<int>(), *off_new = sp->off_new.as<int>();
    {
    // in place: 7 arrays read, 5 written; counting: 7 read
    ProfScope _ps(c, QP_PROF_PUSH, (scatter ? 12.0 : 7.0) * sp->np * sizeof(R) + fields);
    auto kern = scatter ? k_push<R, 0> : k_push<R, 1>;
    hipLaunchKernelGGL(kern, dim3(c->ntiles), dim3(TCELLS), 0, c->stream, g, a, b, off, off_new,
                       ef->rs.as<R>(), bf->rs.as<R>(), psit->rs.as<R>(), c->rplane,
                       (R)sp->qbm, (R)sp->dt, (R)dex, sp->key.as<int>(), sp->slot.as<int>(),
                       sp->bins.as<int>(), sp->err.as<int>(), sp->np, 0);
    }
    QP_LAUNCH_CHECK();
    {
    ProfScope _ps(c, QP_PROF_SORT, 8.0 * (double)c->ntiles * NBINS);
    hipLaunchKernelGGL(k_bin_starts, dim3(c->ncells / 256 + 1), dim3(256), 0, c->stream, g,
                       sp->bins.as<int>(), sp->count.as<int>());
    }
    QP_LAUNCH_CHECK();
    if (scatter) {
        int rc = scan_counts(sp);
        if (rc) return rc;
        {
        ProfScope _ps(c, QP_PROF_SORT, 18.0 * sizeof(R) * (double)sp->np);
        hipLaunchKernelGGL(k_scatter<R>, pgrid(sp->np), dim3(256), 0, c->stream, sp->np, a, b,
                           sp->key.as<int>(), sp->slot.as<int>(), sp->bins.as<int>(), off,
                           c->mx1, c->my1);
        }
        QP_LAUNCH_CHECK();
    } else {
        int rc = scan_counts(sp, &sp->off_new);
        if (rc) return rc;
        {
        // 8 arrays read, 8 written at their sorted places
        ProfScope _ps(c, QP_PROF_SORT, (sp->q_uniform ? 14.0 : 16.0) * sp->np * sizeof(R) + fields);
        auto kern = k_push<R, 2>;
        hipLaunchKernelGGL(kern, dim3(c->ntiles), dim3(TCELLS), 0, c->stream, g, a, b,
                           off, off_new, ef->rs.as<R>(), bf->rs.as<R>(), psit->rs.as<R>(),
                           c->rplane, (R)sp->qbm, (R)sp->dt, (R)dex, nullptr, nullptr,
                           sp->bins.as<int>(), sp->err.as<int>(), sp->np, sp->q_uniform ? 1 : 0);
        }
        QP_LAUNCH_CHECK();
        std::swap(sp->off, sp->off_new);
    }
    for (int k = 0; k < 8; ++k) std::swap(sp->a[k], sp->b[k]);
    return check ? check_err(sp, "pmove") : QP_OK;
}

int species_qdp_into(qp_species *sp, qp_field2d *q)
{
    return qp_species_qdp(sp, q);
}

}  // namespace qp

using namespace qp;

#define SPCHECK(sp) QP_REQUIRE((sp) && (sp)->ctx, QP_ERR_ARG, "%s: null species", __func__)
#define DISPATCH(c, fn, ...) ((c)->prec == QP_FP64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

extern "C" {

int qp_species_create(qp_ctx *c, const qp_species_params *p, double dt, double s,
                      qp_species **out)
{
    QP_REQUIRE(c && p && out, QP_ERR_ARG, "qp_species_create: null argument");
    QP_REQUIRE(p->ppc[0] >= 1 && p->ppc[1] >= 1 && p->q != 0.0 && p->m != 0.0, QP_ERR_ARG,
               "qp_species_create: bad species parameters");
    qp_species *sp = new qp_species();
    sp->ctx = c;
    sp->p = *p;
    sp->qbm = p->q / p->m;
    sp->dt = dt;
    const int64_t np = (int64_t)(c->nx - 2) * (c->ny - 1) * p->ppc[0] * p->ppc[1];
    int rc = den_factor_at(sp, s);
    if (!rc) rc = DISPATCH(c, alloc_particles, sp, np);
    if (!rc) rc = qp_field2d_create(c, 1, 0, &sp->q);
    if (!rc) rc = qp_field2d_create(c, 1, 0, &sp->qn);
    if (!rc) rc = qp_field2d_create(c, 3, 0, &sp->cu);
    if (!rc) rc = qp_field2d_create(c, 2, 0, &sp->dcu);
    if (!rc) rc = qp_field2d_create(c, 3, 0, &sp->amu);
    if (!rc) rc = DISPATCH(c, renew_t, sp);
    if (rc) { qp_species_destroy(sp); return rc; }
    *out = sp;
    return QP_OK;
}

int qp_species_destroy(qp_species *sp)
{
    if (!sp) return QP_OK;
    if (sp->ctx) (void)hipStreamSynchronize(sp->ctx->stream);
    for (int k = 0; k < 8; ++k) { sp->a[k].release(); sp->b[k].release(); }
    sp->key.release(); sp->slot.release(); sp->count.release(); sp->off.release();
    sp->off_new.release();
    sp->bins.release();
    sp->scan_tmp.release(); sp->err.release();
    if (sp->h_err) (void)hipHostFree(sp->h_err);
    if (sp->err_ev) (void)hipEventDestroy(sp->err_ev);
    qp_field2d_destroy(sp->q); qp_field2d_destroy(sp->qn); qp_field2d_destroy(sp->cu);
    qp_field2d_destroy(sp->dcu); qp_field2d_destroy(sp->amu);
    delete sp;
    return QP_OK;
}

int qp_species_renew(qp_species *sp, double s)
{
    SPCHECK(sp);
    int rc = den_factor_at(sp, s);
    if (rc) return rc;
    return DISPATCH(sp->ctx, renew_t, sp);
}

int qp_species_set_longitudinal(qp_species *sp, int n, const double *s, const double *fs,
                                double s_now)
{
    SPCHECK(sp);
    QP_REQUIRE(n == 0 || (n >= 2 && s && fs), QP_ERR_ARG,
               "qp_species_set_longitudinal: %d points (0, or at least 2)", n);
    sp->lp_s.assign(s, s + n);
    sp->lp_fs.assign(fs, fs + n);
    return qp_species_renew(sp, s_now);
}

// init_fdist2d_010 / _011 / _012 (fdist2d_class.f03:318-388, 525-590,
// 713-776): lengths to grid units, the centre relative to the box's lower
// corner; then a reload at s_now
int qp_species_set_profile(qp_species *sp, const qp_species_profile *p, double s_now)
{
    SPCHECK(sp);
    QP_REQUIRE(p, QP_ERR_ARG, "qp_species_set_profile: null profile");
    QP_REQUIRE(p->profile == 0 || p->profile == 10 || p->profile == 11 || p->profile == 12,
               QP_ERR_UNSUPPORTED, "qp_species_set_profile: plasma profile %d (0, 10, 11, 12)",
               p->profile);
    const qp_ctx *c = sp->ctx;
    const double dx = c->dex, dy = c->dey;
    if (p->profile == 10)
        QP_REQUIRE(p->inner_radius[0] > 0 && p->inner_radius[1] > 0 && p->outer_radius[0] > 0 &&
                       p->outer_radius[1] > 0,
                   QP_ERR_ARG, "profile 10: radii must be positive");
    if (p->profile == 11)
        QP_REQUIRE(p->width > 0, QP_ERR_ARG, "profile 11: width must be positive");
    if (p->profile == 12) {
        QP_REQUIRE(p->nr >= 2 && p->r && p->fr, QP_ERR_ARG,
                   "profile 12: at least 2 radial points");
    }
    sp->profile = p->profile;
    sp->pcx = (p->center[0] - c->box[0]) / dx;
    sp->pcy = (p->center[1] - c->box[2]) / dy;
    sp->irx = p->inner_radius[0] / dx;
    sp->iry = p->inner_radius[1] / dy;
    sp->orx = p->outer_radius[0] / dx;
    sp->ory = p->outer_radius[1] / dy;
    sp->r0 = p->radius / dx;
    sp->sigr = p->width / dx;
    sp->pr.clear();
    sp->pfr.clear();
    if (p->profile == 12)
        for (int k = 0; k < p->nr; ++k) {
            sp->pr.push_back(p->r[k] / dx);
            sp->pfr.push_back(p->fr[k]);
        }
    return qp_species_renew(sp, s_now);
}

// qdp_species2d (species2d_class.f03:166-185): this%q = folded deposit;
// q = this%q + q; q = qn + q
int qp_species_qdp(qp_species *sp, qp_field2d *q)
{
    SPCHECK(sp);
    QP_REQUIRE(q && q->dim == 1 && q->state == 0, QP_ERR_STATE, "qdp: q must be real space, dim 1");
    qp_ctx *c = sp->ctx;
    int rc = c->prec == QP_FP64 ? deposit_q_t<double>(sp, sp->q->rs.as<double>())
                                : deposit_q_t<float>(sp, sp->q->rs.as<float>());
    if (rc) return rc;
    sp->q->state = 0;
    sp->q->gcells = 1;
    rc = qp_field2d_add(q, sp->q, q, 0, nullptr, nullptr, nullptr);
    if (!rc) rc = qp_field2d_add(q, sp->qn, q, 0, nullptr, nullptr, nullptr);
    return rc;
}

// qp_species_qdp with the deposit's tile-border halo in the library's static
// scratch array instead of the context's buffer: no allocation of any kind.
// Every call on a GPU shares that GPU's array, so calls on one GPU that
// overlap (other threads, or other streams without ordering) race.  The
// halo is fully rewritten by every deposit before it is read, so nothing
// has to be cleared between calls.
int qp_species_qdp_noalloc(qp_species *sp, qp_field2d *q)
{
    SPCHECK(sp);
    QP_REQUIRE(q && q->dim == 1 && q->state == 0, QP_ERR_STATE,
               "qdp_noalloc: q must be real space, dim 1");
    qp_ctx *c = sp->ctx;
    const size_t need = (size_t)c->ntiles * HALO * c->esize;
    QP_REQUIRE(need <= qp::NOALLOC_BYTES, QP_ERR_UNSUPPORTED,
               "qdp_noalloc: %d tiles need %zu B of scratch; the maximum is %zu B "
               "(%zu tiles in this precision)",
               c->ntiles, need, qp::NOALLOC_BYTES, qp::NOALLOC_BYTES / (HALO * c->esize));
    void *halo = qp::noalloc_scratch(c->device);
    QP_REQUIRE(halo, QP_ERR_UNSUPPORTED, "qdp_noalloc: no scratch address for device %d",
               c->device);
    int rc = c->prec == QP_FP64
                 ? deposit_q_t<double>(sp, sp->q->rs.as<double>(), static_cast<double *>(halo))
                 : deposit_q_t<float>(sp, sp->q->rs.as<float>(), static_cast<float *>(halo));
    if (rc) return rc;
    sp->q->state = 0;
    sp->q->gcells = 1;
    rc = qp_field2d_add(q, sp->q, q, 0, nullptr, nullptr, nullptr);
    if (!rc) rc = qp_field2d_add(q, sp->qn, q, 0, nullptr, nullptr, nullptr);
    return rc;
}

int qp_species_extpsi(qp_species *sp, const qp_field2d *psi, double dex)
{
    SPCHECK(sp);
    QP_REQUIRE(psi && psi->state == 0, QP_ERR_STATE, "extpsi: psi must be real space");
    return DISPATCH(sp->ctx, extpsi_t, sp, psi, dex);
}

// amjdp_species2d (species2d_class.f03:187-218)
int qp_species_amjdp(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                     const qp_field2d *psit, qp_field2d *cu, qp_field2d *amu, qp_field2d *dcu,
                     double dex)
{
    SPCHECK(sp);
    QP_REQUIRE(ef && bf && psit && cu && amu && dcu, QP_ERR_ARG, "amjdp: null field");
    QP_REQUIRE(ef->state == 0 && bf->state == 0 && psit->state == 0, QP_ERR_STATE,
               "amjdp: fields must be real space");
    QP_REQUIRE(ef->dim >= 2 && bf->dim == 3 && cu->dim == 3 && dcu->dim >= 2 && amu->dim == 3,
               QP_ERR_ARG, "amjdp: component counts");
    int rc = DISPATCH(sp->ctx, amjdp_t, sp, ef, bf, psit, dex);
    if (!rc) rc = qp_field2d_add(cu, sp->cu, cu, 0, nullptr, nullptr, nullptr);
    if (!rc) rc = qp_field2d_add(dcu, sp->dcu, dcu, 0, nullptr, nullptr, nullptr);
    if (!rc) rc = qp_field2d_add(amu, sp->amu, amu, 0, nullptr, nullptr, nullptr);
    double d = dex;
    if (!rc) rc = qp_field2d_mult(sp->cu, sp->cu, 0, nullptr, nullptr, &d);
    return rc;
}

int qp_species_push(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                    const qp_field2d *psit, double dex)
{
    SPCHECK(sp);
    QP_REQUIRE(ef && bf && psit && ef->state == 0 && bf->state == 0 && psit->state == 0,
               QP_ERR_STATE, "push: fields must be real space");
    return DISPATCH(sp->ctx, push_t, sp, ef, bf, psit, dex, true);
}

// cbq_species2d (species2d_class.f03:399-418)
int qp_species_cbq(qp_species *sp, qp_field3d *q3, int pos)
{
    SPCHECK(sp);
    const int d1[1] = {1}, d3[1] = {3};
    int rc = qp_field2d_add(sp->q, sp->q, sp->cu, 1, d1, d1, d3);
    if (!rc) rc = qp_field2d_fftrk(sp->q, 1);
    if (!rc) rc = qp_field2d_smooth(sp->q, sp->q);
    if (!rc) rc = qp_field2d_fftkr(sp->q, 1);
    if (!rc && q3) rc = qp_field2d_cb(sp->q, q3, pos, 1, d1, d1);
    return rc;
}

}  // extern "C"

namespace qp {
// the first plane of the species' 3-D charge at a (re)load on the first
// stage (species2d_class.f03:107-113, 151-157): q = the fresh plasma's
// folded deposit (-qn), smoothed, into plane 1
int species_q3_initial(qp_species *sp, qp_field3d *q3)
{
    SPCHECK(sp);
    const int d1[1] = {1};
    double m1 = -1.0;
    int rc = qp_field2d_mult(sp->q, sp->qn, 0, nullptr, nullptr, &m1);
    if (!rc) rc = qp_field2d_fftrk(sp->q, 1);
    if (!rc) rc = qp_field2d_smooth(sp->q, sp->q);
    if (!rc) rc = qp_field2d_fftkr(sp->q, 1);
    if (!rc && q3) rc = qp_field2d_cb(sp->q, q3, 1, 1, d1, d1);
    return rc;
}
}  // namespace qp
extern "C" {

int qp_species_count(const qp_species *sp, int64_t *np)
{
    SPCHECK(sp);
    *np = sp->np;
    return QP_OK;
}

}  // extern "C"

static int host_cell_counts(const qp_species *sp, std::vector<int> &off);

template <typename R>
static int download_t(const qp_species *sp, double *part, int64_t ld)
{
    qp_ctx *c = sp->ctx;
    std::vector<R> h((size_t)sp->np);
    // FP32 positions are relative to their tile's origin: the tile of each
    // particle comes from the cell offsets (the records are sorted by tile)
    std::vector<double> ox, oy;
    if (TileRel<R>::on) {
        std::vector<int> off;
        int rc = host_cell_counts(sp, off);
        if (rc) return rc;
        ox.resize((size_t)sp->np);
        oy.resize((size_t)sp->np);
        for (int t = 0; t < c->ntiles; ++t)
            for (int i = off[(size_t)t * TCELLS]; i < off[(size_t)(t + 1) * TCELLS]; ++i) {
                ox[i] = (double)((t % c->mx1) * TILE);
                oy[i] = (double)((t / c->mx1) * TILE);
            }
    }
    for (int k = 0; k < 8; ++k) {
        QP_HIP(hipMemcpyAsync(h.data(), sp->a[k].p, h.size() * sizeof(R), hipMemcpyDeviceToHost,
                              c->stream));
        QP_HIP(hipStreamSynchronize(c->stream));
        // words 6 and 7 of an FP32 record are held as deviations from 1
        const double base = (P67Dev<R>::on && (k == 5 || k == 6)) ? 1.0 : 0.0;
        if (TileRel<R>::on && k < 2) {
            const std::vector<double> &o = k == 0 ? ox : oy;
            for (int64_t i = 0; i < sp->np; ++i) part[k + 8 * i] = (double)h[i] + o[i];
        } else {
            for (int64_t i = 0; i < sp->np; ++i) part[k + 8 * i] = (double)h[i] + base;
        }
    }
    (void)ld;
    return QP_OK;
}

extern "C" {

int qp_species_download(const qp_species *sp, double *part, int64_t ld)
{
    SPCHECK(sp);
    QP_REQUIRE(part && ld >= sp->np, QP_ERR_ARG, "qp_species_download: bad buffer");
    return sp->ctx->prec == QP_FP64 ? download_t<double>(sp, part, ld)
                                    : download_t<float>(sp, part, ld);
}

}  // extern "C"

template <typename R>
static int upload_t(qp_species *sp, const double *part, int64_t np)
{
    qp_ctx *c = sp->ctx;
    std::vector<R> h((size_t)np);
    // FP32: cell keys from the positions in grid units, and the positions
    // relative to their tile (in FP64, then rounded)
    std::vector<int> hkeys;
    if (TileRel<R>::on) {
        hkeys.resize((size_t)np);
        for (int64_t i = 0; i < np; ++i) {
            const double x = part[8 * i], y = part[1 + 8 * i];
            QP_REQUIRE(x >= 0.0 && x < (double)c->nx && y >= 0.0 && y < (double)c->ny,
                       QP_ERR_CHECK, "particle load: particle position outside the grid");
            hkeys[i] = host_cell_key((int)x, (int)y, c->mx1);
        }
    }
    for (int k = 0; k < 8; ++k) {
        const double base = (P67Dev<R>::on && (k == 5 || k == 6)) ? 1.0 : 0.0;
        if (TileRel<R>::on && k < 2)
            for (int64_t i = 0; i < np; ++i) {
                const double v = part[k + 8 * i];
                h[i] = (R)(v - (double)((int)v & ~(TILE - 1)));
            }
        else
            for (int64_t i = 0; i < np; ++i) h[i] = (R)(part[k + 8 * i] - base);
        QP_HIP(hipMemcpyAsync(sp->a[k].p, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice,
                              c->stream));
        QP_HIP(hipStreamSynchronize(c->stream));
    }
    sp->np = np;
    // a uniform charge (compared as stored) lets the kernels skip q
    bool uniform = true;
    const R q0 = np ? (R)part[7] : R(0);
    for (int64_t i = 1; uniform && i < np; ++i) uniform = (R)part[7 + 8 * i] == q0;
    int rc = rebin_unsorted<R>(sp, TileRel<R>::on ? &hkeys : nullptr);
    if (rc) return rc;
    return set_q_uniform(sp, uniform, (double)q0);
}

extern "C" {

int qp_species_upload(qp_species *sp, const double *part, int64_t np, int64_t ld)
{
    SPCHECK(sp);
    QP_REQUIRE(part && np >= 0 && ld >= np, QP_ERR_ARG, "qp_species_upload: bad buffer");
    QP_REQUIRE((size_t)np * sp->ctx->esize <= sp->a[0].bytes, QP_ERR_ARG,
               "qp_species_upload: %lld particles exceed the capacity", (long long)np);
    return sp->ctx->prec == QP_FP64 ? upload_t<double>(sp, part, np)
                                    : upload_t<float>(sp, part, np);
}

static int host_cell_counts(const qp_species *sp, std::vector<int> &off)
{
    qp_ctx *c = sp->ctx;
    off.resize((size_t)c->ncells + 1);
    QP_HIP(hipMemcpyAsync(off.data(), sp->off.p, off.size() * 4, hipMemcpyDeviceToHost, c->stream));
    QP_HIP(hipStreamSynchronize(c->stream));
    return QP_OK;
}

int qp_species_cell_counts(const qp_species *sp, int32_t *counts)
{
    SPCHECK(sp);
    std::vector<int> off;
    int rc = host_cell_counts(sp, off);
    if (rc) return rc;
    qp_ctx *c = sp->ctx;
    for (int cy = 0; cy < c->ny; ++cy)
        for (int cx = 0; cx < c->nx; ++cx) {
            int k = ((cy / TILE) * c->mx1 + cx / TILE) * TCELLS + (cy % TILE) * TILE + cx % TILE;
            counts[(size_t)cy * c->nx + cx] = off[k + 1] - off[k];
        }
    return QP_OK;
}

int qp_species_tile_counts(const qp_species *sp, int32_t *kpic)
{
    SPCHECK(sp);
    std::vector<int> off;
    int rc = host_cell_counts(sp, off);
    if (rc) return rc;
    for (int t = 0; t < sp->ctx->ntiles; ++t) kpic[t] = off[(t + 1) * TCELLS] - off[t * TCELLS];
    return QP_OK;
}

}  // extern "C"

// entry points used by the slice driver without the per-call error sync
namespace qp {
int species_push_nocheck(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                         const qp_field2d *psit, double dex)
{
    return DISPATCH(sp->ctx, push_t, sp, ef, bf, psit, dex, false);
}
int species_check(qp_species *sp) { return check_err(sp, "pmove"); }

// The push leaves the particles in the alternate arrays and swaps the
// host's buffer pointers (and, when it re-sorts in place, the cell offsets).
// A captured slice repeats the device side only; these let its caller keep
// the host side in step: which buffer holds the particles, and one swap.
const void *species_buffer_key(const qp_species *sp) { return sp->a[0].p; }
void species_flip(qp_species *sp, bool offsets)
{
    if (offsets) std::swap(sp->off, sp->off_new);
    for (int k = 0; k < 8; ++k) std::swap(sp->a[k], sp->b[k]);
}

// The slice loop does not wait for the move check of each push: the error
// word is copied to pinned memory behind the push, and a later slice (or the
// end of the step, which waits) reports it.
int species_check_post(qp_species *sp, int slice)
{
    qp_ctx *c = sp->ctx;
    hipLaunchKernelGGL(k_post_err, dim3(1), dim3(64), 0, c->stream, sp->err.as<int>(), sp->h_err,
                       slice);
    QP_LAUNCH_CHECK();
    QP_HIP(hipEventRecord(sp->err_ev, c->stream));
    sp->err_pending = true;
    return QP_OK;
}

// The move check of the slices posted so far, once their push has run
// (wait: block until it has).  The error names the slice whose push moved a
// particle too far; the slices after it were computed from misplaced
// particles and are invalid, like everything after the reference's stop.
int species_check_poll(qp_species *sp, bool wait)
{
    if (!sp->err_pending) return QP_OK;
    if (wait) QP_HIP(hipEventSynchronize(sp->err_ev));
    else if (hipEventQuery(sp->err_ev) != hipSuccess) return QP_OK;
    sp->err_pending = false;
    const int h = sp->h_err[0], slice = sp->h_err[1];
    if (!h) return QP_OK;
    sp->h_err[0] = sp->h_err[1] = 0;
    QP_HIP(hipMemsetAsync(sp->err.p, 0, sizeof(int), sp->ctx->stream));
    if (h >= (1 << 30))
        set_error("pmove: particle position outside the grid (push of slice %d)", slice);
    else if (h >= (1 << 29))
        set_error("pmove: the push's two passes disagree (push of slice %d)", slice);
    else
        set_error("pmove: particle left its tile by more than one tile (tile %d, push of "
                  "slice %d)", h, slice);
    return QP_ERR_CHECK;
}

// qdp_species2d with the slice's additions fused into the deposit
// (species2d_class.f03:166-185): this%q = deposit; q = this%q + q;
// q = qn + q.  first: q holds nothing yet (the slice loop zeroed it).
template <typename R>
static int qdp_fused_t(qp_species *sp, qp_field2d *q, bool first)
{
    R *qp = q->rs.as<R>();
    DepOut<R> o{};
    o.p[0] = qp;
    o.inner[0] = first ? nullptr : qp;
    o.outer[0] = sp->qn->rs.as<R>();
    o.p2[0] = sp->q->rs.as<R>();
    o.scale = o.scale2 = R(1);
    int rc = deposit_q_out<R>(sp, o);
    sp->q->state = 0;
    sp->q->gcells = 1;
    q->state = 0;
    return rc;
}

int species_qdp_fused(qp_species *sp, qp_field2d *q, bool first)
{
    return DISPATCH(sp->ctx, qdp_fused_t, sp, q, first);
}

// amjdp_species2d plus the slice loop's scaling (species2d_class.f03:187-218,
// simulation_class.f03:682-685): cu = this%cu + cu (likewise dcu, amu),
// this%cu *= dex, and after the last species cu, amu, dcu *= dex.  Of the
// species' own planes only this%cu(3) is kept (its one later reader, cbq).
template <typename R>
static int amjdp_fused_t(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                         const qp_field2d *psit, qp_field2d *cu, qp_field2d *dcu,
                         qp_field2d *amu, double dex, bool first, bool last,
                         const qp_field2d *psi)
{
    const size_t pl = sp->ctx->rplane;
    R *c = cu->rs.as<R>(), *d = dcu->rs.as<R>(), *a = amu->rs.as<R>();
    R *planes[8] = {c, c + pl, c + 2 * pl, d, d + pl, a, a + pl, a + 2 * pl};
    DepOut<R> o{};
    for (int k = 0; k < 8; ++k) {
        o.p[k] = planes[k];
        o.inner[k] = first ? nullptr : planes[k];
    }
    o.p2[2] = sp->cu->rs.as<R>() + 2 * pl;
    o.scale = last ? (R)dex : R(1);
    o.scale2 = (R)dex;
    int rc = amjdp_out<R>(sp, ef, bf, psit, dex, o, psi);
    sp->cu->state = 0;
    cu->state = dcu->state = amu->state = 0;
    return rc;
}

int species_amjdp_fused(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                        const qp_field2d *psit, qp_field2d *cu, qp_field2d *dcu, qp_field2d *amu,
                        double dex, bool first, bool last, const qp_field2d *psi)
{
    return DISPATCH(sp->ctx, amjdp_fused_t, sp, ef, bf, psit, cu, dcu, amu, dex, first, last,
                    psi);
}
}  // namespace qp
