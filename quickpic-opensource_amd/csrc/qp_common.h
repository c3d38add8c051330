// Internal declarations of libqpslice.so (MI355X / gfx950).
//
// Device layouts (row-major, one plane per component):
//   real-space plane  (ny+1) rows x ldr   (row y holds nodes x = 0..nx)
//   k-space plane     (nx+1) rows x ldk   (row kx holds modes ky = 0..ny;
//                     the reference's transposed k order, ufield2d_class.f03:198)
//   3-D stack         (nz+1) real-space planes per component
//   particles         structure of arrays, sorted by a tile-major cell key
//                     key = (ty*mx1 + tx)*256 + ly*16 + lx for cell
//                     (x, y) = (16*tx + lx, 16*ty + ly)
// Real is float (QP_FP32) or double (QP_FP64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/qpslice.h"

namespace qp {

constexpr int TILE = 16;                 // cells per tile edge (part2d_class.f03:99)
constexpr int TCELLS = TILE * TILE;      // cells per tile = threads per particle block
constexpr int HALO = 4 * TILE;           // border nodes of a 17x17 tile block
constexpr double TWOPI_REF = 6.28318530717959;  // the reference's 2*pi literal
constexpr double BEAM_FIX = 17592186044416.0;   // 2^44: beam charge fixed point
constexpr int MAX_BATCH = 8;            // components in one batched transform
constexpr int RES_SLOTS = 64;           // pinned residual slots of the corrector loop
constexpr int FLAG_WORDS = 64;          // converged flags (passes 0..63)

void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

// Scratch of qp_species_qdp_noalloc: the charge deposit's tile-border halo
// (one component) in a static device array, for callers that cannot
// allocate: 2 MiB per GPU, held until the process ends, which holds the
// halo of NOALLOC_TILES tiles of 16 x 16 cells in FP64 (a 1024 x 1024 grid)
// and of twice as many in FP32.
constexpr size_t NOALLOC_TILES = 4096;
constexpr size_t NOALLOC_BYTES = NOALLOC_TILES * HALO * sizeof(double);
constexpr int NOALLOC_DEVICES = 64;
// Looks the array's device address up once for the current device (device);
// qp_ctx_create calls it with its device current.
int noalloc_setup(int device);
// The address noalloc_setup found for device, or null.
void *noalloc_scratch(int device);

#define QP_HIP(call)                                                        \
    do {                                                                    \
        hipError_t _e = (call);                                             \
        if (_e != hipSuccess) return qp::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define QP_LAUNCH_CHECK()                                                   \
    do {                                                                    \
        hipError_t _e = hipGetLastError();                                  \
        if (_e != hipSuccess) return qp::hip_fail(_e, "kernel launch", __FILE__, __LINE__); \
    } while (0)

#define QP_REQUIRE(cond, code, ...)                                         \
    do {                                                                    \
        if (!(cond)) { qp::set_error(__VA_ARGS__); return (code); }         \
    } while (0)

// aligned to its size, so that a pair moves with one 8- or 16-byte access
// (LDS ds_read_b64 / ds_write_b64 with immediate offsets in the transforms)
template <typename T> struct alignas(2 * sizeof(T)) cplx { T x, y; };

// device buffer owned by the library
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int alloc(size_t n);
    void release();
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace qp

namespace qp {
// The corrector's residual: the per-block FP64 partial sums of the corrector
// kernel are added up in block order by one block of the transform launched
// right after it (nblk = 0: nothing to do), which writes the residual
// sqrt(num/den) to h_slot and num, den to out, and sets *flag_out when
// *flag_in is set or the residual is below tol.
struct ResJob {
    const double *partial;
    int nblk;
    const int *flag_in;
    int *flag_out;
    double *h_slot;
    double *out;
    double tol;
};

struct ProfEv {
    hipEvent_t a, b;
    int kind;
    double bytes;
    int tag;
};
}  // namespace qp

struct qp_ctx {
    int indx, indy, indz, nx, ny, nz;
    int prec;             // QP_FP64 / QP_FP32
    int device;
    size_t esize;         // sizeof(Real)
    double box[6];
    double dex, dey, dxi;
    int ldr, ldk;         // row pitches of real / k planes (elements)
    size_t rplane, kplane;  // elements per real / k plane
    int mx1, my1, ntiles, ncells;
    hipStream_t stream;
    qp::DevBuf ffg, ffs;          // Green's function and shape factor [kx][ky], nx*ny
    qp::DevBuf twx, twy;          // twiddles exp(-2 pi i t / 2n), t < 2n
    qp::DevBuf scr_r, scr_k;      // transform scratch, MAX_BATCH planes each
    qp::DevBuf halo;              // deposit tile borders, 8 comps
    qp::DevBuf red;               // reduction partials (double)
    size_t max_lds;               // LDS bytes one workgroup may use
    // kernel-class timing (qp_ctx_profile)
    unsigned prof_on;     // bit k: launches of class k are timed
    int prof_every;       // time one launch in prof_every of a class (0, 1: all)
    int64_t prof_seen[QP_PROF_COUNT];   // launches of each class since the reset
    int prof_tag;         // tag of the launches being enqueued (prof_discard)
    int prof_epoch;       // slices run so far: tags are unique per slice
    std::vector<qp::ProfEv> prof_pending;
    std::vector<hipEvent_t> prof_pool;
    double prof_ms[QP_PROF_COUNT];
    int64_t prof_n[QP_PROF_COUNT];
    double prof_bytes[QP_PROF_COUNT];
    // corrector-loop control.  flags (device ints): word l is set once pass l
    // (or an earlier one) has converged; the kernels of pass l+1 test it and
    // return at once.  h_res (pinned host): residual of pass l in slot
    // l % RES_SLOTS.
    qp::DevBuf flags;
    double *h_res;
    qp::ResJob res_job;   // residual sum carried by the next transform launch
    hipEvent_t pass_ev[2];
    const int *skip;      // non-null while a speculative pass is being enqueued
    int tune[QP_TUNE_COUNT];   // launch-shape variants (qp_ctx_tune)
    // ranorm's state (part3d_lib77.f:1324-1383), which the reference keeps
    // in SAVE'd DATA variables: every seed-0 beam loaded in this context
    // draws from this one stream, beam after beam, as the reference's beams
    // share it in one process
    int32_t rn_state[5];  // r1, r2, r4, r5, iflg
    double rn_r0;
};

struct qp_field2d {
    qp_ctx *ctx;
    int dim;
    int state;   // 0 real, 1 k
    int gcells;
    qp::DevBuf rs, ks;
};

struct qp_field3d {
    qp_ctx *ctx;
    int dim;
    qp::DevBuf buf;   // dim * (nz+1) real planes
};

namespace qp {

// k-space mode grid: arrays [kx][ky] with pitch ldk, cs elements per
// component; tables [kx][ky] with pitch ny (kx < nx, ky < ny).
struct KGrid {
    int nx, ny, ldk;
    size_t cs;
};

// this thread's mode (kx, ky), 0 <= kx <= nx, 0 <= ky <= ny
__device__ __forceinline__ bool kmode(const KGrid &g, int &kx, int &ky)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int n1 = g.ny + 1;
    kx = t / n1;
    ky = t - kx * n1;
    return kx <= g.nx;
}

// dnx = 6.28318530717959/real(nx+nx); dkx = dnx*real(j) (fpois2d_lib77.f:1512)
template <typename R>
__device__ __forceinline__ R dk(int n, int k)
{
    const R d = (R)(TWOPI_REF / (double)(n + n));
    return d * (R)k;
}

inline KGrid kgrid(const qp_ctx *c) { return KGrid{c->nx, c->ny, c->ldk, c->kplane}; }
inline dim3 kgrid_blocks(const qp_ctx *c) { return dim3(((c->nx + 1) * (c->ny + 1) + 255) / 256); }

// ---- transforms (qp_fft.hip) ----
// component (x, y) transform types of fftrk/fftkr kinds; 0 = sine, 1 = cosine
void kind_types(int kind, int dim, int *tx, int *ty);
int fft_r2k(qp_field2d *f, int kind);
int fft_k2r(qp_field2d *f, int kind);
// several fields (at most MAX_BATCH components in all) in one set of launches
int fft_r2k_batch(qp_ctx *c, qp_field2d *const *fs, const int *kinds, int nf);
int fft_k2r_batch(qp_ctx *c, qp_field2d *const *fs, const int *kinds, int nf);

// ---- field kernels (qp_field.hip) ----
int launch_fill(qp_ctx *c, void *p, size_t n, double v);
int copyguard(qp_field2d *f);
int guard_fold(qp_ctx *c, void *planes, int dim);
template <typename R> int bfqp_resid(qp_field2d *cu, qp_field2d *dcu, qp_field2d *amu,
                                     qp_field2d *bxy, double c, double dex, double *num_den);
int bfqp_impl(qp_field2d *cu, qp_field2d *dcu, qp_field2d *amu, qp_field2d *bxy,
              double c, double dex, double *resid_out);

constexpr int BFQP_BLOCKS = 1024;   // blocks of the corrector kernels (FP64 partials)
constexpr int MIX_PARTIALS = 4096;  // offset of the mixing sums in the reduction buffer
__global__ void k_sum_partials(const double *partial, int nblk, double *out);
// Sums of the pairs (p[2i], p[2i+1]), i < n, by the 64 lanes of one wave:
// lane t adds pairs t, t+64, ... in order, then the 64 lane sums are combined
// by a fixed butterfly; every lane gets both sums, the same in every caller.
__device__ __forceinline__ void wave_pair_sum(const double *p, int n, double &a, double &b)
{
    const int t = threadIdx.x & 63;
    a = 0.0;
    b = 0.0;
    for (int i = t; i < n; i += 64) {
        a += p[2 * i];
        b += p[2 * i + 1];
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
}

// the corrector residual from the summed partials
__device__ __forceinline__ double residual_of(double num, double den)
{
    return den > 0.0 ? sqrt(num / den) : (num > 0.0 ? INFINITY : 0.0);
}

// ResJob, run by the first wave of one block
__device__ __forceinline__ void run_res_job(const ResJob &j)
{
    if (j.nblk <= 0 || threadIdx.x >= 64) return;
    const int t = threadIdx.x;
    double a, b;
    wave_pair_sum(j.partial, j.nblk, a, b);
    if (t == 0) {
        const double res = residual_of(a, b);
        j.out[0] = a;
        j.out[1] = b;
        if (j.flag_out)
            *j.flag_out = ((j.flag_in && *j.flag_in) || (j.tol > 0.0 && res < j.tol)) ? 1 : 0;
        if (j.h_slot) {
            *j.h_slot = res;
            __threadfence_system();
        }
    }
}

// ---- fused slice-loop kernels (qp_fused.hip) ----
struct PassFields {
    qp_field2d *cu, *dcu, *amu, *epw, *epwb, *psit, *bxyz;
    qp_field2d *gprev, *fprev;   // corrector history of the mixed iteration (or null)
};
int fused_psi_fxy(qp_field2d *q, qp_field2d *psi, qp_field2d *fxy);
// qp: one corrector pass; otherwise the predictor (which also clears the
// converged flag).  With qp the residual of pass l is written to h_res and,
// when tol > 0 and it is below tol, the converged flag is set.
// seed (predictor only): 1 records B_perp in ehist, 2 also extrapolates
// epw <- 2 epw - ehist first (the corrector seed of QP_TUNE_SEED = 1), 3
// (with ehist2, which keeps the slice before) 3 epw - 3 B(j-2) + B(j-3)
// (QP_TUNE_SEED = 3); with seed_hist (captured slices) the predictor
// reads the count of recorded slices from device memory and extrapolates
// once it is 2 or more (quadratically once it is 3 or more)
int fused_pass(const PassFields &f, double dex, bool qp, int l, double tol, int seed = 0,
               qp_field2d *ehist = nullptr, const int *seed_hist = nullptr,
               qp_field2d *ehist2 = nullptr);
// Anderson mixing (depth 1) of the corrector iterate after pass l (f.gprev,
// f.fprev set; the pass was run by fused_pass with them): unless pass l has
// converged, epw <- g_l - gamma (g_l - g_{l-1}) and B_perp reassembled; mix =
// 0 (first pass, last allowed pass) only keeps g_l for the next pass.
int fused_mix(const PassFields &f, double dex, int l, double tol, bool mix);
int fused_predict_cu(qp_field2d *cu, qp_field2d *dcu, double dxi);
int fused_efield(qp_field2d *fxy, const qp_field2d *bxyz);
// jdev (captured slices): plane lpos - 1 is read from device memory instead;
// ctl (eager slices): the device slice words are set to lpos (the next
// slice) and hist (the seed's recorded slices)
int fused_slice_write(qp_field2d *fxy, const qp_field2d *psit, qp_field2d *bxyz, qp_field2d *psi,
                      qp_field3d *bexyz, qp_field3d *bbxyz, int lpos, double dex, double dd,
                      const int *jdev = nullptr, int *ctl = nullptr, int hist = 0);

// the value a new context gives tuning knob `knob` (qp_ctx_tune)
int tune_default_of(int precision, int knob);

// ---- kernel-class timing ----
int prof_begin(qp_ctx *c, int kind, double bytes);   // returns a slot or -1
void prof_end(qp_ctx *c, int slot);
// forget the timing of pending launches with this tag (a speculative pass
// whose kernels returned at once)
void prof_discard(qp_ctx *c, int tag);

// Brackets one launch with events while profiling is on; bytes is the
// launch's algorithmic traffic.
struct ProfScope {
    qp_ctx *c;
    int slot;
    ProfScope(qp_ctx *cc, int kind, double bytes = 0.0)
        : c(cc), slot((cc->prof_on >> kind) & 1u ? prof_begin(cc, kind, bytes) : -1) {}
    ~ProfScope() { if (slot >= 0) prof_end(c, slot); }
};

// ---- species internals used by the slice driver (qp_particles.hip) ----
int species_qdp_into(qp_species *sp, qp_field2d *q);

}  // namespace qp
