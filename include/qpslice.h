/*
 * qpslice.h -- C ABI of libqpslice.so, the MI355X (gfx950) quasi-static
 * xi-slice solver.
 *
 * The entry points mirror the reference's Fortran-2003 type-bound methods on
 * the hot path, so a Fortran (ISO_C_BINDING), C, C++ or Python caller can
 * replace those classes one for one:
 *   species2d  -> qp_species_*   (species2d_class.f03:37-51)
 *   field2d    -> qp_field2d_*   (field2d_class.f03:38-62)
 *   field3d    -> qp_field3d_*   (field3d_class.f03, ufield3d_class.f03:396-490)
 *   beam3d     -> qp_beam_*      (beam3d_class.f03:35-43)
 *   the xi-slice loop of go_simulation -> qp_sim_* / qp_slice_run
 *                                 (simulation_class.f03:611-765)
 * Conventions kept from the reference: component indices (dim, dim1, dim2,
 * sdim, ddim) and 3-D slice positions (lpos, pos, slice) are 1-based; field
 * states are 0 = real space, 1 = k space; kinds of fftrk/fftkr are 1, 2, 3.
 *
 * All device memory is owned by the library.  Host data crosses only through
 * the upload/download calls, always as double, in Fortran column-major order:
 *   real-space field2d   f(dim, ld1 >= nx+1, ld2 >= ny+1)   (ufield2d rs)
 *   k-space field2d      g(dim, ld1 >= ny+1, ld2 >= nx+1)   (ufield2d ks)
 *   field3d              f(dim, ld1 >= nx+1, ld2 >= ny+1, ld3 >= nz+1)
 *   plasma particles     p(8, np): x, y, px, py, pz, p6, p7, q
 *   beam particles       p(7, np): x, y, z, px, py, pz, q
 * so the reference's own arrays (rf(dim,nx+2,2*ny+1), rf(dim,ny+1,nx+1),
 * rf(dim,nx+2,ny+1,nz+1), part(8,npmax), part(7,npmax)) pass unchanged.
 *
 * Every call returns an int status: QP_OK, one of the reference's particle
 * manager classes irc2(1) = 1..4 (part2d_class.f03:410-579), or a negative
 * error; qp_last_error() describes the last failure of the calling thread.
 * One HIP stream per context; a context and its objects are not thread-safe;
 * distinct contexts are independent.
 */
#ifndef QPSLICE_H
#define QPSLICE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qp_ctx qp_ctx;
typedef struct qp_field2d qp_field2d;
typedef struct qp_field3d qp_field3d;
typedef struct qp_species qp_species;
typedef struct qp_beam qp_beam;
typedef struct qp_sim qp_sim;

enum {
    QP_OK = 0,
    QP_IRC_IHOLE = 1,   /* hole-list overflow  (part2d_class.f03:414)  */
    QP_IRC_PPBUFF = 2,  /* tile buffer overflow (part2d_class.f03:433) */
    QP_IRC_SBUF = 3,    /* send buffer overflow (part2d_class.f03:456) */
    QP_IRC_PPART = 4,   /* tile capacity overflow (part2d_class.f03:527) */
    QP_ERR_ARG = -1,
    QP_ERR_HIP = -2,
    QP_ERR_STATE = -3,  /* field in the wrong space (field2d_class.f03:167) */
    QP_ERR_ALLOC = -4,
    QP_ERR_CHECK = -5,  /* particle outside its tile (PPPCHECK2L) */
    QP_ERR_UNSUPPORTED = -6
};

enum { QP_FP64 = 0, QP_FP32 = 1 };

/* ---------------------------------------------------------------------- */
/* context: the grid descriptors spect2d/spect3d (spect3d_class.f03:16-30) */
/* and the box of simulation%new (simulation_class.f03:165-193)            */
/* ---------------------------------------------------------------------- */
/* box = {xmin, xmax, ymin, ymax, zmin, zmax} in c/wp; grid 2^indx x 2^indy
 * x 2^indz, conducting boundaries (psolver 1), linear interpolation
 * (inorder 1).  precision: QP_FP64 (the reference's DBL build) or QP_FP32
 * (FP32 push/deposit/FFT; FP64 beam positions and reductions). */
int qp_ctx_create(int indx, int indy, int indz, const double box[6],
                  int precision, int device, qp_ctx **out);
int qp_ctx_destroy(qp_ctx *ctx);
int qp_ctx_sync(qp_ctx *ctx);
/* The context's HIP stream (a hipStream_t), on which every call of this
 * library enqueues; callers that compute on fields in place (through
 * qp_field2d_device) enqueue on it too.  No reference counterpart. */
int qp_ctx_stream(const qp_ctx *ctx, void **stream);
int qp_ctx_info(const qp_ctx *ctx, int *nx, int *ny, int *nz, int *precision,
                double *dex, double *dxi);
const char *qp_last_error(void);

/* Kernel timing with HIP events on the context stream, by kernel class.
 * No counterpart on the reference's hot path (its nearest are the wall-clock
 * timers dtimer.c:4-28 and PWTIMERA, fft2d_lib77.f:4518-4521).  While enabled
 * every launch of a class is bracketed by two events; qp_ctx_profile_get
 * synchronises the stream and returns the accumulated time and launch count. */
enum { QP_PROF_FFT = 0, QP_PROF_TRANSPOSE, QP_PROF_KSPACE, QP_PROF_ELEMENTWISE,
       QP_PROF_DEPOSIT_Q, QP_PROF_EXTPSI, QP_PROF_AMJDP, QP_PROF_PUSH,
       QP_PROF_SORT, QP_PROF_BORDER, QP_PROF_BEAM_DEPOSIT, QP_PROF_BEAM_PUSH,
       QP_PROF_COUNT };
int qp_ctx_profile(qp_ctx *ctx, int enable);   /* enabling also resets */
/* the same for the classes in mask only (bit k = class k; 0 = off), so a
 * timed run can carry events on one kernel class at little cost */
int qp_ctx_profile_classes(qp_ctx *ctx, unsigned mask);
/* Time one launch in `every` of each enabled class (default 1: all): the
 * launches of a long run are sampled evenly, and the events' own cost stays
 * small.  The counts and bytes qp_ctx_profile_get returns are those of the
 * timed launches. */
int qp_ctx_profile_every(qp_ctx *ctx, int every);
/* bytes: the algorithmic bytes of those launches (operands read or written
 * once: particle records and field planes), the figure a roofline uses */
int qp_ctx_profile_get(qp_ctx *ctx, int kind, double *total_ms, int64_t *count,
                       double *bytes);

/* Launch-shape tuning, e.g. for an autotuner.  No counterpart in the
 * reference, whose deposit tiling is fixed (16x16 tiles, part2d_class.f03:99).
 * QP_TUNE_DEPOSIT_Q / QP_TUNE_AMJDP select one of the compiled shapes of the
 * charge / current deposit (rows per group, fixed or, for the FP32 current
 * deposit's default, chosen per tile from its particle count; staging chunk;
 * occupancy); shapes differ only in how a cell's particles are split over
 * lanes, so results agree to rounding.  QP_TUNE_SPECULATE = 1 lets qp_slice_run enqueue the
 * deposits of corrector pass l+1 before reading the residual of pass l (they
 * return at once if pass l converged; the default); 0 waits for each residual;
 * 2 also enqueues as many whole passes as the last slice used before the
 * first residual read (the passes after a converged one return at once), so
 * a slice that converges within them costs one host wait.  The passes and
 * results are the same in every mode.
 * QP_TUNE_SEED = 1 seeds the corrector of a slice run
 * with a residual tolerance by linear extrapolation 2 B(xi-1) - B(xi-2)
 * instead of the reference's B(xi-1) (default 0); QP_TUNE_SEED = 2 keeps the
 * B_perp a caller wrote between qp_slice_prepare and qp_slice_finish;
 * QP_TUNE_SEED = 3 extrapolates quadratically, 3 B(xi-1) - 3 B(xi-2) +
 * B(xi-3) (linearly while only two slices are recorded).  The
 * seed changes only how many passes reach the tolerance; fixed-pass runs
 * (tol <= 0) use the reference's seed whatever the knob.
 * QP_TUNE_FFT = 1 folds the transpose between the x and y passes of every
 * transform into the y pass (column-pair reads / writes) instead of a
 * separate LDS-tiled transpose (the FP64 default).  0 (the FP32 default),
 * 2, 3 and 4 keep the transpose and pad the transforms' LDS by one point per
 * 16 points in FP32 or 8 in FP64 (0), or per 8, 32 or 4 points (2, 3, 4);
 * 5 and 6 fold the transpose only in the k -> real or only in the real -> k
 * direction.  Results are identical.
 * QP_TUNE_PUSH selects how the plasma push re-sorts the particles: 0 (default)
 * counts the destination bins in a first pass that stores nothing per
 * particle, then pushes again and stores every particle at its sorted place;
 * 1 pushes in place, keeping each particle's bin and rank, then scatters.
 * Both give the same arrays, bit for bit.
 * QP_TUNE_MIX = 1 accelerates the corrector of runs to a tolerance (tol > 0)
 * by Anderson mixing of depth 1: after pass l >= 2 that has not converged,
 * the next iterate of B_perp is g_l - gamma (g_l - g_{l-1}), g_l the pass's
 * result and gamma = <f_l, f_l - f_{l-1}> / |f_l - f_{l-1}|^2 over the
 * changes f = g - (the pass's input).  The stopping test and the result
 * of the converged pass are the plain iteration's; only the passes to reach
 * the tolerance change.  Default 0 (the reference's iteration).
 * QP_TUNE_STRICT = 1 makes qp_slice_run wait for its slice's particle move
 * check (PPPCHECK2L) and fail at that slice, as the reference stops there.
 * Default 0: the check is read without waiting at a later slice's start or
 * at the end of the step, and its error names the slice whose push failed
 * (the slices after it are invalid).
 * value -1 restores a knob's default. */
enum { QP_TUNE_DEPOSIT_Q = 0, QP_TUNE_AMJDP = 1, QP_TUNE_SPECULATE = 2, QP_TUNE_SEED = 3,
       QP_TUNE_FFT = 4, QP_TUNE_PUSH = 5, QP_TUNE_MIX = 6, QP_TUNE_STRICT = 7, QP_TUNE_COUNT };
int qp_ctx_tune(qp_ctx *ctx, int knob, int value);
/* current value and the number of values of a knob */
int qp_ctx_tune_get(const qp_ctx *ctx, int knob, int *value, int *nvalues);

/* ---------------------------------------------------------------------- */
/* field2d (field2d_class.f03)                                             */
/* ---------------------------------------------------------------------- */
/* init_field2d :81-135 with fftflag = .true.; state 0 or 1 */
int qp_field2d_create(qp_ctx *ctx, int dim, int state, qp_field2d **out);
int qp_field2d_destroy(qp_field2d *f);
int qp_field2d_state(const qp_field2d *f, int *state, int *dim);
/* space: 0 = real-space array, 1 = k-space array; sets the state */
int qp_field2d_upload(qp_field2d *f, int space, const double *host, int ld1,
                      int ld2);
int qp_field2d_download(const qp_field2d *f, int space, double *host, int ld1,
                        int ld2);
/* device storage of one space: element (c, row, col) at
 * ptr + c*comp_stride + row*ld (elements of the context's precision; real:
 * rows y = 0..ny, cols x = 0..nx; k: rows kx, cols ky).  For callers that
 * compute on the fields in place (the learned seed); no reference
 * counterpart. */
int qp_field2d_device(const qp_field2d *f, int space, void **ptr, int *ld,
                      size_t *comp_stride);
int qp_field2d_as(qp_field2d *f, double value);                    /* asc :493 */
int qp_field2d_asa(qp_field2d *f, const qp_field2d *that);         /* asa :624 */
int qp_field2d_fftrk(qp_field2d *f, int kind);                     /* :157-191 */
int qp_field2d_fftkr(qp_field2d *f, int kind);                     /* :193-227 */
int qp_field2d_div(qp_field2d *f, qp_field2d *out);                /* :229-249 */
int qp_field2d_grad(qp_field2d *f, qp_field2d *out);               /* :251-271 */
int qp_field2d_pot(qp_field2d *q, qp_field2d *out);                /* :295-316 */
int qp_field2d_smooth(qp_field2d *q, qp_field2d *out);             /* :341-362 */
int qp_field2d_elf(qp_field2d *q, qp_field2d *out);                /* :364-385 */
int qp_field2d_bf(qp_field2d *cu, qp_field2d *out);                /* :387-408 */
/* bfield_qp :410-432: B_perp update into bxy; dcu is overwritten */
int qp_field2d_bfqp(qp_field2d *cu, qp_field2d *dcu, qp_field2d *amu,
                    qp_field2d *bxy, double c, double dex);
/* sum1/sum2, minus1/minus2 (:665-760): n = 0 means every component;
 * otherwise this(dim(i)) = a1(dim1(i)) op a2(dim2(i)), i = 1..n */
int qp_field2d_add(qp_field2d *f, const qp_field2d *a1, const qp_field2d *a2,
                   int n, const int *dim, const int *dim1, const int *dim2);
int qp_field2d_sub(qp_field2d *f, const qp_field2d *a1, const qp_field2d *a2,
                   int n, const int *dim, const int *dim1, const int *dim2);
/* multiply1/multiply2 (:762-807): n = 0 means every component times
 * value[0]; otherwise this(dim(i)) = a1(dim1(i)) * value(i) */
int qp_field2d_mult(qp_field2d *f, const qp_field2d *a1, int n,
                    const int *dim, const int *dim1, const double *value);
int qp_field2d_ag(qp_field2d *f);                                  /* :454 */
int qp_field2d_cg(qp_field2d *f);                                  /* :434 */
/* copyto :538-560: 3-D(ddim(i), lpos) = this(sdim(i)) */
int qp_field2d_cb(const qp_field2d *f, qp_field3d *f3, int lpos, int n,
                  const int *sdim, const int *ddim);
/* copyadd :586-610: this(ddim(i)) += 3-D(sdim(i), lpos) */
int qp_field2d_ca(qp_field2d *f, const qp_field3d *f3, int lpos, int n,
                  const int *sdim, const int *ddim);

/* ---------------------------------------------------------------------- */
/* field3d (field3d_class.f03, ufield3d_class.f03:74-120)                  */
/* ---------------------------------------------------------------------- */
int qp_field3d_create(qp_ctx *ctx, int dim, qp_field3d **out);
int qp_field3d_destroy(qp_field3d *f);
int qp_field3d_as(qp_field3d *f, double value);
int qp_field3d_upload(qp_field3d *f, const double *host, int ld1, int ld2,
                      int ld3);
int qp_field3d_download(const qp_field3d *f, double *host, int ld1, int ld2,
                        int ld3);
/* one xi plane (lpos 1-based) of one component, real layout (ld1 >= nx+1) */
int qp_field3d_download_plane(const qp_field3d *f, int comp, int lpos,
                              double *host, int ld1, int ld2);

/* ---------------------------------------------------------------------- */
/* species2d (species2d_class.f03) with the uniform profile 0             */
/* (fdist2d_class.f03:168-316)                                             */
/* ---------------------------------------------------------------------- */
typedef struct {
    int ppc[2];       /* particles per cell in x, y */
    double density;   /* n/n0 */
    double q, m;      /* charge and mass (qbm = q/m) */
} qp_species_params;

/* init_species2d :74-118; dt is the xi step (simulation_class.f03:578) */
int qp_species_create(qp_ctx *ctx, const qp_species_params *p, double dt,
                      double s, qp_species **out);
int qp_species_destroy(qp_species *sp);
int qp_species_renew(qp_species *sp, double s);                   /* :138 */
/* the piecewise longitudinal density of profile 0 (fdist2d_class.f03:
 * 190-201, 233-251): the charge of every (re)load at propagation distance s
 * is scaled by fs interpolated at s between the n points s_pts (n = 0: a
 * uniform plasma); reloads the plasma at s_now.  QP_ERR_ARG with the
 * reference's messages when s lies outside the points or the points
 * decrease. */
int qp_species_set_longitudinal(qp_species *sp, int n, const double *s_pts,
                                const double *fs, double s_now);
/* Transverse profiles other than uniform (profile 0): 10, an elliptical
 * annulus (dist2d_010, fdist2d_class.f03:318-523); 11, a Gaussian ring,
 * q weighted by exp(-(r - radius)^2 / (2 width^2)) (dist2d_011, :525-711);
 * 12, a piecewise-linear radial density fr(r) (dist2d_012, :713-950).
 * Lengths in c/wp, center in box coordinates; the longitudinal profile
 * still scales every reload.  Reloads the plasma at s_now.  QP_ERR_ARG
 * with the reference's message when the radial points do not increase. */
typedef struct {
    int profile;               /* 0, 10, 11 or 12 */
    double center[2];          /* 10, 11, 12 */
    double inner_radius[2];    /* 10: x and y semi-axes */
    double outer_radius[2];    /* 10 */
    double radius, width;      /* 11 */
    int nr;                    /* 12: number of radial points */
    const double *r;           /* 12: radii */
    const double *fr;          /* 12: relative densities at r */
} qp_species_profile;
int qp_species_set_profile(qp_species *sp, const qp_species_profile *p, double s_now);
int qp_species_qdp(qp_species *sp, qp_field2d *q);                /* :166 */
/* qp_species_qdp without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The deposit's tile-border halo lives in a
 * static device array of the library (2 MiB per GPU, held until the process
 * ends) instead of the context's buffer, so a grid's halo (64 border nodes
 * per 16 x 16-cell tile) must fit it: up to 4096 tiles in FP64 (1024 x 1024)
 * and 8192 in FP32 (2048 x 1024); larger grids return QP_ERR_UNSUPPORTED.
 * Results equal qp_species_qdp's.  Calls that overlap on one GPU (other
 * threads, or unordered streams) share the array and race. */
int qp_species_qdp_noalloc(qp_species *sp, qp_field2d *q);
int qp_species_extpsi(qp_species *sp, const qp_field2d *psi, double dex);
int qp_species_amjdp(qp_species *sp, const qp_field2d *ef,
                     const qp_field2d *bf, const qp_field2d *psit,
                     qp_field2d *cu, qp_field2d *amu, qp_field2d *dcu,
                     double dex);                                  /* :187 */
int qp_species_push(qp_species *sp, const qp_field2d *ef,
                    const qp_field2d *bf, const qp_field2d *psit,
                    double dex);                                   /* :220 */
int qp_species_cbq(qp_species *sp, qp_field3d *q3, int pos);      /* :399 */
int qp_species_count(const qp_species *sp, int64_t *np);
/* particles in the library's cell order, p(8, ld >= np) */
int qp_species_download(const qp_species *sp, double *part, int64_t ld);
int qp_species_upload(qp_species *sp, const double *part, int64_t np,
                      int64_t ld);
/* per-cell counts, cell = y*nx + x (ny*nx values) */
int qp_species_cell_counts(const qp_species *sp, int32_t *counts);
/* per-tile counts in the reference's 16x16 tile order (kpic, PPDBLKP2L) */
int qp_species_tile_counts(const qp_species *sp, int32_t *kpic);

/* ---------------------------------------------------------------------- */
/* beam3d (beam3d_class.f03) with the tri-Gaussian profile 0              */
/* (fdist3d_class.f03:222-421, PRVDIST32_RANDOM part3d_lib77.f:411-552)    */
/* ---------------------------------------------------------------------- */
typedef struct {
    double center[3];      /* c/wp, in box coordinates */
    double sigma[3];       /* c/wp */
    double sigma_v[3];     /* mc */
    double centroid_x[3];  /* centroid polynomial coefficients */
    double centroid_y[3];
    double gamma, peak, q, m;
    int np[3];             /* npx, npy, npz */
    int quiet;
    int evolution;         /* 0: the beam is not pushed (beam3d_class.f03:170) */
    int64_t npmax;
} qp_beam_params;

/* Beam profile 2 (fdist3d_002, fdist3d_class.f03:652-821): transverse phase
 * space from Twiss parameters, Gaussian in z, loaded by the algorithm of
 * PRVDIST32_TWISS (part3d_lib77.f:2-192): cut-off at 3 sigma in each
 * coordinate and at radius 3 in the normalised transverse pair, momentum
 * p_x = emit_x/sigma_x r - gamma alpha_x/beta_x (x - x0), p_z = gamma +
 * sigma_vz r.  Lengths in c/wp as in the input deck (beta, emittance, centre,
 * sigma_z); sigma_x = sqrt(beta_x emittance_x / gamma). */
typedef struct {
    double center[3];      /* c/wp, in box coordinates */
    double alpha[2];       /* Twiss alpha, x and y */
    double beta[2];        /* Twiss beta, x and y (c/wp) */
    double emittance[2];   /* normalised emittance, x and y (c/wp) */
    double sigma_z;        /* c/wp */
    double sigma_vz;       /* mc */
    double gamma, peak, q, m;
    int np[3];             /* npx, npy, npz */
    int quiet;
    int evolution;
    int64_t npmax;
} qp_beam_twiss_params;

/* init_beam3d :61-93; seed 0 continues the context's ranorm stream, which
 * starts at the reference's DATA state (ranorm's SAVE'd state, shared by
 * every beam the reference loads); other seeds start independent streams.
 * np = {0,0,0} with npmax > 0 makes an empty beam of that capacity, to be
 * filled by qp_beam_upload. */
int qp_beam_create(qp_ctx *ctx, const qp_beam_params *p, double dt,
                   uint64_t seed, qp_beam **out);
/* Beam profile 3 (init_fdist3d_003 + dist3d_003, fdist3d_class.f03:
 * 823-974): np records "z x y vz vx vy q" read from a text file
 * (list-directed), positions in units of 5.32150254e9/sqrt(n0) converted to
 * grid units (x, y about center, z = -z), kept inside the grid.  QP_ERR_ARG
 * "Read beam file error" when the file holds fewer than np records. */
typedef struct {
    const char *file_name;
    double center[3];      /* c/wp, box coordinates (x and y used) */
    double n0;             /* simulation.n0 */
    double q, m;
    int np;                /* records to read */
    int evolution;
    int64_t npmax;         /* 0: np + 2 */
} qp_beam_file_params;
int qp_beam_create_file(qp_ctx *ctx, const qp_beam_file_params *p, double dt,
                        qp_beam **out);
/* Beam profile 100 (init_fdist3d_100 + dist3d_100, fdist3d_class.f03:
 * 976-1166): a quiet-start lattice of ppc(1:3) particles per cell over the
 * range box, p = (0, 0, gamma), charge sign(q)/(ppc product) * np *
 * exp(-((z - zc)^2 + (|r_perp - c| - r1)^2) / (2 r2^2)) in cell units
 * (r1, r2 in x cells): a Gaussian torus. */
typedef struct {
    double center[3];      /* c/wp, box coordinates */
    double range_x[2], range_y[2], range_z[2];   /* c/wp */
    int ppc[3];
    double q, m, np, gamma;
    double r1, r2;         /* c/wp */
    int evolution;
    int64_t npmax;         /* 0: the lattice's count + 2 */
} qp_beam_torus_params;
int qp_beam_create_torus(qp_ctx *ctx, const qp_beam_torus_params *p, double dt,
                         qp_beam **out);
/* init_fdist3d_002 + dist3d_002 (fdist3d_class.f03:652-821); seeds as above */
int qp_beam_create_twiss(qp_ctx *ctx, const qp_beam_twiss_params *p, double dt,
                         uint64_t seed, qp_beam **out);
int qp_beam_destroy(qp_beam *b);
int qp_beam_qdp(qp_beam *b);                                       /* :111 */
int qp_beam_qdpcopy(qp_beam *b, qp_field2d *q, int slice);         /* :136 */
/* drops the particles that leave [1, n-1) and leaves the survivors sorted by
 * cell (z, then y, then x); the reference fills holes from the end instead,
 * so the survivor set is the same and the order is not */
int qp_beam_push(qp_beam *b, const qp_field3d *ef, const qp_field3d *bf,
                 double dex, double dez);                          /* :155 */
int qp_beam_count(const qp_beam *b, int64_t *np);
int qp_beam_download(const qp_beam *b, double *part, int64_t ld);
int qp_beam_upload(qp_beam *b, const double *part, int64_t np, int64_t ld);
int qp_beam_charge_download(const qp_beam *b, double *host, int ld1, int ld2,
                            int ld3);
/* sum of gamma and of the beam charge (FP64 reductions) */
int qp_beam_moments(const qp_beam *b, double *sum_gamma, double *sum_q);

/* ---------------------------------------------------------------------- */
/* the 3-D step and its xi-slice loop (simulation_class.f03:611-765)      */
/* ---------------------------------------------------------------------- */
/* which fields qp_sim_field2d returns */
enum { QP_F_QB = 0, QP_F_QE, QP_F_PSIT, QP_F_PSI, QP_F_DIVVPOT, QP_F_FXY,
       QP_F_BXYZ, QP_F_CU, QP_F_DCU, QP_F_AMU, QP_F_EPW, QP_F_EPWB,
       QP_F_COUNT };

int qp_sim_create(qp_ctx *ctx, qp_species *const *sp, int nsp,
                  qp_beam *const *bm, int nb, int iter, qp_sim **out);
int qp_sim_destroy(qp_sim *sim);
int qp_sim_field2d(qp_sim *sim, int which, qp_field2d **out);
/* bexyz (Ex, Ey, Ez) and bbxyz (Bx, By, Bz) 3-D wake arrays */
int qp_sim_field3d(qp_sim *sim, qp_field3d **bexyz, qp_field3d **bbxyz);
/* optional target of the species charge diagnostic cbq (species index from
 * 0 in the order given to qp_sim_create); without one cbq still runs */
int qp_sim_species_q3(qp_sim *sim, int species, qp_field3d *q3);
/* optional 3-D diagnostics of the slice loop (simulation_class.f03:670-671,
 * 696-698): which = QP_DIAG_PSI, psi (times dex^2) into plane j+1 of a
 * 1-component field; QP_DIAG_CU, the last corrector pass's current (times
 * dex) into a 3-component field; f3 = NULL stops it.  The reference keeps
 * these when its input deck asks for psi or jx/jy/jz output. */
enum { QP_DIAG_PSI = 0, QP_DIAG_CU = 1 };
int qp_sim_diag3d(qp_sim *sim, int which, qp_field3d *f3);
/* start of a 3-D step: beam deposits, zero 2-D fields (pipeline receive on
 * the first stage, ufield2d_class.f03:528-531), slice-1 copies (:639-641) */
int qp_sim_begin_step(qp_sim *sim);
/* one slice j (1-based).  tol <= 0: exactly iter_max corrector passes, as
 * the reference (iter_max = iter).  tol > 0: passes until the relative
 * change of B_perp in k space drops below tol, at most iter_max.  Reports
 * the passes used and the last relative change (FP64). */
int qp_slice_run(qp_sim *sim, int j, int iter_max, double tol,
                 int *iters_used, double *residual);
/* qp_slice_run in two parts, for a caller that sets the corrector's seed
 * itself (the learned seed): prepare runs the slice up to the seed (beam and
 * plasma charge, psi, E_perp, :650-666, 670); finish runs the rest, the
 * species' psi gather (:667-669) inside its first corrector pass.  Between them the caller may overwrite B_perp (QP_F_EPW, k space)
 * through qp_field2d_device; with QP_TUNE_SEED = 2 finish keeps it as the
 * seed.  No counterpart in the reference. */
int qp_slice_prepare(qp_sim *sim, int j);
int qp_slice_finish(qp_sim *sim, int j, int iter_max, double tol, int *iters_used,
                    double *residual);
/* The corrector residuals of the last slice finished with tol > 0, pass 1
 * first (at most 63 and at most max of them; *n gets the count).  A
 * diagnostic with no reference counterpart. */
int qp_slice_residuals(const qp_sim *sim, double *res, int max, int *n);
/* end of the 3-D step: beam push (:752-757) and species renew (:761-764) */
int qp_sim_end_step(qp_sim *sim, double s);
/* its two parts, for a caller that writes diagnostics in between, as the
 * reference does (diag_simulation, :759): the beam push, then the renewal
 * at propagation distance s */
int qp_sim_push_beams(qp_sim *sim);
int qp_sim_renew(qp_sim *sim, double s);

/* ---------------------------------------------------------------------- */
/* batched parameter scan (no counterpart in the reference, which runs one */
/* simulation per MPI job)                                                 */
/* ---------------------------------------------------------------------- */
/* One HIP graph holding one xi slice of each of n simulations (their own
 * contexts, so their own streams; any grids, species, beams and
 * precisions; all on one device).  The slice is qp_slice_run's with these
 * differences: the slice index is a device word of each simulation
 * (qp_sim_begin_step sets it to 1, each captured slice advances it), so the
 * one graph serves every slice; with tol > 0 all iter_max passes are
 * enqueued and the kernels of the passes after the converged one return at
 * once (the stopping test and the result are qp_slice_run's); the particle
 * move check is read once per qp_scan_run; the mixed corrector and the 3-D
 * diagnostics (cbq, psi, current) are not available (QP_ERR_UNSUPPORTED).
 * Create after qp_sim_begin_step; the simulations must have profiling off. */
typedef struct qp_scan qp_scan;
int qp_scan_create(qp_sim *const *sims, int n, int iter_max, double tol, qp_scan **out);
/* nslices slices of every simulation (each continues from its last slice),
 * ordered after the work already enqueued on the simulations' streams and
 * before later work; waits for them, then checks the particle moves */
int qp_scan_run(qp_scan *scan, int nslices);
/* passes[j - 1] = corrector passes of slice j (j = 1..nslices) of one
 * simulation, as the captured slices recorded them */
int qp_scan_passes(const qp_scan *scan, int instance, int *passes, int nslices);
int qp_scan_destroy(qp_scan *scan);

#ifdef __cplusplus
}
#endif
#endif /* QPSLICE_H */
