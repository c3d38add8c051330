// Context, device buffers, error reporting and the solver tables.
#include <cmath>
#include <cstdarg>
#include <cstring>

#include <mutex>

#include "qp_common.h"

// The scratch array of qp_species_qdp_noalloc.  Host code never uses this
// name as an address: each GPU holds its own copy, whose device address
// noalloc_setup looks up with hipGetSymbolAddress and keeps per device.
static __device__ __align__(256) unsigned char scratch[qp::NOALLOC_BYTES];

namespace qp {

static std::mutex g_noalloc_mu;
static void *g_noalloc_ptr[NOALLOC_DEVICES];

int noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_DEVICES) return QP_OK;   // the entry point refuses it
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return QP_OK;
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return hip_fail(e == hipSuccess ? hipErrorInvalidSymbol : e, "hipGetSymbolAddress(scratch)",
                        __FILE__, __LINE__);
    g_noalloc_ptr[device] = ptr;
    return QP_OK;
}

void *noalloc_scratch(int device)
{
    if (device < 0 || device >= NOALLOC_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e),
              what, file, line);
    return QP_ERR_HIP;
}

int DevBuf::alloc(size_t n)
{
    release();
    if (n == 0) return QP_OK;
    hipError_t e = hipMalloc(&p, n);
    if (e != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
        return QP_ERR_ALLOC;
    }
    bytes = n;
    // zeroed on the null stream and waited for: the contexts' streams are
    // non-blocking, so later work on them is not ordered after the null stream
    e = hipMemset(p, 0, n);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return hip_fail(e, "hipMemset", __FILE__, __LINE__);
    return QP_OK;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

static size_t round_up(size_t v, size_t m) { return (v + m - 1) / m * m; }

// Poisson table of MPPOISD22 with isign = 0 (fpois2d_lib77.f:1514-1531):
// g = affp*s/k^2 (affp at k = 0), s = exp(-((kx ax)^2 + (ky ay)^2)/2),
// kx = pi j/nx, ky = pi k/ny, with the reference's 2*pi literal.  Computed in
// double in the reference's operation order, then stored as Real.
static void pois_table(int nx, int ny, std::vector<double> &g, std::vector<double> &s)
{
    const double ax = 0.912871, ay = 0.912871, affp = 1.0;  // field2d_class.f03:129
    const double dnx = TWOPI_REF / (double)(nx + nx);
    const double dny = TWOPI_REF / (double)(ny + ny);
    g.resize((size_t)nx * ny);
    s.resize((size_t)nx * ny);
    for (int j = 0; j < nx; ++j) {
        double dkx = dnx * (double)j;
        double at1 = dkx * dkx;
        double at2 = (dkx * ax) * (dkx * ax);
        for (int k = 0; k < ny; ++k) {
            double dky = dny * (double)k;
            double at3 = dky * dky + at1;
            double at4 = std::exp(-0.5 * ((dky * ay) * (dky * ay) + at2));
            size_t o = (size_t)j * ny + k;
            if (at3 == 0.0) { g[o] = affp; s[o] = 1.0; }
            else { g[o] = affp * at4 / at3; s[o] = at4; }
        }
    }
}

template <typename R>
static int upload_real(DevBuf &b, const std::vector<double> &v)
{
    std::vector<R> h(v.begin(), v.end());
    int rc = b.alloc(h.size() * sizeof(R));
    if (rc) return rc;
    QP_HIP(hipMemcpy(b.p, h.data(), h.size() * sizeof(R), hipMemcpyHostToDevice));
    return QP_OK;
}

// twiddles W[t] = exp(-2 pi i t / m), t = 0..m-1, accurate to double
template <typename R>
static int upload_twiddles(DevBuf &b, int m)
{
    std::vector<cplx<R>> h((size_t)m);
    for (int t = 0; t < m; ++t) {
        long double a = 2.0L * 3.14159265358979323846264338327950288L * t / m;
        h[t].x = (R)cosl(a);
        h[t].y = (R)(-sinl(a));
    }
    int rc = b.alloc(h.size() * sizeof(cplx<R>));
    if (rc) return rc;
    QP_HIP(hipMemcpy(b.p, h.data(), h.size() * sizeof(cplx<R>), hipMemcpyHostToDevice));
    return QP_OK;
}

template <typename R>
static int make_tables(qp_ctx *c)
{
    std::vector<double> g, s;
    pois_table(c->nx, c->ny, g, s);
    int rc = upload_real<R>(c->ffg, g);
    if (!rc) rc = upload_real<R>(c->ffs, s);
    if (!rc) rc = upload_twiddles<R>(c->twx, 2 * c->nx);
    if (!rc) rc = upload_twiddles<R>(c->twy, 2 * c->ny);
    return rc;
}

static hipEvent_t pool_get(qp_ctx *c)
{
    if (!c->prof_pool.empty()) {
        hipEvent_t e = c->prof_pool.back();
        c->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int prof_begin(qp_ctx *c, int kind, double bytes)
{
    if (c->prof_every > 1 && c->prof_seen[kind]++ % c->prof_every != 0) return -1;
    ProfEv ev{pool_get(c), pool_get(c), kind, bytes, c->prof_tag};
    if (!ev.a || !ev.b) return -1;
    if (hipEventRecord(ev.a, c->stream) != hipSuccess) return -1;
    c->prof_pending.push_back(ev);
    return (int)c->prof_pending.size() - 1;
}

void prof_end(qp_ctx *c, int slot)
{
    (void)hipEventRecord(c->prof_pending[slot].b, c->stream);
}

void prof_discard(qp_ctx *c, int tag)
{
    std::vector<ProfEv> keep;
    for (ProfEv &e : c->prof_pending) {
        if (e.tag == tag) {
            c->prof_pool.push_back(e.a);
            c->prof_pool.push_back(e.b);
        } else {
            keep.push_back(e);
        }
    }
    c->prof_pending.swap(keep);
}

static int prof_resolve(qp_ctx *c)
{
    QP_HIP(hipStreamSynchronize(c->stream));
    for (ProfEv &e : c->prof_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
            c->prof_ms[e.kind] += ms;
            c->prof_n[e.kind] += 1;
            c->prof_bytes[e.kind] += e.bytes;
        }
        c->prof_pool.push_back(e.a);
        c->prof_pool.push_back(e.b);
    }
    c->prof_pending.clear();
    return QP_OK;
}

}  // namespace qp

using namespace qp;

// values of each tuning knob (qp_ctx_tune) and the value a context starts
// with: speculative deposits on (bit-identical results, measured 1 % faster
// at 1024^2, 16 ppc), every other knob at its first value, except that FP64
// transforms read / write the y pass as column pairs (no transpose kernel;
// identical results, 8 % less transform time at 1024^2 in FP64, 5 % more in
// FP32)
static const int TUNE_VALUES[QP_TUNE_COUNT] = {4, 15, 3, 4, 7, 2, 2, 2};
static const int TUNE_DEFAULT[QP_TUNE_COUNT] = {0, 0, 1, 0, 0, 0, 0, 0};

namespace qp {
int tune_default_of(int precision, int knob)
{
    if (knob == QP_TUNE_FFT && precision == QP_FP64) return 1;
    return TUNE_DEFAULT[knob];
}
}  // namespace qp

static int tune_default(const qp_ctx *c, int knob) { return tune_default_of(c->prec, knob); }

extern "C" {

const char *qp_last_error(void) { return qp::g_err; }

int qp_ctx_create(int indx, int indy, int indz, const double box[6], int precision,
                  int device, qp_ctx **out)
{
    QP_REQUIRE(out && box, QP_ERR_ARG, "qp_ctx_create: null argument");
    QP_REQUIRE(indx >= 4 && indy >= 4 && indx <= 12 && indy <= 12, QP_ERR_UNSUPPORTED,
               "qp_ctx_create: transverse grid must be 2^4..2^12 (got 2^%d x 2^%d)", indx, indy);
    QP_REQUIRE(indz >= 1 && indz <= 13, QP_ERR_ARG, "qp_ctx_create: bad indz %d", indz);
    QP_REQUIRE(precision == QP_FP64 || precision == QP_FP32, QP_ERR_ARG,
               "qp_ctx_create: bad precision %d", precision);
    QP_HIP(hipSetDevice(device));
    {
        int rc0 = qp::noalloc_setup(device);   // with the context's device current
        if (rc0) return rc0;
    }
    qp_ctx *c = new qp_ctx();
    c->indx = indx; c->indy = indy; c->indz = indz;
    c->nx = 1 << indx; c->ny = 1 << indy; c->nz = 1 << indz;
    c->prec = precision;
    for (int k = 0; k < QP_TUNE_COUNT; ++k) c->tune[k] = tune_default(c, k);
    c->device = device;
    // ranorm's DATA state (part3d_lib77.f:1335-1337)
    c->rn_state[0] = 885098780; c->rn_state[1] = 1824280461;
    c->rn_state[2] = 1396483093; c->rn_state[3] = 55318673; c->rn_state[4] = 0;
    c->rn_r0 = 0.0;
    c->esize = precision == QP_FP64 ? 8 : 4;
    std::memcpy(c->box, box, sizeof(c->box));
    // simulation_class.f03:174-188: dx = (max-min)/2^indx etc.
    c->dex = (box[1] - box[0]) / (double)(1 << indx);
    c->dey = (box[3] - box[2]) / (double)(1 << indy);
    c->dxi = (box[5] - box[4]) / (double)(1 << indz);
    c->ldr = (int)round_up((size_t)c->nx + 1, 64);
    c->ldk = (int)round_up((size_t)c->ny + 1, 64);
    c->rplane = (size_t)(c->ny + 1) * c->ldr;
    c->kplane = (size_t)(c->nx + 1) * c->ldk;
    c->mx1 = c->nx / TILE;
    c->my1 = c->ny / TILE;
    c->ntiles = c->mx1 * c->my1;
    c->ncells = c->nx * c->ny;
    int lds = 0;
    hipError_t e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    c->max_lds = (e == hipSuccess && lds > 0) ? (size_t)lds : 65536;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return hip_fail(e, "hipStreamCreate", __FILE__, __LINE__); }
    int rc = precision == QP_FP64 ? make_tables<double>(c) : make_tables<float>(c);
    size_t pl = c->rplane > c->kplane ? c->rplane : c->kplane;
    if (!rc) rc = c->scr_r.alloc(MAX_BATCH * pl * c->esize);
    if (!rc) rc = c->scr_k.alloc(MAX_BATCH * pl * c->esize);
    if (!rc) rc = c->halo.alloc((size_t)8 * c->ntiles * HALO * c->esize);
    if (!rc) rc = c->red.alloc(sizeof(double) * 8192);
    if (!rc) rc = c->flags.alloc(sizeof(int) * 64);
    if (rc) { qp_ctx_destroy(c); return rc; }
    e = hipHostMalloc((void **)&c->h_res, sizeof(double) * RES_SLOTS, hipHostMallocDefault);
    if (e != hipSuccess) { c->h_res = nullptr; qp_ctx_destroy(c); return hip_fail(e, "hipHostMalloc", __FILE__, __LINE__); }
    for (int i = 0; i < RES_SLOTS; ++i) c->h_res[i] = -1.0;
    for (int i = 0; i < 2; ++i) {
        e = hipEventCreateWithFlags(&c->pass_ev[i], hipEventDisableTiming);
        if (e != hipSuccess) { c->pass_ev[i] = nullptr; qp_ctx_destroy(c); return hip_fail(e, "hipEventCreate", __FILE__, __LINE__); }
    }
    *out = c;
    return QP_OK;
}

int qp_ctx_destroy(qp_ctx *c)
{
    if (!c) return QP_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->ffg.release(); c->ffs.release(); c->twx.release(); c->twy.release();
    c->scr_r.release(); c->scr_k.release(); c->halo.release(); c->red.release();
    c->flags.release();
    if (c->h_res) (void)hipHostFree(c->h_res);
    for (int i = 0; i < 2; ++i)
        if (c->pass_ev[i]) (void)hipEventDestroy(c->pass_ev[i]);
    for (ProfEv &e : c->prof_pending) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
    for (hipEvent_t e : c->prof_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return QP_OK;
}

int qp_ctx_profile_classes(qp_ctx *c, unsigned mask)
{
    QP_REQUIRE(c, QP_ERR_ARG, "qp_ctx_profile_classes: null context");
    int rc = prof_resolve(c);
    if (rc) return rc;
    for (int k = 0; k < QP_PROF_COUNT; ++k) {
        c->prof_ms[k] = 0.0;
        c->prof_n[k] = 0;
        c->prof_bytes[k] = 0.0;
        c->prof_seen[k] = 0;
    }
    c->prof_on = mask & ((1u << QP_PROF_COUNT) - 1u);
    return QP_OK;
}

int qp_ctx_profile_every(qp_ctx *c, int every)
{
    QP_REQUIRE(c && every >= 1, QP_ERR_ARG, "qp_ctx_profile_every: bad argument");
    c->prof_every = every;
    for (int k = 0; k < QP_PROF_COUNT; ++k) c->prof_seen[k] = 0;
    return QP_OK;
}

int qp_ctx_profile(qp_ctx *c, int enable)
{
    return qp_ctx_profile_classes(c, enable ? ~0u : 0u);
}

int qp_ctx_profile_get(qp_ctx *c, int kind, double *total_ms, int64_t *count, double *bytes)
{
    QP_REQUIRE(c && kind >= 0 && kind < QP_PROF_COUNT, QP_ERR_ARG, "qp_ctx_profile_get: bad argument");
    int rc = prof_resolve(c);
    if (rc) return rc;
    if (total_ms) *total_ms = c->prof_ms[kind];
    if (count) *count = c->prof_n[kind];
    if (bytes) *bytes = c->prof_bytes[kind];
    return QP_OK;
}

int qp_ctx_tune(qp_ctx *c, int knob, int value)
{
    QP_REQUIRE(c && knob >= 0 && knob < QP_TUNE_COUNT, QP_ERR_ARG, "qp_ctx_tune: bad knob");
    if (value < 0) value = tune_default(c, knob);
    QP_REQUIRE(value < TUNE_VALUES[knob], QP_ERR_ARG, "qp_ctx_tune: knob %d has %d values", knob,
               TUNE_VALUES[knob]);
    c->tune[knob] = value;
    return QP_OK;
}

int qp_ctx_tune_get(const qp_ctx *c, int knob, int *value, int *nvalues)
{
    QP_REQUIRE(c && knob >= 0 && knob < QP_TUNE_COUNT, QP_ERR_ARG, "qp_ctx_tune_get: bad knob");
    if (value) *value = c->tune[knob];
    if (nvalues) *nvalues = TUNE_VALUES[knob];
    return QP_OK;
}

int qp_ctx_sync(qp_ctx *c)
{
    QP_REQUIRE(c, QP_ERR_ARG, "qp_ctx_sync: null context");
    QP_HIP(hipStreamSynchronize(c->stream));
    return QP_OK;
}

int qp_ctx_stream(const qp_ctx *c, void **stream)
{
    QP_REQUIRE(c && stream, QP_ERR_ARG, "qp_ctx_stream: null argument");
    *stream = (void *)c->stream;
    return QP_OK;
}

int qp_ctx_info(const qp_ctx *c, int *nx, int *ny, int *nz, int *precision, double *dex,
                double *dxi)
{
    QP_REQUIRE(c, QP_ERR_ARG, "qp_ctx_info: null context");
    if (nx) *nx = c->nx;
    if (ny) *ny = c->ny;
    if (nz) *nz = c->nz;
    if (precision) *precision = c->prec;
    if (dex) *dex = c->dex;
    if (dxi) *dxi = c->dxi;
    return QP_OK;
}

}  // extern "C"
