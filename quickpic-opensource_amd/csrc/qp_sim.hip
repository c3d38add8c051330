// The 3-D step and its xi-slice loop on one GPU (simulation_class.f03:611-765).
//
// Field roles and states follow init_sim_fields (simulation_class.f03:267-287):
// qb, qe, psit, psi, div_vpot (1 comp), fxy, dcu, epw, epwb (2), cu, amu,
// bxyz (3); cu and epw start in k space.  Every call below enqueues on the
// context stream; the only host synchronisations are the corrector's stopping
// test (when a tolerance is given) and the particle check at the slice end.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "qp_common.h"

namespace qp {
int species_push_nocheck(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                         const qp_field2d *psit, double dex);
int species_check(qp_species *sp);
const void *species_buffer_key(const qp_species *sp);
void species_flip(qp_species *sp, bool offsets);
int species_check_post(qp_species *sp, int slice);
int species_check_poll(qp_species *sp, bool wait);
int species_qdp_fused(qp_species *sp, qp_field2d *q, bool first);
int beam_slice_assign(qp_beam *b, qp_field2d *q, int slice, const int *jdev = nullptr);
int beam_slice_add_dev(qp_beam *b, qp_field2d *q, const int *jdev);
int species_q3_initial(qp_species *sp, qp_field3d *q3);
int species_amjdp_fused(qp_species *sp, const qp_field2d *ef, const qp_field2d *bf,
                        const qp_field2d *psit, qp_field2d *cu, qp_field2d *dcu, qp_field2d *amu,
                        double dex, bool first, bool last, const qp_field2d *psi);
}  // namespace qp

struct qp_sim {
    qp_ctx *ctx;
    std::vector<qp_species *> sp;
    std::vector<qp_beam *> bm;
    std::vector<qp_field3d *> q3s;   // optional cbq targets, one per species
    qp_field3d *psi3d = nullptr;     // optional diagnostics: psi (dim 1), current (dim 3)
    qp_field3d *cu3d = nullptr;
    int iter;
    qp_field2d *f[QP_F_COUNT];
    qp_field3d *bexyz, *bbxyz;
    qp_field2d *epw_hist;   // B_perp of the slice before the last (extrapolated seed)
    qp_field2d *epw_hist2;  // and of the one before it (quadratic seed)
    qp_field2d *gprev, *fprev;   // last pass's result and change (mixed corrector)
    int hist_slices;        // slices of this step recorded in the history
    int prepared;           // slice whose first part has run (0: none)
    int last_passes;        // corrector passes of the last finished slice
    int pred_passes = 0;    // passes the next slice's corrector enqueues at once
    bool last_adaptive;     // whether its residuals were read back
    // Captured slices (qp_scan): while `graph` is set, prepare and finish
    // enqueue a slice whose index lives in device memory, so one captured
    // slice serves every j.  ctl[0] = j (the slice to run next), ctl[1] =
    // the slices recorded for the extrapolated seed; passes_dev[j] = the
    // corrector passes slice j used.  begin_step resets them.
    qp::DevBuf ctl, passes_dev;
    bool graph = false;
};

using namespace qp;

#define RC(x) do { int _r = (x); if (_r) return _r; } while (0)

namespace {

__global__ void k_ctl_reset(int *ctl)
{
    if (threadIdx.x == 0) {
        ctl[0] = 1;
        ctl[1] = 0;
    }
}

// End of a captured slice: the passes it used (pass l converged when flag
// word l is set; without a tolerance all iter_max ran), then the next slice
__global__ void k_slice_end(int *ctl, const int *flags, int iter_max, int adaptive, int seeding,
                            int *passes, int nz)
{
    if (threadIdx.x != 0) return;
    const int j = ctl[0];
    if (j < 1 || j > nz) return;
    int p = iter_max;
    if (adaptive)
        for (int l = 1; l < iter_max; ++l)
            if (flags[l]) {
                p = l;
                break;
            }
    passes[j] = p;
    ctl[0] = j + 1;
    if (seeding) ctl[1] = ctl[1] + 1;
}

}  // namespace

extern "C" {

int qp_sim_create(qp_ctx *c, qp_species *const *sp, int nsp, qp_beam *const *bm, int nb, int iter,
                  qp_sim **out)
{
    QP_REQUIRE(c && out && nsp >= 0 && nb >= 0 && iter >= 1, QP_ERR_ARG,
               "qp_sim_create: bad argument");
    qp_sim *s = new qp_sim();
    s->ctx = c;
    for (int i = 0; i < nsp; ++i) s->sp.push_back(sp[i]);
    for (int i = 0; i < nb; ++i) s->bm.push_back(bm[i]);
    s->q3s.assign(nsp, nullptr);
    s->iter = iter;
    static const int dims[QP_F_COUNT] = {1, 1, 1, 1, 1, 2, 3, 3, 2, 3, 2, 2};
    static const int st[QP_F_COUNT] = {0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0};
    int rc = QP_OK;
    for (int i = 0; i < QP_F_COUNT && !rc; ++i) rc = qp_field2d_create(c, dims[i], st[i], &s->f[i]);
    if (!rc) rc = qp_field3d_create(c, 3, &s->bexyz);
    if (!rc) rc = qp_field3d_create(c, 3, &s->bbxyz);
    if (!rc) rc = qp_field2d_create(c, 2, 1, &s->epw_hist);
    if (!rc) rc = qp_field2d_create(c, 2, 1, &s->epw_hist2);
    if (!rc) rc = qp_field2d_create(c, 2, 1, &s->gprev);
    if (!rc) rc = qp_field2d_create(c, 2, 1, &s->fprev);
    if (!rc) rc = s->ctl.alloc(4 * sizeof(int));
    if (!rc) rc = s->passes_dev.alloc((size_t)(c->nz + 2) * sizeof(int));
    if (rc) { qp_sim_destroy(s); return rc; }
    *out = s;
    return QP_OK;
}

int qp_sim_destroy(qp_sim *s)
{
    if (!s) return QP_OK;
    for (int i = 0; i < QP_F_COUNT; ++i) qp_field2d_destroy(s->f[i]);
    qp_field3d_destroy(s->bexyz);
    qp_field3d_destroy(s->bbxyz);
    qp_field2d_destroy(s->epw_hist);
    qp_field2d_destroy(s->epw_hist2);
    qp_field2d_destroy(s->gprev);
    qp_field2d_destroy(s->fprev);
    if (s->ctx) (void)hipStreamSynchronize(s->ctx->stream);
    s->ctl.release();
    s->passes_dev.release();
    delete s;
    return QP_OK;
}

int qp_sim_field2d(qp_sim *s, int which, qp_field2d **out)
{
    QP_REQUIRE(s && out && which >= 0 && which < QP_F_COUNT, QP_ERR_ARG, "qp_sim_field2d: bad argument");
    *out = s->f[which];
    return QP_OK;
}

int qp_sim_field3d(qp_sim *s, qp_field3d **bexyz, qp_field3d **bbxyz)
{
    QP_REQUIRE(s, QP_ERR_ARG, "qp_sim_field3d: null sim");
    if (bexyz) *bexyz = s->bexyz;
    if (bbxyz) *bbxyz = s->bbxyz;
    return QP_OK;
}

int qp_sim_species_q3(qp_sim *s, int species, qp_field3d *q3)
{
    QP_REQUIRE(s && species >= 0 && species < (int)s->sp.size(), QP_ERR_ARG,
               "qp_sim_species_q3: bad species index");
    s->q3s[species] = q3;
    // plane 1 holds the smoothed charge of the plasma as loaded
    return q3 ? species_q3_initial(s->sp[species], q3) : QP_OK;
}

int qp_sim_diag3d(qp_sim *s, int which, qp_field3d *f3)
{
    QP_REQUIRE(s && (which == QP_DIAG_PSI || which == QP_DIAG_CU), QP_ERR_ARG,
               "qp_sim_diag3d: bad argument");
    QP_REQUIRE(!f3 || (f3->ctx == s->ctx && f3->dim == (which == QP_DIAG_PSI ? 1 : 3)),
               QP_ERR_ARG, "qp_sim_diag3d: the 3-D field must have %d component(s)",
               which == QP_DIAG_PSI ? 1 : 3);
    (which == QP_DIAG_PSI ? s->psi3d : s->cu3d) = f3;
    return QP_OK;
}

// dim planes of a 2-D field's real space into plane lpos of a 3-D field
// (the copy of cb, field2d_class.f03:538-575, whatever the field's state)
static int copy_planes(qp_ctx *c, const qp_field2d *f, qp_field3d *f3, int lpos)
{
    size_t e = c->esize, pl = c->rplane, st = (size_t)(c->nz + 1) * pl;
    for (int i = 0; i < f3->dim; ++i) {
        char *dst = (char *)f3->buf.p + ((size_t)i * st + (size_t)(lpos - 1) * pl) * e;
        const char *src = (const char *)f->rs.p + (size_t)i * pl * e;
        QP_HIP(hipMemcpyAsync(dst, src, pl * e, hipMemcpyDeviceToDevice, c->stream));
    }
    return QP_OK;
}

int qp_sim_begin_step(qp_sim *s)
{
    QP_REQUIRE(s, QP_ERR_ARG, "qp_sim_begin_step: null sim");
    qp_ctx *c = s->ctx;
    for (qp_beam *b : s->bm) RC(qp_beam_qdp(b));
    // pipeline receive on the first stage zeroes these (ufield2d_class.f03:528-531)
    qp_field2d *cu = s->f[QP_F_CU], *epw = s->f[QP_F_EPW];
    RC(launch_fill(c, cu->ks.p, cu->dim * c->kplane, 0.0));
    cu->state = 1;
    RC(launch_fill(c, epw->ks.p, epw->dim * c->kplane, 0.0));
    epw->state = 1;
    s->hist_slices = 0;
    hipLaunchKernelGGL(k_ctl_reset, dim3(1), dim3(64), 0, c->stream, s->ctl.as<int>());
    QP_LAUNCH_CHECK();
    RC(qp_field2d_as(s->f[QP_F_FXY], 0.0));
    RC(qp_field2d_as(s->f[QP_F_PSIT], 0.0));
    RC(qp_field2d_as(s->f[QP_F_BXYZ], 0.0));
    // slice-1 copies (simulation_class.f03:639-641)
    const int d12[2] = {1, 2}, d1[1] = {1}, d3[1] = {3}, d123[3] = {1, 2, 3};
    RC(qp_field2d_cb(s->f[QP_F_FXY], s->bexyz, 1, 2, d12, d12));
    RC(qp_field2d_cb(s->f[QP_F_PSIT], s->bexyz, 1, 1, d1, d3));
    RC(qp_field2d_cb(s->f[QP_F_BXYZ], s->bbxyz, 1, 3, d123, d123));
    return QP_OK;
}

// The slice follows go_simulation (simulation_class.f03:643-736) call for
// call, with these fusions (each kernel restates the calls it replaces in
// their order):
//  * the plasma deposits write the slice's q (qdp additions) and cu, dcu,
//    amu (amjdp additions and the dex scaling of :682-685) directly;
//  * pot + grad is one k-space kernel; the inverse transforms of fxy and
//    psi, of psit and bxyz, and the forward transforms of cu, dcu, amu are
//    batched;
//  * a corrector pass (bfqp, then div, pot, * -dex for psit, then bf) and
//    the B assembly that follows it (:677-680) are one k-space kernel, so the
//    inverse transform of bxyz that opens the next pass (or follows the loop)
//    runs at the end of this one, beside psit's;
//  * -fxy + B, the predictor cu -= dxi dcu and the slice writes are one
//    kernel each.
// first part of a slice: beam and plasma charge, psi, E_perp, the psi
// gather (:650-670), up to the corrector seed
int qp_slice_prepare(qp_sim *s, int j)
{
    QP_REQUIRE(s, QP_ERR_ARG, "qp_slice_prepare: null sim");
    qp_ctx *c = s->ctx;
    QP_REQUIRE(j >= 1 && j <= c->nz, QP_ERR_ARG, "qp_slice_prepare: slice %d outside 1..%d", j,
               c->nz);
    qp_field2d **F = s->f;
    qp_field2d *qb = F[QP_F_QB], *qe = F[QP_F_QE], *psi = F[QP_F_PSI], *fxy = F[QP_F_FXY];
    qp_field2d *epwb = F[QP_F_EPWB];
    const int nsp = (int)s->sp.size();
    // the move check of an earlier slice, if it has arrived (captured
    // slices leave the error word on the device; qp_scan_run reads it)
    if (!s->graph)
        for (qp_species *sp : s->sp) RC(species_check_poll(sp, false));
    // beam charge (:650-655) and plasma charge (:657-661); the two forward
    // transforms are independent and run as one batch
    const int *jdev = s->graph ? s->ctl.as<int>() : nullptr;
    if (s->bm.empty()) RC(qp_field2d_as(qb, 0.0));
    for (size_t i = 0; i < s->bm.size(); ++i) {
        if (i == 0) RC(beam_slice_assign(s->bm[0], qb, j + 1, jdev));
        else RC(jdev ? beam_slice_add_dev(s->bm[i], qb, jdev) : qp_beam_qdpcopy(s->bm[i], qb, j + 1));
    }
    if (nsp == 0) RC(qp_field2d_as(qe, 0.0));
    for (int i = 0; i < nsp; ++i) RC(species_qdp_fused(s->sp[i], qe, i == 0));
    QP_REQUIRE(qb->state == 0 && qe->state == 0, QP_ERR_STATE, "fftrk: data is in ks");
    {
        qp_field2d *fs[2] = {qb, qe};
        const int kinds[2] = {1, 1};
        RC(fft_r2k_batch(c, fs, kinds, 2));
        qb->state = qe->state = 1;
    }
    // the beam's transverse field (:656), psi and E_perp (:662-666)
    RC(qp_field2d_elf(qb, epwb));
    RC(fused_psi_fxy(qe, psi, fxy));
    {
        qp_field2d *fs[2] = {fxy, psi};
        const int kinds[2] = {1, 1};
        RC(fft_k2r_batch(c, fs, kinds, 2));
        fxy->state = psi->state = 0;
    }
    // the species' psi gather (:667-669) runs inside the first corrector pass
    s->prepared = j;
    return QP_OK;
}

// the rest of the slice: corrector seed, predictor, passes, push, slice writes
int qp_slice_finish(qp_sim *s, int j, int iter_max, double tol, int *iters_used, double *residual)
{
    QP_REQUIRE(s, QP_ERR_ARG, "qp_slice_finish: null sim");
    qp_ctx *c = s->ctx;
    QP_REQUIRE(s->prepared == j, QP_ERR_STATE, "qp_slice_finish: slice %d was not prepared", j);
    QP_REQUIRE(iter_max >= 1, QP_ERR_ARG, "qp_slice_finish: iter_max %d", iter_max);
    s->prepared = 0;
    const double dex = c->dex, dxi = c->dxi;
    qp_field2d **F = s->f;
    qp_field2d *psi = F[QP_F_PSI], *fxy = F[QP_F_FXY];
    qp_field2d *cu = F[QP_F_CU], *dcu = F[QP_F_DCU], *amu = F[QP_F_AMU], *bxyz = F[QP_F_BXYZ];
    qp_field2d *psit = F[QP_F_PSIT], *epw = F[QP_F_EPW], *epwb = F[QP_F_EPWB];
    const int nsp = (int)s->sp.size();
    const double dd = dex * dex;   // psi *= dex^2 (:670) is applied with the slice writes
    // predictor fields from the carried current (:671-675) and the B assembly
    const PassFields pf{cu, dcu, amu, epw, epwb, psit, bxyz, nullptr, nullptr};
    qp_field2d *inv[2] = {psit, bxyz};
    const int inv_kinds[2] = {1, 2};
    qp_field2d *fwd[3] = {cu, dcu, amu};
    const int fwd_kinds[3] = {1, 1, 3};
    const bool adaptive = tol > 0.0;
    const bool graph = s->graph;
    // the mixed corrector (QP_TUNE_MIX) applies only to runs to a tolerance
    // outside captured slices
    const bool mixing = adaptive && !graph && c->tune[QP_TUNE_MIX] == 1;
    PassFields pc = pf;
    if (mixing) {
        pc.gprev = s->gprev;
        pc.fprev = s->fprev;
    }
    // the extrapolated corrector seed is applied by the predictor's kernel:
    // linear (QP_TUNE_SEED = 1) or quadratic (3) in xi
    int seed = 0;
    const bool quad = c->tune[QP_TUNE_SEED] == 3;
    if (adaptive && (c->tune[QP_TUNE_SEED] == 1 || quad)) {
        seed = s->hist_slices >= 3 ? 3 : (s->hist_slices >= 2 ? 2 : 1);
        ++s->hist_slices;
    }
    RC(fused_pass(pf, dex, false, 0, 0.0, seed, s->epw_hist,
                  graph && seed ? s->ctl.as<int>() + 1 : nullptr,
                  quad ? s->epw_hist2 : nullptr));
    RC(fft_k2r_batch(c, inv, inv_kinds, 2));
    psit->state = bxyz->state = 0;
    // corrector passes (:676-710)
    int *flags = c->flags.as<int>();
    // launches of pass l carry tag tag0 + l, unique to this slice, so that
    // the timing of a speculative pass that returned at once is dropped
    // without touching the earlier slices' pending timings
    c->prof_epoch = (c->prof_epoch + 1) & 0xffffff;
    const int tag0 = c->prof_epoch << 6;
    // a pass is its deposits (head) and its field solve (tail)
    auto head = [&](int l) -> int {
        c->prof_tag = tag0 + l;
        if (c->skip) c->skip = flags + (l - 1);
        if (nsp == 0) {
            RC(qp_field2d_as(cu, 0.0));
            RC(qp_field2d_as(dcu, 0.0));
            RC(qp_field2d_as(amu, 0.0));
        }
        // the first pass applies the species' psi gather (extpsi, :667-669)
        // to each particle just before its deposit
        for (int i = 0; i < nsp; ++i)
            RC(species_amjdp_fused(s->sp[i], fxy, bxyz, psit, cu, dcu, amu, dex, i == 0,
                                   i == nsp - 1, l == 1 ? psi : nullptr));
        c->prof_tag = 0;
        return QP_OK;
    };
    auto tail = [&](int l) -> int {
        c->prof_tag = tag0 + l;
        if (c->skip) c->skip = flags + (l - 1);
        RC(fft_r2k_batch(c, fwd, fwd_kinds, 3));
        cu->state = dcu->state = amu->state = 1;
        RC(fused_pass(pc, dex, true, l, tol));
        if (mixing) RC(fused_mix(pc, dex, l, tol, l >= 2 && l < iter_max));
        RC(fft_k2r_batch(c, inv, inv_kinds, 2));
        psit->state = bxyz->state = 0;
        c->prof_tag = 0;
        return QP_OK;
    };
    auto pass = [&](int l) -> int {
        RC(head(l));
        return tail(l);
    };
    int l = 0;
    double res = -1.0;
    if (!adaptive) {
        for (l = 1; l <= iter_max; ++l) RC(pass(l));
        l = iter_max;
    } else if (graph) {
        // every pass is enqueued; once pass l converges (flag word l), the
        // kernels of the later passes return at once and pass the flag on
        QP_REQUIRE(iter_max < FLAG_WORDS, QP_ERR_ARG, "captured slice: iter_max %d >= %d",
                   iter_max, FLAG_WORDS);
        c->skip = flags;
        int rc = QP_OK;
        for (l = 1; l <= iter_max && !rc; ++l) rc = pass(l);
        c->skip = nullptr;
        RC(rc);
        l = iter_max;
    } else if (nsp == 0 || !c->tune[QP_TUNE_SPECULATE] || iter_max >= FLAG_WORDS) {
        // read each residual in turn
        for (l = 1;; ++l) {
            RC(pass(l));
            QP_HIP(hipStreamSynchronize(c->stream));
            res = c->h_res[l % RES_SLOTS];
            if (res < tol || l == iter_max) break;
        }
    } else {
        // The deposits of pass l+1 are enqueued before the residual of pass
        // l is read back, so the device computes them while the host waits
        // and decides; once pass l converges, flag word l is set and those
        // deposits return at once, leaving the fields and particles of the
        // converged pass.  The field solve of pass l+1 is enqueued only when
        // pass l has not converged.  With QP_TUNE_SPECULATE = 2 the first
        // wait comes only after P whole passes, P the last slice's count:
        // the passes after a converged one return at once (the skipped
        // transforms pass the converged flag on), so the host reads P
        // residuals, first pass first, after one wait instead of P.
        struct SkipGuard {
            qp_ctx *c;
            ~SkipGuard() { c->skip = nullptr; }
        } guard{c};
        c->skip = flags;
        const int P = c->tune[QP_TUNE_SPECULATE] == 2
                          ? std::max(1, std::min(iter_max, s->pred_passes)) : 1;
        for (int k = 1; k <= P; ++k) RC(pass(k));
        QP_HIP(hipEventRecord(c->pass_ev[P & 1], c->stream));
        int lo = 1;   // the first pass whose residual has not been read
        for (l = P;; ++l) {
            const bool more = l < iter_max;
            if (more) RC(head(l + 1));
            QP_HIP(hipEventSynchronize(c->pass_ev[l & 1]));
            // passes lo..l ran (those after a converged one returned at once
            // and left their residual slots alone): the first below tol
            int conv = 0;
            for (int k = lo; k <= l && !conv; ++k)
                if (c->h_res[k % RES_SLOTS] < tol) conv = k;
            if (conv || !more) {
                const int last = conv ? conv : l;
                res = c->h_res[last % RES_SLOTS];
                for (int k = last + 1; k <= l + (more ? 1 : 0); ++k) prof_discard(c, tag0 + k);
                l = last;
                break;
            }
            RC(tail(l + 1));
            QP_HIP(hipEventRecord(c->pass_ev[(l + 1) & 1], c->stream));
            lo = l + 1;
        }
        s->pred_passes = l;
    }
    if (l > iter_max) l = iter_max;
    // species charge diagnostic of the last pass (cbq, :692-695): it reads
    // only the species' own q and cu, which the later steps leave untouched;
    // its one product is the 3-D charge, so it runs for species that have one
    for (size_t i = 0; i < s->sp.size(); ++i)
        if (s->q3s[i]) RC(qp_species_cbq(s->sp[i], s->q3s[i], j + 1));
    // the current of the last pass, scaled by dex, into cu3d (:696-698):
    // its real-space planes are kept by the forward transform
    if (s->cu3d) RC(copy_planes(c, cu, s->cu3d, j + 1));
    // final E_perp and predictor current (:721-726); B is already assembled
    RC(fused_efield(fxy, bxyz));
    RC(fused_predict_cu(cu, dcu, dxi));
    // plasma push (:727-730)
    for (qp_species *sp : s->sp) RC(species_push_nocheck(sp, fxy, bxyz, psit, dex));
    // slice j+1 of the 3-D wake (:731-735)
    RC(fused_slice_write(fxy, psit, bxyz, psi, s->bexyz, s->bbxyz, j + 1, dex, dd,
                         graph ? s->ctl.as<int>() : nullptr, graph ? nullptr : s->ctl.as<int>(),
                         s->hist_slices));
    if (graph) {
        hipLaunchKernelGGL(k_slice_end, dim3(1), dim3(64), 0, c->stream, s->ctl.as<int>(), flags,
                           iter_max, adaptive ? 1 : 0, seed ? 1 : 0, s->passes_dev.as<int>(),
                           c->nz);
        QP_LAUNCH_CHECK();
        if (iters_used) *iters_used = -1;
        if (residual) *residual = -1.0;
        return QP_OK;
    }
    // psi * dex^2 into psi3d (:670-671); the slice write applied the scaling
    if (s->psi3d) RC(copy_planes(c, psi, s->psi3d, j + 1));
    for (qp_species *sp : s->sp) RC(species_check_post(sp, j));
    // strict move check: wait for this slice's check before returning
    if (c->tune[QP_TUNE_STRICT])
        for (qp_species *sp : s->sp) RC(species_check_poll(sp, true));
    s->last_passes = l;
    s->last_adaptive = adaptive;
    if (iters_used) *iters_used = l;
    if (residual) *residual = adaptive ? res : -1.0;
    return QP_OK;
}

int qp_slice_residuals(const qp_sim *s, double *res, int max, int *n)
{
    QP_REQUIRE(s && res && n && max >= 0, QP_ERR_ARG, "qp_slice_residuals: bad argument");
    QP_REQUIRE(s->last_adaptive, QP_ERR_STATE,
               "qp_slice_residuals: the last slice ran without a tolerance");
    const int m = s->last_passes < RES_SLOTS - 1 ? s->last_passes : RES_SLOTS - 1;
    *n = m < max ? m : max;
    QP_HIP(hipStreamSynchronize(s->ctx->stream));
    for (int l = 1; l <= *n; ++l) res[l - 1] = s->ctx->h_res[l % RES_SLOTS];
    return QP_OK;
}

int qp_slice_run(qp_sim *s, int j, int iter_max, double tol, int *iters_used, double *residual)
{
    RC(qp_slice_prepare(s, j));
    return qp_slice_finish(s, j, iter_max, tol, iters_used, residual);
}

int qp_sim_push_beams(qp_sim *s)
{
    QP_REQUIRE(s, QP_ERR_ARG, "qp_sim_push_beams: null sim");
    qp_ctx *c = s->ctx;
    for (qp_species *sp : s->sp) RC(species_check_poll(sp, true));
    for (qp_beam *b : s->bm) RC(qp_beam_push(b, s->bexyz, s->bbxyz, c->dex, c->dxi));
    return QP_OK;
}

int qp_sim_renew(qp_sim *s, double snew)
{
    QP_REQUIRE(s, QP_ERR_ARG, "qp_sim_renew: null sim");
    for (size_t i = 0; i < s->sp.size(); ++i) {
        RC(qp_species_renew(s->sp[i], snew));
        if (s->q3s[i]) RC(species_q3_initial(s->sp[i], s->q3s[i]));
    }
    QP_HIP(hipStreamSynchronize(s->ctx->stream));
    return QP_OK;
}

int qp_sim_end_step(qp_sim *s, double snew)
{
    QP_REQUIRE(s, QP_ERR_ARG, "qp_sim_end_step: null sim");
    RC(qp_sim_push_beams(s));
    return qp_sim_renew(s, snew);
}

// ---- batched scan: one HIP graph holding one xi slice of many instances ----

// The plasma push moves the particles between two sets of arrays and swaps
// the host's pointers to them, which a graph cannot repeat: the scan holds
// two graphs, one captured with each buffer parity, and launches the one
// that matches where the particles are (the first species of the instances
// tells), swapping the host pointers after each launch as the push would.
struct qp_scan {
    std::vector<qp_sim *> sims;
    int iter_max = 0;
    double tol = 0.0;
    hipStream_t origin = nullptr;          // the stream the graphs are launched on
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    std::vector<const void *> key0;        // per instance: particle buffer of parity 0
    std::vector<char> offsets;             // per instance: the push swaps cell offsets
};

static void scan_flip(qp_scan *sc)
{
    for (size_t i = 0; i < sc->sims.size(); ++i)
        for (qp_species *sp : sc->sims[i]->sp) species_flip(sp, sc->offsets[i] != 0);
}

int qp_scan_destroy(qp_scan *sc)
{
    if (!sc) return QP_OK;
    if (sc->origin) (void)hipStreamSynchronize(sc->origin);
    for (int p = 0; p < 2; ++p) {
        if (sc->exec[p]) (void)hipGraphExecDestroy(sc->exec[p]);
        if (sc->graph[p]) (void)hipGraphDestroy(sc->graph[p]);
    }
    for (hipEvent_t e : sc->join)
        if (e) (void)hipEventDestroy(e);
    if (sc->fork) (void)hipEventDestroy(sc->fork);
    if (sc->origin) (void)hipStreamDestroy(sc->origin);
    delete sc;
    return QP_OK;
}

// The capture: the origin stream forks to every instance's own stream, each
// instance enqueues one slice there (prepare + finish in graph mode, which
// reads the slice index from its device word and ends with k_slice_end), and
// the streams join back.  Instances are independent branches of the graph.
static int scan_capture(qp_scan *sc, int parity)
{
    QP_HIP(hipStreamBeginCapture(sc->origin, hipStreamCaptureModeRelaxed));
    int rc = QP_OK;
    hipError_t e = hipEventRecord(sc->fork, sc->origin);
    if (e != hipSuccess) rc = hip_fail(e, "hipEventRecord", __FILE__, __LINE__);
    for (size_t i = 0; i < sc->sims.size() && !rc; ++i) {
        qp_sim *s = sc->sims[i];
        qp_ctx *c = s->ctx;
        e = hipStreamWaitEvent(c->stream, sc->fork, 0);
        if (e != hipSuccess) { rc = hip_fail(e, "hipStreamWaitEvent", __FILE__, __LINE__); break; }
        const int hist = s->hist_slices;   // the captured seed counts on the device
        s->graph = true;
        rc = qp_slice_prepare(s, 1);
        if (!rc) rc = qp_slice_finish(s, 1, sc->iter_max, sc->tol, nullptr, nullptr);
        s->graph = false;
        s->prepared = 0;
        s->hist_slices = hist;
        if (rc) break;
        e = hipEventRecord(sc->join[i], c->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(sc->origin, sc->join[i], 0);
        if (e != hipSuccess) rc = hip_fail(e, "capture join", __FILE__, __LINE__);
    }
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(sc->origin, &g);
    if (!rc && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    sc->graph[parity] = g;
    QP_HIP(hipGraphInstantiate(&sc->exec[parity], g, nullptr, nullptr, 0));
    return QP_OK;
}

int qp_scan_create(qp_sim *const *sims, int n, int iter_max, double tol, qp_scan **out)
{
    QP_REQUIRE(sims && out && n >= 1, QP_ERR_ARG, "qp_scan_create: bad argument");
    QP_REQUIRE(iter_max >= 1 && iter_max < FLAG_WORDS, QP_ERR_ARG,
               "qp_scan_create: iter_max %d outside 1..%d", iter_max, FLAG_WORDS - 1);
    for (int i = 0; i < n; ++i) {
        qp_sim *s = sims[i];
        QP_REQUIRE(s && s->ctx, QP_ERR_ARG, "qp_scan_create: null simulation %d", i);
        QP_REQUIRE(s->ctx->device == sims[0]->ctx->device, QP_ERR_ARG,
                   "qp_scan_create: instance %d is on another device", i);
        for (int k = 0; k < i; ++k)
            QP_REQUIRE(sims[k]->ctx != s->ctx, QP_ERR_ARG,
                       "qp_scan_create: instances %d and %d share a context (stream)", k, i);
        bool diag = s->cu3d || s->psi3d;
        for (qp_field3d *q3 : s->q3s) diag = diag || q3;
        QP_REQUIRE(!diag, QP_ERR_UNSUPPORTED,
                   "qp_scan_create: instance %d has 3-D diagnostics (cbq, psi, current); "
                   "captured slices do not write them", i);
        QP_REQUIRE(!s->ctx->prof_on, QP_ERR_STATE,
                   "qp_scan_create: instance %d has kernel timing on", i);
    }
    QP_HIP(hipSetDevice(sims[0]->ctx->device));
    qp_scan *sc = new qp_scan();
    sc->sims.assign(sims, sims + n);
    sc->iter_max = iter_max;
    sc->tol = tol;
    sc->join.assign(n, nullptr);
    int rc = QP_OK;
    hipError_t e = hipStreamCreateWithFlags(&sc->origin, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&sc->fork, hipEventDisableTiming);
    for (int i = 0; i < n && e == hipSuccess; ++i)
        e = hipEventCreateWithFlags(&sc->join[i], hipEventDisableTiming);
    if (e != hipSuccess) rc = hip_fail(e, "qp_scan_create", __FILE__, __LINE__);
    // the instances' pending work (set-up, begin_step) is not part of the graph
    for (int i = 0; i < n && !rc; ++i) {
        e = hipStreamSynchronize(sims[i]->ctx->stream);
        if (e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize", __FILE__, __LINE__);
    }
    sc->key0.assign(n, nullptr);
    sc->offsets.assign(n, 0);
    for (int i = 0; i < n; ++i) {
        if (!sims[i]->sp.empty()) sc->key0[i] = species_buffer_key(sims[i]->sp[0]);
        sc->offsets[i] = sims[i]->ctx->tune[QP_TUNE_PUSH] == 1 ? 0 : 1;
    }
    // each capture swaps the host pointers once (the push's host side), so
    // after the two the host is back where it started
    if (!rc) rc = scan_capture(sc, 0);
    if (!rc) rc = scan_capture(sc, 1);
    if (rc) { qp_scan_destroy(sc); return rc; }
    *out = sc;
    return QP_OK;
}

int qp_scan_run(qp_scan *sc, int nslices)
{
    QP_REQUIRE(sc && sc->exec[0] && sc->exec[1] && nslices >= 0, QP_ERR_ARG,
               "qp_scan_run: bad argument");
    // the buffer parity of every instance (they must agree: an instance
    // advanced by eager slices in between would not match the graphs)
    int parity = -1;
    for (size_t i = 0; i < sc->sims.size(); ++i) {
        if (!sc->key0[i]) continue;
        const int p = species_buffer_key(sc->sims[i]->sp[0]) == sc->key0[i] ? 0 : 1;
        QP_REQUIRE(parity < 0 || p == parity, QP_ERR_STATE,
                   "qp_scan_run: instance %zu's particles are in the other buffer", i);
        parity = p;
    }
    if (parity < 0) parity = 0;
    // the slices left in each instance's step (its device slice word)
    for (size_t i = 0; i < sc->sims.size(); ++i) {
        qp_sim *s = sc->sims[i];
        int j = 0;
        QP_HIP(hipMemcpyAsync(&j, s->ctl.p, sizeof(int), hipMemcpyDeviceToHost, s->ctx->stream));
        QP_HIP(hipStreamSynchronize(s->ctx->stream));
        QP_REQUIRE(j >= 1 && j + nslices - 1 <= s->ctx->nz, QP_ERR_ARG,
                   "qp_scan_run: instance %zu is at slice %d; %d more exceed its %d slices", i, j,
                   nslices, s->ctx->nz);
    }
    // the graph runs after what the instances' streams hold, and they wait
    // for it in turn, so the library's eager calls before and after order
    // with the captured slices
    for (size_t i = 0; i < sc->sims.size(); ++i) {
        QP_HIP(hipEventRecord(sc->join[i], sc->sims[i]->ctx->stream));
        QP_HIP(hipStreamWaitEvent(sc->origin, sc->join[i], 0));
    }
    for (int k = 0; k < nslices; ++k) {
        QP_HIP(hipGraphLaunch(sc->exec[parity], sc->origin));
        scan_flip(sc);
        parity ^= 1;
    }
    QP_HIP(hipEventRecord(sc->fork, sc->origin));
    for (qp_sim *s : sc->sims) QP_HIP(hipStreamWaitEvent(s->ctx->stream, sc->fork, 0));
    QP_HIP(hipStreamSynchronize(sc->origin));
    // the particle move checks of the replayed slices (PPPCHECK2L)
    for (size_t i = 0; i < sc->sims.size(); ++i)
        for (qp_species *sp : sc->sims[i]->sp) {
            int rc = species_check(sp);
            if (rc) {
                const std::string msg = qp_last_error();
                set_error("qp_scan_run: instance %zu: %s", i, msg.c_str());
                return rc;
            }
        }
    return QP_OK;
}

int qp_scan_passes(const qp_scan *sc, int instance, int *passes, int nslices)
{
    QP_REQUIRE(sc && passes && instance >= 0 && instance < (int)sc->sims.size() && nslices >= 0,
               QP_ERR_ARG, "qp_scan_passes: bad argument");
    const qp_sim *s = sc->sims[instance];
    QP_REQUIRE(nslices <= s->ctx->nz, QP_ERR_ARG, "qp_scan_passes: %d slices > nz", nslices);
    std::vector<int> h((size_t)s->ctx->nz + 2);
    QP_HIP(hipMemcpyAsync(h.data(), s->passes_dev.p, h.size() * sizeof(int), hipMemcpyDeviceToHost,
                          s->ctx->stream));
    QP_HIP(hipStreamSynchronize(s->ctx->stream));
    for (int j = 1; j <= nslices; ++j) passes[j - 1] = h[j];
    return QP_OK;
}

}  // extern "C"
