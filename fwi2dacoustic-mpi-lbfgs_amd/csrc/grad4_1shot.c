/*
 * grad4_1shot.c -- the reference's per-shot gradient entry point
 * (header.h:283, called at MBotOpt.c:55,99, never defined there) and its
 * batched form, in C on top of the fwi_hip C ABI.
 *
 * Sequence (the driver this project defines; the oracle implements the
 * same one): forward propagation with receiver recording and frame saving
 * on the GPU -> misfit and adjoint source on the host in FP64 -> exact-
 * adjoint back-propagation, reconstruction and zero-lag imaging on the GPU
 * -> gradient copied back once.  Data sampled every msf->jump steps
 * (fwi_2d.c:195).
 *
 * Host misfit step, per shot, on the data-rate traces: observed and
 * synthetic traces band-passed (grad4_set_bandpass, sfbandpass) and
 * first-break muted (grad4_set_fbt_mute, fbt_muter), then the misfit of
 * grad4_set_misfit (L2 by default; NCC, XCORR1D, MF1D, entropic OT) gives
 * msf->val and the residual, to which the transposed mute and band-pass
 * are applied.  The adjoint source at the propagation rate is, by
 * grad4_set_residual_mode:
 *   0 (default) the residual placed on the sampled steps, zero between them
 *     (the exact transpose of the subsampling: the gradient is exact);
 *   1 the residual upsampled by the natural cubic spline of trcinp
 *     (spline_interp.c:19-38), the reference's resampling of traces between
 *     the data rate and the propagation rate.
 */
#include <limits.h>
#include <math.h>
#include <pthread.h>
#include <unistd.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "fwi_hip.h"
#include "fwi_io.h"
#include "fwi_msf.h"
#include "fwi_plan.h"
#include "fwi_sigproc.h"

/* process-wide contexts, re-created when the geometry or the capacity
 * changes: slot 0 runs grad_all_shots and RWI's incident field, slot 1 RWI's
 * total field */
typedef struct {
    fwi_hip_ctx *ctx;
    long key[8];
    int ck_applied; /* the checkpointing settings reached this context */
    float *tr;      /* page-locked trace buffer of the slot, kept across calls */
    size_t tr_cap;  /* its size in floats */
} CtxSlot;
static CtxSlot g_slot[2];
static char g_msg[512] = "";
static int g_res_mode = 0;
static int g_ck_every = 0, g_ck_storage = 0; /* checkpointing of the batched contexts */

static int fail(int code, const char *fmt, ...);

/* host misfit step configuration */
static struct {
    int kind;
    int offsets;          /* select traces by source-receiver distance */
    float omin, omax;
    par_xcorr1d xc;
    epsdf1 mf;
    char mode[16], reg[16];
    float eot_eps, eot_tor;
    int eot_iter;
    grad4_diw_par diw;
    float rank_reg;
    int bp, bp_phase, bp_nphi, bp_nplo;
    float bp_fhi, bp_flo;
    int mute, mute_early;
    float mute_alpha, mute_tune;
} g_cfg = {.kind = GRAD4_MSF_L2};

static void copy_mode(char *dst, const char *src)
{
    snprintf(dst, 16, "%s", src ? src : "");
}

int grad4_set_misfit(int kind, const void *params, float offset_min, float offset_max)
{
    switch (kind) {
    case GRAD4_MSF_L2:
    case GRAD4_MSF_NCC:
        break;
    case GRAD4_MSF_XCORR1D:
        if (!params) return FWI_HIP_EINVAL;
        g_cfg.xc = *(const par_xcorr1d *)params;
        copy_mode(g_cfg.mode, g_cfg.xc.conv_how_to_slice);
        /* "leading" slicing returns samples 0..nt-1 of obs (*) y, while the
         * derivative of its misfit is samples nt-1..2nt-2 (xcorr_msf.c:166-167) */
        if (strncmp(g_cfg.mode, "middle", 6) != 0) {
            fail(FWI_HIP_EINVAL, "grad4_set_misfit: XCORR1D needs \"middle\" slicing in the gradient driver");
            return FWI_HIP_EINVAL;
        }
        break;
    case GRAD4_MSF_MF1D:
        if (!params) return FWI_HIP_EINVAL;
        g_cfg.mf = *(const epsdf1 *)params;
        copy_mode(g_cfg.mode, g_cfg.mf.conv_how_to_slice);
        copy_mode(g_cfg.reg, g_cfg.mf.reg_type);
        break;
    case GRAD4_MSF_EOT: {
        if (!params) return FWI_HIP_EINVAL;
        const eot_msf_par *e = (const eot_msf_par *)params;
        if (!(e->eps > 0.0f) || e->max_iter < 1) return FWI_HIP_EINVAL;
        g_cfg.eot_eps = e->eps;
        g_cfg.eot_tor = e->tor;
        g_cfg.eot_iter = e->max_iter;
        break;
    }
    case GRAD4_MSF_DIW: {
        if (!params) return FWI_HIP_EINVAL;
        const grad4_diw_par *d = (const grad4_diw_par *)params;
        if (d->nl < 1 || d->nw < 1 || d->nl % 2 != 1 || d->nw % 2 != 1 || !(d->alpha_win > 0.0f)) {
            fail(FWI_HIP_EINVAL, "grad4_set_misfit: DIW needs odd nl, nw >= 1 and alpha_win > 0");
            return FWI_HIP_EINVAL;
        }
        g_cfg.diw = *d;
        break;
    }
    case GRAD4_MSF_SOFTRANK: {
        if (!params || !(*(const float *)params > 0.0f)) {
            fail(FWI_HIP_EINVAL, "grad4_set_misfit: soft ranking needs a regularisation > 0");
            return FWI_HIP_EINVAL;
        }
        g_cfg.rank_reg = *(const float *)params;
        break;
    }
    default:
        return FWI_HIP_EINVAL;
    }
    g_cfg.kind = kind;
    g_cfg.offsets = offset_min <= offset_max;
    g_cfg.omin = offset_min;
    g_cfg.omax = offset_max;
    return FWI_HIP_OK;
}

int grad4_set_checkpointing(int every, int storage)
{
    if (every < 0 || storage < 0 || storage > 2) return FWI_HIP_EINVAL;
    g_ck_every = every;
    g_ck_storage = storage;
    g_slot[0].ck_applied = g_slot[1].ck_applied = 0;
    return FWI_HIP_OK;
}

int grad4_set_bandpass(int on, int phase, int nphi, int nplo, float fhi, float flo)
{
    if (on && flo < 0.0f) return FWI_HIP_EINVAL;
    g_cfg.bp = on != 0;
    g_cfg.bp_phase = phase != 0;
    g_cfg.bp_nphi = nphi;
    g_cfg.bp_nplo = nplo;
    g_cfg.bp_fhi = fhi;
    g_cfg.bp_flo = flo;
    return FWI_HIP_OK;
}

int grad4_set_fbt_mute(int on, float alpha, float tune, int early_gone)
{
    g_cfg.mute = on != 0;
    g_cfg.mute_alpha = alpha;
    g_cfg.mute_tune = tune;
    g_cfg.mute_early = early_gone != 0;
    return FWI_HIP_OK;
}

int grad4_set_residual_mode(int mode)
{
    if (mode < 0 || mode > 1) return FWI_HIP_EINVAL;
    g_res_mode = mode;
    return FWI_HIP_OK;
}

static int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    return code;
}

const char *grad4_last_error(void) { return g_msg; }

/* wall-clock milliseconds of the phases of the last grad_all_shots call */
#define GRAD4_NPHASE 5
static double g_phase_ms[GRAD4_NPHASE];

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

int grad4_phase_ms(double *ms, int n)
{
    for (int i = 0; ms && i < n && i < GRAD4_NPHASE; i++) ms[i] = g_phase_ms[i];
    return GRAD4_NPHASE;
}

/* the page-locked trace buffer of slot `which`, at least n floats: the
 * forward's traces come back and the adjoint source goes out through it at
 * the pinned-copy rate, and its pages are touched once per process */
static float *slot_traces(int which, size_t n)
{
    CtxSlot *sl = &g_slot[which];
    if (n < 1) n = 1;
    if (sl->tr_cap < n) {
        if (sl->tr) fwi_hip_host_free(sl->tr);
        sl->tr = NULL;
        sl->tr_cap = 0;
        void *p = NULL;
        if (fwi_hip_host_alloc(&p, sizeof(float) * n)) {
            fail(FWI_HIP_ENOMEM, "pinned host memory for %zu trace samples: %s", n, fwi_hip_last_error(NULL));
            return NULL;
        }
        sl->tr = (float *)p;
        sl->tr_cap = n;
    }
    return sl->tr;
}

/* the context of slot `which` for this geometry and capacity, with the
 * checkpointing settings applied */
static int ensure_ctx(int which, const fwi_base *b, int nshots, int nt, int nr)
{
    CtxSlot *sl = &g_slot[which];
    long key[8] = {(long)b->nx, (long)b->nz, (long)b->npml, (long)nshots, (long)nt, (long)nr, 0, 0};
    memcpy(&key[6], &b->dx, sizeof(float));
    memcpy((char *)&key[6] + 4, &b->dz, sizeof(float));
    memcpy(&key[7], &b->dt, sizeof(float));
    if (!(sl->ctx && memcmp(key, sl->key, sizeof(key)) == 0)) {
        if (sl->ctx) fwi_hip_destroy(sl->ctx);
        sl->ctx = NULL;
        fwi_hip_desc d;
        memset(&d, 0, sizeof(d));
        d.grid.nx = (int)b->nx;
        d.grid.nz = (int)b->nz;
        d.grid.npml = (int)b->npml;
        d.grid.dx = b->dx;
        d.grid.dz = b->dz;
        d.grid.dt = b->dt;
        d.grid.max_f = 15.0f;    /* fwi_sys.h MAX_F  */
        d.grid.max_v = 6000.0f;  /* fwi_sys.h MAX_V  */
        d.grid.coef_r = 0.001f;  /* fwi_sys.h COEF_R */
        d.max_shots = nshots;
        d.max_nt = nt;
        d.max_nr = nr;
        int rc = fwi_hip_create(&d, &sl->ctx);
        if (rc) return fail(rc, "fwi_hip_create: %s", fwi_hip_last_error(NULL));
        memcpy(sl->key, key, sizeof(key));
        sl->ck_applied = 0;
    }
    if (!sl->ck_applied) {
        int rc = fwi_hip_set_checkpointing(sl->ctx, g_ck_every, g_ck_storage);
        if (rc) return fail(rc, "set_checkpointing: %s", fwi_hip_last_error(sl->ctx));
        sl->ck_applied = 1;
    }
    return FWI_HIP_OK;
}

/* rows of a slab a[n2][n1] */
static float **slab_rows(float *a, int n1, int n2)
{
    float **r = (float **)malloc(sizeof(float *) * (n2 ? n2 : 1));
    if (r)
        for (int i = 0; i < n2; i++) r[i] = a + (size_t)i * n1;
    return r;
}

/* per-thread working traces of the host misfit step (observed, synthetic,
 * adjoint source), kept across shots and calls so that their pages are
 * touched once per process, not once per shot */
#define MISFIT_MAX_THREADS 64
static struct {
    float *ob, *sy, *rs;
    size_t cap;
} g_scr[MISFIT_MAX_THREADS];

static int misfit_scratch(int tid, size_t n)
{
    if (n < 1) n = 1;
    if (g_scr[tid].cap >= n) return FWI_HIP_OK;
    free(g_scr[tid].ob);
    free(g_scr[tid].sy);
    free(g_scr[tid].rs);
    g_scr[tid].ob = (float *)malloc(sizeof(float) * n);
    g_scr[tid].sy = (float *)malloc(sizeof(float) * n);
    g_scr[tid].rs = (float *)malloc(sizeof(float) * n);
    g_scr[tid].cap = (g_scr[tid].ob && g_scr[tid].sy && g_scr[tid].rs) ? n : 0;
    return g_scr[tid].cap ? FWI_HIP_OK : fail(FWI_HIP_ENOMEM, "host memory");
}

/* the working set of soft_ranking for n samples (NULL without memory) */
static soft_ranking_vars *rank_vars(int n)
{
    soft_ranking_vars *v = (soft_ranking_vars *)calloc(1, sizeof(soft_ranking_vars));
    if (!v) return NULL;
    soft_ranking_init(n, v);
    if (v->sorted && v->sol && v->soft_rank && v->hard_rank && v->perm && v->target) return v;
    soft_ranking_free(v);
    return NULL;
}

/* The host misfit step of one shot, on thread tid's working traces.  t holds
 * the shot's nr traces at the propagation rate (nt samples each) and
 * receives the adjoint source. */
static int host_misfit(fwi_plan *pl, float *t, int nt, float dt_fine, int tid, int tm)
{
    fwi_msf *m = pl->msf;
    const int nr = (int)m->nr, nt0 = (int)m->nt, jump = (int)m->jump;
    const size_t n = (size_t)nr * nt0;
    if (misfit_scratch(tid, n)) return FWI_HIP_ENOMEM;
    float *ob = g_scr[tid].ob, *sy = g_scr[tid].sy, *rs = g_scr[tid].rs;
    memset(rs, 0, sizeof(float) * (n ? n : 1));
    float **obr = slab_rows(ob, nt0, nr), **syr = slab_rows(sy, nt0, nr), **rsr = slab_rows(rs, nt0, nr);
    unsigned char *drop = NULL; /* traces outside the offset window */
    int rc = FWI_HIP_OK;
    if (!ob || !sy || !rs || !obr || !syr || !rsr) {
        rc = fail(FWI_HIP_ENOMEM, "host memory");
        goto done;
    }
    /* synthetic at the data rate (fwi_2d.c:195); t is [nr][nt], or [nt][nr] when tm */
    for (int r = 0; r < nr; r++)
        for (int i0 = 0; i0 < nt0; i0++) {
            const size_t it = (size_t)i0 * jump;
            const float v = tm ? t[it * nr + r] : t[(size_t)r * nt + it];
            m->syn[r][i0] = v;
            sy[(size_t)r * nt0 + i0] = v;
            ob[(size_t)r * nt0 + i0] = m->obs0[r][i0];
        }
    if (g_cfg.bp) {
        if (sfbandpass(ob, nt0, nr, m->dt, g_cfg.bp_phase, g_cfg.bp_nphi, g_cfg.bp_nplo, g_cfg.bp_fhi, g_cfg.bp_flo,
                       ob) ||
            sfbandpass(sy, nt0, nr, m->dt, g_cfg.bp_phase, g_cfg.bp_nphi, g_cfg.bp_nplo, g_cfg.bp_fhi, g_cfg.bp_flo,
                       sy)) {
            rc = fail(FWI_HIP_EINVAL, "sfbandpass: %s", fwi_sigproc_last_error());
            goto done;
        }
    }
    if (g_cfg.mute) {
        fbt_muter(m->fbt, ob, nt0, nr, g_cfg.mute_alpha, m->dt, g_cfg.mute_tune, g_cfg.mute_early);
        fbt_muter(m->fbt, sy, nt0, nr, g_cfg.mute_alpha, m->dt, g_cfg.mute_tune, g_cfg.mute_early);
    }
    float *sr = g_cfg.offsets ? pl->sou->xyz_sr : NULL;
    /* the offset window applies to every misfit: traces outside it are left
     * out (zeroed here, since MF1D and EOT take no offsets) */
    if (sr) {
        drop = (unsigned char *)calloc(nr ? nr : 1, 1);
        if (!drop) {
            rc = fail(FWI_HIP_ENOMEM, "host memory");
            goto done;
        }
        for (int r = 0; r < nr; r++) {
            const double dx = (double)sr[2 + 2 * r] - sr[0], dz = (double)sr[3 + 2 * r] - sr[1];
            const double d = sqrt(dx * dx + dz * dz);
            if (d < g_cfg.omin || d > g_cfg.omax) {
                drop[r] = 1;
                memset(ob + (size_t)r * nt0, 0, sizeof(float) * nt0);
                memset(sy + (size_t)r * nt0, 0, sizeof(float) * nt0);
            }
        }
    }
    float val[2] = {0.0f, 0.0f};
    switch (g_cfg.kind) {
    case GRAD4_MSF_NCC:
        rc = NCCnorm(obr, syr, nt0, nr, sr, g_cfg.omin, g_cfg.omax, val, rsr);
        break;
    case GRAD4_MSF_XCORR1D: {
        par_xcorr1d p = g_cfg.xc;
        p.nr = nr;
        p.nt = nt0;
        p.conv_how_to_slice = g_cfg.mode;
        rc = XCORR1D(obr, syr, sr, g_cfg.omin, g_cfg.omax, &p, val, rsr);
        break;
    }
    case GRAD4_MSF_MF1D: {
        epsdf1 p = g_cfg.mf;
        p.nr = nr;
        p.nt = nt0;
        p.conv_how_to_slice = g_cfg.mode;
        p.reg_type = g_cfg.reg;
        rc = MF1D(obr, syr, &p, val, rsr);
        break;
    }
    case GRAD4_MSF_EOT: {
        eot_msf_par p;
        memset(&p, 0, sizeof(p));
        p.nt = nt0;
        p.nr = nr;
        p.max_iter = g_cfg.eot_iter;
        p.eps = g_cfg.eot_eps;
        p.tor = g_cfg.eot_tor;
        p.dt = m->dt;
        rc = entropy_ot_2d_trace_by_trace(obr, syr, &p, val, rsr);
        break;
    }
    case GRAD4_MSF_DIW: {
        const grad4_diw_par d = g_cfg.diw;
        float range[2] = {g_cfg.omin, g_cfg.omax};
        rc = lcc_diw_msf(obr, syr, sr, range, d.eps, nt0, nr, d.nl, d.nw, d.alpha_win, d.sim_barrier, d.lag_barrier,
                         val, rsr);
        break;
    }
    case GRAD4_MSF_SOFTRANK: {
        soft_ranking_vars *par[2] = {rank_vars(nt0), rank_vars(nt0)};
        const int have = par[0] && par[1];
        if (have)
            rc = soft_ranking_2d_trace_by_trace(obr, syr, sr, g_cfg.omin, g_cfg.omax, par, g_cfg.rank_reg, nt0, nr,
                                                val, rsr);
        else
            rc = fail(FWI_HIP_ENOMEM, "host memory");
        soft_ranking_free(par[0]);
        soft_ranking_free(par[1]);
        if (rc && have) rc = fail(rc, "misfit: %s", fwi_msf_rank_last_error());
        if (rc) goto done;
        break;
    }
    default:
        rc = L2norm(obr, syr, nt0, nr, sr, g_cfg.omin, g_cfg.omax, val, rsr);
    }
    if (rc && g_cfg.kind == GRAD4_MSF_DIW) {
        rc = fail(rc, "misfit: %s", fwi_msf_rank_last_error());
        goto done;
    }
    if (rc) {
        rc = fail(rc, "misfit: %s", fwi_msf_last_error());
        goto done;
    }
    m->val[0] = val[0];
    m->val[1] = val[1];
    if (drop)
        for (int r = 0; r < nr; r++)
            if (drop[r]) memset(rs + (size_t)r * nt0, 0, sizeof(float) * nt0);
    /* make the residual the derivative of val[1] with respect to syn:
     * MF1D returns it with respect to 100 syn (lsq_decon_cg.c:183-184,377);
     * XCORR1D's middle window sits one lag off for even nt
     * (xcorr_msf.c:43-46,105-106), which shifts its residual one sample */
    if (g_cfg.kind == GRAD4_MSF_MF1D)
        for (size_t i = 0; i < n; i++) rs[i] *= 100.0f;
    if (g_cfg.kind == GRAD4_MSF_XCORR1D && nt0 % 2 == 0)
        for (int r = 0; r < nr; r++) {
            float *x = rs + (size_t)r * nt0;
            memmove(x + 1, x, sizeof(float) * (nt0 - 1));
            x[0] = 0.0f;
        }
    /* transposed conditioning of the residual: mute (diagonal), then F^T */
    if (g_cfg.mute) fbt_muter(m->fbt, rs, nt0, nr, g_cfg.mute_alpha, m->dt, g_cfg.mute_tune, g_cfg.mute_early);
    if (g_cfg.bp)
        sfbandpass_adjoint(rs, nt0, nr, m->dt, g_cfg.bp_phase, g_cfg.bp_nphi, g_cfg.bp_nplo, g_cfg.bp_fhi,
                           g_cfg.bp_flo, rs);
    /* adjoint source at the propagation rate */
    if (g_res_mode == 0 || jump == 1) {
        for (int r = 0; r < nr; r++) /* residual on the samples, zero between them */
            for (int it = 0; it < nt; it++) {
                const float v = (it % jump == 0) ? rs[(size_t)r * nt0 + it / jump] : 0.0f;
                t[tm ? (size_t)it * nr + r : (size_t)r * nt + it] = v;
            }
    } else {
        const size_t nfine = (size_t)nt * nr;
        float *dst = tm ? (float *)malloc(sizeof(float) * (nfine > 0 ? nfine : 1)) : t;
        float **tr = dst ? slab_rows(dst, nt, nr) : NULL;
        if (!tr) {
            if (tm) free(dst);
            rc = fail(FWI_HIP_ENOMEM, "host memory");
            goto done;
        }
        int e = fwi_trcinp(rsr, nt0, m->dt, tr, nt, dt_fine, nr);
        free(tr);
        if (e) rc = fail(e, "trcinp: %s", fwi_io_last_error());
        if (tm) {
            for (int r = 0; !e && r < nr; r++)
                for (int it = 0; it < nt; it++) t[(size_t)it * nr + r] = dst[(size_t)r * nt + it];
            free(dst);
        }
    }
done:
    free(drop);
    free(obr);
    free(syr);
    free(rsr);
    return rc;
}

/* the host misfit step of every shot, shots spread over a few threads
 * (FWI_HOST_THREADS, default: the online cores, at most 16) */
typedef struct {
    fwi_plan **plans;
    float *tr;
    const long long *off;
    int nshots, nthreads, t, rc;
    float dt;
    int tm;
} MisfitJob;

static void *misfit_worker(void *p)
{
    MisfitJob *j = (MisfitJob *)p;
    for (int s = j->t; s < j->nshots && !j->rc; s += j->nthreads)
        j->rc = host_misfit(j->plans[s], j->tr + j->off[s], (int)j->plans[s]->grid->base->nt, j->dt, j->t, j->tm);
    return NULL;
}

/* tr holds the shots' traces one after the other, shot s as nr x nt with
 * its own plan's nt */
/* the misfit threads: FWI_HOST_THREADS, else the online cores, at most 16 */
static int misfit_threads(int nshots)
{
    int nth = 16;
    const char *env = getenv("FWI_HOST_THREADS");
    if (env && atoi(env) > 0) nth = atoi(env);
    else {
        const long cpu = sysconf(_SC_NPROCESSORS_ONLN);
        if (cpu > 0 && cpu < nth) nth = (int)cpu;
    }
    if (nth > nshots) nth = nshots;
    if (nth < 1) nth = 1;
    if (nth > MISFIT_MAX_THREADS) nth = MISFIT_MAX_THREADS;
    return nth;
}

/* tm: the shots' traces are time-major ([nt][nr], fwi_hip_forward_tm) */
static int host_misfit_all(fwi_plan **plans, int nshots, float *tr, float dt, int tm)
{
    long long *off = (long long *)malloc(sizeof(long long) * (nshots + 1));
    if (!off) return fail(FWI_HIP_ENOMEM, "host memory");
    off[0] = 0;
    for (int s = 0; s < nshots; s++) off[s + 1] = off[s] + (long long)plans[s]->msf->nr * (long long)plans[s]->grid->base->nt;
    const int nth = misfit_threads(nshots);
    MisfitJob jobs[MISFIT_MAX_THREADS];
    pthread_t th[MISFIT_MAX_THREADS];
    int made[MISFIT_MAX_THREADS];
    int rc = FWI_HIP_OK;
    for (int t = 0; t < nth; t++) {
        jobs[t] = (MisfitJob){plans, tr, off, nshots, nth, t, FWI_HIP_OK, dt, tm};
        made[t] = t > 0 && pthread_create(&th[t], NULL, misfit_worker, &jobs[t]) == 0;
    }
    for (int t = 0; t < nth; t++) /* the calling thread takes share 0 and any share without a thread */
        if (!made[t]) misfit_worker(&jobs[t]);
    for (int t = 1; t < nth; t++)
        if (made[t]) pthread_join(th[t], NULL);
    for (int t = 0; t < nth; t++)
        if (jobs[t].rc && !rc) rc = jobs[t].rc;
    free(off);
    return rc;
}

/* checks shared by the drivers: data rate (fwi_2d.c:195), one grid */
/* Each plan is one shot with its own time axis (the reference runs one rank
 * per shot, each with its own fwi_base): shots may differ in nt and in
 * their receivers, and share the grid, the propagation dt and the CPML. */
static int check_shots(const fwi_base *b, fwi_specs *const *grids, fwi_msf *const *msfs, int nshots, int *nrmax,
                       int *ntmax)
{
    if (b->free_surface != 1) return fail(FWI_HIP_EINVAL, "only the free-surface grid is supported");
    *nrmax = 0;
    *ntmax = 0;
    for (int s = 0; s < nshots; s++) {
        const fwi_msf *m = msfs[s];
        const fwi_base *bs = grids[s]->base;
        if ((int)m->nr > *nrmax) *nrmax = (int)m->nr;
        if ((int)bs->nt > *ntmax) *ntmax = (int)bs->nt;
        if (bs->nt < 1 || (m->nt - 1) * m->jump != bs->nt - 1)
            return fail(FWI_HIP_EINVAL, "shot %d: (nt0-1)*jump != nt-1 (%zu, %zu, %zu)", s, m->nt, m->jump, bs->nt);
        /* the data rate must be jump x the propagation rate (fwi_2d.c:195): the
         * filters and mutes run at msf.dt, the residual is resampled at base.dt */
        if (fabs((double)m->dt - (double)bs->dt * (double)m->jump) > 1e-3)
            return fail(FWI_HIP_EINVAL, "shot %d: msf.dt %g != jump %zu x base.dt %g", s, m->dt, m->jump, bs->dt);
        if (bs->nx != b->nx || bs->nz != b->nz || bs->npml != b->npml || bs->dx != b->dx || bs->dz != b->dz ||
            bs->dt != b->dt || bs->free_surface != b->free_surface)
            return fail(FWI_HIP_EINVAL, "shot %d: grid (nx, nz, npml, dx, dz, dt) differs from shot 0", s);
    }
    return FWI_HIP_OK;
}

/* imaging window from a plan's fwi_grad: every jump-th step with
 * start_t <= k dt <= end_t (end_t <= 0: to the last step) */
static int set_window(fwi_hip_ctx *ctx, const fwi_grad *gr, const fwi_base *b)
{
    const int jump = gr->jump > 0 ? (int)gr->jump : 1;
    const int k0 = gr->start_t > 0.0f ? (int)ceil((double)gr->start_t / b->dt - 1e-9) : 0;
    const int k1 = gr->end_t > 0.0f ? (int)floor((double)gr->end_t / b->dt + 1e-9) : INT_MAX;
    const int rc = k1 >= k0 ? fwi_hip_set_imaging(ctx, jump, k0, k1) : FWI_HIP_EINVAL;
    if (rc) return fail(rc, "imaging window: start_t=%g end_t=%g jump=%zu", gr->start_t, gr->end_t, gr->jump);
    return FWI_HIP_OK;
}

/* the fwi_hip shot table of the plans' sources and receivers */
typedef struct {
    fwi_hip_shot *sh;
    int **cells;
    int n;
    long long ntr; /* trace samples at the propagation rate */
} ShotTable;

static void shot_table_free(ShotTable *t)
{
    if (t->cells)
        for (int s = 0; s < 2 * t->n; s++) free(t->cells[s]);
    free(t->cells);
    free(t->sh);
    memset(t, 0, sizeof(*t));
}

static int shot_table(ShotTable *t, fwi_specs *const *grids, fwi_sou *const *sous, fwi_msf *const *msfs,
                      int nshots)
{
    memset(t, 0, sizeof(*t));
    t->n = nshots;
    t->sh = (fwi_hip_shot *)calloc(nshots, sizeof(fwi_hip_shot));
    t->cells = (int **)calloc(2 * nshots, sizeof(int *));
    if (!t->sh || !t->cells) {
        shot_table_free(t);
        return fail(FWI_HIP_ENOMEM, "host memory");
    }
    for (int s = 0; s < nshots; s++) {
        const fwi_sou *so = sous[s];
        const int nr = (int)msfs[s]->nr;
        t->cells[2 * s] = (int *)malloc(sizeof(int) * (nr ? nr : 1));
        t->cells[2 * s + 1] = (int *)malloc(sizeof(int) * (nr ? nr : 1));
        if (!t->cells[2 * s] || !t->cells[2 * s + 1]) {
            shot_table_free(t);
            return fail(FWI_HIP_ENOMEM, "host memory");
        }
        for (int r = 0; r < nr; r++) { /* ixyz_sr layout, fwi_2d.c:366-371 */
            t->cells[2 * s][r] = so->ixyz_sr[2 + 2 * r];
            t->cells[2 * s + 1][r] = so->ixyz_sr[2 + 2 * r + 1];
        }
        const int nt = (int)grids[s]->base->nt;
        t->sh[s].nt = nt;
        t->sh[s].isx = so->ixyz_sr[0];
        t->sh[s].isz = so->ixyz_sr[1];
        t->sh[s].nr = nr;
        t->sh[s].irx = t->cells[2 * s];
        t->sh[s].irz = t->cells[2 * s + 1];
        t->sh[s].wav = so->wav0;
        t->ntr += (long long)nr * nt;
    }
    return FWI_HIP_OK;
}

/* model, imaging window and shots of one context */
static int ctx_setup(fwi_hip_ctx *ctx, const float *vel, const fwi_grad *gr, const fwi_base *b, const ShotTable *t)
{
    int rc = fwi_hip_set_model(ctx, vel);
    if (rc) return fail(rc, "set_model: %s", fwi_hip_last_error(ctx));
    if ((rc = set_window(ctx, gr, b))) return rc;
    rc = fwi_hip_set_shots(ctx, t->n, t->sh);
    if (rc) return fail(rc, "set_shots: %s", fwi_hip_last_error(ctx));
    return FWI_HIP_OK;
}

/* The L2 misfit of plain traces (no band-pass, no mute, data at the
 * propagation rate) needs only the samples of one time step at a time, so it
 * runs on each block of trace rows as fwi_hip_forward_tm hands it over,
 * beside the propagation of the later blocks: msf->syn gets the rows, the
 * rows become the residual syn - obs0 in place (0 on traces outside the
 * offset window), and the per-shot sums of 1/2 res^2 (FP64) become val. */
typedef struct {
    fwi_plan **plans;
    float *tr;
    const long long *off;
    int nshots, nth;
    double *J;
    unsigned char **drop; /* per shot: traces outside the offset window, or NULL */
} L2Stream;

typedef struct {
    const L2Stream *st;
    int k0, k1, t;
} L2Job;

static void *l2_rows(void *p)
{
    const L2Job *j = (const L2Job *)p;
    const L2Stream *st = j->st;
    for (int s = j->t; s < st->nshots; s += st->nth) {
        fwi_msf *m = st->plans[s]->msf;
        const int nr = (int)m->nr, k1 = j->k1 < (int)m->nt ? j->k1 : (int)m->nt;
        const unsigned char *dr = st->drop ? st->drop[s] : NULL;
        float *t = st->tr + st->off[s];
        double acc = 0.0;
        for (int k = j->k0; k < k1; k++) {
            float *row = t + (size_t)k * nr;
            for (int r = 0; r < nr; r++) {
                const float v = row[r];
                m->syn[r][k] = v;
                const double d = (dr && dr[r]) ? 0.0 : (double)v - m->obs0[r][k];
                row[r] = (float)d;
                acc += 0.5 * d * d;
            }
        }
        st->J[s] += acc;
    }
    return NULL;
}

static int l2_block(void *user, int k0, int k1)
{
    const L2Stream *st = (const L2Stream *)user;
    L2Job jobs[MISFIT_MAX_THREADS];
    pthread_t th[MISFIT_MAX_THREADS];
    int made[MISFIT_MAX_THREADS];
    for (int t = 0; t < st->nth; t++) {
        jobs[t] = (L2Job){st, k0, k1, t};
        made[t] = t > 0 && pthread_create(&th[t], NULL, l2_rows, &jobs[t]) == 0;
    }
    for (int t = 0; t < st->nth; t++)
        if (!made[t]) l2_rows(&jobs[t]);
    for (int t = 1; t < st->nth; t++)
        if (made[t]) pthread_join(th[t], NULL);
    return 0;
}

/* whether the batch's misfit step is the plain L2 of l2_block */
static int l2_streams(fwi_plan **plans, int nshots)
{
    if (g_cfg.kind != GRAD4_MSF_L2 || g_cfg.bp || g_cfg.mute) return 0;
    for (int s = 0; s < nshots; s++)
        if (plans[s]->msf->jump != 1 || plans[s]->msf->nt != plans[s]->grid->base->nt) return 0;
    return 1;
}

int grad_all_shots(fwi_plan **plans, int nshots, int k)
{
    (void)k;
    if (!plans || nshots < 1 || !plans[0]) return fail(FWI_HIP_EINVAL, "grad_all_shots: no plans");
    double tm[GRAD4_NPHASE + 1];
    memset(g_phase_ms, 0, sizeof(g_phase_ms));
    tm[0] = now_ms();
    const fwi_base *b = plans[0]->grid->base;
    const int nx = (int)b->nx, nz = (int)b->nz;
    fwi_specs **grids = (fwi_specs **)malloc(sizeof(void *) * nshots);
    fwi_msf **msfs = (fwi_msf **)malloc(sizeof(void *) * nshots);
    fwi_sou **sous = (fwi_sou **)malloc(sizeof(void *) * nshots);
    ShotTable t = {0};
    float *tr = NULL;
    double *grad = NULL;
    int nrmax = 0, ntmax = 0, rc = (grids && msfs && sous) ? FWI_HIP_OK : fail(FWI_HIP_ENOMEM, "host memory");
    for (int s = 0; !rc && s < nshots; s++) {
        grids[s] = plans[s]->grid;
        msfs[s] = plans[s]->msf;
        sous[s] = plans[s]->sou;
    }
    if (!rc) rc = check_shots(b, grids, msfs, nshots, &nrmax, &ntmax);
    if (!rc) rc = ensure_ctx(0, b, nshots, ntmax, nrmax);
    if (!rc) rc = shot_table(&t, grids, sous, msfs, nshots);
    fwi_hip_ctx *ctx = g_slot[0].ctx;
    if (!rc) rc = ctx_setup(ctx, &plans[0]->model->md[0][0], plans[0]->grad, b, &t);
    if (!rc) {
        grad = (double *)malloc(sizeof(double) * nx * nz);
        if (!grad) rc = fail(FWI_HIP_ENOMEM, "host memory");
        else if (!(tr = slot_traces(0, (size_t)t.ntr))) rc = FWI_HIP_ENOMEM;
    }
    /* traces time-major in blocks beside the propagation; the plain L2
     * misfit runs on each block as it arrives, any other after the pass */
    const int stream_l2 = !rc && l2_streams(plans, nshots);
    L2Stream l2 = {plans, tr, NULL, nshots, misfit_threads(nshots), NULL, NULL};
    long long *toff = NULL; /* packed trace offsets */
    if (stream_l2) {
        toff = (long long *)malloc(sizeof(long long) * (nshots + 1));
        if (toff) {
            toff[0] = 0;
            for (int s = 0; s < nshots; s++)
                toff[s + 1] = toff[s] + (long long)plans[s]->msf->nr * (long long)plans[s]->grid->base->nt;
        }
        l2.off = toff;
        l2.J = (double *)calloc(nshots, sizeof(double));
        if (!toff) rc = fail(FWI_HIP_ENOMEM, "host memory");
        l2.drop = g_cfg.offsets ? (unsigned char **)calloc(nshots, sizeof(unsigned char *)) : NULL;
        if (!l2.J || (g_cfg.offsets && !l2.drop)) rc = fail(FWI_HIP_ENOMEM, "host memory");
        for (int s = 0; !rc && l2.drop && s < nshots; s++) {
            const int nr = (int)plans[s]->msf->nr;
            const float *sr = plans[s]->sou->xyz_sr;
            if (!(l2.drop[s] = (unsigned char *)calloc(nr ? nr : 1, 1))) rc = fail(FWI_HIP_ENOMEM, "host memory");
            for (int r = 0; !rc && r < nr; r++) {
                const double dx = (double)sr[2 + 2 * r] - sr[0], dz = (double)sr[3 + 2 * r] - sr[1];
                const double d = sqrt(dx * dx + dz * dz);
                l2.drop[s][r] = d < g_cfg.omin || d > g_cfg.omax;
            }
        }
    }
    tm[1] = now_ms();
    if (!rc && (rc = fwi_hip_forward_tm(ctx, tr, stream_l2 ? l2_block : NULL, &l2)))
        rc = fail(rc, "forward: %s", fwi_hip_last_error(ctx));
    tm[2] = now_ms();
    if (!rc && stream_l2)
        for (int s = 0; s < nshots; s++) plans[s]->msf->val[0] = plans[s]->msf->val[1] = (float)l2.J[s];
    else if (!rc)
        rc = host_misfit_all(plans, nshots, tr, b->dt, 1);
    tm[3] = now_ms();
    if (!rc && (rc = fwi_hip_set_residual_tm(ctx, tr))) rc = fail(rc, "set_residual: %s", fwi_hip_last_error(ctx));
    tm[4] = now_ms();
    if (!rc && (rc = fwi_hip_adjoint(ctx, grad, NULL))) rc = fail(rc, "adjoint: %s", fwi_hip_last_error(ctx));
    if (!rc)
        for (long i = 0; i < (long)nx * nz; i++) plans[0]->grad->body[i] = (float)grad[i];
    tm[5] = now_ms();
    for (int i = 0; i < GRAD4_NPHASE; i++) g_phase_ms[i] = tm[i + 1] - tm[i];
    for (int s = 0; l2.drop && s < nshots; s++) free(l2.drop[s]);
    free(l2.drop);
    free(l2.J);
    free(toff);
    shot_table_free(&t);
    free(grids);
    free(msfs);
    free(sous);
    free(grad);
    return rc;
}

void grad4_1shot(fwi_plan *plan, int k)
{
    fwi_plan *one[1] = {plan};
    int rc = grad_all_shots(one, 1, k);
    if (rc) {
        /* the reference has no error channel here: report and poison the misfit */
        fprintf(stderr, "grad4_1shot: %s\n", g_msg);
        if (plan && plan->msf && plan->msf->val) plan->msf->val[0] = plan->msf->val[1] = NAN;
    }
}

/* ------------------------------------------------------------------------ */
/* reflection waveform inversion, difference form (fwi_plan.h): the incident
 * field in m0 on context slot 0, the perturbed field in m0 + m1 on slot 1 */
int grad_all_shots_rfl(rwi_plan **plans, int nshots, int k, float *grad_m1)
{
    (void)k;
    if (!plans || nshots < 1 || !plans[0]) return fail(FWI_HIP_EINVAL, "grad_all_shots_rfl: no plans");
    const fwi_base *b = plans[0]->grid->base;
    const int nx = (int)b->nx, nz = (int)b->nz;
    const long nb = (long)nx * nz;
    fwi_specs **grids = (fwi_specs **)malloc(sizeof(void *) * nshots);
    fwi_msf **msfs = (fwi_msf **)malloc(sizeof(void *) * nshots);
    fwi_sou **sous = (fwi_sou **)malloc(sizeof(void *) * nshots);
    fwi_plan *shell = (fwi_plan *)calloc(nshots, sizeof(fwi_plan)); /* the host misfit step's view */
    fwi_plan **shells = (fwi_plan **)malloc(sizeof(void *) * nshots);
    ShotTable t = {0};
    float *tr0 = NULL, *tr1 = NULL, *vtot = NULL;
    double *g0 = NULL, *g1 = NULL;
    int nrmax = 0, ntmax = 0,
        rc = (grids && msfs && sous && shell && shells) ? FWI_HIP_OK : fail(FWI_HIP_ENOMEM, "host memory");
    for (int s = 0; !rc && s < nshots; s++) {
        grids[s] = plans[s]->grid;
        msfs[s] = plans[s]->msf;
        sous[s] = plans[s]->sou;
        shell[s].grid = plans[s]->grid;
        shell[s].msf = plans[s]->msf;
        shell[s].sou = plans[s]->sou;
        shell[s].grad = plans[s]->grad;
        shells[s] = &shell[s];
    }
    if (!rc) rc = check_shots(b, grids, msfs, nshots, &nrmax, &ntmax);
    if (!rc) rc = ensure_ctx(0, b, nshots, ntmax, nrmax);
    if (!rc) rc = ensure_ctx(1, b, nshots, ntmax, nrmax);
    if (!rc) rc = shot_table(&t, grids, sous, msfs, nshots);
    if (!rc) {
        tr0 = slot_traces(0, (size_t)t.ntr);
        tr1 = slot_traces(1, (size_t)t.ntr);
        vtot = (float *)malloc(sizeof(float) * nb);
        g0 = (double *)malloc(sizeof(double) * nb);
        g1 = (double *)malloc(sizeof(double) * nb);
        if (!tr0 || !tr1 || !vtot || !g0 || !g1) rc = fail(FWI_HIP_ENOMEM, "host memory");
    }
    fwi_hip_ctx *c0 = g_slot[0].ctx, *c1 = g_slot[1].ctx;
    if (!rc) {
        const float *m0 = &plans[0]->model->m0[0][0], *m1 = &plans[0]->model->m1[0][0];
        for (long i = 0; i < nb; i++) vtot[i] = m0[i] + m1[i];
        rc = ctx_setup(c0, m0, plans[0]->grad, b, &t);
        if (!rc) rc = ctx_setup(c1, vtot, plans[0]->grad, b, &t);
    }
    if (!rc && (rc = fwi_hip_forward(c0, tr0))) rc = fail(rc, "forward (incident): %s", fwi_hip_last_error(c0));
    if (!rc && (rc = fwi_hip_forward(c1, tr1))) rc = fail(rc, "forward (perturbed): %s", fwi_hip_last_error(c1));
    if (!rc) {
        for (long long i = 0; i < t.ntr; i++) tr1[i] -= tr0[i]; /* the reflected data */
        rc = host_misfit_all(shells, nshots, tr1, b->dt, 0); /* tr1 <- adjoint source r */
    }
    if (!rc && (rc = fwi_hip_set_residual(c1, tr1))) rc = fail(rc, "set_residual: %s", fwi_hip_last_error(c1));
    if (!rc && (rc = fwi_hip_adjoint(c1, g1, NULL))) rc = fail(rc, "adjoint (perturbed): %s", fwi_hip_last_error(c1));
    if (!rc) {
        for (long long i = 0; i < t.ntr; i++) tr0[i] = -tr1[i];
        if ((rc = fwi_hip_set_residual(c0, tr0))) rc = fail(rc, "set_residual: %s", fwi_hip_last_error(c0));
    }
    if (!rc && (rc = fwi_hip_adjoint(c0, g0, NULL))) rc = fail(rc, "adjoint (incident): %s", fwi_hip_last_error(c0));
    if (!rc) {
        for (long i = 0; i < nb; i++) plans[0]->grad->body[i] = (float)(g0[i] + g1[i]);
        if (grad_m1)
            for (long i = 0; i < nb; i++) grad_m1[i] = (float)g1[i];
    }
    shot_table_free(&t);
    free(grids);
    free(msfs);
    free(sous);
    free(shell);
    free(shells);
    free(vtot);
    free(g0);
    free(g1);
    return rc;
}

void grad4_1shot_rfl(rwi_plan *plan, int k)
{
    rwi_plan *one[1] = {plan};
    int rc = grad_all_shots_rfl(one, 1, k, NULL);
    if (rc) {
        fprintf(stderr, "grad4_1shot_rfl: %s\n", g_msg);
        if (plan && plan->msf && plan->msf->val) plan->msf->val[0] = plan->msf->val[1] = NAN;
    }
}

/* ------------------------------------------------------------------------ */
/* row-pointer arrays as alloc.c:3-16: a[n2][n1] over one zeroed slab */
static float **alloc2d(size_t n1, size_t n2)
{
    float **a = (float **)calloc(n2 ? n2 : 1, sizeof(float *));
    size_t n = n1 * n2;
    float *s = (float *)calloc(n ? n : 1, sizeof(float));
    if (!a || !s) {
        free(a);
        free(s);
        return NULL;
    }
    for (size_t i = 0; i < n2; i++) a[i] = s + i * n1;
    if (n2 == 0) a[0] = s;
    return a;
}

static void free2d(float **a)
{
    if (!a) return;
    free(a[0]);
    free(a);
}

/* grid, misfit, source and gradient members shared by the FWI and RWI
 * plans (fwi_2d.c:148-165,194-224,380; rwi_init the same, fwi_2d.c:402-515) */
typedef struct {
    fwi_specs *grid;
    fwi_msf *msf;
    fwi_sou *sou;
    fwi_grad *grad;
} PlanCommon;

static void common_free(PlanCommon *c)
{
    if (c->msf) {
        free2d(c->msf->obs0);
        free2d(c->msf->obs);
        free2d(c->msf->syn);
        free(c->msf->val);
        free(c->msf->fbt);
        free(c->msf->ifbt);
        free(c->msf);
    }
    if (c->sou) {
        free(c->sou->ixyz_sr);
        free(c->sou->xyz_sr);
        free(c->sou->wav0);
        free(c->sou->wav);
        free(c->sou);
    }
    if (c->grad) {
        free(c->grad->body);
        free(c->grad);
    }
    if (c->grid) {
        free(c->grid->base);
        free(c->grid->f);
        free(c->grid->b);
        free(c->grid);
    }
}

static int common_alloc(PlanCommon *c, int nx, int nz, int npml, int nt, int jump, int nr, float dx, float dz,
                        float dt)
{
    memset(c, 0, sizeof(*c));
    if (nx < 1 || nz < 1 || npml < 0 || nt < 1 || jump < 1 || nr < 0 || (nt - 1) % jump != 0) return 0;
    const fwi_base base = {.dim = 2, .mm = 2, .npml = (size_t)npml, .nx = (size_t)nx, .nz = (size_t)nz,
                           .nt = (size_t)nt, .nt_mem = 1, .free_surface = 1, .dx = dx, .dz = dz, .dt = dt};
    fwi_base *bp = (fwi_base *)malloc(sizeof(fwi_base));
    fwi_shape *gb = (fwi_shape *)calloc(1, sizeof(fwi_shape)), *gf = (fwi_shape *)calloc(1, sizeof(fwi_shape));
    c->grid = (fwi_specs *)calloc(1, sizeof(fwi_specs));
    if (!bp || !gb || !gf || !c->grid) {
        free(bp);
        free(gb);
        free(gf);
        free(c->grid);
        c->grid = NULL;
        return 0;
    }
    memcpy(bp, &base, sizeof(base));
    gb->mnx = gb->pnx = nx;            /* fwi_2d.c:148-152 */
    gb->mnz = gb->pnz = nz;
    gb->unx = nx + 1; gb->unz = nz;
    gb->vnx = nx; gb->vnz = nz + 1;
    gf->mnx = gf->pnx = nx + 2 * npml; /* fwi_2d.c:154-157 */
    gf->mnz = gf->pnz = nz + npml;
    gf->unx = gf->pnx + 1; gf->unz = gf->pnz;
    gf->vnx = gf->pnx; gf->vnz = gf->pnz + 1;
    c->grid->base = bp;
    c->grid->f = gf;
    c->grid->b = gb;
    const int nt0 = (nt - 1) / jump + 1;
    const fwi_msf msf = {.nt = (size_t)nt0, .nr = (size_t)nr, .jump = (size_t)jump, .dt = dt * jump};
    c->msf = (fwi_msf *)malloc(sizeof(fwi_msf));
    if (c->msf) {
        memcpy(c->msf, &msf, sizeof(msf));
        c->msf->obs0 = alloc2d(nt0, nr);
        c->msf->obs = alloc2d(nt0, nr);
        c->msf->syn = alloc2d(nt0, nr);
        c->msf->val = (float *)calloc(2, sizeof(float));
        c->msf->fbt = (float *)calloc(nr ? nr : 1, sizeof(float));
        c->msf->ifbt = (int *)calloc(nr ? nr : 1, sizeof(int));
    }
    c->sou = (fwi_sou *)calloc(1, sizeof(fwi_sou));
    if (c->sou) {
        c->sou->iarea = 1.0 / (dx * dz);
        c->sou->ixyz_sr = (int *)calloc((size_t)(1 + nr) * 2, sizeof(int));
        c->sou->xyz_sr = (float *)calloc((size_t)(1 + nr) * 2, sizeof(float));
        c->sou->wav0 = (float *)calloc(nt, sizeof(float));
        c->sou->wav = (float *)calloc(nt, sizeof(float));
    }
    const fwi_grad gr = {.jump = 1};
    c->grad = (fwi_grad *)malloc(sizeof(fwi_grad));
    if (c->grad) {
        memcpy(c->grad, &gr, sizeof(gr));
        c->grad->body = (float *)calloc((size_t)nx * nz, sizeof(float));
    }
    const int ok = c->msf && c->msf->obs0 && c->msf->obs && c->msf->syn && c->msf->val && c->msf->fbt &&
                   c->msf->ifbt && c->sou && c->sou->ixyz_sr && c->sou->xyz_sr && c->sou->wav0 && c->sou->wav &&
                   c->grad && c->grad->body;
    if (!ok) common_free(c);
    return ok;
}

fwi_plan *fwi_plan_alloc(int nx, int nz, int npml, int nt, int jump, int nr, float dx, float dz, float dt)
{
    PlanCommon c;
    if (!common_alloc(&c, nx, nz, npml, nt, jump, nr, dx, dz, dt)) return NULL;
    fwi_plan *p = (fwi_plan *)calloc(1, sizeof(fwi_plan));
    fwi_model *m = (fwi_model *)calloc(1, sizeof(fwi_model));
    if (!p || !m) {
        free(p);
        free(m);
        common_free(&c);
        return NULL;
    }
    p->grid = c.grid;
    p->msf = c.msf;
    p->sou = c.sou;
    p->grad = c.grad;
    p->model = m;
    m->md = alloc2d(nx, nz);
    m->kp = alloc2d(nx, nz);
    m->ekp = alloc2d(c.grid->f->mnx, c.grid->f->mnz);
    m->btry = (float *)calloc(nx, sizeof(float));
    m->ibtry = (int *)calloc(nx, sizeof(int));
    if (!m->md || !m->kp || !m->ekp || !m->btry || !m->ibtry) {
        fwi_plan_free(p);
        return NULL;
    }
    return p;
}

rwi_plan *rwi_plan_alloc(int nx, int nz, int npml, int nt, int jump, int nr, float dx, float dz, float dt)
{
    PlanCommon c;
    if (!common_alloc(&c, nx, nz, npml, nt, jump, nr, dx, dz, dt)) return NULL;
    rwi_plan *p = (rwi_plan *)calloc(1, sizeof(rwi_plan));
    rwi_model *m = (rwi_model *)calloc(1, sizeof(rwi_model));
    if (!p || !m) {
        free(p);
        free(m);
        common_free(&c);
        return NULL;
    }
    p->grid = c.grid;
    p->msf = c.msf;
    p->sou = c.sou;
    p->grad = c.grad;
    p->model = m;
    const size_t fnx = c.grid->f->mnx, fnz = c.grid->f->mnz; /* rwi_init, fwi_2d.c:438-452 */
    m->m0 = alloc2d(nx, nz);
    m->m1 = alloc2d(nx, nz);
    m->kp_0 = alloc2d(nx, nz);
    m->kp_total = alloc2d(nx, nz);
    m->ekp_0 = alloc2d(fnx, fnz);
    m->ekp_total = alloc2d(fnx, fnz);
    m->btry = (float *)calloc(nx, sizeof(float));
    m->ibtry = (int *)calloc(nx, sizeof(int));
    if (!m->m0 || !m->m1 || !m->kp_0 || !m->kp_total || !m->ekp_0 || !m->ekp_total || !m->btry || !m->ibtry) {
        rwi_plan_free(p);
        return NULL;
    }
    return p;
}

void rwi_plan_free(rwi_plan *p)
{
    if (!p) return;
    if (p->model) {
        free2d(p->model->m0);
        free2d(p->model->m1);
        free2d(p->model->kp_0);
        free2d(p->model->kp_total);
        free2d(p->model->ekp_0);
        free2d(p->model->ekp_total);
        free(p->model->btry);
        free(p->model->ibtry);
        free(p->model);
    }
    PlanCommon c = {p->grid, p->msf, p->sou, p->grad};
    common_free(&c);
    free(p);
}

int fwi_plan_set_grad(fwi_plan *p, float perc, float alpha_grad, int rectx1, int rectz1, int repeat1, int rectx2,
                      int rectz2, int repeat2)
{
    if (!p || !p->grad || rectx1 < 0 || rectz1 < 0 || repeat1 < 0 || rectx2 < 0 || rectz2 < 0 || repeat2 < 0)
        return FWI_HIP_EINVAL;
    const fwi_grad *o = p->grad;
    const fwi_grad gr = {.id = o->id, .end_t = o->end_t, .start_t = o->start_t, .alpha_grad = alpha_grad,
                         .perc = perc, .jump = o->jump, .rectx1 = (size_t)rectx1, .rectz1 = (size_t)rectz1,
                         .repeat1 = (size_t)repeat1, .rectx2 = (size_t)rectx2, .rectz2 = (size_t)rectz2,
                         .repeat2 = (size_t)repeat2, .body = o->body};
    memcpy(p->grad, &gr, sizeof(gr));
    return FWI_HIP_OK;
}

void fwi_plan_free(fwi_plan *p)
{
    if (!p) return;
    if (p->model) {
        free2d(p->model->md);
        free2d(p->model->kp);
        free2d(p->model->ekp);
        free(p->model->btry);
        free(p->model->ibtry);
        free(p->model);
    }
    PlanCommon c = {p->grid, p->msf, p->sou, p->grad};
    common_free(&c);
    free(p);
}
