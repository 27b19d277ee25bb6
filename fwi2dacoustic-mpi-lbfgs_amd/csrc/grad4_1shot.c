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

#include "fwi_hip.h"
#include "fwi_io.h"
#include "fwi_msf.h"
#include "fwi_plan.h"
#include "fwi_sigproc.h"

static fwi_hip_ctx *g_ctx = NULL;
static long g_key[8];
static char g_msg[512] = "";
static int g_res_mode = 0;
static int g_ck_every = 0, g_ck_storage = 0, g_ck_applied = 0; /* checkpointing of the batched context */

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
    g_ck_applied = 0;
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

/* one context per process, re-created when the geometry or the capacity changes */
static int ensure_ctx(const fwi_base *b, int nshots, int nt, int nr)
{
    long key[8] = {(long)b->nx, (long)b->nz, (long)b->npml, (long)nshots, (long)nt, (long)nr, 0, 0};
    memcpy(&key[6], &b->dx, sizeof(float));
    memcpy((char *)&key[6] + 4, &b->dz, sizeof(float));
    memcpy(&key[7], &b->dt, sizeof(float));
    if (g_ctx && memcmp(key, g_key, sizeof(key)) == 0) return FWI_HIP_OK;
    if (g_ctx) fwi_hip_destroy(g_ctx);
    g_ctx = NULL;
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
    int rc = fwi_hip_create(&d, &g_ctx);
    if (rc) return fail(rc, "fwi_hip_create: %s", fwi_hip_last_error(NULL));
    memcpy(g_key, key, sizeof(key));
    g_ck_applied = 0;
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

/* The host misfit step of one shot.  t holds the shot's nr traces at the
 * propagation rate (nt samples each) and receives the adjoint source. */
static int host_misfit(fwi_plan *pl, float *t, int nt, float dt_fine)
{
    fwi_msf *m = pl->msf;
    const int nr = (int)m->nr, nt0 = (int)m->nt, jump = (int)m->jump;
    const size_t n = (size_t)nr * nt0;
    float *ob = (float *)malloc(sizeof(float) * (n ? n : 1)), *sy = (float *)malloc(sizeof(float) * (n ? n : 1));
    float *rs = (float *)calloc(n ? n : 1, sizeof(float));
    float **obr = slab_rows(ob, nt0, nr), **syr = slab_rows(sy, nt0, nr), **rsr = slab_rows(rs, nt0, nr);
    unsigned char *drop = NULL; /* traces outside the offset window */
    int rc = FWI_HIP_OK;
    if (!ob || !sy || !rs || !obr || !syr || !rsr) {
        rc = fail(FWI_HIP_ENOMEM, "host memory");
        goto done;
    }
    /* synthetic at the data rate (fwi_2d.c:195) */
    for (int r = 0; r < nr; r++)
        for (int i0 = 0; i0 < nt0; i0++) {
            const float v = t[(size_t)r * nt + (size_t)i0 * jump];
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
    default:
        rc = L2norm(obr, syr, nt0, nr, sr, g_cfg.omin, g_cfg.omax, val, rsr);
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
            for (int it = 0; it < nt; it++)
                t[(size_t)r * nt + it] = (it % jump == 0) ? rs[(size_t)r * nt0 + it / jump] : 0.0f;
    } else {
        float **tr = slab_rows(t, nt, nr);
        if (!tr) {
            rc = fail(FWI_HIP_ENOMEM, "host memory");
            goto done;
        }
        int e = fwi_trcinp(rsr, nt0, m->dt, tr, nt, dt_fine, nr);
        free(tr);
        if (e) rc = fail(e, "trcinp: %s", fwi_io_last_error());
    }
done:
    free(drop);
    free(ob);
    free(sy);
    free(rs);
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
    int nshots, nt, nthreads, t, rc;
    float dt;
} MisfitJob;

static void *misfit_worker(void *p)
{
    MisfitJob *j = (MisfitJob *)p;
    for (int s = j->t; s < j->nshots && !j->rc; s += j->nthreads)
        j->rc = host_misfit(j->plans[s], j->tr + j->off[s], j->nt, j->dt);
    return NULL;
}

static int host_misfit_all(fwi_plan **plans, int nshots, float *tr, int nt, float dt)
{
    long long *off = (long long *)malloc(sizeof(long long) * (nshots + 1));
    if (!off) return fail(FWI_HIP_ENOMEM, "host memory");
    off[0] = 0;
    for (int s = 0; s < nshots; s++) off[s + 1] = off[s] + (long long)plans[s]->msf->nr * nt;
    int nth = 16;
    const char *env = getenv("FWI_HOST_THREADS");
    if (env && atoi(env) > 0) nth = atoi(env);
    else {
        const long cpu = sysconf(_SC_NPROCESSORS_ONLN);
        if (cpu > 0 && cpu < nth) nth = (int)cpu;
    }
    if (nth > nshots) nth = nshots;
    if (nth < 1) nth = 1;
    if (nth > 64) nth = 64;
    MisfitJob jobs[64];
    pthread_t th[64];
    int made[64];
    int rc = FWI_HIP_OK;
    for (int t = 0; t < nth; t++) {
        jobs[t] = (MisfitJob){plans, tr, off, nshots, nt, nth, t, FWI_HIP_OK, dt};
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

int grad_all_shots(fwi_plan **plans, int nshots, int k)
{
    (void)k;
    if (!plans || nshots < 1 || !plans[0]) return fail(FWI_HIP_EINVAL, "grad_all_shots: no plans");
    const fwi_base *b = plans[0]->grid->base;
    if (b->free_surface != 1) return fail(FWI_HIP_EINVAL, "only the free-surface grid is supported");
    const int nx = (int)b->nx, nz = (int)b->nz, nt = (int)b->nt;
    int nrmax = 0;
    for (int s = 0; s < nshots; s++) {
        const fwi_msf *m = plans[s]->msf;
        if ((int)m->nr > nrmax) nrmax = (int)m->nr;
        if ((m->nt - 1) * m->jump != b->nt - 1) /* fwi_2d.c:195 */
            return fail(FWI_HIP_EINVAL, "shot %d: (nt0-1)*jump != nt-1 (%zu, %zu, %zu)", s, m->nt, m->jump, b->nt);
        if (plans[s]->grid->base->nx != b->nx || plans[s]->grid->base->nz != b->nz)
            return fail(FWI_HIP_EINVAL, "shot %d: grid differs from shot 0", s);
    }
    int rc = ensure_ctx(b, nshots, nt, nrmax);
    if (rc) return rc;
    if (!g_ck_applied) {
        rc = fwi_hip_set_checkpointing(g_ctx, g_ck_every, g_ck_storage);
        if (rc) return fail(rc, "set_checkpointing: %s", fwi_hip_last_error(g_ctx));
        g_ck_applied = 1;
    }
    rc = fwi_hip_set_model(g_ctx, &plans[0]->model->md[0][0]);
    if (rc) return fail(rc, "set_model: %s", fwi_hip_last_error(g_ctx));
    {
        /* imaging window from plans[0]->grad: every jump-th step with
         * start_t <= k dt <= end_t (end_t <= 0: to the last step) */
        const fwi_grad *gr = plans[0]->grad;
        const int jump = gr->jump > 0 ? (int)gr->jump : 1;
        const int k0 = gr->start_t > 0.0f ? (int)ceil((double)gr->start_t / b->dt - 1e-9) : 0;
        const int k1 = gr->end_t > 0.0f ? (int)floor((double)gr->end_t / b->dt + 1e-9) : INT_MAX;
        rc = k1 >= k0 ? fwi_hip_set_imaging(g_ctx, jump, k0, k1) : FWI_HIP_EINVAL;
        if (rc) return fail(rc, "imaging window: start_t=%g end_t=%g jump=%zu", gr->start_t, gr->end_t, gr->jump);
    }

    fwi_hip_shot *sh = (fwi_hip_shot *)calloc(nshots, sizeof(fwi_hip_shot));
    int **cells = (int **)calloc(2 * nshots, sizeof(int *));
    long long ntr = 0;
    for (int s = 0; s < nshots; s++) {
        const fwi_sou *so = plans[s]->sou;
        const int nr = (int)plans[s]->msf->nr;
        cells[2 * s] = (int *)malloc(sizeof(int) * (nr ? nr : 1));
        cells[2 * s + 1] = (int *)malloc(sizeof(int) * (nr ? nr : 1));
        for (int r = 0; r < nr; r++) { /* ixyz_sr layout, fwi_2d.c:366-371 */
            cells[2 * s][r] = so->ixyz_sr[2 + 2 * r];
            cells[2 * s + 1][r] = so->ixyz_sr[2 + 2 * r + 1];
        }
        sh[s].nt = nt;
        sh[s].isx = so->ixyz_sr[0];
        sh[s].isz = so->ixyz_sr[1];
        sh[s].nr = nr;
        sh[s].irx = cells[2 * s];
        sh[s].irz = cells[2 * s + 1];
        sh[s].wav = so->wav0;
        ntr += (long long)nr * nt;
    }
    float *tr = (float *)malloc(sizeof(float) * (ntr ? ntr : 1));
    double *grad = (double *)malloc(sizeof(double) * nx * nz);
    rc = fwi_hip_set_shots(g_ctx, nshots, sh);
    if (rc) rc = fail(rc, "set_shots: %s", fwi_hip_last_error(g_ctx));
    if (!rc) {
        rc = fwi_hip_forward(g_ctx, tr);
        if (rc) rc = fail(rc, "forward: %s", fwi_hip_last_error(g_ctx));
    }
    if (!rc) rc = host_misfit_all(plans, nshots, tr, nt, b->dt);
    if (!rc) {
        if (!rc) {
            rc = fwi_hip_set_residual(g_ctx, tr);
            if (rc) rc = fail(rc, "set_residual: %s", fwi_hip_last_error(g_ctx));
        }
    }
    if (!rc) {
        rc = fwi_hip_adjoint(g_ctx, grad, NULL);
        if (rc) rc = fail(rc, "adjoint: %s", fwi_hip_last_error(g_ctx));
    }
    if (!rc)
        for (long i = 0; i < (long)nx * nz; i++) plans[0]->grad->body[i] = (float)grad[i];
    for (int s = 0; s < 2 * nshots; s++) free(cells[s]);
    free(cells);
    free(sh);
    free(tr);
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
/* row-pointer arrays as alloc.c:3-16: a[n2][n1] over one zeroed slab */
static float **alloc2d(size_t n1, size_t n2)
{
    float **a = (float **)calloc(n2 ? n2 : 1, sizeof(float *));
    size_t n = n1 * n2;
    float *s = (float *)calloc(n ? n : 1, sizeof(float));
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

fwi_plan *fwi_plan_alloc(int nx, int nz, int npml, int nt, int jump, int nr, float dx, float dz, float dt)
{
    if (nx < 1 || nz < 1 || npml < 0 || nt < 1 || jump < 1 || nr < 0 || (nt - 1) % jump != 0) return NULL;
    fwi_plan *p = (fwi_plan *)calloc(1, sizeof(fwi_plan));
    const fwi_base base = {.dim = 2, .mm = 2, .npml = (size_t)npml, .nx = (size_t)nx, .nz = (size_t)nz,
                           .nt = (size_t)nt, .nt_mem = 1, .free_surface = 1, .dx = dx, .dz = dz, .dt = dt};
    fwi_base *bp = (fwi_base *)malloc(sizeof(fwi_base));
    memcpy(bp, &base, sizeof(base));
    fwi_shape *gb = (fwi_shape *)calloc(1, sizeof(fwi_shape)), *gf = (fwi_shape *)calloc(1, sizeof(fwi_shape));
    gb->mnx = gb->pnx = nx;            /* fwi_2d.c:148-152 */
    gb->mnz = gb->pnz = nz;
    gb->unx = nx + 1; gb->unz = nz;
    gb->vnx = nx; gb->vnz = nz + 1;
    gf->mnx = gf->pnx = nx + 2 * npml; /* fwi_2d.c:154-157 */
    gf->mnz = gf->pnz = nz + npml;
    gf->unx = gf->pnx + 1; gf->unz = gf->pnz;
    gf->vnx = gf->pnx; gf->vnz = gf->pnz + 1;
    p->grid = (fwi_specs *)calloc(1, sizeof(fwi_specs));
    p->grid->base = bp;
    p->grid->f = gf;
    p->grid->b = gb;
    p->model = (fwi_model *)calloc(1, sizeof(fwi_model));
    p->model->md = alloc2d(nx, nz);
    p->model->kp = alloc2d(nx, nz);
    p->model->ekp = alloc2d(gf->mnx, gf->mnz);
    p->model->btry = (float *)calloc(nx, sizeof(float));
    p->model->ibtry = (int *)calloc(nx, sizeof(int));
    const int nt0 = (nt - 1) / jump + 1;
    const fwi_msf msf = {.nt = (size_t)nt0, .nr = (size_t)nr, .jump = (size_t)jump, .dt = dt * jump};
    p->msf = (fwi_msf *)malloc(sizeof(fwi_msf));
    memcpy(p->msf, &msf, sizeof(msf));
    p->msf->obs0 = alloc2d(nt0, nr);
    p->msf->obs = alloc2d(nt0, nr);
    p->msf->syn = alloc2d(nt0, nr);
    p->msf->val = (float *)calloc(2, sizeof(float));
    p->msf->fbt = (float *)calloc(nr ? nr : 1, sizeof(float));
    p->msf->ifbt = (int *)calloc(nr ? nr : 1, sizeof(int));
    p->sou = (fwi_sou *)calloc(1, sizeof(fwi_sou));
    p->sou->iarea = 1.0 / (dx * dz);
    p->sou->ixyz_sr = (int *)calloc((size_t)(1 + nr) * 2, sizeof(int));
    p->sou->xyz_sr = (float *)calloc((size_t)(1 + nr) * 2, sizeof(float));
    p->sou->wav0 = (float *)calloc(nt, sizeof(float));
    p->sou->wav = (float *)calloc(nt, sizeof(float));
    const fwi_grad gr = {.jump = 1};
    p->grad = (fwi_grad *)malloc(sizeof(fwi_grad));
    memcpy(p->grad, &gr, sizeof(gr));
    p->grad->body = (float *)calloc((size_t)nx * nz, sizeof(float));
    return p;
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
    free2d(p->model->md);
    free2d(p->model->kp);
    free2d(p->model->ekp);
    free(p->model->btry);
    free(p->model->ibtry);
    free(p->model);
    free2d(p->msf->obs0);
    free2d(p->msf->obs);
    free2d(p->msf->syn);
    free(p->msf->val);
    free(p->msf->fbt);
    free(p->msf->ifbt);
    free(p->msf);
    free(p->sou->ixyz_sr);
    free(p->sou->xyz_sr);
    free(p->sou->wav0);
    free(p->sou->wav);
    free(p->sou);
    free(p->grad->body);
    free(p->grad);
    free(p->grid->base);
    free(p->grid->f);
    free(p->grid->b);
    free(p->grid);
    free(p);
}
