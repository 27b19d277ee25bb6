/*
 * grad4_1shot.c -- the reference's per-shot gradient entry point
 * (header.h:283, called at MBotOpt.c:55,99, never defined there) and its
 * batched form, in C on top of the fwi_hip C ABI.
 *
 * Sequence (the driver this project defines; the oracle implements the
 * same one): forward propagation with receiver recording and frame saving
 * on the GPU -> L2 misfit and adjoint source on the host in FP64 ->
 * exact-adjoint back-propagation, reconstruction and zero-lag imaging on the
 * GPU -> gradient copied back once.  Data sampled every msf->jump steps
 * (fwi_2d.c:195).  The adjoint source at the propagation rate is, by
 * grad4_set_residual_mode:
 *   0 (default) the residual placed on the sampled steps, zero between them
 *     (the exact transpose of the subsampling: the gradient is exact);
 *   1 the residual upsampled by the natural cubic spline of trcinp
 *     (spline_interp.c:19-38), the reference's resampling of traces between
 *     the data rate and the propagation rate.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fwi_hip.h"
#include "fwi_io.h"
#include "fwi_plan.h"

static fwi_hip_ctx *g_ctx = NULL;
static long g_key[8];
static char g_msg[512] = "";
static int g_res_mode = 0;

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
    return FWI_HIP_OK;
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
    rc = fwi_hip_set_model(g_ctx, &plans[0]->model->md[0][0]);
    if (rc) return fail(rc, "set_model: %s", fwi_hip_last_error(g_ctx));

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
    if (!rc) {
        /* host misfit step: sample every jump, L2, adjoint source at the fine rate */
        long long o = 0;
        for (int s = 0; s < nshots && !rc; s++) {
            fwi_msf *m = plans[s]->msf;
            const int nr = (int)m->nr, nt0 = (int)m->nt, jump = (int)m->jump;
            double J = 0.0;
            for (int r = 0; r < nr; r++) {
                float *t = tr + o + (long long)r * nt;
                for (int i0 = 0; i0 < nt0; i0++) {
                    const float syn = t[(long long)i0 * jump];
                    m->syn[r][i0] = syn;
                    const float res = syn - m->obs0[r][i0];
                    J += 0.5 * (double)res * (double)res;
                }
                if (g_res_mode == 0 || jump == 1) {
                    for (int it = nt - 1; it >= 0; it--) /* residual on the samples, zero between them */
                        t[it] = (it % jump == 0) ? t[it] - m->obs0[r][it / jump] : 0.0f;
                } else {
                    float *r0 = (float *)malloc(sizeof(float) * nt0);
                    if (!r0) { rc = fail(FWI_HIP_ENOMEM, "host memory"); break; }
                    for (int i0 = 0; i0 < nt0; i0++) r0[i0] = m->syn[r][i0] - m->obs0[r][i0];
                    float *rows0[1] = {r0}, *rows1[1] = {t};
                    int e = fwi_trcinp(rows0, nt0, m->dt, rows1, nt, b->dt, 1);
                    free(r0);
                    if (e) { rc = fail(e, "trcinp: %s", fwi_io_last_error()); break; }
                }
            }
            m->val[0] = (float)J;
            m->val[1] = (float)J;
            o += (long long)nr * nt;
        }
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
