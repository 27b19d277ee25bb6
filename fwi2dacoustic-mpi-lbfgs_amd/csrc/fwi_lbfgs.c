/*
 * fwi_lbfgs.c -- the L-BFGS inversion loop of MBotOpt.c:65-221 on one
 * process, in FP64, and fwi_invert: that loop over batched GPU gradients.
 *
 * MBotOpt.c runs the optimiser on rank 0 and keeps the other ranks in step
 * with 1-int messages (tags lbfgs.h:80-83) and a broadcast of the model
 * before every gradient (MBotOpt.c:98); here the evaluation callback is the
 * whole batched gradient, so none of that is needed.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fwi_gradprec.h"
#include "fwi_hip.h"
#include "fwi_lbfgs.h"

static char g_lb_msg[512] = "";

static int lb_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_lb_msg, sizeof(g_lb_msg), fmt, ap);
    va_end(ap);
    return code;
}

const char *fwi_lbfgs_last_error(void) { return g_lb_msg; }

void fwi_lbfgs_default_param(lbfgs_parameter_t *p)
{
    /* lbfgs.h:77-79 */
    p->m = 12;
    p->epsilon = 1e-9f;
    p->max_iterations = 8;
    p->max_linesearch = 6;
    p->ftol = 1e-7f;
    p->wolfe = 0.98f;
}

static double dotd(const double *a, const double *b, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int fwi_lbfgs(int n, float *x, fwi_lbfgs_evaluate evaluate, fwi_lbfgs_progress progress, void *user,
              const lbfgs_parameter_t *param, int *status)
{
    if (n < 1 || !x || !evaluate || !param || param->m < 1 || param->max_linesearch < 1)
        return lb_fail(FWI_HIP_EINVAL, "fwi_lbfgs: bad arguments");
    const int m = param->m;
    const double dec = 0.5, inc = 2.1; /* MBotOpt.c:11 */
    const size_t nn = (size_t)n;
    double *buf = (double *)malloc(sizeof(double) * nn * (5 + 2 * (size_t)m));
    float *xf = (float *)malloc(sizeof(float) * nn), *g = (float *)malloc(sizeof(float) * nn);
    double *alpha = (double *)calloc(m, sizeof(double)), *ys_m = (double *)calloc(m, sizeof(double));
    double *step_arr = (double *)calloc((size_t)param->max_iterations + 3, sizeof(double));
    int rc = FWI_HIP_OK, st = LBFGS_SUCCESS;
    if (!buf || !xf || !g || !alpha || !ys_m || !step_arr) {
        rc = lb_fail(FWI_HIP_ENOMEM, "fwi_lbfgs: host memory");
        goto done;
    }
    double *xd = buf, *xp = xd + nn, *gd = xp + nn, *gp = gd + nn, *d = gp + nn;
    double *S = d + nn, *Y = S + nn * m;

    for (size_t i = 0; i < nn; i++) {
        xd[i] = x[i];
        xf[i] = x[i];
    }
    float fx;
    if (evaluate(user, xf, g, &fx, 0)) {
        rc = lb_fail(FWI_HIP_EINVAL, "fwi_lbfgs: evaluation failed at the start");
        goto done;
    }
    for (size_t i = 0; i < nn; i++) {
        gd[i] = g[i];
        d[i] = -gd[i]; /* MBotOpt.c:67 */
    }
    if (progress) progress(user, 0, fx, 0.0, 0, xf, g, NULL, n);
    double step = 0.8; /* MBotOpt.c:71 */
    int k = 1, end = 0;
    for (;;) {
        memcpy(xp, xd, sizeof(double) * nn);
        memcpy(gp, gd, sizeof(double) * nn);
        const double dginit = dotd(gd, d, n), finit = fx, dgtest = param->ftol * dginit;
        int count = 0;
        for (;;) { /* MBotOpt.c:91-159 */
            for (size_t i = 0; i < nn; i++) {
                xd[i] = xp[i] + step * d[i];
                xf[i] = (float)xd[i];
            }
            if (evaluate(user, xf, g, &fx, k)) {
                rc = lb_fail(FWI_HIP_EINVAL, "fwi_lbfgs: evaluation failed at iteration %d", k);
                goto done;
            }
            for (size_t i = 0; i < nn; i++) gd[i] = g[i];
            ++count;
            double width;
            if (fx > finit + step * dgtest) {
                width = dec; /* Armijo */
            } else {
                const double dg = dotd(gd, d, n);
                if (dg < param->wolfe * dginit) width = inc;        /* curvature */
                else if (dg > -param->wolfe * dginit) width = dec;  /* strong Wolfe */
                else break;
            }
            if (param->max_linesearch <= count) {
                st = LBFGSERR_MAXIMUMLINESEARCH;
                goto done;
            }
            step *= width;
        }
        if (progress) progress(user, k, fx, step, count, xf, g, d, n);
        const double xnorm = sqrt(dotd(xd, xd, n)), gnorm = sqrt(dotd(gd, gd, n));
        step_arr[k] = step;
        if (gnorm / xnorm <= param->epsilon) { /* MBotOpt.c:166 */
            st = LBFGS_SUCCESS;
            goto done;
        }
        if (param->max_iterations != 0 && param->max_iterations < k + 1) { /* MBotOpt.c:177 */
            st = LBFGSERR_MAXIMUMITERATION;
            goto done;
        }
        /* MBotOpt.c:188-213: new pair, two-loop recursion */
        double *s = S + nn * end, *y = Y + nn * end;
        for (size_t i = 0; i < nn; i++) {
            s[i] = xd[i] - xp[i];
            y[i] = gd[i] - gp[i];
        }
        const double ys = dotd(y, s, n), yy = dotd(y, y, n);
        ys_m[end] = ys;
        const int bound = m <= k ? m : k;
        ++k;
        end = (end + 1) % m;
        for (size_t i = 0; i < nn; i++) d[i] = -gd[i];
        int j = end;
        for (int i = 0; i < bound; i++) {
            j = (j + m - 1) % m;
            alpha[j] = dotd(S + nn * j, d, n) / ys_m[j];
            const double *yj = Y + nn * j;
            for (size_t l = 0; l < nn; l++) d[l] -= alpha[j] * yj[l];
        }
        const double sc = ys / yy;
        for (size_t l = 0; l < nn; l++) d[l] *= sc;
        for (int i = 0; i < bound; i++) {
            const double beta = dotd(Y + nn * j, d, n) / ys_m[j];
            const double *sj = S + nn * j;
            for (size_t l = 0; l < nn; l++) d[l] += (alpha[j] - beta) * sj[l];
            j = (j + 1) % m;
        }
        step = k < 4 ? 0.2 : 0.5 * (step_arr[k - 1] + step_arr[k - 2]); /* MBotOpt.c:214 */
    }
done:
    if (!rc)
        for (size_t i = 0; i < nn; i++) x[i] = (float)buf[i];
    if (status) *status = st;
    free(buf);
    free(xf);
    free(g);
    free(alpha);
    free(ys_m);
    free(step_arr);
    return rc;
}

/* ------------------------------------------------------------------------ */

typedef struct {
    fwi_plan **plans;
    int nshots, precondition;
    const float *true_model;
    FILE *log;
    double true_norm;
} Inv;

static int inv_evaluate(void *u, const float *x, float *g, float *fx, int k)
{
    Inv *v = (Inv *)u;
    fwi_plan *p0 = v->plans[0];
    const size_t n = p0->grid->base->nx * p0->grid->base->nz;
    memcpy(&p0->model->md[0][0], x, sizeof(float) * n);
    int rc = grad_all_shots(v->plans, v->nshots, k);
    if (rc) {
        lb_fail(rc, "grad_all_shots: %s", grad4_last_error());
        return rc;
    }
    if (v->precondition && (rc = grad_precond_new5(p0->grad, p0->grid->b, p0->model->ibtry))) {
        lb_fail(rc, "grad_precond_new5: %s", fwi_gradprec_last_error());
        return rc;
    }
    double f = 0.0;
    for (int s = 0; s < v->nshots; s++) f += v->plans[s]->msf->val[1];
    *fx = (float)f;
    memcpy(g, p0->grad->body, sizeof(float) * n);
    return 0;
}

static void inv_progress(void *u, int k, double fx, double step, int trials, const float *x, const float *g,
                         const double *d, int n)
{
    (void)d;
    Inv *v = (Inv *)u;
    if (!v->log) return;
    double gn = 0.0;
    for (int i = 0; i < n; i++) gn += (double)g[i] * g[i];
    fprintf(v->log, "{\"k\": %d, \"misfit\": %.9g, \"step\": %.6g, \"trials\": %d, \"gnorm\": %.9g", k, fx, step, trials,
            sqrt(gn));
    if (v->true_model) {
        double e = 0.0;
        for (int i = 0; i < n; i++) {
            const double r = (double)x[i] - v->true_model[i];
            e += r * r;
        }
        fprintf(v->log, ", \"model_rel_l2\": %.9g", sqrt(e) / v->true_norm);
    }
    fprintf(v->log, "}\n");
    fflush(v->log);
}

int fwi_invert(fwi_plan **plans, int nshots, const lbfgs_parameter_t *param, int precondition,
               const float *true_model, const char *log_path, int *status)
{
    if (!plans || nshots < 1 || !plans[0] || !param) return lb_fail(FWI_HIP_EINVAL, "fwi_invert: bad arguments");
    fwi_plan *p0 = plans[0];
    const int n = (int)(p0->grid->base->nx * p0->grid->base->nz);
    Inv v = {plans, nshots, precondition, true_model, NULL, 0.0};
    if (true_model) {
        for (int i = 0; i < n; i++) v.true_norm += (double)true_model[i] * true_model[i];
        v.true_norm = sqrt(v.true_norm);
        if (v.true_norm == 0.0) v.true_model = NULL;
    }
    if (log_path && !(v.log = fopen(log_path, "a"))) return lb_fail(FWI_HIP_EINVAL, "fwi_invert: cannot open %s", log_path);
    float *x = (float *)malloc(sizeof(float) * n);
    int rc = FWI_HIP_OK;
    if (!x) rc = lb_fail(FWI_HIP_ENOMEM, "fwi_invert: host memory");
    if (!rc) {
        memcpy(x, &p0->model->md[0][0], sizeof(float) * n);
        rc = fwi_lbfgs(n, x, inv_evaluate, inv_progress, &v, param, status);
        if (!rc) memcpy(&p0->model->md[0][0], x, sizeof(float) * n);
    }
    free(x);
    if (v.log) fclose(v.log);
    return rc;
}
