/*
 * fwi_gradprec.c -- gradient post-processing in FP64: triangle smoothing
 * (triangle_smooth.c) and grad_precond_new5 (grad_prec.c).
 *
 * The triangle filter of half-width r is two running sums of length r
 * (a box of r samples convolved with itself: weights 1..r..1, total r^2);
 * here each output is evaluated directly as the weighted sum over the
 * reflected input, with prefix sums: for output i, the box sum of the
 * reflected sequence over [i-r+1, i] convolved again over the next box.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fwi_gradprec.h"
#include "fwi_hip.h"

static char g_gp_msg[512] = "";

static int gp_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_gp_msg, sizeof(g_gp_msg), fmt, ap);
    va_end(ap);
    return code;
}

const char *fwi_gradprec_last_error(void) { return g_gp_msg; }

/* index of position p (may lie outside [0, n)) after the even reflections
 * p -> -1-p and p -> 2n-1-p; one fold suffices for |p| < 2n */
static int fold(int p, int n)
{
    if (p < 0) p = -1 - p;
    if (p >= n) p = 2 * n - 1 - p;
    return p;
}

/* y[i] = sum_{k=-(r-1)}^{r-1} (r - |k|) x[fold(i+k)] / r^2 , work: 3n+4r doubles */
static void tri_line(int r, int n, const double *x, double *y, double *work)
{
    const int pad = r - 1, m = n + 2 * pad;
    double *ext = work, *box = work + m; /* reflected input; box sums of width r */
    for (int j = 0; j < m; j++) ext[j] = x[fold(j - pad, n)];
    /* box[j] = sum ext[j-r+1 .. j], j >= r-1 */
    double s = 0.0;
    for (int j = 0; j < m; j++) {
        s += ext[j];
        if (j >= r) s -= ext[j - r];
        box[j] = s;
    }
    /* triangle at i = sum of box over [i+pad-(r-1)+ (r-1) ...]: output i is
     * centred at ext index i+pad, so it is the sum of box[i+pad .. i+pad+r-1] */
    const double inv = 1.0 / ((double)r * r);
    double t = 0.0;
    for (int j = pad; j < pad + r && j < m; j++) t += box[j];
    for (int i = 0; i < n; i++) {
        y[i] = t * inv;
        const int a = i + pad, b = i + pad + r; /* slide the window [a, b) by one */
        if (b < m) t += box[b];
        t -= box[a];
    }
}

int triangle(int nr, int n12, float *uu, float *vv)
{
    if (!uu || !vv || n12 < 1 || nr < 1 || nr > n12 + 1)
        return gp_fail(FWI_HIP_EINVAL, "triangle: need 1 <= nr <= n12 + 1 (nr=%d n12=%d)", nr, n12);
    double *w = (double *)malloc(sizeof(double) * (2 * (size_t)n12 + 2 * (n12 + 2 * (size_t)nr)));
    if (!w) return gp_fail(FWI_HIP_ENOMEM, "triangle: host memory");
    double *x = w, *y = w + n12, *work = w + 2 * n12;
    for (int i = 0; i < n12; i++) x[i] = uu[i];
    tri_line(nr, n12, x, y, work);
    for (int i = 0; i < n12; i++) vv[i] = (float)y[i];
    free(w);
    return FWI_HIP_OK;
}

/* smoothing of a double field f[n2][n1], nt passes per axis */
static int tri2_d(int rect1, int rect2, int nt, int n1, int n2, double *f)
{
    if (nt > 0 && (rect1 < 1 || rect2 < 1 || rect1 > n1 + 1 || rect2 > n2 + 1))
        return gp_fail(FWI_HIP_EINVAL, "triangle2: need 1 <= rect <= n + 1 (rect1=%d n1=%d rect2=%d n2=%d)", rect1,
                       n1, rect2, n2);
    const int nm = n1 > n2 ? n1 : n2, rm = rect1 > rect2 ? rect1 : rect2;
    double *w = (double *)malloc(sizeof(double) * (2 * (size_t)nm + 2 * (nm + 2 * (size_t)rm)));
    if (!w) return gp_fail(FWI_HIP_ENOMEM, "triangle2: host memory");
    double *x = w, *y = w + nm, *work = w + 2 * nm;
    for (int it = 0; it < nt; it++)
        for (int i2 = 0; i2 < n2; i2++) {
            double *row = f + (size_t)i2 * n1;
            tri_line(rect1, n1, row, y, work);
            memcpy(row, y, sizeof(double) * n1);
        }
    for (int it = 0; it < nt; it++)
        for (int i1 = 0; i1 < n1; i1++) {
            for (int i2 = 0; i2 < n2; i2++) x[i2] = f[(size_t)i2 * n1 + i1];
            tri_line(rect2, n2, x, y, work);
            for (int i2 = 0; i2 < n2; i2++) f[(size_t)i2 * n1 + i1] = y[i2];
        }
    free(w);
    return FWI_HIP_OK;
}

int triangle2(int rect1, int rect2, int nt, int n1, int n2, float *uu, float *vv)
{
    if (!uu || !vv || n1 < 1 || n2 < 1 || nt < 0) return gp_fail(FWI_HIP_EINVAL, "triangle2: bad arguments");
    const size_t n = (size_t)n1 * n2;
    double *f = (double *)malloc(sizeof(double) * n);
    if (!f) return gp_fail(FWI_HIP_ENOMEM, "triangle2: host memory");
    for (size_t i = 0; i < n; i++) f[i] = uu[i];
    int rc = tri2_d(rect1, rect2, nt, n1, n2, f);
    if (!rc)
        for (size_t i = 0; i < n; i++) vv[i] = (float)f[i];
    free(f);
    return rc;
}

int grad_precond_new5(fwi_grad *g, fwi_shape *bw, int *bottom)
{
    if (!g || !g->body || !bw || !bottom || bw->mnx < 1 || bw->mnz < 1)
        return gp_fail(FWI_HIP_EINVAL, "grad_precond_new5: bad arguments");
    const int nx = (int)bw->mnx, nz = (int)bw->mnz;
    const size_t n = (size_t)nx * nz;
    double *f = (double *)malloc(sizeof(double) * n);
    if (!f) return gp_fail(FWI_HIP_ENOMEM, "grad_precond_new5: host memory");
    double vmin = g->body[0], vmax = g->body[0];
    for (size_t i = 0; i < n; i++) {
        f[i] = g->body[i];
        if (f[i] < vmin) vmin = f[i];
        if (f[i] > vmax) vmax = f[i];
    }
    const double pc = g->perc;
    if (pc != 0.0) /* grad_prec.c:24-35 */
        for (size_t i = 0; i < n; i++) {
            if (f[i] > vmax * pc) f[i] = vmax * pc;
            if (f[i] < vmin * pc) f[i] = vmin * pc;
        }
    int rc = tri2_d((int)g->rectx1, (int)g->rectx1, (int)g->repeat1, nx, nz, f); /* grad_prec.c:37 */
    if (!rc) {
        for (int iz = 0; iz < nz; iz++) /* grad_prec.c:39-42 */
            for (int ix = 0; ix < nx; ix++)
                if (iz < bottom[ix]) f[(size_t)iz * nx + ix] = 0.0;
        rc = tri2_d((int)g->rectx2, (int)g->rectx2, (int)g->repeat2, nx, nz, f); /* grad_prec.c:44 */
    }
    if (!rc)
        for (size_t i = 0; i < n; i++) g->body[i] = (float)(f[i] * g->alpha_grad); /* grad_prec.c:46-48 */
    free(f);
    return rc;
}
