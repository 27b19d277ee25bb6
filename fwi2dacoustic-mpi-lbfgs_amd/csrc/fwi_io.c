/*
 * fwi_io.c -- the input side of the reference's fwi_init and its data-rate
 * bridge, for C host code on top of the fwi_hip ABI:
 *
 *   fwi_read_f32 / fwi_write_f32   read_from_file / write_to_file
 *                                  (read_write.c:4-56), returning an error
 *                                  code where the reference calls exit()
 *   fwi_plan_read_inputs           the file reads and index arithmetic of
 *                                  fwi_init (fwi_2d.c:320-371) with the
 *                                  file names of fwi_files_path.h:7-18
 *   rwi_plan_read_inputs           the same for rwi_init (fwi_2d.c:637-688),
 *                                  with the two models m0, m1
 *   fwi_trcinp                     trcinp (spline_interp.c:19-38): natural
 *                                  cubic spline from nt0 samples at dt0 to
 *                                  nt samples at dt, end samples copied
 *
 * The spline is written from the definition of the natural cubic spline
 * (second derivative zero at both ends; tridiagonal system for the knot
 * curvatures, solved by elimination), evaluated in double on the interval
 * [x_i, x_{i+1}) holding the abscissa, as the reference's GSL cspline call
 * does.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fwi_hip.h"
#include "fwi_plan.h"

static char g_io_msg[512] = "";

static int io_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_io_msg, sizeof(g_io_msg), fmt, ap);
    va_end(ap);
    return code;
}

const char *fwi_io_last_error(void) { return g_io_msg; }

int fwi_read_f32(const char *path, float *dst, size_t n, int exact)
{
    if (!path || (!dst && n)) return io_fail(FWI_HIP_EINVAL, "read: NULL argument");
    FILE *fp = fopen(path, "rb");
    if (!fp) return io_fail(FWI_HIP_EINVAL, "error opening %s", path);
    const size_t got = fread(dst, sizeof(float), n, fp);
    if (got != n) {
        const int err = ferror(fp);
        fclose(fp);
        return io_fail(FWI_HIP_EINVAL, "%s reading %s: %zu of %zu floats", err ? "error" : "end of file", path, got,
                       n);
    }
    if (exact && fgetc(fp) != EOF) { /* read_write.c:23-31: nothing may be left */
        fclose(fp);
        return io_fail(FWI_HIP_EINVAL, "elements left in %s after %zu floats", path, n);
    }
    fclose(fp);
    return FWI_HIP_OK;
}

int fwi_write_f32(const char *path, const float *src, size_t n)
{
    if (!path || (!src && n)) return io_fail(FWI_HIP_EINVAL, "write: NULL argument");
    FILE *fp = fopen(path, "wb");
    if (!fp) return io_fail(FWI_HIP_EINVAL, "error opening %s for writing", path);
    const size_t put = fwrite(src, sizeof(float), n, fp);
    const int cl = fclose(fp);
    if (put != n || cl != 0) return io_fail(FWI_HIP_EINVAL, "error writing %s: %zu of %zu floats", path, put, n);
    return FWI_HIP_OK;
}

/* The reads and index arithmetic fwi_init (fwi_2d.c:320-371) and rwi_init
 * (fwi_2d.c:637-688) share: data, geometry, first breaks, sea bottom and
 * wavelet, then the bottom/first-break indices and the source and receiver
 * cells. */
static int read_shot_inputs(const fwi_base *b, fwi_msf *m, fwi_sou *so, float *btry, int *ibtry, const char *common,
                            int ishot, const char *wavelet, int shots_avail, int fbt_avail, float wav_scale)
{
    const size_t nx = b->nx, nr = m->nr, nt0 = m->nt, dim = b->dim;
    if (dim != 2) return io_fail(FWI_HIP_EINVAL, "read_inputs: dim=%zu (2-D only)", dim);
    if (!btry || !ibtry || !m->fbt || !m->ifbt)
        return io_fail(FWI_HIP_EINVAL, "read_inputs: plan lacks btry/ibtry/fbt/ifbt (use fwi_plan_alloc)");
    char path[600];
    int rc;
    /* fwi_files_path.h:7-18 */
    if (shots_avail) {
        snprintf(path, sizeof(path), "%sinput/shot2ms_%d.bin", common, ishot);
        if ((rc = fwi_read_f32(path, &m->obs0[0][0], nr * nt0, 1))) return rc;
    }
    snprintf(path, sizeof(path), "%sinput/sr_location_%d.bin", common, ishot);
    if ((rc = fwi_read_f32(path, so->xyz_sr, dim * (nr + 1), 1))) return rc;
    memset(m->fbt, 0, sizeof(float) * nr);
    if (fbt_avail) {
        snprintf(path, sizeof(path), "%sinput/fbt_%d.bin", common, ishot);
        if ((rc = fwi_read_f32(path, m->fbt, nr, 1))) return rc;
    }
    snprintf(path, sizeof(path), "%sinput/bottom.bin", common);
    if ((rc = fwi_read_f32(path, btry, nx, 1))) return rc;
    if (wavelet)
        snprintf(path, sizeof(path), "%s", wavelet);
    else
        snprintf(path, sizeof(path), "%sinput/wavelet_1ms.bin", common);
    if ((rc = fwi_read_f32(path, so->wav0, b->nt, 1))) return rc;
    for (size_t i = 0; i < b->nt; i++) so->wav0[i] *= wav_scale; /* fwi_2d.c:353 */

    /* fwi_2d.c:357-371: float quotients, C conversions and round() */
    for (size_t i = 0; i < nx; i++) ibtry[i] = btry[i] / b->dz - 0.5;
    for (size_t i = 0; i < nr; i++) m->ifbt[i] = m->fbt[i] / m->dt;
    so->ixyz_sr[0] = (int)round(so->xyz_sr[0] / b->dx);
    so->ixyz_sr[1] = (int)round(so->xyz_sr[1] / b->dz - 0.5);
    for (size_t i = 0; i < nr; i++) {
        so->ixyz_sr[dim + 2 * i] = (int)round(so->xyz_sr[dim + 2 * i] / b->dx);
        so->ixyz_sr[dim + 2 * i + 1] = (int)round(so->xyz_sr[dim + 2 * i + 1] / b->dz - 0.5);
    }
    return FWI_HIP_OK;
}

int fwi_plan_read_inputs(fwi_plan *p, const char *common, int ishot, const char *wavelet, int shots_avail,
                         int fbt_avail, float wav_scale)
{
    if (!p || !common || !p->grid || !p->model || !p->msf || !p->sou)
        return io_fail(FWI_HIP_EINVAL, "read_inputs: NULL argument");
    const fwi_base *b = p->grid->base;
    fwi_model *md = p->model;
    if (!md->md) return io_fail(FWI_HIP_EINVAL, "read_inputs: plan lacks the model");
    char path[600];
    snprintf(path, sizeof(path), "%sinput/initial_model.bin", common);
    int rc = fwi_read_f32(path, &md->md[0][0], b->nz * b->nx, 1);
    if (rc) return rc;
    return read_shot_inputs(b, p->msf, p->sou, md->btry, md->ibtry, common, ishot, wavelet, shots_avail, fbt_avail,
                            wav_scale);
}

int rwi_plan_read_inputs(rwi_plan *p, const char *common, int ishot, const char *init_mdl_0, const char *init_mdl_1,
                         const char *wavelet, int shots_avail, int fbt_avail, float wav_scale)
{
    if (!p || !common || !p->grid || !p->model || !p->msf || !p->sou)
        return io_fail(FWI_HIP_EINVAL, "rwi_read_inputs: NULL argument");
    if (!init_mdl_0 || !init_mdl_1)
        return io_fail(FWI_HIP_EINVAL, "rwi_read_inputs: the paths of the background (m0) and perturbation (m1) "
                                       "models are required");
    const fwi_base *b = p->grid->base;
    rwi_model *md = p->model;
    if (!md->m0 || !md->m1) return io_fail(FWI_HIP_EINVAL, "rwi_read_inputs: plan lacks m0/m1 (use rwi_plan_alloc)");
    /* fwi_2d.c:667-668: m0, then m1, each [nz][nx] */
    int rc = fwi_read_f32(init_mdl_0, &md->m0[0][0], b->nz * b->nx, 1);
    if (!rc) rc = fwi_read_f32(init_mdl_1, &md->m1[0][0], b->nz * b->nx, 1);
    if (rc) return rc;
    return read_shot_inputs(b, p->msf, p->sou, md->btry, md->ibtry, common, ishot, wavelet, shots_avail, fbt_avail,
                            wav_scale);
}

/* natural cubic spline through (x[i], y[i]), i < n: curvature c[i] (c = y''/2) */
static void spline_natural(const double *x, const double *y, int n, double *c, double *work)
{
    c[0] = c[n - 1] = 0.0;
    if (n < 3) return;
    /* rows i = 1 .. n-2: h[i-1] c[i-1] + 2 (h[i-1] + h[i]) c[i] + h[i] c[i+1]
     *                    = 3 ((y[i+1]-y[i])/h[i] - (y[i]-y[i-1])/h[i-1])          */
    const int m = n - 2;
    double *diag = work, *rhs = work + m;
    for (int k = 0; k < m; k++) {
        const int i = k + 1;
        const double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
        diag[k] = 2.0 * (h0 + h1);
        rhs[k] = 3.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
    }
    for (int k = 1; k < m; k++) { /* elimination; off-diagonal of row k is h = x[k+1]-x[k] */
        const double off = x[k + 1] - x[k];
        const double f = off / diag[k - 1];
        diag[k] -= f * off;
        rhs[k] -= f * rhs[k - 1];
    }
    c[m] = rhs[m - 1] / diag[m - 1];
    for (int k = m - 2; k >= 0; k--) c[k + 1] = (rhs[k] - (x[k + 2] - x[k + 1]) * c[k + 2]) / diag[k];
}

static double spline_eval(const double *x, const double *y, const double *c, int n, double xv, int *lo)
{
    int i = *lo; /* interval with x[i] <= xv < x[i+1], clamped to [0, n-2] */
    if (i < 0 || i > n - 2 || xv < x[i] || xv >= x[i + 1]) {
        int a = 0, bnd = n - 1;
        while (bnd - a > 1) {
            const int mid = (a + bnd) / 2;
            if (x[mid] > xv) bnd = mid;
            else a = mid;
        }
        i = a;
    }
    *lo = i;
    const double h = x[i + 1] - x[i], dx = xv - x[i];
    const double bcoef = (y[i + 1] - y[i]) / h - h * (c[i + 1] + 2.0 * c[i]) / 3.0;
    const double dcoef = (c[i + 1] - c[i]) / (3.0 * h);
    return y[i] + dx * (bcoef + dx * (c[i] + dx * dcoef));
}

int fwi_trcinp(float **obs0, int nt0, float dt0, float **obs, int nt, float dt, int howmany)
{
    if (!obs0 || !obs || nt0 < 2 || nt < 2 || howmany < 0 || !(dt0 > 0) || !(dt > 0))
        return io_fail(FWI_HIP_EINVAL, "trcinp: bad arguments");
    double *x0 = (double *)malloc(sizeof(double) * nt0), *y0 = (double *)malloc(sizeof(double) * nt0);
    double *c = (double *)malloc(sizeof(double) * nt0), *work = (double *)malloc(sizeof(double) * 2 * nt0);
    double *x1 = (double *)malloc(sizeof(double) * nt);
    if (!x0 || !y0 || !c || !work || !x1) {
        free(x0); free(y0); free(c); free(work); free(x1);
        return io_fail(FWI_HIP_ENOMEM, "trcinp: out of host memory");
    }
    /* spline_interp.c:21-25: abscissae are float products stored as double */
    for (int i = 0; i < nt; i++) x1[i] = (float)(i * dt);
    for (int i = 0; i < nt0; i++) x0[i] = (float)(i * dt0);
    for (int r = 0; r < howmany; r++) {
        for (int i = 0; i < nt0; i++) y0[i] = obs0[r][i];
        spline_natural(x0, y0, nt0, c, work);
        int lo = 0;
        for (int i = 1; i < nt - 1; i++) obs[r][i] = (float)spline_eval(x0, y0, c, nt0, x1[i], &lo);
        obs[r][0] = (float)y0[0]; /* spline_interp.c:13: end samples copied */
        obs[r][nt - 1] = (float)y0[nt0 - 1];
    }
    free(x0); free(y0); free(c); free(work); free(x1);
    return FWI_HIP_OK;
}
