/*
 * fwi_msf.c -- the host misfit pool in FP64: L2, normalised cross-correlation,
 * XCORR1D (xcorr_msf.c), MF1D (lsq_decon_cg.c) and entropic OT (entropy_ot.h
 * declarations), on FP32 traces.
 *
 * The reference computes its correlations with FFTW in FP32 on FFT lengths
 * s0 >= 2 nt - 1; such a circular correlation equals the linear one, so
 * here every correlation is a linear convolution evaluated by an FP64
 * radix-2 FFT of the next power of two (Conv below), and the padding
 * parameters only have to keep the reference's lengths long enough.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fwi_hip.h"
#include "fwi_msf.h"

static char g_msf_msg[512] = "";

static int msf_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msf_msg, sizeof(g_msf_msg), fmt, ap);
    va_end(ap);
    return code;
}

const char *fwi_msf_last_error(void) { return g_msf_msg; }

/* ------------------------------------------------------------------------ */
/* FP64 radix-2 FFT                                                          */

typedef struct {
    int n;
    double *wr, *wi; /* exp(-2 pi i k / n), k < n/2 */
    int *rev;        /* bit reversal of 0..n-1 */
} Fft;

static void fft_free(Fft *f)
{
    free(f->wr);
    free(f->wi);
    free(f->rev);
    memset(f, 0, sizeof(*f));
}

static int fft_init(Fft *f, int n)
{
    memset(f, 0, sizeof(*f));
    int lg = 0;
    while ((1 << lg) < n) lg++;
    if ((1 << lg) != n) return FWI_HIP_EINVAL;
    f->n = n;
    f->wr = (double *)malloc(sizeof(double) * (n / 2 + 1));
    f->wi = (double *)malloc(sizeof(double) * (n / 2 + 1));
    f->rev = (int *)malloc(sizeof(int) * n);
    if (!f->wr || !f->wi || !f->rev) {
        fft_free(f);
        return FWI_HIP_ENOMEM;
    }
    for (int k = 0; k < n / 2; k++) {
        const double a = -2.0 * M_PI * (double)k / (double)n;
        f->wr[k] = cos(a);
        f->wi[k] = sin(a);
    }
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int b = 0; b < lg; b++)
            if (i & (1 << b)) r |= 1 << (lg - 1 - b);
        f->rev[i] = r;
    }
    return FWI_HIP_OK;
}

/* in-place transform; inv != 0: conjugate twiddles, no 1/n scaling */
static void fft_run(const Fft *f, double *re, double *im, int inv)
{
    const int n = f->n;
    for (int i = 0; i < n; i++) {
        const int j = f->rev[i];
        if (j > i) {
            double t = re[i]; re[i] = re[j]; re[j] = t;
            t = im[i]; im[i] = im[j]; im[j] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len)
            for (int k = 0; k < half; k++) {
                const double wr = f->wr[k * step], wi = inv ? -f->wi[k * step] : f->wi[k * step];
                const int a = s + k, b = a + half;
                const double xr = re[b] * wr - im[b] * wi, xi = re[b] * wi + im[b] * wr;
                re[b] = re[a] - xr;
                im[b] = im[a] - xi;
                re[a] += xr;
                im[a] += xi;
            }
    }
}

/* ------------------------------------------------------------------------ */
/* linear convolution with a fixed operand a[0..na)                          */

typedef struct {
    Fft f;
    int na;
    double *ar, *ai; /* spectrum of a */
    double *xr, *xi; /* work */
} Conv;

static void conv_free(Conv *c)
{
    fft_free(&c->f);
    free(c->ar);
    free(c->ai);
    free(c->xr);
    free(c->xi);
    memset(c, 0, sizeof(*c));
}

/* room for full convolutions of length up to len */
static int conv_init(Conv *c, int len)
{
    memset(c, 0, sizeof(*c));
    int n = 1;
    while (n < len) n <<= 1;
    int rc = fft_init(&c->f, n);
    if (rc) return rc;
    c->ar = (double *)malloc(sizeof(double) * n);
    c->ai = (double *)malloc(sizeof(double) * n);
    c->xr = (double *)malloc(sizeof(double) * n);
    c->xi = (double *)malloc(sizeof(double) * n);
    if (!c->ar || !c->ai || !c->xr || !c->xi) {
        conv_free(c);
        return FWI_HIP_ENOMEM;
    }
    return FWI_HIP_OK;
}

static void conv_set(Conv *c, const double *a, int na)
{
    const int n = c->f.n;
    c->na = na;
    for (int i = 0; i < n; i++) {
        c->ar[i] = i < na ? a[i] : 0.0;
        c->ai[i] = 0.0;
    }
    fft_run(&c->f, c->ar, c->ai, 0);
}

/* out[i] = (a * x)[off + i], i < nout, with (a * x)[k] = sum_j a[k-j] x[j];
 * needs na + nx - 1 <= n (no wrap-around) */
static void conv_apply(Conv *c, const double *x, int nx, int off, int nout, double *out)
{
    const int n = c->f.n;
    for (int i = 0; i < n; i++) {
        c->xr[i] = i < nx ? x[i] : 0.0;
        c->xi[i] = 0.0;
    }
    fft_run(&c->f, c->xr, c->xi, 0);
    for (int i = 0; i < n; i++) {
        const double r = c->ar[i] * c->xr[i] - c->ai[i] * c->xi[i];
        const double m = c->ar[i] * c->xi[i] + c->ai[i] * c->xr[i];
        c->xr[i] = r;
        c->xi[i] = m;
    }
    fft_run(&c->f, c->xr, c->xi, 1);
    const double s = 1.0 / n;
    for (int i = 0; i < nout; i++) {
        const int k = off + i;
        out[i] = (k >= 0 && k < n) ? c->xr[k] * s : 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* shared helpers                                                            */

static int keep_trace(const float *sr, float mn, float mx, int r)
{
    if (!sr) return 1;
    const double dx = (double)sr[2 + 2 * r] - sr[0], dz = (double)sr[3 + 2 * r] - sr[1];
    const double d = sqrt(dx * dx + dz * dz);
    return !(d < mn || d > mx);
}

static int rows_ok(float **a, int nr)
{
    if (!a) return 0;
    for (int r = 0; r < nr; r++)
        if (!a[r]) return 0;
    return 1;
}

static double dot(const double *a, const double *b, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* the reference's mode test: strncmp over n characters */
static int mode_is(const char *s, const char *m, size_t n) { return s && strncmp(s, m, n) == 0; }

/* ------------------------------------------------------------------------ */
/* L2 and NCC                                                                */

int L2norm(float **obs, float **syn, int nt, int nr, float *sr_xyz, float min, float max, float val[], float **res)
{
    if (nt < 0 || nr < 0 || !rows_ok(obs, nr) || !rows_ok(syn, nr) || !rows_ok(res, nr) || !val)
        return msf_fail(FWI_HIP_EINVAL, "L2norm: bad arguments");
    double J = 0.0;
    for (int r = 0; r < nr; r++) {
        const int keep = keep_trace(sr_xyz, min, max, r);
        for (int i = 0; i < nt; i++) {
            const double d = keep ? (double)syn[r][i] - obs[r][i] : 0.0;
            res[r][i] = (float)d;
            J += 0.5 * d * d;
        }
    }
    val[0] = val[1] = (float)J;
    return FWI_HIP_OK;
}

int NCCnorm(float **obs, float **syn, int nt, int nr, float *sr_xyz, float min, float max, float val[],
            float **res)
{
    if (nt < 0 || nr < 0 || !rows_ok(obs, nr) || !rows_ok(syn, nr) || !rows_ok(res, nr) || !val)
        return msf_fail(FWI_HIP_EINVAL, "NCCnorm: bad arguments");
    /* traces with less than NCC_FLOOR of the gather's largest trace energy
     * (observed or synthetic) hold no usable waveform, only rounding noise
     * that the normalisation would amplify by 1/|s| */
    const double NCC_FLOOR = 1e-8;
    double smax = 0.0, omax = 0.0;
    for (int r = 0; r < nr; r++) {
        double ss = 0.0, oo = 0.0;
        for (int i = 0; i < nt; i++) {
            ss += (double)syn[r][i] * syn[r][i];
            oo += (double)obs[r][i] * obs[r][i];
        }
        if (ss > smax) smax = ss;
        if (oo > omax) omax = oo;
    }
    double J = 0.0, L2 = 0.0;
    for (int r = 0; r < nr; r++) {
        double ss = 0.0, oo = 0.0, so = 0.0;
        for (int i = 0; i < nt; i++) {
            const double s = syn[r][i], o = obs[r][i];
            ss += s * s;
            oo += o * o;
            so += s * o;
            L2 += (s - o) * (s - o);
        }
        if (!keep_trace(sr_xyz, min, max, r) || ss == 0.0 || oo == 0.0 || ss < NCC_FLOOR * smax ||
            oo < NCC_FLOOR * omax) {
            for (int i = 0; i < nt; i++) res[r][i] = 0.0f;
            continue;
        }
        const double ns = sqrt(ss), no = sqrt(oo), c = so / (ns * no);
        J += 1.0 - c;
        for (int i = 0; i < nt; i++) res[r][i] = (float)(-(obs[r][i] / (ns * no) - c * syn[r][i] / ss));
    }
    val[0] = (float)L2;
    val[1] = (float)J;
    return FWI_HIP_OK;
}

/* ------------------------------------------------------------------------ */
/* XCORR1D                                                                   */

int XCORR1D(float **obs, float **syn, float *sr_xyz, float min, float max, void *mpars, float *misfit,
            float **res)
{
    const par_xcorr1d *p = (const par_xcorr1d *)mpars;
    if (!p || !misfit) return msf_fail(FWI_HIP_EINVAL, "XCORR1D: NULL parameters");
    const int nt = p->nt, nr = p->nr;
    if (nt < 1 || nr < 0 || p->npad4Ab < 0 || !rows_ok(obs, nr) || !rows_ok(syn, nr) || !rows_ok(res, nr))
        return msf_fail(FWI_HIP_EINVAL, "XCORR1D: bad arguments (nt=%d nr=%d npad4Ab=%d)", nt, nr, p->npad4Ab);
    const int middle = mode_is(p->conv_how_to_slice, "middle", 6);
    if (!middle && !mode_is(p->conv_how_to_slice, "leading", 6)) /* xcorr_msf.c:35 */
        return msf_fail(FWI_HIP_EINVAL, "XCORR1D: conv_how_to_slice must be \"middle\" or \"leading\"");
    const int off = middle ? nt / 2 : 0; /* slice of the full length-(2nt-1) correlation */

    Conv cv;
    int rc = conv_init(&cv, 2 * nt - 1);
    if (rc) return msf_fail(rc, "XCORR1D: host memory");
    double *T = (double *)malloc(sizeof(double) * nt), *a = (double *)malloc(sizeof(double) * nt);
    double *s = (double *)malloc(sizeof(double) * nt), *x = (double *)malloc(sizeof(double) * nt);
    double *y = (double *)malloc(sizeof(double) * nt);
    if (!T || !a || !s || !x || !y) {
        rc = msf_fail(FWI_HIP_ENOMEM, "XCORR1D: host memory");
        goto done;
    }
    for (int i = 0; i < nt; i++) T[i] = middle ? fabs((double)(i - nt / 2)) : (double)i;
    double m0 = 0.0, m1 = 0.0;
    for (int r = 0; r < nr; r++) {
        double e1 = 0.0, e2 = 0.0;
        for (int i = 0; i < nt; i++) {
            res[r][i] = 0.0f;
            e1 += fabs((double)obs[r][i]);
            e2 += fabs((double)syn[r][i]);
        }
        if (e2 / (e1 + 1e-12) > 10.0 || e1 / (e2 + 1e-12) > 10.0 || !keep_trace(sr_xyz, min, max, r)) continue;
        /* x = (reversed obs) * syn, sliced: the correlation of syn with obs */
        for (int i = 0; i < nt; i++) {
            a[i] = obs[r][nt - 1 - i];
            s[i] = syn[r][i];
            m0 += fabs(a[i] - s[i]);
        }
        conv_set(&cv, a, nt);
        conv_apply(&cv, s, nt, off, nt, x);
        double num = 0.0, den = 0.0;
        for (int i = 0; i < nt; i++) {
            num += T[i] * x[i] * T[i] * x[i];
            den += x[i] * x[i];
        }
        if (den == 0.0) continue;
        const double alpha = num / den;
        m1 += alpha;
        for (int i = 0; i < nt; i++) y[i] = 2.0 * (T[i] * T[i] * x[i] - alpha * x[i]) / den;
        /* res = obs * y, sliced the same way */
        for (int i = 0; i < nt; i++) a[i] = obs[r][i];
        conv_set(&cv, a, nt);
        conv_apply(&cv, y, nt, off, nt, x);
        for (int i = 0; i < nt; i++) res[r][i] = (float)x[i];
    }
    misfit[0] = (float)m0;
    misfit[1] = (float)m1;
done:
    conv_free(&cv);
    free(T);
    free(a);
    free(s);
    free(x);
    free(y);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* MF1D                                                                      */

typedef struct {
    int nt, nm, middle, diff;
    double eps;
    Conv ca, cat; /* a and reversed a */
    double *tmp;  /* nt + nm */
} Mf;

/* w0[0..nt) = A w: (a * w) sliced at nm/2 (middle) or 0 */
static void mf_A(Mf *m, const double *w, double *w0)
{
    conv_apply(&m->ca, w, m->nm, m->middle ? m->nm / 2 : 0, m->nt, w0);
}

/* z[0..nm) = A^T r: (reversed a * r) sliced at nt - nm/2 - 1 (middle) or nt - 1 */
static void mf_At(Mf *m, const double *r, double *z)
{
    conv_apply(&m->cat, r, m->nt, m->middle ? m->nt - m->nm / 2 - 1 : m->nt - 1, m->nm, z);
}

/* r1 = D w */
static void mf_D(const Mf *m, const double *w, double *r1)
{
    const int n = m->nm;
    if (m->diff) {
        for (int i = 0; i + 1 < n; i++) r1[i] = m->eps * (w[i + 1] - w[i]);
        r1[n - 1] = m->eps * w[n - 1];
    } else {
        for (int i = 0; i < n; i++) r1[i] = m->eps * w[i];
    }
}

/* z += D^T r1 */
static void mf_Dt_add(const Mf *m, const double *r1, double *z)
{
    const int n = m->nm;
    if (m->diff) {
        if (n == 1) {
            z[0] += m->eps * r1[0];
            return;
        }
        z[0] -= m->eps * r1[0];
        for (int i = 1; i + 1 < n; i++) z[i] += m->eps * (r1[i - 1] - r1[i]);
        z[n - 1] += m->eps * (r1[n - 1] + r1[n - 2]);
    } else {
        for (int i = 0; i < n; i++) z[i] += m->eps * r1[i];
    }
}

int MF1D(float **obs, float **syn, void *mpars, float *misfit, float **res)
{
    const epsdf1 *p = (const epsdf1 *)mpars;
    if (!p || !misfit) return msf_fail(FWI_HIP_EINVAL, "MF1D: NULL parameters");
    const int nt = p->nt, nr = p->nr, nm = p->nmodel, niter = p->niter;
    if (nt < 1 || nr < 0 || nm < 1 || p->npad4Ax < 0 || p->npad4Atr < 0 || !rows_ok(obs, nr) || !rows_ok(syn, nr) ||
        !rows_ok(res, nr))
        return msf_fail(FWI_HIP_EINVAL, "MF1D: bad arguments (nt=%d nr=%d nmodel=%d)", nt, nr, nm);
    Mf m;
    memset(&m, 0, sizeof(m));
    m.nt = nt;
    m.nm = nm;
    m.middle = mode_is(p->conv_how_to_slice, "middle", 6);
    if (!m.middle && !mode_is(p->conv_how_to_slice, "leading", 6)) /* lsq_decon_cg.c:141 */
        return msf_fail(FWI_HIP_EINVAL, "MF1D: conv_how_to_slice must be \"middle\" or \"leading\"");
    m.diff = mode_is(p->reg_type, "diff", 3);
    if (!m.diff && !mode_is(p->reg_type, "ide", 3)) /* lsq_decon_cg.c:143 */
        return msf_fail(FWI_HIP_EINVAL, "MF1D: reg_type must be \"diff\" or \"ide\"");
    if (m.middle && nm % 2 == 0)
        return msf_fail(FWI_HIP_EINVAL, "MF1D: \"middle\" needs an odd nmodel (got %d)", nm);
    if (m.middle && nm / 2 > nt - 1)
        return msf_fail(FWI_HIP_EINVAL, "MF1D: \"middle\" needs nmodel/2 < nt (nmodel=%d nt=%d)", nm, nt);
    m.eps = p->eps0;
    int rc = conv_init(&m.ca, nt + nm - 1);
    if (!rc) rc = conv_init(&m.cat, 2 * nt - 1);
    const int nb = nt > nm ? nt : nm;
    double *buf = (double *)calloc((size_t)nb * 16, sizeof(double));
    if (rc || !buf) {
        conv_free(&m.ca);
        conv_free(&m.cat);
        free(buf);
        return msf_fail(rc ? rc : FWI_HIP_ENOMEM, "MF1D: host memory");
    }
    double *T = buf, *a = T + nb, *ar = a + nb, *b = ar + nb, *w = b + nb, *r0 = w + nb, *r1 = r0 + nb;
    double *z = r1 + nb, *pp = z + nb, *w0 = pp + nb, *w1 = w0 + nb, *c = w1 + nb, *y0 = c + nb, *y1 = y0 + nb;
    double *q0 = y1 + nb, *q1 = q0 + nb;
    for (int i = 0; i < nm; i++) T[i] = p->Tdepressor * (m.middle ? fabs((double)(i - nm / 2)) : (double)i);
    double m0 = 0.0, m1 = 0.0;
    for (int ir = 0; ir < nr; ir++)
        for (int i = 0; i < nt; i++) {
            const double d = (double)obs[ir][i] - syn[ir][i];
            m0 += d * d;
        }
    for (int ir = 0; ir < nr; ir++) {
        for (int i = 0; i < nt; i++) {
            res[ir][i] = 0.0f;
            a[i] = 100.0 * obs[ir][i]; /* lsq_decon_cg.c:183-184 */
            b[i] = 100.0 * syn[ir][i];
        }
        for (int i = 0; i < nt; i++) ar[i] = a[nt - 1 - i];
        conv_set(&m.ca, a, nt);
        conv_set(&m.cat, ar, nt);

        /* CGLS for w: min |A w - b|^2 + |D w|^2, from w = 0 (lsq_decon_cg.c:187-274) */
        for (int i = 0; i < nm; i++) w[i] = 0.0;
        for (int i = 0; i < nt; i++) r0[i] = b[i];
        for (int i = 0; i < nm; i++) r1[i] = 0.0;
        mf_At(&m, r0, z);
        mf_Dt_add(&m, r1, z);
        for (int i = 0; i < nm; i++) pp[i] = z[i];
        for (int it = 0; it < niter; it++) {
            const double rn0 = sqrt(dot(r0, r0, nt) + dot(r1, r1, nm));
            mf_A(&m, pp, w0);
            mf_D(&m, pp, w1);
            const double num = dot(z, z, nm), den = dot(w0, w0, nt) + dot(w1, w1, nm);
            if (num == 0.0 || den == 0.0) break; /* converged exactly: the reference would divide 0/0 */
            const double al = num / den;
            for (int i = 0; i < nm; i++) w[i] += al * pp[i];
            for (int i = 0; i < nt; i++) r0[i] -= al * w0[i];
            for (int i = 0; i < nm; i++) r1[i] -= al * w1[i];
            mf_At(&m, r0, z);
            mf_Dt_add(&m, r1, z);
            const double be = dot(z, z, nm) / num;
            for (int i = 0; i < nm; i++) pp[i] = z[i] + be * pp[i];
            const double rn1 = sqrt(dot(r0, r0, nt) + dot(r1, r1, nm));
            if (fabs(rn1 - rn0) < p->tor0) break;
        }

        /* penalty and its derivative with respect to w (lsq_decon_cg.c:280-290) */
        double num = 0.0, den = 0.0;
        for (int i = 0; i < nm; i++) {
            num += T[i] * w[i] * T[i] * w[i];
            den += w[i] * w[i];
        }
        if (den == 0.0) continue;
        const double alpha = num / den;
        m1 += alpha;
        for (int i = 0; i < nm; i++) c[i] = 2.0 * (T[i] * T[i] * w[i] - alpha * w[i]) / den;

        /* CGNE for [y0; y1] with A^T y0 + D^T y1 = c, from y = 0 (lsq_decon_cg.c:297-360) */
        for (int i = 0; i < nt; i++) y0[i] = 0.0;
        for (int i = 0; i < nm; i++) y1[i] = 0.0;
        double *rr = z; /* c - (A^T y0 + D^T y1) = c */
        for (int i = 0; i < nm; i++) rr[i] = c[i];
        mf_A(&m, rr, w0); /* search direction [w0; w1] */
        mf_D(&m, rr, w1);
        for (int it = 0; it < niter; it++) {
            const double rn0 = sqrt(dot(rr, rr, nm));
            const double nn = rn0 * rn0, dd = dot(w0, w0, nt) + dot(w1, w1, nm);
            if (nn == 0.0 || dd == 0.0) break;
            const double al = nn / dd;
            for (int i = 0; i < nt; i++) y0[i] += al * w0[i];
            for (int i = 0; i < nm; i++) y1[i] += al * w1[i];
            mf_At(&m, w0, pp);
            mf_Dt_add(&m, w1, pp);
            for (int i = 0; i < nm; i++) rr[i] -= al * pp[i];
            const double rr2 = dot(rr, rr, nm), be = rr2 / nn;
            mf_A(&m, rr, q0);
            mf_D(&m, rr, q1);
            for (int i = 0; i < nt; i++) w0[i] = be * w0[i] + q0[i];
            for (int i = 0; i < nm; i++) w1[i] = be * w1[i] + q1[i];
            if (fabs(sqrt(rr2) - rn0) < p->tor0) break;
        }
        for (int i = 0; i < nt; i++) res[ir][i] = (float)y0[i];
    }
    misfit[0] = (float)m0;
    misfit[1] = (float)m1;
    conv_free(&m.ca);
    conv_free(&m.cat);
    free(buf);
    return FWI_HIP_OK;
}

/* ------------------------------------------------------------------------ */
/* entropic OT                                                               */

int eot_msf_init(eot_msf_par *par)
{
    if (!par || par->nt < 1) return msf_fail(FWI_HIP_EINVAL, "eot_msf_init: bad parameters");
    par->M = NULL;
    par->u = (float *)calloc(par->nt, sizeof(float));
    par->v = (float *)calloc(par->nt, sizeof(float));
    if (!par->u || !par->v) {
        eot_msf_destroy(par);
        return msf_fail(FWI_HIP_ENOMEM, "eot_msf_init: host memory");
    }
    return FWI_HIP_OK;
}

void eot_msf_destroy(eot_msf_par *par)
{
    if (!par) return;
    free(par->u);
    free(par->v);
    par->u = par->v = NULL;
}

int eot_log_msf_init(eot_msf_par *par, eot_log_msf_par *exd)
{
    if (!par || !exd || par->nt < 1) return msf_fail(FWI_HIP_EINVAL, "eot_log_msf_init: bad parameters");
    exd->M = NULL;
    exd->u = exd->v = NULL;
    exd->logp = exd->logq = exd->delta_u = exd->delta_v = NULL;
    return FWI_HIP_OK;
}

/* Workspace of one Sinkhorn solve on n points. */
typedef struct {
    int n, band;      /* scaling form: kernel half-width, exp(-C/eps) below e^-700 (double underflow) dropped */
    double eps, dt;
    double *k;        /* k[d] = exp(-(d dt)^2 / eps), d <= band */
    double *p, *q;    /* normalised marginals */
    double *f, *g;    /* potentials (log form) or scalings u, v (scaling form) */
    double *t;        /* work */
} Sk;

static void sk_free(Sk *s)
{
    free(s->k);
    free(s->p);
    free(s->q);
    free(s->f);
    free(s->g);
    free(s->t);
    memset(s, 0, sizeof(*s));
}

static int sk_init(Sk *s, int n, double eps, double dt)
{
    memset(s, 0, sizeof(*s));
    s->n = n;
    s->eps = eps;
    s->dt = dt;
    double b = sqrt(700.0 * eps) / dt;
    s->band = b >= n ? n - 1 : (int)b;
    s->k = (double *)malloc(sizeof(double) * (s->band + 1));
    s->p = (double *)malloc(sizeof(double) * n);
    s->q = (double *)malloc(sizeof(double) * n);
    s->f = (double *)malloc(sizeof(double) * n);
    s->g = (double *)malloc(sizeof(double) * n);
    s->t = (double *)malloc(sizeof(double) * n);
    if (!s->k || !s->p || !s->q || !s->f || !s->g || !s->t) {
        sk_free(s);
        return FWI_HIP_ENOMEM;
    }
    for (int d = 0; d <= s->band; d++) {
        const double c = d * dt;
        s->k[d] = exp(-c * c / eps);
    }
    return FWI_HIP_OK;
}

/* y = K x (K symmetric Toeplitz, banded) */
static void sk_kmul(const Sk *s, const double *x, double *y)
{
    const int n = s->n, B = s->band;
    for (int i = 0; i < n; i++) {
        const int lo = i - B < 0 ? 0 : i - B, hi = i + B > n - 1 ? n - 1 : i + B;
        double acc = 0.0;
        for (int j = lo; j <= hi; j++) acc += s->k[i > j ? i - j : j - i] * x[j];
        y[i] = acc;
    }
}

/* y_i = -eps log sum_j exp((x_j - C_ij) / eps), over all j, max-shifted */
static void sk_lse(const Sk *s, const double *x, double *y)
{
    const int n = s->n;
    const double e = s->eps, dt = s->dt;
    for (int i = 0; i < n; i++) {
        double mx = -INFINITY;
        for (int j = 0; j < n; j++) {
            const double c = (i - j) * dt, v = x[j] - c * c;
            if (v > mx) mx = v;
        }
        double acc = 0.0;
        for (int j = 0; j < n; j++) {
            const double c = (i - j) * dt;
            acc += exp((x[j] - c * c - mx) / e);
        }
        y[i] = -(mx + e * log(acc));
    }
}

/* Sinkhorn between s->p and s->q (normalised, positive).  On return s->f,
 * s->g hold the dual potentials and *W the entropic cost; the return value
 * is the iteration count, or -1 when the scaling form underflowed. */
static int sk_solve(Sk *s, int max_iter, double tor, int logform, double *W)
{
    const int n = s->n;
    const double e = s->eps;
    int it = 0;
    double err = INFINITY;
    if (!logform) {
        double *u = s->f, *v = s->g, *kv = s->t;
        for (int i = 0; i < n; i++) v[i] = 1.0;
        for (it = 0; it < max_iter && err >= tor; it++) {
            sk_kmul(s, v, kv);
            for (int i = 0; i < n; i++) u[i] = s->p[i] / kv[i];
            sk_kmul(s, u, kv);
            for (int i = 0; i < n; i++) v[i] = s->q[i] / kv[i];
            sk_kmul(s, v, kv);
            err = 0.0;
            for (int i = 0; i < n; i++) err += fabs(u[i] * kv[i] - s->p[i]);
            if (!isfinite(err)) return -1;
        }
        for (int i = 0; i < n; i++) {
            if (!(u[i] > 0.0) || !(v[i] > 0.0) || !isfinite(u[i]) || !isfinite(v[i])) return -1;
            u[i] = e * log(u[i]);
            v[i] = e * log(v[i]);
        }
    } else {
        double *f = s->f, *g = s->g, *t = s->t;
        for (int i = 0; i < n; i++) g[i] = 0.0;
        for (it = 0; it < max_iter && err >= tor; it++) {
            sk_lse(s, g, t);
            for (int i = 0; i < n; i++) f[i] = t[i] + e * log(s->p[i]);
            sk_lse(s, f, t);
            for (int i = 0; i < n; i++) g[i] = t[i] + e * log(s->q[i]);
            /* row sums of the coupling exp((f_i + g_j - C_ij) / eps): exp((f_i - t_i) / eps) */
            sk_lse(s, g, t);
            err = 0.0;
            for (int i = 0; i < n; i++) err += fabs(exp((f[i] - t[i]) / e) - s->p[i]);
        }
    }
    /* W = <f, p> + <g, q> - eps * (total mass of the coupling, 1 at convergence) */
    double fp = 0.0, gq = 0.0;
    for (int i = 0; i < n; i++) {
        fp += s->f[i] * s->p[i];
        gq += s->g[i] * s->q[i];
    }
    *W = fp + gq - e;
    return it;
}

/* one trace: entropic OT between p_raw and q_raw (non-negative, not both
 * zero mass); grad[i] = dW/dp_raw[i] */
static int eot_trace(const double *p_raw, const double *q_raw, Sk *s, int max_iter, double tor, int logform,
                     double *W, double *grad)
{
    const int n = s->n;
    double sp = 0.0, sq = 0.0;
    for (int i = 0; i < n; i++) {
        if (!(p_raw[i] > 0.0) || !(q_raw[i] > 0.0)) return FWI_HIP_EINVAL;
        sp += p_raw[i];
        sq += q_raw[i];
    }
    for (int i = 0; i < n; i++) {
        s->p[i] = p_raw[i] / sp;
        s->q[i] = q_raw[i] / sq;
    }
    /* the scaling form is the fast path; where exp(-C/eps) underflows it
     * falls back to the log-domain iterations (same objective) */
    if (sk_solve(s, max_iter, tor, logform, W) < 0 && (logform || sk_solve(s, max_iter, tor, 1, W) < 0))
        return FWI_HIP_ERANGE;
    double fp = 0.0;
    for (int i = 0; i < n; i++) fp += s->f[i] * s->p[i];
    for (int i = 0; i < n; i++) grad[i] = (s->f[i] - fp) / sp;
    return FWI_HIP_OK;
}

static int eot_args(const eot_msf_par *par, const char *who)
{
    if (!par || par->nt < 1 || par->max_iter < 0 || !(par->eps > 0.0f) || !(par->dt > 0.0f))
        return msf_fail(FWI_HIP_EINVAL, "%s: bad parameters", who);
    return FWI_HIP_OK;
}

static void store_potentials(eot_msf_par *par, const Sk *s)
{
    for (int i = 0; i < s->n; i++) {
        if (par->u) par->u[i] = (float)s->f[i];
        if (par->v) par->v[i] = (float)s->g[i];
    }
}

static int eot_one(float *p, float *q, eot_msf_par *par, int logform, float *msf, float *res, const char *who)
{
    int rc = eot_args(par, who);
    if (rc) return rc;
    if (!p || !q || !msf || !res) return msf_fail(FWI_HIP_EINVAL, "%s: NULL argument", who);
    const int n = par->nt;
    Sk s;
    if ((rc = sk_init(&s, n, par->eps, par->dt))) return msf_fail(rc, "%s: host memory", who);
    double *pr = (double *)malloc(sizeof(double) * n * 3);
    if (!pr) {
        sk_free(&s);
        return msf_fail(FWI_HIP_ENOMEM, "%s: host memory", who);
    }
    double *qr = pr + n, *gr = qr + n, W = 0.0;
    for (int i = 0; i < n; i++) {
        pr[i] = p[i];
        qr[i] = q[i];
    }
    rc = eot_trace(pr, qr, &s, par->max_iter, par->tor, logform, &W, gr);
    if (rc == FWI_HIP_EINVAL) rc = msf_fail(rc, "%s: p and q must be positive", who);
    else if (rc) rc = msf_fail(rc, "%s: the Sinkhorn iterations left the floating-point range", who);
    else {
        msf[0] = (float)W;
        for (int i = 0; i < n; i++) res[i] = (float)gr[i];
        store_potentials(par, &s);
    }
    free(pr);
    sk_free(&s);
    return rc;
}

int entropy_ot(float *p, float *q, eot_msf_par *par, float *msf, float *res)
{
    return eot_one(p, q, par, 0, msf, res, "entropy_ot");
}

int entropy_ot_log(float *p, float *q, eot_msf_par *par, eot_log_msf_par *par_exd, float *msf, float *res)
{
    (void)par_exd;
    return eot_one(p, q, par, 1, msf, res, "entropy_ot_log");
}

static int eot_2d(float **obs, float **syn, eot_msf_par *par, double bar, int logform, float *msf, float **res,
                  const char *who)
{
    int rc = eot_args(par, who);
    if (rc) return rc;
    const int n = par->nt, nr = par->nr;
    if (nr < 0 || !rows_ok(obs, nr) || !rows_ok(syn, nr) || !rows_ok(res, nr) || !msf)
        return msf_fail(FWI_HIP_EINVAL, "%s: bad arguments", who);
    Sk s;
    if ((rc = sk_init(&s, n, par->eps, par->dt))) return msf_fail(rc, "%s: host memory", who);
    double *pr = (double *)malloc(sizeof(double) * n * 3);
    if (!pr) {
        sk_free(&s);
        return msf_fail(FWI_HIP_ENOMEM, "%s: host memory", who);
    }
    double *qr = pr + n, *gr = qr + n, J = 0.0, L2 = 0.0;
    for (int r = 0; r < nr && !rc; r++) {
        double omax = 0.0;
        for (int i = 0; i < n; i++) {
            const double o = obs[r][i], d = (double)syn[r][i] - o;
            L2 += d * d;
            if (o * o > omax) omax = o * o;
            res[r][i] = 0.0f;
        }
        if (omax == 0.0) continue; /* no observed energy: trace skipped */
        const double eta = bar * omax;
        for (int i = 0; i < n; i++) {
            const double sv = syn[r][i], o = obs[r][i];
            pr[i] = sv * sv + eta;
            qr[i] = o * o + eta;
        }
        double W = 0.0;
        rc = eot_trace(pr, qr, &s, par->max_iter, par->tor, logform, &W, gr);
        if (rc) {
            rc = msf_fail(rc, "%s: trace %d: the Sinkhorn iterations left the floating-point range", who, r);
            break;
        }
        J += W;
        for (int i = 0; i < n; i++) res[r][i] = (float)(2.0 * syn[r][i] * gr[i]);
        if (r == nr - 1) store_potentials(par, &s);
    }
    if (!rc) {
        msf[0] = (float)L2;
        msf[1] = (float)J;
    }
    free(pr);
    sk_free(&s);
    return rc;
}

int entropy_ot_2d_trace_by_trace(float **obs, float **syn, eot_msf_par *par, float *msf, float **res)
{
    return eot_2d(obs, syn, par, 1e-3, 0, msf, res, "entropy_ot_2d_trace_by_trace");
}

int entropy_ot_log_2d_trace_by_trace(float **obs, float **syn, eot_msf_par *par, eot_log_msf_par *par_exd,
                                     float *msf, float **res)
{
    const double bar = (par_exd && par_exd->bar > 0.0f) ? par_exd->bar : 1e-3;
    return eot_2d(obs, syn, par, bar, 1, msf, res, "entropy_ot_log_2d_trace_by_trace");
}
