/*
 * fwi_sigproc.c -- host trace conditioning: the reference's Butterworth
 * band-pass (butterworth_filter.c) and mutes/tapers (mute_fun.c), evaluated
 * in FP64 on FP32 traces.
 *
 * Butterworth: the digital filter is the bilinear transform of an analog
 * Butterworth prototype prewarped at the cut-off, factored into second-order
 * sections (plus one first-order section for an odd pole count), the
 * cascade butterworth_filter.c:32-117 evaluates.  With w = 2 pi fc (fc in
 * cycles per sample), section j of n poles has
 *     s_j = sin(pi (2j+1) / (2n)) sin w,   f = sin^2(w/2) (low) or cos^2(w/2) (high)
 *     y[i] = f/(1+s_j) * (x[i] +- 2 x[i-1] + x[i-2] + 2 cos w / f * y[i-1] - (1-s_j)/f * y[i-2])
 * and the first-order section y[i] = (x[i] +- x[i-1] - (c-1 or 1-c) y[i-1]) / (1+c)
 * with c = cot(w/2) (low) or tan(w/2) (high).  Sections run one after the
 * other over the whole trace, the first-order one first, all from rest.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fwi_hip.h"
#include "fwi_sigproc.h"

/* the value pi takes in butterworth_filter.c and mute_fun.c: fwi_sys.h:1,
 * the last definition included there */
#define REF_PI 3.141592654

static char g_sp_msg[512] = "";

static int sp_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_sp_msg, sizeof(g_sp_msg), fmt, ap);
    va_end(ap);
    return code;
}

const char *fwi_sigproc_last_error(void) { return g_sp_msg; }

/* ------------------------------------------------------------------------ */
/* Butterworth cascade                                                       */

typedef struct {
    int low;     /* 1: low-pass, 0: high-pass */
    int poles;
    double a1;   /* shared first feedback coefficient of the biquads, -2 cos w / f */
    double g1, c1; /* first-order section: gain and feedback (odd pole count) */
    double *g, *a2; /* per biquad: gain f/(1+s_j) and second feedback (1-s_j)/f */
} Butter;

static int butter_make(Butter *b, int low, double fc, int poles)
{
    const double w = 2.0 * REF_PI * fc, sw = sin(w), cw = cos(w);
    const int nb = poles / 2;
    b->low = low;
    b->poles = poles;
    b->g = (double *)malloc(sizeof(double) * (nb ? nb : 1));
    b->a2 = (double *)malloc(sizeof(double) * (nb ? nb : 1));
    if (!b->g || !b->a2) {
        free(b->g);
        free(b->a2);
        return FWI_HIP_ENOMEM;
    }
    if (poles & 1) {
        const double c = low ? (1.0 + cw) / sw : sw / (1.0 + cw);
        b->g1 = 1.0 / (1.0 + c);
        b->c1 = low ? 1.0 - c : c - 1.0;
    }
    double f = low ? sin(0.5 * w) : cos(0.5 * w);
    f *= f;
    for (int j = 0; j < nb; j++) {
        const double s = sin(REF_PI * (2 * j + 1) / (2 * poles)) * sw;
        b->g[j] = f / (1.0 + s);
        b->a2[j] = (1.0 - s) / f;
    }
    b->a1 = -2.0 * cw / f;
    return FWI_HIP_OK;
}

static void butter_free(Butter *b)
{
    free(b->g);
    free(b->a2);
}

static void butter_run(const Butter *b, double *x, int n)
{
    const double sgn = b->low ? 1.0 : -1.0;
    if (b->poles & 1) {
        double xp = 0.0, yp = 0.0;
        for (int i = 0; i < n; i++) {
            const double xi = x[i];
            yp = (xi + sgn * xp - b->c1 * yp) * b->g1;
            xp = xi;
            x[i] = yp;
        }
    }
    for (int j = 0; j < b->poles / 2; j++) {
        const double g = b->g[j], a2 = b->a2[j], a1 = b->a1;
        double x1 = 0.0, x2 = 0.0, y1 = 0.0, y2 = 0.0;
        for (int i = 0; i < n; i++) {
            const double xi = x[i];
            const double y = (xi + sgn * 2.0 * x1 + x2 - a1 * y1 - a2 * y2) * g;
            x2 = x1;
            x1 = xi;
            y2 = y1;
            y1 = y;
            x[i] = y;
        }
    }
}

static void reverse(double *x, int n)
{
    for (int i = 0, j = n - 1; i < j; i++, j--) {
        const double t = x[i];
        x[i] = x[j];
        x[j] = t;
    }
}

/* one filter, causal or forward-then-backward */
static void butter_pass(const Butter *b, double *x, int n, int zero_phase)
{
    butter_run(b, x, n);
    if (zero_phase) {
        reverse(x, n);
        butter_run(b, x, n);
        reverse(x, n);
    }
}

static int bandpass(float *indat, int nt, int nr, float dt, bool phase, int nphi, int nplo, float fhi, float flo,
                    float *outdat, int adjoint);

int sfbandpass(float *indat, int nt, int nr, float dt, bool phase, int nphi, int nplo, float fhi, float flo,
               float *outdat)
{
    return bandpass(indat, nt, nr, dt, phase, nphi, nplo, fhi, flo, outdat, 0);
}

int sfbandpass_adjoint(float *indat, int nt, int nr, float dt, bool phase, int nphi, int nplo, float fhi, float flo,
                       float *outdat)
{
    return bandpass(indat, nt, nr, dt, phase, nphi, nplo, fhi, flo, outdat, 1);
}

/* The cascade from rest is a lower-triangular Toeplitz matrix L per filter,
 * reversal R.  Causal: F = L_hi L_lo, and since L^T = R L R and the L
 * commute, F^T = R F R.  Zero phase: each filter is the symmetric
 * R L R L, but the low- and high-cut passes do not commute, so
 * F^T = F_lo F_hi: the same passes in the opposite order.                  */
static int bandpass(float *indat, int nt, int nr, float dt, bool phase, int nphi, int nplo, float fhi, float flo,
                    float *outdat, int adjoint)
{
    if (!indat || !outdat || nt < 0 || nr < 0) return sp_fail(FWI_HIP_EINVAL, "sfbandpass: bad arguments");
    if (flo < 0.0f) return sp_fail(FWI_HIP_EINVAL, "sfbandpass: negative flo=%g", flo);
    const double eps = 0.0001; /* butterworth_filter.c:134 */
    const double lo = (double)flo * dt, hi = (double)fhi * dt;
    if (lo > hi) fprintf(stderr, "sfbandpass: need flo < fhi, got flo=%g, fhi=%g\n", flo, fhi);
    if (hi > 0.5) fprintf(stderr, "sfbandpass: need fhi < Nyquist, got fhi=%g, Nyquist=%g\n", fhi, 0.5 / dt);
    if (nplo < 1) nplo = 1;
    if (nplo > 1 && !phase) nplo /= 2;
    if (nphi < 1) nphi = 1;
    if (nphi > 1 && !phase) nphi /= 2;
    Butter blo, bhi;
    const int use_lo = lo > eps, use_hi = hi < 0.5 - eps;
    int rc = FWI_HIP_OK;
    if (use_lo && (rc = butter_make(&blo, 0, lo, nplo))) return sp_fail(rc, "sfbandpass: host memory");
    if (use_hi && (rc = butter_make(&bhi, 1, hi, nphi))) {
        if (use_lo) butter_free(&blo);
        return sp_fail(rc, "sfbandpass: host memory");
    }
    double *tr = (double *)malloc(sizeof(double) * (nt ? nt : 1));
    if (!tr) rc = sp_fail(FWI_HIP_ENOMEM, "sfbandpass: host memory");
    for (int r = 0; r < nr && !rc; r++) {
        const float *src = indat + (size_t)r * nt;
        for (int i = 0; i < nt; i++) tr[i] = src[i];
        if (!adjoint) {
            if (use_lo) butter_pass(&blo, tr, nt, !phase);
            if (use_hi) butter_pass(&bhi, tr, nt, !phase);
        } else if (phase) {
            reverse(tr, nt);
            if (use_lo) butter_run(&blo, tr, nt);
            if (use_hi) butter_run(&bhi, tr, nt);
            reverse(tr, nt);
        } else {
            if (use_hi) butter_pass(&bhi, tr, nt, 1);
            if (use_lo) butter_pass(&blo, tr, nt, 1);
        }
        float *dst = outdat + (size_t)r * nt;
        for (int i = 0; i < nt; i++) dst[i] = (float)tr[i];
    }
    free(tr);
    if (use_lo) butter_free(&blo);
    if (use_hi) butter_free(&bhi);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* mutes and tapers                                                          */

int fbt_muter(float *fbt, float *obs, int nt, int nr, float alpha, float dt, float tune, bool early_gone)
{
    if (!fbt || !obs || nt < 0 || nr < 0) return sp_fail(FWI_HIP_EINVAL, "fbt_muter: bad arguments");
    for (int r = 0; r < nr; r++) {
        const double t1 = (double)fbt[r] + tune;
        float *o = obs + (size_t)r * nt;
        for (int i = 0; i < nt; i++) {
            double t = i * (double)dt - t1;
            if (!early_gone) t = -t;
            if (t <= 0.0) {
                const double e = alpha * t;
                o[i] = e > -10.0 ? (float)(o[i] * exp(e)) : 0.0f;
            }
        }
    }
    return FWI_HIP_OK;
}

int win_apply(float *win, float *obs, int nt, int nr, float dt, float alpha, float *wobs)
{
    if (!win || !obs || !wobs || nt < 0 || nr < 0) return sp_fail(FWI_HIP_EINVAL, "win_apply: bad arguments");
    if (wobs != obs) memcpy(wobs, obs, sizeof(float) * (size_t)nt * nr);
    for (int r = 0; r < nr; r++) {
        const double w0 = win[2 * r], w1 = win[2 * r + 1];
        float *o = wobs + (size_t)r * nt;
        for (int i = 0; i < nt; i++) {
            const double t = (i + 1) * (double)dt;
            double v = o[i];
            const double a = t - w0, b = w1 - t; /* distance inside the window: >0 */
            if (a <= 0.0) v = alpha * a >= -20.0 ? v * exp(alpha * a) : 0.0;
            if (b <= 0.0) v = alpha * b >= -20.0 ? v * exp(alpha * b) : 0.0;
            o[i] = (float)v;
        }
    }
    return FWI_HIP_OK;
}

/* source-receiver distance of receiver r (sr = sx, sz, rx0, rz0, ...) */
static double sr_dist(const float *sr, int r)
{
    const double dx = (double)sr[2 + 2 * r] - sr[0], dz = (double)sr[3 + 2 * r] - sr[1];
    return sqrt(dx * dx + dz * dz);
}

static int taper_args(float **in, float *sr, int nr, int nt, const char *who)
{
    if (!in || !sr || nr < 0 || nt < 0) return sp_fail(FWI_HIP_EINVAL, "%s: bad arguments", who);
    for (int r = 0; r < nr; r++)
        if (!in[r] && nt) return sp_fail(FWI_HIP_EINVAL, "%s: NULL trace %d", who, r);
    return FWI_HIP_OK;
}

static int tanh_taper(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, float K, int upper)
{
    int rc = taper_args(in, sr, nr, nt, upper ? "upper_off_srcood" : "lower_off_srcood");
    if (rc) return rc;
    for (int r = 0; r < nr; r++) {
        const double t1 = t0 + sr_dist(sr, r) / v0;
        for (int i = 0; i < nt; i++) {
            const double th = tanh(K * (i * (double)dt - t1));
            in[r][i] = (float)(in[r][i] * (upper ? 0.5 * (1.0 + th) : 1.0 - th));
        }
    }
    return FWI_HIP_OK;
}

int upper_off_srcood(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, float K, float **out)
{
    (void)out;
    return tanh_taper(in, sr, nr, nt, dt, t0, v0, K, 1);
}

int lower_off_srcood(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, float K, float **out)
{
    (void)out;
    return tanh_taper(in, sr, nr, nt, dt, t0, v0, K, 0);
}

/* cosine ramp of N samples ending (upper) or starting (lower) at t1; zero
 * beyond the ramp on the far side; quadratic moveout when quad != 0 */
static int cos_taper(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, int N, int upper,
                     int quad, const char *who)
{
    int rc = taper_args(in, sr, nr, nt, who);
    if (rc) return rc;
    const double len = N * (double)dt;
    for (int r = 0; r < nr; r++) {
        const double d = sr_dist(sr, r);
        const double t1 = quad ? sqrt((double)t0 * t0 + d * d / ((double)v0 * v0)) : t0 + d / v0;
        const double te = upper ? t1 - len : t1 + len; /* far end of the ramp */
        for (int i = 0; i < nt; i++) {
            const double t = i * (double)dt;
            const int inside = upper ? t < t1 : t > t1;
            if (!inside) continue;
            const int on_ramp = upper ? t > te : t < te;
            in[r][i] = on_ramp ? (float)(in[r][i] * 0.5 * (1.0 - cos(REF_PI * (t - te) / len))) : 0.0f;
        }
    }
    return FWI_HIP_OK;
}

int linear_upper_off(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, int N)
{
    return cos_taper(in, sr, nr, nt, dt, t0, v0, N, 1, 0, "linear_upper_off");
}

int linear_lower_off(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, int N)
{
    return cos_taper(in, sr, nr, nt, dt, t0, v0, N, 0, 0, "linear_lower_off");
}

int quadratic_upper_off(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, int N)
{
    return cos_taper(in, sr, nr, nt, dt, t0, v0, N, 1, 1, "quadratic_upper_off");
}

int quadratic_lower_off(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, int N)
{
    return cos_taper(in, sr, nr, nt, dt, t0, v0, N, 0, 1, "quadratic_lower_off");
}
