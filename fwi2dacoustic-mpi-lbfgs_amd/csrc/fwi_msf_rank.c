/*
 * fwi_msf_rank.c -- two members of the host misfit pool: dynamic warping by
 * local cross-correlation (lcc_diw.c, lcc_dtw.h) and soft ranking
 * (soft_ranking.c, soft_ranking.h).  FP64 inside, FP32 traces, row pointers
 * obs[ir][it] as the rest of the pool (fwi_msf.h).
 *
 * The reference holds the building blocks of both (the correlation cube,
 * the accumulations and backtracking of the warping path; the argsort and
 * the isotonic regression) and declares the misfits themselves,
 * lcc_diw_msf (lcc_dtw.h:3) and soft_ranking / soft_ranking_msf /
 * soft_ranking_2d_trace_by_trace (soft_ranking.h:52-56), without bodies.
 * The misfits here are defined from their methods:
 *
 *   dynamic warping: the lag path tau(t) maximising the summed local
 *     correlation of syn and obs under the 5-point slope constraint; the
 *     objective is the L2 misfit of syn against obs warped along that path,
 *     1/2 sum (syn(t) - obs(t + tau(t)))^2, whose derivative with the path
 *     held fixed (the path is piecewise constant in syn) is the residual.
 *
 *   soft ranking (Blondel, Teboul, Berthet, Djolonga, "Fast differentiable
 *     sorting and ranking", ICML 2020): r(a) = the Euclidean projection of
 *     a / reg onto the permutahedron of (n, n-1, ..., 1), computed as
 *     z - v with v the decreasing isotonic regression of sort(z) - w; r
 *     tends to the ascending ranks as reg -> 0 and to (n+1)/2 as reg -> inf.
 *     The objective is 1/2 |r(syn) - r(obs)|^2; the Jacobian of the
 *     projection is the identity minus the averaging over the isotonic
 *     blocks (in sorted order), which gives the residual.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fwi_hip.h"
#include "fwi_msf.h"

static char g_rank_msg[512] = "";

static int rk_fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_rank_msg, sizeof(g_rank_msg), fmt, ap);
    va_end(ap);
    return code;
}

const char *fwi_msf_rank_last_error(void) { return g_rank_msg; }

static int keep(const float *sr, const float *range, int r)
{
    if (!sr || !range) return 1;
    const double dx = (double)sr[2 + 2 * r] - sr[0], dz = (double)sr[3 + 2 * r] - sr[1];
    const double d = sqrt(dx * dx + dz * dz);
    return d >= range[0] && d <= range[1];
}

/* ======================================================================
 * dynamic warping by local cross-correlation
 * ====================================================================== */

/* lcc_diw.c:14-25: a Gaussian of n samples centred at n/2 */
void gaussian_win(size_t n, float delta, float *win)
{
    const long h = (long)n / 2;
    for (long i = 0; i < (long)n; i++) {
        const double t = (double)(i - h) / delta;
        win[i] = (float)exp(-0.5 * t * t);
    }
}

#define DIW_BARRIER (-1e30f) /* columns 0 and nl+1 of the cubes: never on a path */

/* lcc_diw.c:242-319.  err[ir][it][il], il = 1..nl: the correlation of the
 * syn window centred at it with the obs window centred at it + il - 1 - nl/2
 * (both tapered by gaussian_win(nw, alpha_win), zero outside the trace),
 * normalised by the windows' energies (+1e-12), for it from the first to
 * before the last sample where syn > eps.  Columns 0 and nl+1 hold
 * DIW_BARRIER (the reference leaves them as allocated); samples outside
 * that span, and every sample of a trace outside the offset range, hold 0. */
int lcc_err(float **obs, float **syn, float *sr_xyz, float offset_range[], float eps, int nt, int nr, int nl, int nw,
            float alpha_win, float ***err)
{
    if (!obs || !syn || !err || nt < 1 || nr < 0 || nl < 1 || nw < 1)
        return rk_fail(FWI_HIP_EINVAL, "lcc_err: bad arguments");
    if (nw % 2 != 1 || nl % 2 != 1) return rk_fail(FWI_HIP_EINVAL, "lcc_err: nw and nl must be odd");
    const int nw2 = nw / 2, nl2 = nl / 2;
    double *win = (double *)malloc(sizeof(double) * nw);
    float *winf = (float *)malloc(sizeof(float) * nw);
    double *b = (double *)calloc((size_t)nt + nw, sizeof(double));
    double *a = (double *)calloc((size_t)nt + nw + nl, sizeof(double));
    if (!win || !winf || !b || !a) {
        free(win);
        free(winf);
        free(b);
        free(a);
        return rk_fail(FWI_HIP_ENOMEM, "lcc_err: host memory");
    }
    gaussian_win((size_t)nw, alpha_win, winf);
    for (int i = 0; i < nw; i++) win[i] = winf[i];
    for (int ir = 0; ir < nr; ir++) {
        for (int it = 0; it < nt; it++) {
            err[ir][it][0] = err[ir][it][nl + 1] = DIW_BARRIER;
            for (int il = 1; il <= nl; il++) err[ir][it][il] = 0.0f;
        }
        if (!keep(sr_xyz, offset_range, ir)) continue;
        int it0 = -1, it1 = -1;
        for (int it = 0; it < nt; it++)
            if (syn[ir][it] > eps) {
                it0 = it;
                break;
            }
        for (int it = nt - 1; it >= 0; it--)
            if (syn[ir][it] > eps) {
                it1 = it;
                break;
            }
        if (it0 < 0) continue;
        /* b[k] = syn[k - nw2], a[k] = obs[k - nw2 - nl2] (zero padded) */
        memset(b, 0, sizeof(double) * ((size_t)nt + nw));
        memset(a, 0, sizeof(double) * ((size_t)nt + nw + nl));
        for (int it = 0; it < nt; it++) {
            b[it + nw2] = syn[ir][it];
            a[it + nw2 + nl2] = obs[ir][it];
        }
        for (int it = it0; it < it1; it++)
            for (int il = 1; il <= nl; il++) {
                double sab = 0.0, saa = 0.0, sbb = 0.0;
                for (int iw = 0; iw < nw; iw++) {
                    const double x = b[it + iw] * win[iw], y = a[it + iw + il - 1] * win[iw];
                    sab += x * y;
                    saa += y * y;
                    sbb += x * x;
                }
                err[ir][it][il] = (float)(sab / (sqrt(saa * sbb) + 1e-12));
            }
    }
    free(win);
    free(winf);
    free(b);
    free(a);
    return FWI_HIP_OK;
}

static double max3(double x, double y, double z)
{
    const double m = x > y ? x : y;
    return m > z ? m : z;
}

/* The accumulations of the similarity err[nt][nl+2] along lag paths (the
 * best path ending at (i, j)); columns 0 and nl+1 of ddd hold DIW_BARRIER.
 *   3-point (lcc_diw.c:36-46, 77-87): the lag moves by at most one sample
 *     per time step;
 *   5-point (lcc_diw.c:48-75, 89-115): by at most one sample per two time
 *     steps, the intermediate sample taken at the half lag's neighbour.
 * _fw accumulates from t = 0 (the best path ending at i), _bw from t =
 * nt-1 (the best path starting at i).                                     */
void ac_3pts_fw(float **err, float **ddd, int nl, int nt)
{
    for (int i = 0; i < nt; i++) ddd[i][0] = ddd[i][nl + 1] = DIW_BARRIER;
    for (int j = 1; j <= nl; j++) ddd[0][j] = err[0][j];
    for (int i = 1; i < nt; i++)
        for (int j = 1; j <= nl; j++)
            ddd[i][j] = (float)((double)err[i][j] + max3(ddd[i - 1][j], ddd[i - 1][j - 1], ddd[i - 1][j + 1]));
}

void ac_3pts_bw(float **err, float **ddd, int nl, int nt)
{
    for (int i = 0; i < nt; i++) ddd[i][0] = ddd[i][nl + 1] = DIW_BARRIER;
    for (int j = 1; j <= nl; j++) ddd[nt - 1][j] = err[nt - 1][j];
    for (int i = nt - 2; i >= 0; i--)
        for (int j = 1; j <= nl; j++)
            ddd[i][j] = (float)((double)err[i][j] + max3(ddd[i + 1][j], ddd[i + 1][j + 1], ddd[i + 1][j - 1]));
}

void ac_5pts_fw(float **err, float **ddd, int nl, int nt)
{
    for (int i = 0; i < nt; i++) ddd[i][0] = ddd[i][nl + 1] = DIW_BARRIER;
    for (int j = 1; j <= nl; j++) ddd[0][j] = err[0][j];
    if (nt > 1)
        for (int j = 1; j <= nl; j++)
            ddd[1][j] = (float)((double)err[1][j] + max3(ddd[0][j], ddd[0][j - 1], ddd[0][j + 1]));
    for (int i = 2; i < nt; i++)
        for (int j = 1; j <= nl; j++) {
            const double up = (double)ddd[i - 2][j - 1] + err[i - 1][j - 1];
            const double dn = (double)ddd[i - 2][j + 1] + err[i - 1][j + 1];
            ddd[i][j] = (float)((double)err[i][j] + max3(up, ddd[i - 1][j], dn));
        }
}

void ac_5pts_bw(float **err, float **ddd, int nl, int nt)
{
    for (int i = 0; i < nt; i++) ddd[i][0] = ddd[i][nl + 1] = DIW_BARRIER;
    for (int j = 1; j <= nl; j++) ddd[nt - 1][j] = err[nt - 1][j];
    if (nt > 1)
        for (int j = 1; j <= nl; j++)
            ddd[nt - 2][j] = (float)((double)err[nt - 2][j] + max3(ddd[nt - 1][j], ddd[nt - 1][j + 1], ddd[nt - 1][j - 1]));
    for (int i = nt - 3; i >= 0; i--)
        for (int j = 1; j <= nl; j++) {
            const double up = (double)ddd[i + 2][j - 1] + err[i + 1][j - 1];
            const double dn = (double)ddd[i + 2][j + 1] + err[i + 1][j + 1];
            ddd[i][j] = (float)((double)err[i][j] + max3(up, ddd[i + 1][j], dn));
        }
}

static int argmax_row(const float *r, int n)
{
    int k = 0;
    for (int j = 1; j < n; j++)
        if (r[j] > r[k]) k = j;
    return k;
}

/* one 3-point step from column ll on row d: up one, stay, or down one,
 * preferring to stay on ties with the upper neighbour and to go down when
 * the lower neighbour is at least as good as both (lcc_diw.c:129-143) */
static float step3(const float *d, float ll)
{
    const int l = (int)ll;
    const float up = d[l + 1], st = d[l], dn = d[l - 1];
    if ((up > st ? up : st) > dn) return up > st ? ll + 1.0f : ll;
    return ll - 1.0f;
}

/* The paths (lag columns, half columns on 5-point diagonals), lcc_diw.c:117-233.
 *   bt_fw_3pts: on a backward 3-point accumulation, from the best start;
 *   bt_fw_5pts: on a backward 5-point accumulation, from the best start;
 *   bt_bw_5pts: on a forward 5-point accumulation, from the best end.     */
void bt_fw_3pts(float **ddd, int nl, int nt, float *path)
{
    path[0] = (float)argmax_row(ddd[0], nl + 2);
    for (int i = 1; i < nt; i++) path[i] = step3(ddd[i], path[i - 1]);
}

void bt_fw_5pts(float **ddd, float **err, int nl, int nt, float *path)
{
    path[0] = (float)argmax_row(ddd[0], nl + 2);
    int i = 1;
    while (i < nt - 1) {
        const int l = (int)path[i - 1];
        const double dn = (double)ddd[i + 1][l - 1] + err[i][l - 1];
        const double st = ddd[i][l];
        const double up = (double)ddd[i + 1][l + 1] + err[i][l + 1];
        if ((up > st ? up : st) > dn) {
            if (up > st) {
                path[i] = path[i - 1] + 0.5f;
                path[i + 1] = path[i - 1] + 1.0f;
                i += 2;
            } else {
                path[i] = path[i - 1];
                i += 1;
            }
        } else {
            path[i] = path[i - 1] - 0.5f;
            path[i + 1] = path[i - 1] - 1.0f;
            i += 2;
        }
    }
    if (i == nt - 1) path[i] = step3(ddd[i], path[i - 1]);
}

void bt_bw_5pts(float **ddd, float **err, int nl, int nt, float *path)
{
    int i = nt - 1;
    path[i] = (float)argmax_row(ddd[i], nl + 2);
    i -= 1;
    while (i >= 1) {
        const int l = (int)path[i + 1];
        const double dn = (double)ddd[i - 1][l - 1] + err[i][l - 1];
        const double st = ddd[i][l];
        const double up = (double)ddd[i - 1][l + 1] + err[i][l + 1];
        if ((up > st ? up : st) > dn) {
            if (up > st) {
                path[i] = path[i + 1] + 0.5f;
                path[i - 1] = path[i + 1] + 1.0f;
                i -= 2;
            } else {
                path[i] = path[i + 1];
                i -= 1;
            }
        } else {
            path[i] = path[i + 1] - 0.5f;
            path[i - 1] = path[i + 1] - 1.0f;
            i -= 2;
        }
    }
    if (i == 0) path[0] = step3(ddd[0], path[1]);
}

/* obs at the fractional sample t (linear between samples, 0 outside) */
static double sample_at(const float *x, int nt, double t)
{
    const double f = floor(t);
    const long i = (long)f;
    const double w = t - f;
    const double x0 = (i >= 0 && i < nt) ? x[i] : 0.0, x1 = (i + 1 >= 0 && i + 1 < nt) ? x[i + 1] : 0.0;
    return (1.0 - w) * x0 + w * x1;
}

/* lcc_diw_msf (lcc_dtw.h:3, declared without a body): per trace in the
 * offset range, the local correlation cube (lcc_err), its forward 5-point
 * accumulation and the path backtracked from the best end (bt_bw_5pts),
 * lag tau(t) = path(t) - 1 - nl/2 samples.  A sample counts when the
 * similarity on the path is at least sim_barrier and |tau| <= lag_barrier
 * samples:  msf[0] += 1/2 tau^2 (the traveltime misfit, samples^2),
 * msf[1] += 1/2 (syn(t) - obs(t + tau))^2 (obs linear between samples),
 * res(t) = syn(t) - obs(t + tau(t)) (0 elsewhere): the gradient of msf[1]
 * with the path held fixed.                                               */
int lcc_diw_msf(float **obs, float **syn, float *sy_xy, float offset_range[], float eps, int nt, int nr, int nl,
                int nw, float alpha_win, float sim_barrier, float lag_barrier, float *msf, float **res)
{
    if (!obs || !syn || !msf || !res || nt < 1 || nr < 0 || nl < 1 || nw < 1)
        return rk_fail(FWI_HIP_EINVAL, "lcc_diw_msf: bad arguments");
    const int m = nl + 2, nl2 = nl / 2;
    float *ebuf = (float *)malloc(sizeof(float) * (size_t)nt * m), *dbuf = (float *)malloc(sizeof(float) * (size_t)nt * m);
    float **erow = (float **)malloc(sizeof(float *) * nt), **drow = (float **)malloc(sizeof(float *) * nt);
    float *path = (float *)malloc(sizeof(float) * nt);
    int rc = (ebuf && dbuf && erow && drow && path) ? FWI_HIP_OK : rk_fail(FWI_HIP_ENOMEM, "lcc_diw_msf: host memory");
    double J0 = 0.0, J1 = 0.0;
    for (int it = 0; !rc && it < nt; it++) {
        erow[it] = ebuf + (size_t)it * m;
        drow[it] = dbuf + (size_t)it * m;
    }
    for (int ir = 0; !rc && ir < nr; ir++) {
        memset(res[ir], 0, sizeof(float) * nt);
        if (!keep(sy_xy, offset_range, ir)) continue;
        float **cube = erow;
        rc = lcc_err(obs + ir, syn + ir, NULL, NULL, eps, nt, 1, nl, nw, alpha_win, &cube);
        if (rc) break;
        ac_5pts_fw(erow, drow, nl, nt);
        bt_bw_5pts(drow, erow, nl, nt, path);
        for (int it = 0; it < nt; it++) {
            const double tau = (double)path[it] - 1.0 - nl2;
            const double s = erow[it][(int)floor(path[it])];
            if (s < sim_barrier || fabs(tau) > lag_barrier) continue;
            const double r = (double)syn[ir][it] - sample_at(obs[ir], nt, it + tau);
            J0 += 0.5 * tau * tau;
            J1 += 0.5 * r * r;
            res[ir][it] = (float)r;
        }
    }
    if (!rc) {
        msf[0] = (float)J0;
        msf[1] = (float)J1;
    }
    free(ebuf);
    free(dbuf);
    free(erow);
    free(drow);
    free(path);
    return rc;
}

/* ======================================================================
 * soft ranking
 * ====================================================================== */

typedef struct {
    double v;
    int i;
} KeyIdx;

static int cmp_key(const void *x, const void *y)
{
    const KeyIdx *a = (const KeyIdx *)x, *b = (const KeyIdx *)y;
    if (a->v < b->v) return -1;
    if (a->v > b->v) return 1;
    return (a->i > b->i) - (a->i < b->i); /* equal values keep their order */
}

/* soft_ranking.c:21-66: perm (1-based) sorts a ascending (order 0) or
 * descending (1), rank[i] (1-based) is the place of a[i], sorted_a = a in
 * that order.  Equal values keep their index order here (the reference's
 * qsort leaves their order unspecified).                                  */
int sorting_flt_argsort(const int n, const float *a, const enum sortOrder_enum order, int *perm, int *rank,
                        float *sorted_a)
{
    if (n < 0 || (n > 0 && (!a || !perm || !rank || !sorted_a))) return rk_fail(FWI_HIP_EINVAL, "argsort: bad arguments");
    KeyIdx *k = (KeyIdx *)malloc(sizeof(KeyIdx) * (n ? n : 1));
    if (!k) return rk_fail(FWI_HIP_ENOMEM, "argsort: host memory");
    for (int i = 0; i < n; i++) {
        k[i].v = order == SORT_DESCENDING ? -(double)a[i] : (double)a[i];
        k[i].i = i;
    }
    qsort(k, (size_t)n, sizeof(KeyIdx), cmp_key);
    for (int i = 0; i < n; i++) {
        perm[i] = k[i].i + 1;
        rank[k[i].i] = i + 1;
        sorted_a[i] = a[k[i].i];
    }
    free(k);
    return FWI_HIP_OK;
}

/* soft_ranking.c:68-121: sol = argmin |sol - y|^2 over non-increasing
 * sequences, by pooling adjacent violators (FP64 here).  On return
 * target[i] = the last index of the block starting at i (and of its last
 * element, the first), as the reference leaves it.                        */
static int iso_dec(int n, const double *y, double *sol, int *target)
{
    double *c = (double *)malloc(sizeof(double) * (n ? n : 1)), *sums = (double *)malloc(sizeof(double) * (n ? n : 1));
    if (!c || !sums) {
        free(c);
        free(sums);
        return rk_fail(FWI_HIP_ENOMEM, "isotonic_l2: host memory");
    }
    for (int i = 0; i < n; i++) {
        target[i] = i;
        c[i] = 1.0;
        sol[i] = sums[i] = y[i];
    }
    int i = 0;
    while (i < n) {
        int k = target[i] + 1;
        if (k == n) break;
        if (sol[i] > sol[k]) { /* already decreasing into the next block */
            i = k;
            continue;
        }
        double sy = sums[i], sc = c[i];
        for (;;) { /* absorb blocks while the sequence does not decrease */
            const double prev = sol[k];
            sy += sums[k];
            sc += c[k];
            k = target[k] + 1;
            if (k == n || prev > sol[k]) {
                sol[i] = sy / sc;
                sums[i] = sy;
                c[i] = sc;
                target[i] = k - 1;
                target[k - 1] = i;
                if (i > 0) i = target[i - 1]; /* the previous block may now violate */
                break;
            }
        }
    }
    for (i = 0; i < n;) {
        const int k = target[i] + 1;
        for (int j = i + 1; j < k; j++) sol[j] = sol[i];
        i = k;
    }
    free(c);
    free(sums);
    return FWI_HIP_OK;
}

int isotonic_l2(int n, float *y, float *sol, int *target)
{
    if (n < 0 || (n > 0 && (!y || !sol || !target))) return rk_fail(FWI_HIP_EINVAL, "isotonic_l2: bad arguments");
    if (n == 0) return FWI_HIP_OK;
    double *yd = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    if (!yd) return rk_fail(FWI_HIP_ENOMEM, "isotonic_l2: host memory");
    double *sd = yd + n;
    for (int i = 0; i < n; i++) yd[i] = y[i];
    const int rc = iso_dec(n, yd, sd, target);
    for (int i = 0; !rc && i < n; i++) sol[i] = (float)sd[i];
    free(yd);
    return rc;
}

void soft_ranking_init(int n, soft_ranking_vars *base)
{
    const size_t m = n > 0 ? (size_t)n : 1;
    base->sorted = (float *)malloc(sizeof(float) * m);
    base->sol = (float *)malloc(sizeof(float) * m);
    base->soft_rank = (float *)malloc(sizeof(float) * m);
    base->hard_rank = (int *)malloc(sizeof(int) * m);
    base->perm = (int *)malloc(sizeof(int) * m);
    base->target = (int *)malloc(sizeof(int) * m);
}

/* frees the arrays and base itself (soft_ranking.c:136-146) */
void soft_ranking_free(soft_ranking_vars *base)
{
    if (!base) return;
    free(base->sorted);
    free(base->sol);
    free(base->soft_rank);
    free(base->hard_rank);
    free(base->perm);
    free(base->target);
    free(base);
}

/* soft_ranking (soft_ranking.h:54, no body): soft_rank = the projection of
 * a / reg onto the permutahedron of w = (n, ..., 1).  With perm sorting a
 * descending, v = isotonic_l2(sort(z) - w) and soft_rank[perm[k]] =
 * z[perm[k]] - v[k]; hard_rank the ascending ranks (1 = smallest), sol the
 * isotonic solution v, target its blocks, sorted the descending a.        */
int soft_ranking(int n, float *a, float regularization_strength, soft_ranking_vars *base)
{
    if (n < 0 || (n > 0 && (!a || !base)) || !(regularization_strength > 0.0f))
        return rk_fail(FWI_HIP_EINVAL, "soft_ranking: bad arguments (reg must be > 0)");
    if (n == 0) return FWI_HIP_OK;
    const double reg = regularization_strength;
    int *rk = base->hard_rank;
    int rc = sorting_flt_argsort(n, a, SORT_DESCENDING, base->perm, rk, base->sorted);
    if (rc) return rc;
    double *d = (double *)malloc(sizeof(double) * n), *v = (double *)malloc(sizeof(double) * n);
    if (!d || !v) {
        free(d);
        free(v);
        return rk_fail(FWI_HIP_ENOMEM, "soft_ranking: host memory");
    }
    for (int k = 0; k < n; k++) d[k] = (double)base->sorted[k] / reg - (double)(n - k); /* sort(z) - w */
    rc = iso_dec(n, d, v, base->target);
    for (int k = 0; !rc && k < n; k++) {
        const int i = base->perm[k] - 1;
        base->soft_rank[i] = (float)((double)a[i] / reg - v[k]);
        base->sol[k] = (float)v[k];
    }
    for (int i = 0; !rc && i < n; i++) rk[i] = n + 1 - rk[i]; /* descending places -> ascending ranks */
    free(d);
    free(v);
    return rc;
}

/* the transposed Jacobian of soft_rank(a) applied to g (in place):
 * (g - P^T B P g) / reg with P the descending sort and B the block averages */
static void soft_rank_vjp(int n, const soft_ranking_vars *b, double reg, double *g, double *tmp)
{
    for (int k = 0; k < n; k++) tmp[k] = g[b->perm[k] - 1];
    for (int i = 0; i < n;) {
        const int e = b->target[i] + 1;
        double m = 0.0;
        for (int k = i; k < e; k++) m += tmp[k];
        m /= (double)(e - i);
        for (int k = i; k < e; k++) tmp[k] = m;
        i = e;
    }
    for (int k = 0; k < n; k++) g[b->perm[k] - 1] -= tmp[k];
    for (int i = 0; i < n; i++) g[i] /= reg;
}

/* soft_ranking_msf (soft_ranking.h:55, no body): msf[0] = 1/2 |r(syn) -
 * r(obs)|^2 with r the soft ranks, res = its gradient with respect to syn. */
int soft_ranking_msf(float *obs, float *syn, int n, float reg, soft_ranking_vars *obs_ranking,
                     soft_ranking_vars *syn_ranking, float *msf, float *res)
{
    if (!msf || n < 0 || (n > 0 && (!obs || !syn || !res || !obs_ranking || !syn_ranking)))
        return rk_fail(FWI_HIP_EINVAL, "soft_ranking_msf: bad arguments");
    msf[0] = 0.0f;
    if (n == 0) return FWI_HIP_OK;
    int rc = soft_ranking(n, obs, reg, obs_ranking);
    if (!rc) rc = soft_ranking(n, syn, reg, syn_ranking);
    if (rc) return rc;
    double *g = (double *)malloc(sizeof(double) * n), *tmp = (double *)malloc(sizeof(double) * n);
    if (!g || !tmp) {
        free(g);
        free(tmp);
        return rk_fail(FWI_HIP_ENOMEM, "soft_ranking_msf: host memory");
    }
    double J = 0.0;
    for (int i = 0; i < n; i++) {
        g[i] = (double)syn_ranking->soft_rank[i] - (double)obs_ranking->soft_rank[i];
        J += 0.5 * g[i] * g[i];
    }
    soft_rank_vjp(n, syn_ranking, reg, g, tmp);
    for (int i = 0; i < n; i++) res[i] = (float)g[i];
    msf[0] = (float)J;
    free(g);
    free(tmp);
    return FWI_HIP_OK;
}

/* soft_ranking_2d_trace_by_trace (soft_ranking.h:56, no body): per trace
 * in [min, max] of offset, soft_ranking_msf on the time samples with the
 * working sets par[0] (obs) and par[1] (syn), each initialised for nt:
 * msf[1] = sum of the per-trace misfits, msf[0] = sum (obs - syn)^2, res
 * the gradient of msf[1] (0 on the traces left out).                     */
int soft_ranking_2d_trace_by_trace(float **obs, float **syn, float *sr_xyz, float min, float max,
                                   soft_ranking_vars *par[], float reg, int nt, int nr, float msf[], float **res)
{
    if (!obs || !syn || !par || !par[0] || !par[1] || !msf || !res || nt < 0 || nr < 0)
        return rk_fail(FWI_HIP_EINVAL, "soft_ranking_2d_trace_by_trace: bad arguments");
    const float range[2] = {min, max};
    double J = 0.0, D = 0.0;
    for (int ir = 0; ir < nr; ir++) {
        memset(res[ir], 0, sizeof(float) * nt);
        if (min <= max && !keep(sr_xyz, range, ir)) continue;
        float m = 0.0f;
        const int rc = soft_ranking_msf(obs[ir], syn[ir], nt, reg, par[0], par[1], &m, res[ir]);
        if (rc) return rc;
        J += m;
        for (int it = 0; it < nt; it++) {
            const double d = (double)obs[ir][it] - syn[ir][it];
            D += d * d;
        }
    }
    msf[0] = (float)D;
    msf[1] = (float)J;
    return FWI_HIP_OK;
}
