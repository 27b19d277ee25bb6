/*
 * fwi_msf.h -- the host misfit pool: misfit values and adjoint sources of
 * synthetic traces syn against observed traces obs, in FP64 on FP32 traces,
 * under the reference's names and argument meanings.
 *
 * Traces are row pointers obs[ir][it] (alloc.c layout), nr traces of nt
 * samples.  Every function returns 0 or a FWI_HIP_E* code of fwi_hip.h (the
 * reference functions return void and exit() on a bad mode string);
 * fwi_msf_last_error() has the message.  misfit[0] is a diagnostic,
 * misfit[1] the objective whose gradient with respect to syn is res (up to
 * the reference's own scalings, stated per function).  sr_xyz is the
 * sr_location layout (source x,z then receiver x,z pairs, fwi_2d.c:366-371);
 * a trace whose source-receiver distance lies outside [min, max] contributes
 * nothing and gets res = 0 (xcorr_msf.c:69-78); sr_xyz = NULL keeps every
 * trace.
 */
#ifndef FWI_MSF_H
#define FWI_MSF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* header.h:309-315 (same layout) */
typedef struct _par_xcorr1d {
    int nr, nt, npad4Ab;
    char *conv_how_to_slice; /* "middle" or "leading" */
} par_xcorr1d;

/* header.h:287-305 (same layout) */
typedef struct _epsdf1 {
    int nr, nt, nmodel, npad4Ax, npad4Atr, niter;
    char *conv_how_to_slice, *reg_type; /* "middle"/"leading"; "diff"/"ide" */
    float tor0, eps0, Tdepressor;
} epsdf1;

/* entropy_ot.h:3-16 (same layouts) */
typedef struct _eot_msf_par {
    int nt, nr, max_iter;
    float **M, *u, *v;
    float eps, tor, dt;
} eot_msf_par;

typedef struct _eot_log_msf_par {
    float bar;
    float **M, *u, *v;
    float *logp, *logq, *delta_u, *delta_v, vmax0;
} eot_log_msf_par;

/* L2norm, msf_pool.h:1 (declared only in the reference):
 * res = syn - obs, misfit[0] = misfit[1] = 1/2 sum res^2. */
int L2norm(float **obs, float **syn, int nt, int nr, float *sr_xyz, float min, float max, float val[], float **res);

/* Normalised zero-lag cross-correlation (no body in the reference; used by
 * the inversion with this misfit): per trace c = <s,o> / (|s| |o|),
 * misfit[1] = sum (1 - c), res = -(o / (|s||o|) - c s / |s|^2);
 * misfit[0] = sum (s - o)^2.  Traces whose energy |s|^2 or |o|^2 is zero or
 * below 1e-8 of the gather's largest are skipped (no usable waveform: the
 * normalisation would amplify their rounding noise by 1/|s|).               */
int NCCnorm(float **obs, float **syn, int nt, int nr, float *sr_xyz, float min, float max, float val[],
            float **res);

/* XCORR1D, xcorr_msf.c:4-181.  Per trace, x = the cross-correlation of syn
 * with obs on nt lags (conv_how_to_slice "middle": lags 1-nt+nt/2 ..
 * nt/2+... centred at index nt/2; "leading": the first nt lags of the full
 * correlation), T[i] = |i - nt/2| (middle) or i (leading),
 * misfit[1] += sum (T x)^2 / sum x^2, res = obs (*) 2 (T^2 x - alpha x) / sum x^2
 * sliced the same way; misfit[0] += sum |obs[nt-1-i] - syn[i]|.  Traces whose
 * L1 norms differ by more than a factor 10 are skipped (xcorr_msf.c:77), and
 * so (unlike the reference, which divides by zero there) are traces with
 * sum x^2 = 0.  The FFT length (npad4Ab) does not change the result; it
 * must be >= 0.                                                               */
int XCORR1D(float **obs, float **syn, float *sr_xyz, float min, float max, void *mpars, float *misfit,
            float **res);

/* MF1D, lsq_decon_cg.c:82-392 (adaptive matching filter).  Per trace, with
 * a = 100 obs, b = 100 syn, A w = a (*) w sliced to nt samples and D the
 * eps0-scaled first difference ("diff") or identity ("ide"): w from niter
 * CGLS steps on min |A w - b|^2 + |D w|^2 (stop when the residual norm moves
 * by < tor0), T[i] = Tdepressor |i - nmodel/2| (middle) or Tdepressor i,
 * misfit[1] += |T w|^2 / |w|^2, and res = y0 from CGNE steps on
 * [A^T D^T] [y0; y1] = 2 (T^2 w - alpha w) / |w|^2 (the derivative of the
 * penalty with respect to b = 100 syn, as the reference returns it);
 * misfit[0] = sum (obs - syn)^2.  "middle" needs an odd nmodel (the
 * reference writes one element past its buffer otherwise).                   */
int MF1D(float **obs, float **syn, void *mpars, float *misfit, float **res);

/* Entropic optimal transport between two non-negative traces on the time
 * axis (entropy_ot.h:18-30; declared only in the reference).  Cost
 * C_ij = ((i-j) dt)^2, regulariser eps: W = min_P <P,C> + eps sum P (log P - 1)
 * over couplings of p/|p|_1 and q/|q|_1, by Sinkhorn iterations (at most
 * max_iter, until the row-marginal L1 error < tor) in FP64.
 *   entropy_ot: misfit[0] = W, res = dW/dp (through the normalisation).
 *   entropy_ot_2d_trace_by_trace: per trace p = syn^2 + eta, q = obs^2 + eta
 *     with eta = 1e-3 max(obs^2) of that trace; misfit[1] = sum W,
 *     misfit[0] = sum (obs - syn)^2, res = dW/dsyn.
 *   The _log forms run the same iterations on the log potentials (stable for
 *     small eps) and take eta = bar max(obs^2) when bar > 0.  The plain forms
 *     iterate on the scalings exp(f/eps), exp(g/eps) with the banded kernel
 *     exp(-C/eps) and fall back to the log form where that underflows.
 * eot_msf_init allocates u, v (nt floats: the final potentials f, g after a
 * call) and leaves M NULL: the cost is evaluated on the fly as a banded
 * Toeplitz kernel.                                                           */
int eot_msf_init(eot_msf_par *par);
void eot_msf_destroy(eot_msf_par *par);
int eot_log_msf_init(eot_msf_par *par, eot_log_msf_par *par_exd);
int entropy_ot(float *p, float *q, eot_msf_par *par, float *msf, float *res);
int entropy_ot_2d_trace_by_trace(float **obs, float **syn, eot_msf_par *par, float *msf, float **res);
int entropy_ot_log(float *p, float *q, eot_msf_par *par, eot_log_msf_par *par_exd, float *msf, float *res);
int entropy_ot_log_2d_trace_by_trace(float **obs, float **syn, eot_msf_par *par, eot_log_msf_par *par_exd,
                                     float *msf, float **res);

const char *fwi_msf_last_error(void);

/* ---------------------------------------------------------------------------
 * Dynamic warping by local cross-correlation (lcc_diw.c, lcc_dtw.h).
 * Cubes err / ddd are row pointers [it][il] with nl + 2 columns; columns 0 and
 * nl + 1 hold -1e30 (never on a path), lags il = 1 .. nl stand for the shifts
 * il - 1 - nl/2 samples of obs against syn.
 */
/* lcc_diw.c:8-19: win[i] = exp(-(i - n/2)^2 / (2 delta^2)), i = 0 .. n-1 */
void gaussian_win(size_t n, float delta, float *win);
/* lcc_diw.c:242-319: err[ir][it][il] = the normalised correlation of the
 * gaussian_win(nw, alpha_win)-tapered windows of syn centred at it and of obs
 * centred at it + il - 1 - nl/2, for it from the first to before the last
 * sample of syn above eps; 0 elsewhere and on traces outside offset_range
 * (sr_xyz = NULL or offset_range = NULL keeps every trace).  nw and nl must
 * be odd (the reference exit()s otherwise); FWI_HIP_EINVAL here. */
int lcc_err(float **obs, float **syn, float *sr_xyz, float offset_range[], float eps, int nt, int nr, int nl, int nw,
            float alpha_win, float ***err);
/* lcc_diw.c:30-115: accumulation of err along the best lag path ending
 * (_fw) or starting (_bw) at each cell; 3-point: the lag moves by at most one
 * sample per step; 5-point: by at most one per two steps. */
void ac_3pts_fw(float **err, float **ddd, int nl, int nt);
void ac_3pts_bw(float **err, float **ddd, int nl, int nt);
void ac_5pts_fw(float **err, float **ddd, int nl, int nt);
void ac_5pts_bw(float **err, float **ddd, int nl, int nt);
/* lcc_diw.c:117-240: path[it] = the lag column of the best path (half
 * columns in 5-point diagonals), from the best start of a backward
 * accumulation (bt_fw_*) or the best end of a forward one (bt_bw_5pts). */
void bt_fw_3pts(float **ddd, int nl, int nt, float *path);
void bt_fw_5pts(float **ddd, float **err, int nl, int nt, float *path);
void bt_bw_5pts(float **ddd, float **err, int nl, int nt, float *path);
/* lcc_dtw.h:3 (declared only in the reference): per trace in the offset
 * range, the 5-point path of the forward accumulation of lcc_err, lag
 * tau(t) = path(t) - 1 - nl/2 samples; where the similarity on the path is
 * >= sim_barrier and |tau| <= lag_barrier: misfit[0] += 1/2 tau^2 (samples^2),
 * misfit[1] += 1/2 (syn(t) - obs(t + tau))^2 (obs linear between samples),
 * res(t) = syn(t) - obs(t + tau(t)) (0 elsewhere), the gradient of misfit[1]
 * with the path held fixed. */
int lcc_diw_msf(float **obs, float **syn, float *sy_xy, float offset_range[], float eps, int nt, int nr, int nl,
                int nw, float alpha_win, float sim_barrier, float lag_barrier, float *msf, float **res);

/* ---------------------------------------------------------------------------
 * Soft ranking (soft_ranking.c, soft_ranking.h), after Blondel et al., "Fast
 * differentiable sorting and ranking" (ICML 2020).
 */
/* soft_ranking.h:7-11 (same values) */
enum sortOrder_enum { SORT_ASCENDING = 0, SORT_DESCENDING = 1 };
/* soft_ranking.h:27-33 (same layout) */
typedef struct _soft_ranking_vars {
    float *sorted, *sol;
    float *soft_rank;
    int *hard_rank, *perm, *target;
} soft_ranking_vars;
/* soft_ranking.c:21-66: perm (1-based) sorts a in the order, rank[i]
 * (1-based) = the place of a[i] in it, sorted_a = a in that order; equal
 * values keep their index order. */
int sorting_flt_argsort(const int n, const float *a, const enum sortOrder_enum order, int *perm, int *rank,
                        float *sorted_a);
/* soft_ranking.c:68-121: sol = the non-increasing sequence nearest y in L2
 * (pool adjacent violators, FP64 sums); target[first] = last index of each
 * block and target[last] = its first. */
int isotonic_l2(int n, float *y, float *sol, int *target);
/* soft_ranking.c:123-146: the working arrays of n entries; free releases
 * them and base itself (base must come from malloc). */
void soft_ranking_init(int n, soft_ranking_vars *base);
void soft_ranking_free(soft_ranking_vars *base);
/* soft_ranking.h:54 (declared only): soft_rank = the Euclidean projection of
 * a / reg onto the permutahedron of (n, ..., 1): tends to the ascending ranks
 * (1 = smallest) as reg -> 0 and to (n + 1)/2 as reg -> inf.  hard_rank = the
 * ascending ranks, perm / sorted = the descending order, sol / target = the
 * isotonic solution and its blocks.  reg must be > 0. */
int soft_ranking(int n, float *a, float regularization_strength, soft_ranking_vars *base);
/* soft_ranking.h:55 (declared only): msf[0] = 1/2 |r(syn) - r(obs)|^2 with
 * r = soft_ranking(., reg); res = its gradient with respect to syn. */
int soft_ranking_msf(float *obs, float *syn, int n, float reg, soft_ranking_vars *obs_ranking,
                     soft_ranking_vars *syn_ranking, float *msf, float *res);
/* soft_ranking.h:56 (declared only): soft_ranking_msf per trace in [min, max]
 * of offset (min > max keeps every trace), working sets par[0] (obs) and
 * par[1] (syn) initialised for nt; msf[1] = the summed misfit, msf[0] =
 * sum (obs - syn)^2, res its gradient (0 on the traces left out). */
int soft_ranking_2d_trace_by_trace(float **obs, float **syn, float *sr_xyz, float min, float max,
                                   soft_ranking_vars *par[], float reg, int nt, int nr, float msf[], float **res);
const char *fwi_msf_rank_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
