/*
 * fwi_lbfgs.h -- the reference's L-BFGS inversion loop (MBotOpt.c:76-221,
 * lbfgs.h:1-84) as a single-process host library in FP64: the model is
 * broadcast to nobody and the per-rank control messages disappear (every
 * shot of an iteration runs in one batched gradient on the GPU).  Returns
 * 0 or a FWI_HIP_E* code of fwi_hip.h; fwi_lbfgs_last_error() has the
 * message.
 */
#ifndef FWI_LBFGS_H
#define FWI_LBFGS_H

#include "fwi_plan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* lbfgs.h:1-6 */
enum { LBFGS_SUCCESS, LBFGSERR_MAXIMUMLINESEARCH, LBFGSERR_MAXIMUMITERATION };

/* lbfgs.h:63-70 (same layout); defaults lbfgs.h:77-79: {12, 1e-9, 8, 6, 1e-7, 0.98} */
typedef struct {
    int m;
    float epsilon;
    int max_iterations;
    int max_linesearch;
    float ftol;
    float wolfe;
} lbfgs_parameter_t;

void fwi_lbfgs_default_param(lbfgs_parameter_t *param);

/* Objective and gradient at x (n floats): *fx and g[n].  k is the iteration
 * counter MBotOpt.c passes to grad4_1shot.  Non-zero return aborts.       */
typedef int (*fwi_lbfgs_evaluate)(void *user, const float *x, float *g, float *fx, int k);

/* Called after every accepted step (and once for the start, k = 0):
 * iteration k, objective, accepted step length, line-search trials, the
 * model, its gradient and the next search direction d (FP64).  May be NULL. */
typedef void (*fwi_lbfgs_progress)(void *user, int k, double fx, double step, int trials, const float *x,
                                   const float *g, const double *d, int n);

/* MBotOpt.c:65-221: d = -g, step 0.8; per iteration a backtracking /
 * extrapolating line search on x = xp + step d with the Armijo test
 * f <= f0 + step ftol <g0,d> (else step *= 0.5), the curvature test
 * <g,d> >= wolfe <g0,d> (else step *= 2.1) and the strong Wolfe test
 * <g,d> <= -wolfe <g0,d> (else step *= 0.5), at most max_linesearch trials;
 * stop when |g|/|x| <= epsilon or after max_iterations; then the two-loop
 * recursion over the last m pairs (s, y) scaled by <y,s>/<y,y>, and the next
 * initial step 0.2 for the first three iterations, else the mean of the
 * last two accepted steps.  x (in/out) holds the final model; *status gets
 * LBFGS_SUCCESS / LBFGSERR_MAXIMUMLINESEARCH / LBFGSERR_MAXIMUMITERATION.
 * The iterates and the (s, y) memory are FP64; x is rounded to FP32 for each
 * evaluation.                                                                */
int fwi_lbfgs(int n, float *x, fwi_lbfgs_evaluate evaluate, fwi_lbfgs_progress progress, void *user,
              const lbfgs_parameter_t *param, int *status);

/* Full-waveform inversion of the model plans[0]->model->md with all shots
 * of plans batched on the GPU per evaluation (grad_all_shots, then
 * grad_precond_new5 when precondition != 0, with plans[0]->grad and
 * model->ibtry as the sea bottom): f = sum over shots of msf->val[1].
 * log_path != NULL appends one JSON line per iteration: k, misfit, step,
 * line-search trials, |g|, and when true_model != NULL the model's relative
 * L2 error against it.                                                       */
int fwi_invert(fwi_plan **plans, int nshots, const lbfgs_parameter_t *param, int precondition,
               const float *true_model, const char *log_path, int *status);

const char *fwi_lbfgs_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
