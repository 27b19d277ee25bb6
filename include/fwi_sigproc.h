/*
 * fwi_sigproc.h -- host trace conditioning beside the wave-equation core:
 * the Butterworth band-pass and the mutes/tapers of the reference, under the
 * reference's names and argument meanings, computed in FP64 on FP32 traces.
 *
 * Traces are either one slab obs[ix*nt + it] (nr traces of nt samples) or
 * row pointers in[ix][it], as each reference function takes them.  Every
 * function returns 0 or a FWI_HIP_E* code of fwi_hip.h (the reference
 * functions return void and exit() on bad input); fwi_sigproc_last_error()
 * has the message.  No function allocates device memory.
 */
#ifndef FWI_SIGPROC_H
#define FWI_SIGPROC_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* sfbandpass, butterworth_filter.c:127-200 (declared header.h:364-366):
 * band-pass of nr traces indat[j*nt + i] into outdat (may alias indat).
 * flo, fhi in Hz, dt in s.  Low cut when flo*dt > 1e-4 (a high-pass of
 * nplo poles), high cut when fhi*dt < 0.5-1e-4 (a low-pass of nphi poles).
 * phase = true: causal, one pass.  phase = false: zero phase, each filter
 * run forward and then backward in time with its pole count halved
 * (butterworth_filter.c:159-163).  flo < 0 is an error (the reference prints
 * and exits, :141-144); flo > fhi and fhi above Nyquist only warn, as there. */
int sfbandpass(float *indat, int nt, int nr, float dt, bool phase, int nphi, int nplo, float fhi, float flo,
               float *outdat);

/* The transpose of sfbandpass with the same arguments (no counterpart in
 * the reference; for adjoint sources of filtered misfits): <F x, y> = <x, F^T y>. */
int sfbandpass_adjoint(float *indat, int nt, int nr, float dt, bool phase, int nphi, int nplo, float fhi, float flo,
                       float *outdat);

/* fbt_muter, mute_fun.c:5-23: per trace ix, tt = it*dt - (fbt[ix] + tune),
 * negated unless early_gone; where tt <= 0 the sample is scaled by
 * exp(alpha*tt), or zeroed once alpha*tt <= -10.  obs[ix*nt + it] in place. */
int fbt_muter(float *fbt, float *obs, int nt, int nr, float alpha, float dt, float tune, bool early_gone);

/* win_apply, mute_fun.c:25-59: wobs = obs tapered outside the window
 * [win[2ix], win[2ix+1]] (s) of each trace by exp(alpha*distance), zeroed
 * beyond alpha*distance < -20.  The reference's running sums ss += dt and
 * tt -= dt are evaluated as the exact products (it+1)*dt.                   */
int win_apply(float *win, float *obs, int nt, int nr, float dt, float alpha, float *wobs);

/* The offset-dependent tapers of mute_fun.c:62-190.  sr = source x,z then
 * receiver x,z pairs (the sr_location layout, fwi_2d.c:366-371); in[ix][it]
 * is modified in place (the reference's out argument is unused, :62-90).
 *   upper_off_srcood  in *= (1 + tanh(K (t - t0 - d/v0))) / 2
 *   lower_off_srcood  in *= 1 - tanh(K (t - t0 - d/v0))            (sic: no 1/2)
 *   linear_upper_off  zero before t1 - N dt, cosine ramp up to t1 = t0 + d/v0
 *   linear_lower_off  cosine ramp down after t1, zero after t1 + N dt
 *   quadratic_*_off   the same about t1 = sqrt(t0^2 + d^2/v0^2)
 * d is the source-receiver distance; pi is the reference's 3.141592654
 * (fwi_sys.h:1).                                                             */
int upper_off_srcood(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, float K, float **out);
int lower_off_srcood(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, float K, float **out);
int linear_upper_off(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, int N);
int linear_lower_off(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, int N);
int quadratic_upper_off(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, int N);
int quadratic_lower_off(float **in, float *sr, int nr, int nt, float dt, float t0, float v0, int N);

const char *fwi_sigproc_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
