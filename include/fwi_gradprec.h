/*
 * fwi_gradprec.h -- gradient post-processing on the host in FP64 (FP32 in
 * and out): the reference's triangle smoothing (triangle_smooth.c, declared
 * header.h:276-278) and grad_precond_new5 (grad_prec.c:5-52, header.h:358),
 * under their names and argument meanings.  Every function returns 0 or a
 * FWI_HIP_E* code of fwi_hip.h; fwi_gradprec_last_error() has the message.
 */
#ifndef FWI_GRADPREC_H
#define FWI_GRADPREC_H

#include "fwi_plan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* triangle, triangle_smooth.c:36-68: vv = uu convolved with the triangle
 * of half-width nr (weights 1, 2, .., nr, .., 2, 1 over nr^2), the parts that
 * fall outside [0, n12) folded back by the even reflection x -> -1-x and
 * x -> 2 n12 - 1 - x.  Needs 1 <= nr <= n12 + 1. */
int triangle(int nr, int n12, float *uu, float *vv);

/* triangle2, triangle_smooth.c:70-110: uu[n2][n1] smoothed nt times along
 * the fast axis with rect1, then nt times along the slow axis with rect2,
 * into vv (uu is not modified; vv may alias uu). */
int triangle2(int rect1, int rect2, int nt, int n1, int n2, float *uu, float *vv);

/* grad_precond_new5, grad_prec.c:5-52, on g->body[mnz][mnx] in place:
 *   clip to [min*perc, max*perc] when perc != 0 (min/max of the gradient);
 *   triangle2 with rectx1 on both axes, repeat1 times (the reference passes
 *   rectx1 for the slow axis too, grad_prec.c:37; rectz1 is unused);
 *   zero above the sea bottom: iz < bottom[ix];
 *   triangle2 with rectx2 on both axes, repeat2 times (grad_prec.c:44);
 *   scale by alpha_grad.                                                     */
int grad_precond_new5(fwi_grad *g, fwi_shape *bw, int *bottom);

const char *fwi_gradprec_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
