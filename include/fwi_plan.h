/*
 * fwi_plan.h -- the reference's plan structs and its per-shot gradient entry
 * point, for C host code shaped like MBotOpt.c.
 *
 * The struct layouts are those of the reference's header.h:21-185 (member
 * order, types and constness kept, because MBotOpt.c-style callers pass
 * these structs by pointer): fwi_base, fwi_shape, fwi_specs, fwi_model,
 * fwi_msf, fwi_state, pml_base, pml_profile, fwi_fdcoe, fwi_sou, fwi_grad,
 * fwi_files_path, fwi_plan, and the reflection-WI structs rwi_model,
 * rwi_files_path, rwi_plan (header.h:53-58,144-156,187-220).  Only the
 * GSL/FFTW/OpenMP includes of header.h:7-14 are gone: nothing here needs
 * them.
 *
 * grad4_1shot is declared at header.h:283 and called at MBotOpt.c:55,99 but
 * has no body in the reference.  This library provides it on top of
 * fwi_hip.h.  What it reads and writes:
 *   in:  grid->base (nx, nz, npml, nt, dx, dz, dt; free_surface must be 1),
 *        model->md[nz][nx] = velocity (m/s), sou->ixyz_sr (source then
 *        receiver cells, fwi_2d.c:363-371), sou->wav0[nt] (the scaled
 *        wavelet fwi_init reads, fwi_2d.c:352-353), msf->nr, msf->nt,
 *        msf->jump, msf->obs0[nr][nt0] (the observed gather fwi_init reads,
 *        fwi_2d.c:342; the host conditioning steps the reference applies to
 *        it are not run here);
 *   out: msf->syn[nr][nt0] (synthetics every jump steps), msf->val[0] =
 *        msf->val[1] = 1/2 sum (syn - obs)^2 (L2 misfit, FP64 sum),
 *        grad->body[nz*nx] = dJ/dmd (exact gradient of the discrete misfit,
 *        the CPML extension of the model held fixed).
 * The iteration number k is accepted and unused, as in MBotOpt.c.
 */
#ifndef FWI_PLAN_H
#define FWI_PLAN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    const size_t dim, mm, npml, nx, nz, nt, nt_mem, free_surface;
    const float dx, dz, dt;
} fwi_base;

typedef struct {
    size_t mnx, mnz;
    size_t pnx, pnz;
    size_t unx, unz;
    size_t vnx, vnz;
} fwi_shape;

typedef struct {
    fwi_base *base;
    fwi_shape *f, *b;
} fwi_specs;

typedef struct {
    float **md, **kp, **ekp, *btry, alpha_btry;
    int *ibtry;
} fwi_model;

typedef struct {
    float **m0, **m1, **kp_total, **kp_0, **ekp_total, **ekp_0, *btry, alpha_btry;
    int *ibtry;
} rwi_model;

typedef struct {
    const size_t nt, nr, jump, rectx, rectt, times;
    const int alpha_fbt0, alpha_fbt1;
    const float dr, dt, os0, os1, alpha_msf;
    float **obs0, **obs, **obs_t, **syn, **syn0, **syn_t, *fbt, *val;
    int ios0, ios1, *ifbt;
} fwi_msf;

typedef struct {
    float **pold, **pnow, **pnew, **u, **v, **ux, **vz;
} fwi_state;

typedef struct {
    const size_t dim;
    const float maxF, maxV, coefR;
} pml_base;

typedef struct {
    int npml;
    float *a0[2], *b0[2], *a1[2], *b1[2];
    float **u0, **u1, **v0, **v1, **ux0, **ux1, **vz0, **vz1;
} pml_profile;

typedef struct {
    float c0x, c1x, ix;
    float c0z, c1z, iz;
    float iz840, iz24;
} fwi_fdcoe;

typedef struct {
    float iarea, *xyz_sr;
    int *ixyz_sr;
    float *wav0, *wav, **res;
} fwi_sou;

typedef struct {
    size_t id;
    const float end_t, start_t, alpha_grad, perc;
    const size_t jump;
    const size_t rectx1, rectz1, repeat1;
    const size_t rectx2, rectz2, repeat2;
    float *body;
} fwi_grad;

typedef struct {
    char *common;
    char *shots;
    char *fbt;
    char *xyz_sr;
    char *init_mdl;
    char *btry;
    char *wavelet;
} fwi_files_path;

typedef struct {
    char *common;
    char *shots;
    char *fbt;
    char *xyz_sr;
    char *init_mdl_0;
    char *init_mdl_1;
    char *btry;
    char *wavelet;
} rwi_files_path;

typedef struct {
    fwi_specs *grid;
    fwi_model *model;
    fwi_state *fw;
    fwi_state *bw;
    fwi_sou *sou;
    fwi_msf *msf;
    pml_profile *pml;
    fwi_fdcoe *coe;
    fwi_files_path *path;
    fwi_grad *grad;
    struct {
        char path_bd_wr[300];
        float ***bc_ud, ***bc_lr;
        int offset;
    };
} fwi_plan;

typedef struct {
    fwi_specs *grid;
    rwi_model *model;
    fwi_state *fw_inc;
    fwi_state *bw_inc;
    fwi_state *fw_per;
    fwi_state *bw_per;
    fwi_sou *sou;
    fwi_msf *msf;
    pml_profile *pml_inc;
    pml_profile *pml_per;
    fwi_fdcoe *coe;
    rwi_files_path *path;
    fwi_grad *grad;
    struct {
        char path_bd0_wr[300];
        char path_bd1_wr[300];
        float ***bc_ud_inc, ***bc_lr_inc;
        float ***bc_ud_per, ***bc_lr_per;
        int offset;
    };
} rwi_plan;

/* header.h:283 -- one shot: forward, L2 misfit, adjoint, gradient. */
void grad4_1shot(fwi_plan *plan, int k);

/* All shots of an iteration in one batched run (replaces the one-rank-per-
 * shot loop of MBotOpt.c:51).  The plans share the model of plans[0];
 * each plan gets its own syn and val; the gradient summed over the shots
 * goes to plans[0]->grad->body.  Like one rank per shot, each plan keeps its
 * own time axis (grid->base->nt with its msf->nt, jump and dt) and its own
 * receivers; the grid, the propagation dt and the CPML must be those of
 * plans[0].  Returns 0 or a FWI_HIP_E* code. */
int grad_all_shots(fwi_plan **plans, int nshots, int k);

/* header.h:285 -- reflection waveform inversion, one shot.  Declared but
 * not defined by the reference; rwi_init (fwi_2d.c:398-719) sets up its
 * inputs: a background velocity m0, a second model m1, incident and
 * perturbed states with their own CPML memory and frame buffers, one source
 * and receiver set and one misfit.  This library's driver, the difference
 * form of RWI:
 *   p_inc   propagates in model->m0 (the incident field),
 *   p_per   propagates in m0 + m1 (model->m1 is the velocity perturbation
 *           that holds the reflectors),
 *   syn     = R(p_per - p_inc) every jump steps (the reflected data),
 *   misfit  the host misfit step of grad4_1shot (grad4_set_misfit and the
 *           trace conditioning) on syn against msf->obs0 (reflection data),
 *   grad    grad->body = dJ/dm0: the gradient through p_per (adjoint source
 *           r, model m0 + m1) plus the gradient through p_inc (adjoint
 *           source -r, model m0), both exact discrete gradients.
 * Reads grid->base, model->m0/m1, sou (as grad4_1shot), msf->obs0; writes
 * msf->syn, msf->val, grad->body.  kp_*, ekp_*, the states, the CPML
 * profiles and the frame buffers of the plan are not used: the device
 * holds its own. */
void grad4_1shot_rfl(rwi_plan *plan, int k);

/* Every shot of an RWI iteration in one batched run (two device contexts:
 * the incident and the perturbed fields).  The model is plans[0]'s; the
 * gradient summed over shots goes to plans[0]->grad->body, and when grad_m1
 * is not NULL, dJ/dm1 (the gradient through p_per alone) to grad_m1[nz*nx].
 * Returns 0 or a FWI_HIP_E* code (grad4_last_error). */
int grad_all_shots_rfl(rwi_plan **plans, int nshots, int k, float *grad_m1);

/* An RWI plan with what grad4_1shot_rfl reads and writes (the fields
 * rwi_init fills from its files, fwi_2d.c:398-719, without the files and
 * the host states).  Free with rwi_plan_free. */
rwi_plan *rwi_plan_alloc(int nx, int nz, int npml, int nt, int jump, int nr, float dx, float dz, float dt);
void rwi_plan_free(rwi_plan *plan);

/* Message of the last failure of grad4_1shot / grad_all_shots (and of the
 * RWI forms). */
const char *grad4_last_error(void);

/* Wall-clock milliseconds of the phases of the last grad_all_shots (or
 * grad4_1shot) call, written to ms[0..n-1] (no counterpart in the
 * reference): [0] set-up (context, model, shots), [1] forward pass with the
 * traces copied to the host, [2] host misfit step, [3] adjoint source
 * copied to the device, [4] adjoint pass with the gradient copied back.
 * Returns the number of phases (5). */
int grad4_phase_ms(double *ms, int n);

/* The host misfit step of grad4_1shot / grad_all_shots (fwi_msf.h,
 * fwi_sigproc.h), process-wide settings.  The reference leaves this step
 * inside its missing grad4_1shot; the misfit functions are its pool
 * (msf_pool.h, xcorr_msf.c, lsq_decon_cg.c, entropy_ot.h).
 *   kind: GRAD4_MSF_* below; params: par_xcorr1d * (XCORR1D), epsdf1 *
 *   (MF1D), eot_msf_par * (EOT: eps, tor, max_iter used), grad4_diw_par *
 *   (DIW), float * (SOFTRANK: the regularisation), else ignored.
 *   The nr/nt fields of params are set per shot from msf->nr and msf->nt.
 *   Traces whose source-receiver distance (sou->xyz_sr) lies outside
 *   [offset_min, offset_max] are left out; offset_min > offset_max keeps all.
 * msf->val gets the misfit's (diagnostic, objective) pair, and the gradient
 * is that of val[1]: MF1D's residual (a derivative with respect to 100 syn)
 * is scaled by 100, XCORR1D's (one sample early for even nt) is shifted
 * back, and XCORR1D's "leading" slicing, whose residual is not a derivative
 * of its misfit, is refused.  The default is L2: val[0] = val[1] =
 * 1/2 sum (syn - obs)^2, residual syn - obs.                                 */
#define GRAD4_MSF_L2 0
#define GRAD4_MSF_NCC 1
#define GRAD4_MSF_XCORR1D 2
#define GRAD4_MSF_MF1D 3
#define GRAD4_MSF_EOT 4
#define GRAD4_MSF_DIW 5      /* lcc_diw_msf; params: grad4_diw_par * */
#define GRAD4_MSF_SOFTRANK 6 /* soft_ranking_2d_trace_by_trace; params: float * (reg > 0) */
/* the arguments of lcc_diw_msf (lcc_dtw.h:3) beyond the traces: odd window
 * nw and lag count nl (samples), syn threshold eps, Gaussian width alpha_win
 * (samples), and the similarity / |lag| (samples) limits of the samples
 * that count */
typedef struct _grad4_diw_par {
    int nl, nw;
    float eps, alpha_win, sim_barrier, lag_barrier;
} grad4_diw_par;
int grad4_set_misfit(int kind, const void *params, float offset_min, float offset_max);

/* Band-pass (sfbandpass, butterworth_filter.c:127-200) of the observed and
 * synthetic data-rate traces before the misfit, and its transpose on the
 * residual.  on = 0 switches it off (the default). */
int grad4_set_bandpass(int on, int phase, int nphi, int nplo, float fhi, float flo);

/* Checkpointed backward pass of the batched context (fwi_hip_set_checkpointing):
 * every = 0 (default) boundary-saving reconstruction; storage 0/1/2 =
 * FP32/FP16/BF16 snapshots.  The host misfit step runs the shots over
 * FWI_HOST_THREADS threads (default: the online cores, at most 16). */
int grad4_set_checkpointing(int every, int storage);

/* First-break mute (fbt_muter, mute_fun.c:5-23) with msf->fbt, on the
 * observed and synthetic traces and on the residual.  on = 0: off. */
int grad4_set_fbt_mute(int on, float alpha, float tune, int early_gone);

/* A plan holding exactly what grad4_1shot reads and writes, with the
 * reference's row-pointer arrays (alloc.c:3-48) and shapes (fwi_2d.c:148-
 * 165, 179-222, 380): for callers without fwi_init's input files.
 * nt0 = (nt - 1) / jump + 1 samples per trace (fwi_2d.c:195).  Free with
 * fwi_plan_free. */
fwi_plan *fwi_plan_alloc(int nx, int nz, int npml, int nt, int jump, int nr, float dx, float dz, float dt);
void fwi_plan_free(fwi_plan *plan);

/* The gradient post-processing parameters of plan->grad (header.h:119-127,
 * const there; fwi_init sets them from its inputs) read by
 * grad_precond_new5: clip percentage, final scale, and the two smoothing
 * passes (rect, repeat).  Returns 0 or FWI_HIP_EINVAL. */
int fwi_plan_set_grad(fwi_plan *plan, float perc, float alpha_grad, int rectx1, int rectz1, int repeat1, int rectx2,
                      int rectz2, int repeat2);

#ifdef __cplusplus
}
#endif
#endif
