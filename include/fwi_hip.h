/*
 * fwi_hip.h -- C ABI of the MI355X (gfx950) wave-equation core.
 *
 * Plain C: pointers, sizes and ints only.  Every call returns 0 on success
 * and a FWI_HIP_E* code otherwise (fwi_hip_last_error() has the message);
 * nothing calls exit().  One context per thread; a context owns its device
 * buffers and runs on one HIP stream.
 *
 * Layouts are the reference's: 2-D fields row-major with x fastest
 * (alloc.c:3-12), physical grid nz x nx, propagation grid (nz+npml) x
 * (nx+2 npml) under the free surface (fwi_2d.c:148-165).
 *
 * Replaces, in the reference:
 *   the body of grad4_1shot (declared header.h:283, called MBotOpt.c:55,99,
 *   never defined): fwi_hip_forward + misfit + fwi_hip_adjoint below, and
 *   grad4_1shot itself in fwi_plan.h;
 *   the per-step kernels a_lap_opt_2d_pml (fwi_2d.c:761), a_lap_opt_2d
 *   (fwi_2d.c:722), time_integral (fwi_2d.c:866), write_bdr / read_bdr
 *   (fwi_2d.c:882,919) and cpml_coe (fwi_2d.c:110): the fwi_hip_op_* /
 *   fwi_hip_*_bdr / fwi_hip_cpml_coe entry points, on device buffers.
 */
#ifndef FWI_HIP_H
#define FWI_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FWI_HIP_OK 0
#define FWI_HIP_EINVAL 1  /* bad argument or shape */
#define FWI_HIP_ENOMEM 2  /* allocation failed or the HBM budget is exceeded */
#define FWI_HIP_EHIP 3    /* a HIP runtime call failed */
#define FWI_HIP_ESTATE 4  /* call out of order (e.g. adjoint before forward) */
#define FWI_HIP_ERANGE 5  /* a host iteration left the floating-point range (e.g. Sinkhorn scaling underflow) */

typedef struct fwi_hip_ctx fwi_hip_ctx;

/* Grid of one model.  Replaces fwi_base (header.h:21-27) and the CPML meta
 * pml_base (header.h:84-89, set at fwi_2d.c:257). Free surface on top.     */
typedef struct {
    int nx, nz;     /* physical grid (fwi_base.nx, .nz)                     */
    int npml;       /* CPML cells left, right and below (fwi_base.npml)    */
    float dx, dz, dt;
    float max_f, max_v, coef_r; /* MAX_F, MAX_V, COEF_R (fwi_sys.h)         */
} fwi_hip_grid;

typedef struct {
    fwi_hip_grid grid;
    int max_shots;       /* batch capacity (one MPI rank per shot in the reference) */
    int max_nt;          /* longest shot (fwi_base.nt)                              */
    int max_nr;          /* most receivers in a shot (fwi_msf.nr)                   */
    int tile_z;          /* rows per wave tile, >= 8 (0 = automatic)                */
    int shots_per_group; /* shots per workgroup 1..4 (0 = automatic)                */
    size_t hbm_budget;   /* bytes the context may allocate (0 = no limit)           */
    void *stream;        /* hipStream_t to run on (NULL = the context makes one)    */
} fwi_hip_desc;

/* One shot.  Replaces fwi_sou (header.h:109-117): cells as computed at
 * fwi_2d.c:363-371 (physical grid), wavelet already scaled (fwi_2d.c:353). */
typedef struct {
    int nt;
    int isx, isz;
    int nr;
    const int *irx, *irz;
    const float *wav;
} fwi_hip_shot;

/* Buffers of the caller (model, traces, observed data, residual, gradient)
 * may be host memory (page-locked is fastest) or device memory of the
 * context's GPU: the copies take either.  Every call below returns after its
 * copies into the caller's buffers are complete; device buffers the caller
 * fills on another stream must be complete before the call.                */
int fwi_hip_create(const fwi_hip_desc *desc, fwi_hip_ctx **out);
void fwi_hip_destroy(fwi_hip_ctx *ctx);
const char *fwi_hip_last_error(const fwi_hip_ctx *ctx); /* ctx may be NULL */
size_t fwi_hip_bytes(const fwi_hip_ctx *ctx);           /* device bytes held */

/* Velocity model vel[nz][nx] (m/s; host or device memory, checked and
 * extended from a host copy).  ekp = vel^2 dt^2 extended into the CPML by
 * edge replication (model_ext, fwi_2d.c:67-107).                            */
int fwi_hip_set_model(fwi_hip_ctx *ctx, const float *vel);

/* The host check fwi_hip_set_model runs first (no device work): every
 * velocity finite and > 0, and none above the stability limit of the
 * leapfrog step with the staggered pair of fwi_fdcoe (fwi_2d.c:302-316) and
 * the free-surface rows (fwi_2d.c:805-807,848),
 * v dt sqrt(4 (c0 - c1)^2 / dx^2 + 6.0195791 / dz^2) <= 2 (the second term
 * is the largest eigenvalue of the z operator, a mode bound to the free
 * surface).  The reference checks neither
 * and runs an unstable model into overflow.  FWI_HIP_EINVAL names the first
 * offending cell in fwi_hip_last_error(NULL) and *cell (-1 when the model
 * passes; cell may be NULL). */
int fwi_hip_check_model(const fwi_hip_grid *g, const float *vel, long long *cell);

/* The shots of the next forward/adjoint pair (receiver cells must be unique
 * within a shot).  Traces of a batch are packed shot after shot, each
 * [nr][nt] with time contiguous (fwi_msf.syn, header.h:66).  A batch equal
 * to the one the context holds (every nt, cell, receiver and wavelet sample)
 * keeps the device tables, strip lists and captured graphs.                 */
int fwi_hip_set_shots(fwi_hip_ctx *ctx, int nshots, const fwi_hip_shot *shots);

/* Forward propagation of every shot in one launch per time step.  Receiver
 * traces stay on the device; copied to `traces` (host) when non-NULL.       */
int fwi_hip_forward(fwi_hip_ctx *ctx, float *traces);

/* fwi_hip_forward with the traces handed over time-major, each shot
 * [nt][nr] at the same packed offset (the layout the steps record), in
 * blocks: in the graph form of a boundary-saving pass each block of
 * graph-steps steps goes to traces_tm (page-locked memory, fwi_hip_host_alloc,
 * lets the copy run beside the next block) while later blocks propagate, and
 * fn(user, k0, k1), when given, runs on the calling thread as soon as rows
 * k0 <= k < k1 of every shot are on the host, blocks in increasing order;
 * other forms hand over all rows once, fn(user, 0, kmax).  A nonzero return
 * of fn stops further calls; the pass completes and that value is returned.
 * No counterpart in the reference, whose grad4_1shot holds its traces on the
 * host from the start: this is the pipelined form of fwi_hip_forward for the
 * host misfit step (grad_all_shots).  traces_tm may be NULL only without fn. */
typedef int (*fwi_hip_block_fn)(void *user, int k0, int k1);
int fwi_hip_forward_tm(fwi_hip_ctx *ctx, float *traces_tm, fwi_hip_block_fn fn, void *user);

/* Observed traces (host, packed like the synthetics), kept on the device. */
int fwi_hip_set_observed(fwi_hip_ctx *ctx, const float *obs);
/* L2 misfit on the device: residual = syn - obs, *misfit = 1/2 sum res^2
 * accumulated in FP64 (msf_pool.h:1 declares L2norm without a body).      */
int fwi_hip_misfit_l2(fwi_hip_ctx *ctx, double *misfit);
/* Adjoint source computed elsewhere (host misfit pool), packed like obs. */
int fwi_hip_set_residual(fwi_hip_ctx *ctx, const float *res);
/* The adjoint source time-major, packed like fwi_hip_forward_tm's traces.
 * Not copied here: the next fwi_hip_adjoint uploads it, block by block
 * beside its steps in the graph form of a one-chunk boundary-saving pass
 * (else whole, first), so res_tm must stay valid and unchanged until that
 * call returns (page-locked memory lets the copies overlap the steps).    */
int fwi_hip_set_residual_tm(fwi_hip_ctx *ctx, const float *res_tm);

/* Back-propagation of the residual (exact transpose of the forward),
 * source-field reconstruction from the saved frames, zero-lag imaging
 * summed over all shots in FP64.  grad[nz][nx] = dJ/dvel, gk[nz][nx] =
 * ekp * dJ/dekp; either may be NULL (then it stays on the device).  The
 * reconstruction consumes the forward's final wavefield: a second adjoint
 * needs a new fwi_hip_forward (FWI_HIP_ESTATE otherwise).                    */
int fwi_hip_adjoint(fwi_hip_ctx *ctx, double *grad, double *gk);

/* Time steps per captured HIP graph (default 100; 0 = plain launches).
 * The kernels of a block of steps read the step index from a device
 * counter, so one executable graph is replayed for every block.           */
int fwi_hip_set_graph(fwi_hip_ctx *ctx, int steps);

/* Kernel form of the forward, adjoint and reconstruction steps: 0 = 2-D
 * tile kernels (default), 1 = z-marching kernels.  Traces agree bit for
 * bit; gradients differ only in the FP64 summation order over shots.
 * (No counterpart in the reference, which has one CPU loop nest.)        */
int fwi_hip_set_kernels(fwi_hip_ctx *ctx, int kind);
/* Imaging window of the adjoint pass (fwi_grad's jump / start_t / end_t,
 * header.h:119-127, read by grad4_1shot at header.h:283): the zero-lag
 * image sums only the reconstruction steps k with k0 <= k <= k1 and
 * (k - k0) % jump == 0.  Default: every step (jump 1, k0 0, k1 INT_MAX). */
int fwi_hip_set_imaging(fwi_hip_ctx *ctx, int jump, int k0, int k1);

/* Checkpointed backward pass instead of the boundary-saving reconstruction
 * (every = 0, the default).  With every = K > 0 the forward pass keeps a
 * snapshot of p (both time levels) and of the CPML memory before every K-th
 * step, and no boundary frames; the adjoint pass restores the snapshots
 * segment by segment from the last, recomputes the segment's K forward
 * steps storing X on the physical grid, and runs its adjoint steps imaging
 * mu X from that store.  storage: 0 = FP32, 1 = FP16, 2 = BF16 snapshots:
 * in 16 bit, p^k and the step difference p^k - p^{k-1}, each shot's field
 * scaled by its maximum magnitude (the CPML memory stays FP32; all
 * arithmetic stays FP32, the image FP64).  Extra device memory:
 * ceil(nt/K) x 2 x shots x (pnz x ld) snapshot values + K x shots x nx x nz
 * floats of X; the boundary frames (2 x 6 (nx + nz) floats per step and
 * shot) are not allocated.  Tile kernels only.                              */
int fwi_hip_set_checkpointing(fwi_hip_ctx *ctx, int every, int storage);

/* Binomial checkpointing of the backward pass, in the manner of Revolve
 * (Griewank and Walther's schedule), instead of the boundary frames or the
 * uniform snapshots of set_checkpointing.  snaps > 0: that many snapshot
 * slots of the forward state (p at both time levels and the CPML memory of
 * every shot; FP32, or FP16/BF16 as set_checkpointing stores them), one of
 * them the initial state, plus one slot for the state before the last step.
 * The forward pass takes the schedule's first snapshots; the adjoint pass
 * restores and advances states so that each step's X is recomputed from the
 * nearest snapshot, with the fewest forward steps the binomial bound allows
 * (about r nt for nt <= C(snaps + r, snaps)).  snaps = -1: chosen from the
 * HBM budget of the descriptor at each forward pass: the boundary frames
 * when they fit beside the context, else as many slots as fit (at most 128;
 * FWI_HIP_ENOMEM below 3).  snaps = 0: off.  Tile kernels only.  Calling
 * set_checkpointing switches this off, and this switches off its uniform
 * snapshots.  A new forward pass is needed before the next adjoint.         */
int fwi_hip_set_revolve(fwi_hip_ctx *ctx, int snaps, int storage);

/* The backward pass the last forward prepared: *method 0 boundary frames,
 * 1 uniform snapshots, 2 binomial checkpoints with *snaps slots; and the
 * forward steps the last adjoint pass recomputed.  Any pointer may be NULL. */
int fwi_hip_backward_plan(const fwi_hip_ctx *ctx, int *method, int *snaps, long long *recomputed);

/* The backward method a context with descriptor d takes under
 * set_revolve(ctx, -1, storage) within d->hbm_budget, and the device bytes
 * it then holds, for max_shots shots of max_nt steps with one receiver row
 * each (no device call): *method 0 (boundary frames) or 2 (binomial, *snaps
 * slots).  FWI_HIP_ENOMEM when neither fits.  No counterpart in the
 * reference, which saves its frames to disk (write_bdr, fwi_2d.c:882-918). */
int fwi_hip_plan_memory(const fwi_hip_desc *d, int storage, int *method, int *snaps, size_t *bytes);

/* The binomial schedule for nt steps and snaps slots (no device call):
 * *nacts actions, the first min(*nacts, cap) written to acts as (op, a, b,
 * slot): op 0 advance the state from step a to step b, 1 snapshot the state
 * at step a into slot, 2 restore slot (the state at step a), 3 turn at a (a
 * forward step from the state at a, then the adjoint step a).  The state at
 * step 0 is in slot 0.  *forward_steps: the forward steps of the advances
 * and turns.                                                                 */
int fwi_hip_revolve_schedule(int nt, int snaps, int *acts, int cap, int *nacts, long long *forward_steps);

/* Tile kernels: the general strips of a step (CPML, edges, free surface,
 * frames, sources, receivers) on a second stream beside the interior launch
 * (on = 1, the default), joined before the next step; off: one stream.
 * Results are bit-identical.                                                */
int fwi_hip_set_concurrent(fwi_hip_ctx *ctx, int on);

/* With the general strips on the second stream: at most `wgs` workgroups of
 * the interior launches (forward, fused backward) per CU, by reserving unused
 * LDS in each, so that a general workgroup fits beside them (0: no cap), in
 * both passes (defaults 6 forward, 4 backward; set_pass_shape sets one pass).
 * Results are bit-identical.                                                */
int fwi_hip_set_interior_wgs(fwi_hip_ctx *ctx, int wgs);

/* Backward step of the tile kernels fused into one kernel (on = 1, the
 * default): per shot group and 64 x 8 tile of the padded grid, adjoint step
 * k then reconstruction + imaging step k, the imaging taking mu^{k+1} from
 * the adjoint's registers and ekp read once per tile.  on = 0: the separate
 * adjoint and reconstruction kernels.  The image is bit-identical.  The
 * checkpointed backward pass (set_checkpointing) keeps separate kernels.
 * One launch is one backward time step of the gradient driver grad4_1shot
 * (declared at header.h:283, defined nowhere): the adjoint operator built
 * from a_lap_opt_2d_pml (fwi_2d.c:761) and the reconstruction built from
 * a_lap_opt_2d (fwi_2d.c:722) + read_bdr (fwi_2d.c:919).                    */
int fwi_hip_set_bwd_fused(fwi_hip_ctx *ctx, int on);

/* Backward pass of the separate tile kernels (set_bwd_fused 0) with the
 * adjoint step k and the reconstruction + imaging step k+1 launched side by
 * side on two streams (they touch disjoint data: the adjoint writes only the
 * mu buffer the imaging step does not read).  The image is bit-identical.
 * on = 0: one stream, adjoint then imaging per step.                        */
int fwi_hip_set_bwd_overlap(fwi_hip_ctx *ctx, int on);

/* Shots per backward pass of the tile kernels.  With n > 0 the adjoint pass
 * runs the backward time loop once per chunk of n shots (rounded up to whole
 * reconstruction groups), so that a chunk's wavefields stay in the device's
 * last-level cache from one step to the next; 0: all shots in one pass;
 * -1 (the default): chunks when the four wavefields of the chunks in flight
 * (set_bwd_lanes) fit in 200 MB and a step of them still has 10M
 * cell-updates or more (16 of 64 shots per chunk at 1024x320 with two in
 * flight), else one pass.  The image is bit-identical to the single pass.   */
int fwi_hip_set_bwd_chunk(fwi_hip_ctx *ctx, int shots);

/* Backward chunks in flight: 2 (the default) runs chunks in pairs, the second
 * chunk's steps on streams of its own beside the first's, so that each fills
 * the other's waits between steps; 1: one chunk after another.  Pairs run in
 * the graph form of the fused backward step with concurrent general strips
 * (set_graph > 0, set_bwd_fused 1, set_concurrent 1, set_bwd_mixed 0);
 * otherwise chunks run one after another.  The image is bit-identical.
 * No counterpart in the reference (its shots run one per MPI rank).        */
int fwi_hip_set_bwd_lanes(fwi_hip_ctx *ctx, int lanes);

/* Rows per wave of the fused backward step's general side (the adjoint
 * strips and reconstruction items of the tiles that are not interior) when
 * the backward pass runs 4-row strips: 4, or 8 (the general side on blocks
 * of 8 rows, run by the 8-row kernels; a block's two 4-row tiles stay on the
 * interior side when both are interior).  Taller general strips pay each
 * wave's fixed work (class, coefficients, buffer set-up) once per 8 rows.
 * Ignored with 8-row strips and with set_bwd_mixed 1.  The image is
 * bit-identical.  No counterpart in the reference.                         */
int fwi_hip_set_bwd_general_rows(fwi_hip_ctx *ctx, int rows);

/* Shots per workgroup of the tile reconstruction + imaging kernel (default
 * 8): the group's image sum meets the FP64 accumulator once per step.     */
int fwi_hip_set_rec_group(fwi_hip_ctx *ctx, int shots);

/* Tile shape of one pass (0: the forward pass and the recompute of the
 * checkpointed backward pass; 1: the backward pass): rows per wave of the
 * tile kernels, 4 or 8, and the interior workgroups per CU beside the
 * general strips (as set_interior_wgs, which sets both passes).  Defaults:
 * 4 rows in both passes, 6 forward and 4 backward workgroups per CU.
 * A workgroup of 4 waves covers 64 x 4*rows cells; shorter strips take fewer
 * registers and re-read more halo rows.  Traces, adjoint fields and images
 * are bit-identical across the shapes.  No counterpart in the reference (its
 * loops have no tiles); the autotuner picks it per workload shape.          */
int fwi_hip_set_pass_shape(fwi_hip_ctx *ctx, int pass, int rows, int interior_wgs);

/* The forward step of the tile kernels as one launch (on = 1, the default):
 * the interior and general strips' workgroups in one grid, the general ones
 * first on each XCD, instead of two launches on two streams joined every
 * step (on = 0).  Results are bit-identical.  No counterpart in the
 * reference (its loops run one after another).                             */
int fwi_hip_set_fwd_mixed(fwi_hip_ctx *ctx, int on);

/* The fused backward step of the tile kernels as one launch (on = 1): the
 * interior items, the other tiles' adjoint strips and their reconstruction
 * items in one grid, the general ones first on each XCD, instead of three
 * streams joined every step (on = 0).  The reconstruction of the other tiles
 * then images one step late (mu^{k+2} X^{k+1} at step k, X kept one step in
 * device memory), so that it does not wait for the adjoint of its cells; the
 * image is bit-identical.  No counterpart in the reference.                 */
int fwi_hip_set_bwd_mixed(fwi_hip_ctx *ctx, int on);

/* Wait for all work of the context. */
int fwi_hip_synchronize(fwi_hip_ctx *ctx);

/* Timing with HIP events on the context stream.  mode 0: off.  mode 1: an
 * event pair around every launch, plain launches (no graphs, no second
 * stream); fwi_hip_kernel_time returns the summed milliseconds and launch
 * count of kernel `which` (0 forward step, 1 adjoint step or, fused, the
 * whole backward step, 2 reconstruction step).  mode 2: an event pair
 * around each whole pass, run as usual (graphs, second stream);
 * fwi_hip_kernel_time returns the summed milliseconds and time steps of
 * the forward (0) and backward (1) passes.                                  */
int fwi_hip_set_timing(fwi_hip_ctx *ctx, int mode);
int fwi_hip_kernel_time(fwi_hip_ctx *ctx, int which, double *ms, long long *launches);

/* Device state for inspection (host copy, dense rows):
 *   what 0/1: p buffer 0/1 [pnz][pnx]   2/3: adjoint buffer 0/1 [pnz][pnx]
 *   4/5: boundary / operator frames of the shot [nt][6 nx + 6 nz]          */
int fwi_hip_read_state(fwi_hip_ctx *ctx, int what, int shot, float *out);

/* ---- device memory for callers of the per-step entry points ------------
 * (no counterpart in the reference, which has no device).  kind: 1 host to
 * device, 2 device to host, 3 device to device.  Synchronous.             */
int fwi_hip_device_count(int *n);
int fwi_hip_dmalloc(void **ptr, size_t bytes);
int fwi_hip_dfree(void *ptr);
/* Page-locked host memory (no counterpart in the reference): traces passed to
 * fwi_hip_forward / fwi_hip_set_residual from such a buffer cross PCIe at the
 * pinned-copy rate, without staging.  Freed with fwi_hip_host_free.       */
int fwi_hip_host_alloc(void **ptr, size_t bytes);
int fwi_hip_host_free(void *ptr);
int fwi_hip_memcpy(void *dst, const void *src, size_t bytes, int kind);

/* ---- per-step entry points on caller-owned device buffers (dense rows) --
 * Shapes follow fwi_init (fwi_2d.c:230-296) for the f-grid p[pnz][pnx],
 * u[pnz][pnx+1], v[pnz+1][pnx], ux/vz[pnz][pnx], CPML memory
 * u0,u1[pnz][npml+1], v1[npml+1][pnx], ux0,ux1[pnz][npml], vz1[npml][pnx].
 * stream: hipStream_t or NULL (default stream).  They return after
 * enqueueing (except fwi_hip_op_pml's table upload, which is synchronous). */

/* a_lap_opt_2d_pml, fwi_2d.c:761-864 (free surface) */
int fwi_hip_op_pml(const fwi_hip_grid *g, const float *p, float *u, float *v, float *ux, float *vz,
                   float *u0, float *u1, float *v1, float *ux0, float *ux1, float *vz1, void *stream);
/* fwi_hip_op_pml, allocating nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The CPML
 * tables go to a static device array of the library instead of a fresh
 * allocation: 6 (2 npml + 1) floats, at most 3096 bytes (npml <= 64, larger
 * npml is refused with FWI_HIP_EINVAL), one copy per GPU held until the
 * process ends.  Every call on a GPU rewrites that array, so calls that
 * overlap on one GPU (other streams or threads) race.  Same results as
 * fwi_hip_op_pml, bit for bit; returns after enqueueing on `stream`.       */
int fwi_hip_op_pml_noalloc(const fwi_hip_grid *g, const float *p, float *u, float *v, float *ux, float *vz,
                           float *u0, float *u1, float *v1, float *ux0, float *ux1, float *vz1, void *stream);
/* a_lap_opt_2d, fwi_2d.c:722-759, on the physical grid p[nz][nx] */
int fwi_hip_op_interior(const fwi_hip_grid *g, const float *p, float *u, float *v, float *ux, float *vz,
                        void *stream);
/* time_integral, fwi_2d.c:866-880 */
int fwi_hip_time_integral(float *pnew, const float *pnow, const float *pold, const float *pxx,
                          const float *pzz, const float *ekp, size_t m, size_t n, void *stream);
/* one step of write_bdr (fwi_2d.c:891-910): ud[6][nx], lr[nz][6] from the f-grid field */
int fwi_hip_write_bdr(const fwi_hip_grid *g, float *ud, float *lr, const float *pf, void *stream);
/* one step of read_bdr (fwi_2d.c:940-957) into a physical-grid field pb[nz][nx] */
int fwi_hip_read_bdr(const fwi_hip_grid *g, const float *ud, const float *lr, float *pb, void *stream);
/* write_bdr_fs_no (fwi_2d.c:962-997): one step of write_bdr from a field
 * without the free surface, pf[nz + 2 npml][nx + 2 npml] (CPML above the
 * physical grid too), same strip layout.  The reference declares the
 * matching propagator (a_lap_opt_2d_pml_fs_no, header.h:318-328) without a
 * body; the propagators here keep the free surface. */
int fwi_hip_write_bdr_fs_no(const fwi_hip_grid *g, float *ud, float *lr, const float *pf, void *stream);
/* model_ext_fs_no (fwi_2d.c:29-65, host): model[n2][n1] into the middle of
 * emodel[en2][en1] with margins (en1 - n1)/2 and (en2 - n2)/2 on every side
 * filled with the nearest edge or corner value.  Uneven margins are refused
 * (the reference reads past its model then). */
int fwi_hip_model_ext_fs_no(const float *model, int n1, int n2, float *emodel, int en1, int en2);
/* cpml_coe, fwi_2d.c:110-137 (host): a[2 npml + 1], b[2 npml + 1] */
int fwi_hip_cpml_coe(float *a, float *b, int npml, float dx, float dt, float max_f, float max_v, float coef_r);

#ifdef __cplusplus
}
#endif
#endif
