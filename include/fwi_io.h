/*
 * fwi_io.h -- input files and the data-rate bridge on the host side of the
 * wave-equation core (plain C; every call returns 0 or a FWI_HIP_E* code of
 * fwi_hip.h, fwi_io_last_error() has the message; nothing calls exit()).
 */
#ifndef FWI_IO_H
#define FWI_IO_H

#include <stddef.h>

#include "fwi_plan.h"

#ifdef __cplusplus
extern "C" {
#endif

/* read_from_file, read_write.c:4-34: n floats from a raw float32 file; with
 * exact != 0 the file must hold exactly n floats (the reference's check=1). */
int fwi_read_f32(const char *path, float *dst, size_t n, int exact);
/* write_to_file, read_write.c:36-56 */
int fwi_write_f32(const char *path, const float *src, size_t n);

/* The file reads and index arithmetic of fwi_init, fwi_2d.c:320-371, into a
 * plan from fwi_plan_alloc.  Files (fwi_files_path.h:7-18, common ends in '/'):
 *   <common>input/shot2ms_<ishot>.bin      obs0[nr][nt0]  (if shots_avail)
 *   <common>input/sr_location_<ishot>.bin  source x,z then receiver x,z (m)
 *   <common>input/fbt_<ishot>.bin          fbt[nr]        (if fbt_avail)
 *   <common>input/initial_model.bin        md[nz][nx]
 *   <common>input/bottom.bin               btry[nx]
 *   wavelet (NULL: <common>input/wavelet_1ms.bin)  wav0[nt], times wav_scale
 * Then ibtry = btry/dz - 0.5 and ifbt = fbt/dt0 (C truncation), and the
 * source/receiver cells ix = round(x/dx), iz = round(z/dz - 0.5) (C round,
 * half away from zero).                                                       */
int fwi_plan_read_inputs(fwi_plan *plan, const char *common, int ishot, const char *wavelet, int shots_avail,
                         int fbt_avail, float wav_scale);

/* The input side of rwi_init (fwi_2d.c:637-688) for a plan of rwi_plan_alloc:
 * the files of fwi_plan_read_inputs except the initial model, and the
 * background and perturbation models m0[nz][nx], m1[nz][nx] from the paths
 * init_mdl_0 and init_mdl_1 (required: the reference reads them through
 * INIT_MDL_0 / INIT_MDL_1, which its fwi_files_path.h does not define); the
 * same index arithmetic.  0 or FWI_HIP_EINVAL (message in
 * fwi_io_last_error) where the reference exit()s.                            */
int rwi_plan_read_inputs(rwi_plan *plan, const char *common, int ishot, const char *init_mdl_0,
                         const char *init_mdl_1, const char *wavelet, int shots_avail, int fbt_avail, float wav_scale);

/* trcinp, spline_interp.c:19-38: each of the howmany rows obs0[r][0..nt0) on
 * the abscissae i*dt0 is resampled to obs[r][0..nt) at i*dt by the natural
 * cubic spline; the end samples are copied (spline_interp.c:13).  Rows are
 * row pointers, as alloc.c builds them.                                       */
int fwi_trcinp(float **obs0, int nt0, float dt0, float **obs, int nt, float dt, int howmany);

/* Adjoint source of grad4_1shot / grad_all_shots from data sampled every
 * jump steps: 0 = residual on the sampled steps (exact transpose of the
 * subsampling, the default), 1 = residual upsampled by fwi_trcinp.           */
int grad4_set_residual_mode(int mode);

const char *fwi_io_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
