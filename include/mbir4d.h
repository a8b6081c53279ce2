/*
 * mbir4d.h -- drop-in replacement for the MBIR_4D library entry point.
 *
 * Replaces: src/include/mbir4d.h:34 (reference), defined at
 * src/MBIR_4D/XT_Engine.c:81.  Same symbol, argument order, argument meaning,
 * preconditions, return codes (0 / -1), ownership (projections and weights
 * are free()d after the first stage is set up; proj_angles, proj_times and
 * recon_times are free()d on success) and output files in the current
 * directory (object_time_%d.bin, proj_offset.bin, variance_estimate.bin,
 * mag_update.bin, status.bin, projs_runtime.bin, weights_runtime.bin, and
 * cost.bin when the MAP cost is enabled).
 *
 * Differences a caller can observe, all compatible supersets:
 *   - *object receives a malloc()ed float array
 *     [recon_num][proj_rows][proj_cols][proj_cols] holding the finest stage;
 *     the reference leaves it unset (XT_Engine.c:178-179).
 *   - MPI is gone: the whole scan is reconstructed by one process on one GPU,
 *     so the "nodes" of the reference's preconditions is 1.
 *   - restart=1 re-runs the stage recorded in status.bin, as the
 *     reference does (XT_Engine.c:143-157), but starts it from that stage's
 *     own files (object_time_*.bin, mag_update.bin, proj_offset.bin,
 *     variance_estimate.bin) as they are, where the reference interpolates
 *     them as if they came from the stage before (a behavioural
 *     difference: the reference's interpolation is mis-sized there).
 *
 * Link: -ltimbir_hip (libtimbir_hip.so) instead of -lmbir4d.
 */
#ifndef MBIR4D_H_
#define MBIR4D_H_
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

int reconstruct(float **object, float *projections, float *weights,
                float *proj_angles, float *proj_times, float *recon_times,
                int32_t proj_rows, int32_t proj_cols, int32_t proj_num,
                int32_t recon_num, float vox_wid, float rot_center,
                float sig_s, float sig_t, float c_s, float c_t,
                float convg_thresh, int32_t remove_rings, int32_t quad_convex,
                float huber_delta, float huber_T, uint8_t restart,
                FILE *debug_msg_ptr);

#ifdef __cplusplus
}
#endif
#endif
