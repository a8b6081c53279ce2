/*
 * orc.h -- CPU oracle for the TIMBIR MBIR_4D ICD/NHICD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm (adityamnk/timbir, src/MBIR_4D) used as the checker for the HIP
 * implementation in timbir_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it.  Nothing in the product links it.
 *
 * The library is compiled twice from the same source: liborc_f32.so keeps the
 * reference's float arrays (Real_arr_t = float, XT_Constants.h:43) and
 * liborc_f64.so stores every array in double (the "FP64 ICD oracle").  Scalar
 * arithmetic is double in both, as in the reference (Real_t, XT_Constants.h:42).
 *
 * Pinning status of each part is recorded in oracle/README.md.
 */
#ifndef ORC_H
#define ORC_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- schedules -------------------------------------------------------- */
/* ORC_SCHED_REFERENCE: the reference's own order: Park-Miller random voxel
 *   lists per (frame, z-block), z-blocks derived from a thread count exactly
 *   as XT_Init.c:389-398, randomized quickselect for NHICD (XT_NHICD.c).
 * ORC_SCHED_COLOUR: the schedule the HIP path runs: one z-block per slice,
 *   colour (frame + slice) & 1, a keyed Feistel permutation per
 *   (iteration, frame, slice) chunked in permutation-index space, and an exact
 *   deterministic order statistic for NHICD.  Same fixed point. */
#define ORC_SCHED_REFERENCE 0
#define ORC_SCHED_COLOUR 1

/* Options that the reference fixes at compile time or never exposes. */
typedef struct {
  int32_t schedule;         /* ORC_SCHED_* */
  int32_t num_threads;      /* reference schedule: stands for omp_get_max_threads() */
  int32_t compute_cost;     /* 1: behave as a build without -DNO_COST_CALCULATE */
  int32_t abort_on_cost_increase; /* XT_ICD_update.c:1185 when cost is on */
  int32_t max_iterations;   /* MAX_NUM_ITERATIONS (XT_Constants.h:86) if <= 0 */
  int32_t write_files;      /* write the reference's .bin outputs into cwd */
  int32_t no_nhicd;         /* NO_NHICD (XT_Constants.h:91) */
  uint32_t order_seed;      /* colour schedule permutation seed */
  int32_t hicd_order;       /* colour schedule: a HICD chunk's pixels in
                               permutation order (0) or in raster order (1) */
} orc_options;

/* ---- RNG: randlib.c:119-174 -------------------------------------------- */
uint32_t orc_pm_next(uint32_t state);                /* 16807*s mod (2^31-1) */
void orc_random2_stream(uint32_t seed, int64_t n, double *out);

/* ---- permutation used by the colour schedule ----------------------------- */
uint32_t orc_perm_key(uint32_t seed, int32_t iter, int32_t frame, int32_t slice);
uint32_t orc_perm(uint32_t n, uint32_t N, uint32_t key);
uint32_t orc_perm_inv(uint32_t n, uint32_t N, uint32_t key);

/* ---- system model ------------------------------------------------------ */
/* XT_Profile.c:47-176 (DetectorResponseProfile) -> H_r[N_p][65] */
void orc_detector_response(const float *angles, int32_t N_p, double delta_xy,
                           double delta_r, double OffsetR, double beam_width,
                           float *H_r);
/* XT_AMatrix.c:50-110 (calcAMatrixColumnforAngle) */
int32_t orc_amatrix_column(const float *H_r_row, double cosv, double sinv,
                           int32_t row, int32_t col, double x0, double y0,
                           double delta_xy, double R0, double RMax,
                           double delta_r, int32_t N_r, double OffsetR,
                           double *values, int32_t *index);

/* ---- qGGMRF prior: XT_ICD_update.c:83-129 ------------------------------- */
double orc_qggmrf_value(double delta, double sigma_q, double sigma_q_p, double c);
double orc_qggmrf_derivative(double delta, double sigma_q, double sigma_q_p, double c);
double orc_qggmrf_second_derivative(double sigma_q, double c);

/* ---- stage problem ------------------------------------------------------ */
typedef struct orc_problem orc_problem;

/* Builds one multi-resolution stage exactly as initStatStructures +
 * initDynStructures (XT_Init.c:211-669) would, for a single rank.
 * projections / weights are in the reference input layout
 * [proj_num][proj_rows][proj_cols] and are transposed internally
 * (XT_Init.c:298-305).  var_est divides the weights (XT_Init.c:327-330).
 * stage_idx chooses initICD / updateProjOffset / updateVar / initMagUpMap
 * (XT_Init.c:230-277) with stage 0 the coarsest; prev_mult_z is the previous
 * stage's z multiplier (ignored for stage 0). */
orc_problem *orc_problem_new(const float *projections, const float *weights,
                             const float *proj_angles, const float *proj_times,
                             const float *recon_times, int32_t proj_rows,
                             int32_t proj_cols, int32_t proj_num,
                             int32_t recon_num, float vox_wid, float rot_center,
                             float sig_s, float sig_t, float c_s, float c_t,
                             float convg_thresh, int32_t remove_rings,
                             int32_t quad_convex, float huber_delta,
                             float huber_T, int32_t mult_xy, int32_t mult_z,
                             int32_t stage_idx, int32_t prev_mult_z,
                             double var_est, const orc_options *opt);
void orc_problem_free(orc_problem *p);

/* The pixels of cell (f, z)'s HICD chunk of permutation indices [n0, n1) in
 * the colour schedule's visit order (p's hicd_order); returns the count.
 * xs and ys hold N_x * N_y entries. */
int32_t orc_colour_chunk(const orc_problem *p, int32_t iter, int32_t f, int32_t z, int32_t n0,
                         int32_t n1, int32_t *xs, int32_t *ys);

/* Scalars of the stage, for tests.  ints: N_p N_r N_t N_x N_y N_z N_time
 * mult_xy mult_z num_z_blocks updateProjOffset updateVar initICD U(masked)
 * ; dbl: delta_r delta_t R0 RMax T0 x0 y0 delta_xy delta_z OffsetR
 * radius_obj sigS_q sigS_qp sigT_q sigT_qp c_s c_t tfilt alpha T delta
 * var_est delta_recon sfilt[27] */
void orc_problem_info(const orc_problem *p, int32_t *ints, double *dbls);

/* Raw array access (element type is float in liborc_f32, double in
 * liborc_f64, except H_r which is always float and b/mask which are uint8). */
void *orc_obj(orc_problem *p);       /* [N_time][N_z][N_y][N_x] */
void *orc_err(orc_problem *p);       /* [N_p][N_r][N_t] e */
void *orc_wgt(orc_problem *p);       /* [N_p][N_r][N_t] w (divided by var_est) */
void *orc_proj(orc_problem *p);      /* [N_p][N_r][N_t] y */
uint8_t *orc_sel(orc_problem *p);    /* [N_p][N_r][N_t] ProjSelect b */
void *orc_off(orc_problem *p);       /* [N_r][N_t] d */
float *orc_hr(orc_problem *p);       /* [N_p][65] */
uint8_t *orc_mask(orc_problem *p);   /* [N_y][N_x] */
float *orc_mag(orc_problem *p);      /* [N_time][num_z_blocks][N_y][N_x] */
const int32_t *orc_view_off(orc_problem *p); /* [N_time+1] */
const int32_t *orc_view_idx(orc_problem *p); /* views of each frame */
double orc_var_est(const orc_problem *p);
void orc_set_var_est(orc_problem *p, double v);
int64_t orc_zero_count(const orc_problem *p);
int64_t orc_update_count(const orc_problem *p);  /* voxel updates performed */
/* RNG state (random2) and per-cell selection counts, for known-answer tests */
uint32_t orc_rng_state(const orc_problem *p);
void orc_set_rng_state(orc_problem *p, uint32_t s);
void orc_set_select_counts(orc_problem *p, int32_t n);

/* XT_ICD_update.c:655-716 */
void orc_forward_project(orc_problem *p, void *out /* [N_p][N_r][N_t] */);
void orc_init_error_sino(orc_problem *p);
/* XT_ICD_update.c:186-319; returns total cost, fills parts if non-NULL */
double orc_cost(orc_problem *p, double *forward, double *prior);
/* XT_ICD_update.c:919-927 dispatch on remove_rings */
void orc_offset_update(orc_problem *p, int32_t type);
/* XT_ICD_update.c:720-751 */
void orc_variance_update(orc_problem *p);
/* XT_ICD_update.c:134-183 for the voxel (frame, slice, y, x) given the
 * current object, with theta1/theta2 supplied by the caller */
double orc_surrogate_update(orc_problem *p, int32_t frame, int32_t slice,
                            int32_t y, int32_t x, double theta1, double theta2);
/* XT_VoxUpdate.c:43-115 for one voxel; returns the new value */
double orc_voxel_update(orc_problem *p, int32_t frame, int32_t slice,
                        int32_t y, int32_t x);
/* XT_VoxUpdate.c:117-281 over an explicit list of (x,y) lines */
double orc_update_voxels(orc_problem *p, int32_t frame, int32_t z_begin,
                         int32_t z_end, int32_t n, const int32_t *xs,
                         const int32_t *ys, int32_t iter, int32_t block);
/* XT_ICD_update.c:936-1105: one ICD iteration.  Returns 1 when the
 * reference's convergence rule fires; *ratio gets the percentage. */
int32_t orc_icd_iteration(orc_problem *p, int32_t iter, double *ratio);
/* One colour phase of updateVoxelsTimeSlices (XT_ICD_update.c:977-1053):
 * phase 0/1, HICD chunk `chunk` (hicd=1) or the NHICD lists (hicd=0).
 * Returns the summed |x| of the updated voxels.  For the colour schedule
 * the per-iteration selection counts must be set (orc_set_select_counts). */
double orc_run_phase(orc_problem *p, int32_t phase, int32_t iter, int32_t chunk, int32_t hicd);
/* XT_NHICD.c:164-216 for every (frame, block) */
void orc_vsc_select(orc_problem *p);
int32_t orc_nhicd_count(const orc_problem *p, int32_t frame, int32_t block,
                        int32_t *xs, int32_t *ys);
/* XT_ICD_update.c:323-651 (initObject, initICD 2/3 and MagUpdateMap) from a
 * previous stage's object [N_time][Nz'][Ny/2][Nx/2] and map. */
void orc_init_from_coarse(orc_problem *p, const void *prev_obj, int32_t prev_nz,
                          const float *prev_mag, int32_t prev_blocks);
/* ICD_BackProject loop (XT_ICD_update.c:1108-1244) after the object is set.
 * cost_out (may be NULL) receives cost per iteration, index 0 = initial. */
int32_t orc_stage_run(orc_problem *p, int32_t max_iter, double *cost_out,
                      double *ratio_out, int32_t *iters);

/* Whole reconstruct() (XT_Engine.c:81-203) with the options above.  Writes
 * the reference's output files into cwd when opt->write_files.  *object_out
 * receives malloc'd [recon_num][proj_rows][proj_cols][proj_cols] (f32 lib:
 * float, f64 lib: double). stage_costs receives final cost per stage. */
int orc_reconstruct(void **object_out, const float *projections,
                    const float *weights, const float *proj_angles,
                    const float *proj_times, const float *recon_times,
                    int32_t proj_rows, int32_t proj_cols, int32_t proj_num,
                    int32_t recon_num, float vox_wid, float rot_center,
                    float sig_s, float sig_t, float c_s, float c_t,
                    float convg_thresh, int32_t remove_rings,
                    int32_t quad_convex, float huber_delta, float huber_T,
                    const orc_options *opt, double *final_cost,
                    int32_t *total_iters);
void orc_free(void *ptr);

#ifdef __cplusplus
}
#endif
#endif
