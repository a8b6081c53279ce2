/*
 * timbir_dev.h -- C ABI of the GPU-resident MBIR_4D hot path (libtimbir_hip).
 *
 * One tb_stage is one multi-resolution stage of reconstruct() held in the
 * GPU's memory: error sinogram, weights, offsets, object, magnitude-update
 * map and the detector response table.  Each entry point below replaces one
 * reference function (cited as file:line under adityamnk/timbir/src/MBIR_4D);
 * reconstruct() in mbir4d.h is built from them.  Every function returns 0 on
 * success and -1 on failure (tb_last_error() then describes it), the
 * reference's check_error convention (XT_Debug.h:22).
 *
 * Arrays crossing this boundary are host pointers in the reference's own
 * layouts:
 *   raw projections / weights : float [proj_num][proj_rows][proj_cols]
 *   object                    : float [N_time][N_z][N_y][N_x]
 *   sinogram-shaped outputs   : float [N_p][N_r][N_t]  (XT_Init.c:298-305)
 *   offsets                   : float [N_r][N_t]
 *   magnitude-update map      : float [N_time][N_z][N_y][N_x] (one z-block
 *                               per slice)
 * No torch or HIP type appears here.
 */
#ifndef TIMBIR_DEV_H_
#define TIMBIR_DEV_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct tb_stage tb_stage;

/* Options the reference fixes with compile-time macros (XT_Constants.h,
 * src/MBIR_4D/Makefile:37).  Defaults: zeros except max_iterations = 1000,
 * write_files = 1 and verbose = 1. */
typedef struct {
  int32_t compute_cost;           /* 1 = build without -DNO_COST_CALCULATE */
  int32_t abort_on_cost_increase; /* XT_ICD_update.c:1185 */
  int32_t max_iterations;         /* MAX_NUM_ITERATIONS, XT_Constants.h:86 */
  int32_t no_nhicd;               /* NO_NHICD, XT_Constants.h:91 */
  int32_t write_files;            /* write the .bin outputs (default 1) */
  uint32_t order_seed;            /* seed of the per-cell voxel order */
  int32_t verbose;                /* [INFO] lines on the debug stream */
  int32_t kernel;                 /* voxel-update kernel: 0 = the fastest
                                     that applies (LDS-resident multi-
                                     wavefront, then multi-wavefront with e
                                     in global memory, then wide-column
                                     multi-wavefront, then pass-based,
                                     then one-wavefront, then one-voxel),
                                     1 = the one-voxel
                                     kernel only, 2 = the one-wavefront LDS
                                     kernel where it fits, 3 = the
                                     multi-wavefront LDS kernel where it
                                     fits, 4 = the multi-wavefront kernel
                                     with e in global memory where it
                                     applies, 5 = the wide-column
                                     multi-wavefront kernels (columns of
                                     up to 17 bins, views in passes when
                                     there are many) where they apply,
                                     6 = the pass kernel that holds its
                                     passes in registers alone, 7 = the
                                     pass kernel that walks any number of
                                     passes in groups alone, 8 = the
                                     deep-pipeline multi-wavefront kernel
                                     with e in global memory where it
                                     applies (3- and 5-bin columns), 9 =
                                     the batched finest-stage kernel
                                     (3-bin columns, one detector row per
                                     slice, up to 315 views per cell), 10 =
                                     the split-cell kernel (each cell's
                                     views over several workgroups of one
                                     cooperative launch; kernel 0 takes it
                                     for colours of at most 32 cells) */
  int32_t views_per_lane;         /* views per theta-wave lane of the
                                     multi-wavefront LDS kernel: 0 = the
                                     fewest waves that keep one wavefront
                                     per SIMD, else 1 or 2 */
  int32_t init_fbp;               /* 1 = reconstruct() starts stage 0 from
                                     tb_init_fbp instead of zero (default 0,
                                     the reference's start) */
  int32_t no_column_table;        /* 1 = build no A-column table: the
                                     kernels that read it are skipped, the
                                     wide-column and pass-based ones
                                     evaluate the same values (default 0) */
  int32_t theta_fp;               /* theta1 / theta2 sums of the voxel
                                     update (XT_VoxUpdate.c:61-92, double
                                     in the reference): 0 = theta1 in FP64,
                                     theta2 in FP32 (default); 1 = both in
                                     FP32 with compensated (TwoSum)
                                     accumulation; 2 = both in plain FP32.
                                     1 and 2 exist in the multi-wavefront
                                     global-e kernel for 3- and 5-bin
                                     columns (the two finest stages); a
                                     stage it does not serve fails */
  int32_t row_lookahead;          /* voxels whose (e, w) rows are in
                                     flight while one is solved, in the
                                     deep-pipeline kernel: 0 = 3
                                     (default), 1, 3 or 7 */
  int32_t autotune;               /* 1 = with kernel 0, use the kernel
                                     configuration the autotuner measured
                                     fastest for the stage's shape, timing
                                     it on first use (default 0) */
  int32_t split_segments;         /* workgroups per cell of the split-cell
                                     kernel: 0 = as many as the compute
                                     units allow (default), 1 = never use
                                     it (kernel 0 takes the other
                                     families), else at most this many
                                     (2 to 64) */
  int32_t split_spin_limit;       /* polls of one cross-workgroup wait of
                                     the split-cell kernel before the wait
                                     is abandoned and the call fails:
                                     0 = 2^22 (about a second, default) */
  int32_t hicd_order;             /* the order a HICD chunk's pixels are
                                     visited in: each chunk holds the pixels
                                     of one range of the cell's keyed random
                                     permutation (as the reference's chunks
                                     hold ranges of its random list,
                                     XT_ICD_update.c:983-986); 0 = in
                                     permutation order (default), 1 = in
                                     raster order (neighbouring voxels share
                                     cache lines: faster on large cells) */
} tb_options;

void tb_get_options(tb_options *opt);
void tb_set_options(const tb_options *opt);
const char *tb_last_error(void);

/* initStatStructures + initDynStructures (XT_Init.c:211-669),
 * DetectorResponseProfile (XT_Profile.c:129-176) and the object mask
 * (XT_ICD_update.c:1140-1149) for one stage.  Uploads the raw scan; weights
 * are divided by var_est.  stage_idx counts from the coarsest stage (0);
 * prev_mult_z is the previous stage's z multiplier. */
int tb_stage_create(tb_stage **out, const float *projections,
                    const float *weights, const float *proj_angles,
                    const float *proj_times, const float *recon_times,
                    int32_t proj_rows, int32_t proj_cols, int32_t proj_num,
                    int32_t recon_num, float vox_wid, float rot_center,
                    float sig_s, float sig_t, float c_s, float c_t,
                    float convg_thresh, int32_t remove_rings,
                    int32_t quad_convex, float huber_delta, float huber_T,
                    int32_t mult_xy, int32_t mult_z, int32_t stage_idx,
                    int32_t prev_mult_z, double var_est);
void tb_stage_destroy(tb_stage *st);
/* Ends the cross-workgroup waits of the split-cell kernel launch running on
 * the stage (from any host thread, through a stream of its own): its
 * workgroups stop waiting, finish with no further steps, and the call that
 * launched it fails ("k_icd_split: a cross-workgroup wait timed out"); the
 * stage's object and error sinogram are then invalid and the stage is to be
 * destroyed.  A split-cell launch that starts later clears the request;
 * other stages are not affected.  No counterpart in the reference (its threads' waits
 * are OpenMP barriers). */
int tb_stage_abort(tb_stage *st);

/* The multi-resolution schedule reconstruct() runs (XT_Engine.c:111-141,
 * one node): writes (mult_xy, mult_z) of each stage, coarse to fine, into
 * arrays of at least 8 entries and returns the number of stages (-1 on
 * sizes that are not positive).  Needs no GPU. */
int32_t tb_multires_schedule(int32_t proj_rows, int32_t proj_cols, int32_t *mult_xy,
                             int32_t *mult_z);

/* ints: N_p N_r N_t N_x N_y N_z N_time mult_xy mult_z num_z_blocks
 * updateProjOffset updateVar initICD U ; dbls as the oracle's
 * orc_problem_info (delta_r .. delta_recon, then 27 filter weights). */
int tb_stage_info(const tb_stage *st, int32_t *ints, double *dbls);

/* Object / offset / map transfer: initObject (XT_ICD_update.c:541-651) and
 * write_ObjectProjOff2TiffBinPerIter (XT_IOMisc.c:238-271). */
int tb_set_object(tb_stage *st, const float *obj);
int tb_get_object(tb_stage *st, float *obj);
int tb_set_offset(tb_stage *st, const float *d);
int tb_get_offset(tb_stage *st, float *d);
int tb_set_magmap(tb_stage *st, const float *mag);
int tb_get_magmap(tb_stage *st, float *mag);
/* initObject initICD = 2 / 3 (XT_ICD_update.c:323-651) from the previous
 * stage's object and map held by the caller (host arrays, as the reference
 * reads them back from object_time_*.bin / mag_update.bin): bilinear x2 in
 * x-y (and z when mult_z halves). */
int tb_init_from_coarse(tb_stage *st, const float *prev_obj, int32_t prev_nz,
                        const float *prev_mag);
/* The same handoff from a previous stage still in HBM (initObject without the
 * file round trip of XT_Engine.c:156-182): same values, bit for bit. */
int tb_init_from_stage(tb_stage *st, tb_stage *prev);
int tb_get_error_sino(tb_stage *st, float *e);  /* [N_p][N_r][N_t] */
int tb_get_weights(tb_stage *st, float *w);     /* [N_p][N_r][N_t] */
int tb_get_detector_response(tb_stage *st, float *H_r); /* [N_p][65] */
double tb_get_var_est(const tb_stage *st);

/* initErrorSinogam (XT_ICD_update.c:655-716): e = y - A x - d with the
 * forward projection of forward_project_voxel_AMat1D (XT_ForwardProject.c:8-25). */
int tb_init_error_sino(tb_stage *st);
/* tb_init_error_sino for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same result, bit for bit;
 * calls no hipMalloc / hipFree: the forward projector's slice-interleaved
 * object copy goes to a static per-GPU array of 64 MiB, so the stage must
 * satisfy N_time x ceil(N_z / 8) x N_y x N_x x 32 B <= 64 MiB (-1 and an
 * error message otherwise).  Runs on the device the stage was created on.
 * Calls on one GPU share that array: calls that overlap in time on one GPU
 * (stages on different streams or host threads) race and must be
 * serialised by the caller. */
int tb_init_error_sino_static(tb_stage *st);
/* Forward projection alone (no y, no d): out = A x, [N_p][N_r][N_t]. */
int tb_forward_project(tb_stage *st, float *out);
/* Filtered back-projection of the stage's measurements into its object:
 * y - d averaged over each slice's detector rows, Ram-Lak filter at the bin
 * pitch, each frame from its own views (pi / views weight), linear
 * interpolation, zero outside the mask.  No counterpart in the reference,
 * whose stage 0 starts from OBJECT_INIT_VAL = 0 (XT_ICD_update.c:549-553);
 * reconstruct() uses it for stage 0 when tb_options.init_fbp is set. */
int tb_init_fbp(tb_stage *st);

/* computeCost (XT_ICD_update.c:186-319), FP64 reductions. */
int tb_cost(tb_stage *st, double *total, double *forward, double *prior);

/* update_Sinogram_Offset (XT_ICD_update.c:919-927) with the given type
 * (1 unconstrained :870, 2 zero-mean :817, 3 patch-constrained with no
 * patches :753). */
int tb_offset_update(tb_stage *st, int32_t type);
/* update_variance_parameter (XT_ICD_update.c:720-751). */
int tb_variance_update(tb_stage *st);

/* compute_voxel_update_Atten (XT_VoxUpdate.c:43-115) for a single voxel,
 * gathering its neighbourhood as updateVoxels_Atten does (:166-233).
 * Returns the new value in *value. */
int tb_voxel_update(tb_stage *st, int32_t frame, int32_t slice, int32_t y,
                    int32_t x, float *value);
/* updateVoxels (XT_VoxUpdate.c:284-289) for every cell of one colour:
 * HICD chunk `chunk` (0..9) of the iteration's voxel order, or the NHICD
 * lines when nhicd != 0.  Cells (frame f, slice z) of colour (f + z) & 1. */
int tb_update_colour(tb_stage *st, int32_t colour, int32_t iter,
                     int32_t chunk, int32_t nhicd);
/* VSC_based_Voxel_Line_Select (XT_NHICD.c:164-216). */
int tb_vsc_select(tb_stage *st);
/* NHICD list of cell (frame, slice), raster order; returns the count. */
int tb_nhicd_list(tb_stage *st, int32_t frame, int32_t slice, int32_t *xs,
                  int32_t *ys);
/* updateVoxelsTimeSlices (XT_ICD_update.c:936-1105): one ICD iteration.
 * *converged = the reference's convergence flag, *ratio its percentage. */
int tb_icd_iteration(tb_stage *st, int32_t iter, int32_t *converged,
                     double *ratio);
/* ICD_BackProject's loop (XT_ICD_update.c:1108-1244) after the object is
 * set: error sinogram, iterations until the reference's rule (or
 * max_iter), cost per iteration when enabled.  costs (max_iter+1 entries)
 * may be NULL. */
int tb_stage_run(tb_stage *st, int32_t max_iter, double *costs,
                 int32_t *iters);

/* Accounting of the GPU work since the last reset, measured with HIP
 * events on the stage's stream: the voxel-update kernel time (ms), the
 * voxel updates performed (zero-skipped voxels excluded), and the
 * algorithmic bytes they moved (12 B per touched A-matrix entry: e read,
 * w read, e write).  Other kernels' time in other_ms.  lds_launches counts
 * the voxel-update launches that ran an LDS-resident kernel, global_launches
 * those that ran a multi-wavefront kernel with e in global memory (the
 * others ran the one-voxel-at-a-time kernel); family_launches splits them
 * by kernel.  The launch counters restart at tb_reset_stats, the table
 * counters do not. */
/* Voxel-update kernel families, the index of tb_stats.family_launches. */
enum {
  TB_K_ONE_VOXEL = 0, /* k_update_voxels: one voxel at a time */
  TB_K_LDS = 1,       /* k_icd_lds: one wavefront per cell, (e, w) in LDS */
  TB_K_MW = 2,        /* k_icd_mw: several wavefronts, (e, w) in LDS */
  TB_K_MWG = 3,       /* k_icd_mwg: several wavefronts, e in global memory */
  TB_K_WIDE = 4,      /* k_icd_wide: columns of up to 17 bins */
  TB_K_WPASS = 5,     /* k_icd_wpass: views in passes held in registers */
  TB_K_WLOOP = 6,     /* k_icd_wloop: views in groups of passes */
  TB_K_MWD = 7,       /* k_icd_mwd: as k_icd_mwg, rows of the next voxels
                         in flight */
  TB_K_BAT = 8,       /* k_icd_bat: the finest stage in batches of voxels */
  TB_K_SPLIT = 9,     /* k_icd_split: each cell's views over several workgroups */
  TB_KERNEL_FAMILIES = 12
};
typedef struct {
  double voxel_ms;
  double other_ms;
  int64_t voxel_launches;
  int64_t updates;
  int64_t zero_skips;
  int64_t entries;
  double alg_bytes;
  int64_t lds_launches;
  int64_t global_launches;
  int64_t family_launches[TB_KERNEL_FAMILIES]; /* launches per TB_K_* */
  int64_t column_builds; /* A-column tables built since the stage was created */
  int32_t column_bins;   /* bins per record of the table held now (0: none) */
  int32_t reserved;
} tb_stats;
int tb_get_stats(tb_stage *st, tb_stats *out);

/* Batches of independent problems (no counterpart in the reference, whose
 * concurrency unit is one problem's cells, XT_ICD_update.c:977-991): the
 * stages of many scans advanced one ICD iteration at a time in lockstep on
 * one stream.  Each voxel-update phase is one launch per kernel
 * configuration over the cells of every member that uses it (members whose
 * kernel has no multi-problem form launch their own); a member's kernels,
 * arguments and order are those of tb_icd_iteration, so its results are
 * bit for bit the same.  With TB_BATCH_GRAPH the launches of iterations 2
 * on are captured once in a HIP graph and replayed while the members and
 * their kernel choices stay the same.  Voxel-kernel time is not split per
 * member (tb_stats.voxel_ms stays 0 for batched launches); the batch keeps
 * its own totals. */
typedef struct tb_batch tb_batch;
enum { TB_BATCH_GRAPH = 1 };
typedef struct {
  int64_t iterations;    /* tb_batch_iteration calls */
  int64_t launches;      /* voxel-update launches made */
  int64_t captures;      /* graphs captured */
  int64_t graph_replays; /* iterations run by replaying a graph */
  int64_t groups;        /* multi-problem launch groups of the last call */
  int64_t solo;          /* members launching their own kernel (last call) */
  double ms;             /* GPU time of the iterations' launches */
} tb_batch_stats;
int tb_batch_create(tb_batch **out, int32_t flags);
void tb_batch_destroy(tb_batch *b);
/* One ICD iteration (tb_icd_iteration) of each of the n stages; converged
 * and ratio (n entries each, may be NULL) receive each member's stop flag
 * and ratio.  The stages must be distinct and on the batch's device. */
int tb_batch_iteration(tb_batch *b, tb_stage *const *stages, int32_t n, int32_t iter,
                       int32_t *converged, double *ratio);
int tb_batch_get_stats(const tb_batch *b, tb_batch_stats *out);

/* Autotuner (no counterpart in the reference): per stage shape, the kernel
 * option (2-9) and views per lane whose HICD colour phase ran fastest on
 * this GPU, timed on the stage's own data with every trial undone.
 * tb_autotune tunes the stage's shape (or returns the cached choice);
 * tb_autotune_entries copies up to cap cached entries and returns how many
 * there are; tb_autotune_put adds one (a saved cache); tb_autotune_clear
 * empties the cache.  With tb_options.autotune, launches of kernel 0 use
 * it. */
typedef struct {
  int32_t shape[8]; /* mult_xy, mult_z, views per frame, N_r, N_x, N_y, N_z, N_time */
  int32_t kernel;   /* tb_options.kernel of the fastest configuration */
  int32_t views_per_lane;
  float ms;         /* its time for one HICD colour phase */
  int32_t reserved;
} tb_tune_entry;
int tb_autotune(tb_stage *st, tb_tune_entry *out);
int tb_autotune_entries(tb_tune_entry *out, int32_t cap);
int tb_autotune_put(const tb_tune_entry *e);
void tb_autotune_clear(void);
int tb_reset_stats(tb_stage *st);
int tb_synchronize(tb_stage *st);

/* Device address of the stage's object, float [N_time][N_z][N_y][N_x] in
 * HBM (*bytes its size), valid until tb_stage_destroy.  For initialisers
 * that write a starting object on the GPU without a host round trip, as
 * initObject's initICD = 1 path reads one without interpolation
 * (XT_ICD_update.c:558-571); work on the stage's data must be ordered after
 * it with tb_synchronize. */
int tb_stage_object_device(tb_stage *st, void **ptr, uint64_t *bytes);

/* ---- learned initialiser: kernels of a 3D U-Net (tb_unet.hip) ----------
 * No counterpart in the reference (its first stage starts from zero,
 * XT_ICD_update.c:549-553).  Pointers are device addresses (tb_dev_malloc);
 * `stream` is a hipStream_t of this library's runtime (NULL: the default
 * stream); every call is asynchronous on it.
 * Activations and activation gradients are bf16 [n][d][h][w][c] (channels
 * last, c a multiple of 32); convolution weights bf16 [27][cout][cin] with
 * tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1); biases, weight gradients
 * and optimiser state fp32.  Return 0, or -1 with tb_nn_last_error(). */
const char *tb_nn_last_error(void);
/* Device memory from this library's HIP runtime, for the buffers of these
 * kernels: allocate / free, blocking copies (kind 0 host to device, 1
 * device to host, 2 device to device), zero fill on a stream, and a wait
 * for all work on the device. */
int tb_dev_malloc(uint64_t bytes, void **ptr);
int tb_dev_free(void *ptr);
int tb_dev_copy(void *dst, const void *src, uint64_t bytes, int32_t kind);
int tb_dev_zero(void *ptr, uint64_t bytes, void *stream);
int tb_dev_synchronize(void);
/* 3x3x3 convolution with zero padding on the matrix cores (bf16 x bf16,
 * fp32 accumulate): out = [ReLU](bias + conv(in, w)); out fp32 when
 * out_f32, else bf16.  bias may be NULL. */
int tb_nn_conv3d(const void *in, const void *w, const float *bias, void *out, int32_t n,
                 int32_t d, int32_t h, int32_t wd, int32_t cin, int32_t cout, int32_t relu,
                 int32_t out_f32, void *stream);
/* wt[t][ci][co] = w[26 - t][co][ci]: with it tb_nn_conv3d maps an output
 * gradient to the input gradient. */
int tb_nn_conv3d_flip(const void *w, void *wt, int32_t cin, int32_t cout, void *stream);
/* gw[tap][co][ci] = sum over voxels of gout[v][co] in[v + off(tap)][ci]
 * (fp32); part holds slices x 27 x cin x cout floats of partial sums, added
 * in a fixed order. */
int tb_nn_conv3d_wgrad(const void *in, const void *gout, float *gw, float *part,
                       int32_t slices, int32_t n, int32_t d, int32_t h, int32_t wd,
                       int32_t cin, int32_t cout, void *stream);
/* out[c] = sum over nvox voxels of g[v][c] (bias gradient); part holds
 * slices x c floats. */
int tb_nn_channel_sum(const void *g, float *out, float *part, int32_t slices, int64_t nvox,
                      int32_t c, void *stream);
/* g[i] = 0 where the stored ReLU output y[i] is not positive. */
int tb_nn_relu_bwd(const void *y, void *g, int64_t count, void *stream);
/* 2x2x2 average pooling of [n][d][h][wd][c] (backward: in = the pooled
 * gradient, out = the full-size gradient). */
int tb_nn_avgpool2(const void *in, void *out, int32_t n, int32_t d, int32_t h, int32_t wd,
                   int32_t c, int32_t backward, void *stream);
/* out = [nearest x2 upsampling of a (ca channels), b (cb channels)] at size
 * n x d x h x wd, and its backward (ga at half size, gb at full size). */
int tb_nn_up2_concat(const void *a, const void *b, void *out, int32_t n, int32_t d, int32_t h,
                     int32_t wd, int32_t ca, int32_t cb, void *stream);
int tb_nn_up2_concat_bwd(const void *g, void *ga, void *gb, int32_t n, int32_t d, int32_t h,
                         int32_t wd, int32_t ca, int32_t cb, void *stream);
/* a[i] += b[i] (bf16). */
int tb_nn_add(void *a, const void *b, int64_t count, void *stream);
/* Network input from fp32 volumes: channel 0 = scale * fbp[v], channel 1 =
 * scale * mean[v mod per_sample] (mean may be NULL), channels 2-31 zero. */
int tb_nn_make_input(const float *fbp, const float *mean, void *x, int64_t nvox,
                     int64_t per_sample, float scale, void *stream);
/* mean[q] = (1 / nframes) sum_f x[f * per + q] (fp32). */
int tb_nn_frame_mean(const float *x, float *mean, int64_t nframes, int64_t per, void *stream);
/* Output head: out[v] = fbp[v] + (b + sum_c w[c] h[v][c]) / scale (fp32),
 * and its backward (gh bf16; part = slices x (c + 1) partial sums of the
 * weight and bias gradients, added with tb_nn_sum_parts). */
int tb_nn_head(const void *h, const float *w, const float *b, const float *fbp, float *out,
               int64_t nvox, int32_t c, float scale, void *stream);
int tb_nn_head_bwd(const void *h, const float *w, const float *gout, void *gh, float *part,
                   int32_t slices, int64_t nvox, int32_t c, float scale, void *stream);
/* Mean squared error: gout[v] = 2 (out[v] - target[v]) / nvox, part[b] the
 * fp64 sums of squares of `blocks` blocks. */
int tb_nn_mse(const float *out, const float *target, float *gout, double *part, int32_t blocks,
              int64_t nvox, void *stream);
/* out[i] = sum over s < slices of part[s * count + i], in order. */
int tb_nn_sum_parts(const float *part, float *out, int64_t count, int32_t slices, void *stream);
/* Adam step `step` (from 1) on fp32 parameters p with gradient g and moments
 * m, v; wb (may be NULL) receives the bf16 copy. */
int tb_nn_adam(float *p, const float *g, float *m, float *v, void *wb, int64_t count, float lr,
               float b1, float b2, float eps, int32_t step, void *stream);
int tb_nn_to_bf16(const float *a, void *b, int64_t count, void *stream);

#ifdef __cplusplus
}
#endif
#endif
