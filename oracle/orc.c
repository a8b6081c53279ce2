/*
 * orc.c -- CPU oracle for the TIMBIR MBIR_4D ICD/NHICD path.
 *
 * TEST INFRASTRUCTURE ONLY (see orc.h).  Every function below cites the
 * reference file:line it restates.  Compile with -DORC_ARR_T=float for the
 * reference-faithful oracle or -DORC_ARR_T=double for the FP64 ICD oracle.
 */
#define _GNU_SOURCE
#include "orc.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef ORC_ARR_T
#define ORC_ARR_T float
#endif
typedef ORC_ARR_T arr_t;

/* XT_Constants.h:62-102 */
#define ORC_P 1.2
#define ORC_Q 2.0
#define ORC_PROFILE_RES 1536
#define ORC_BEAM_RES 512
#define ORC_NBINS 64
#define ORC_ALPHA 1.5
#define ORC_MAX_ITERS 1000
#define ORC_ZINGER_OFF_T 100000.0
#define ORC_ZINGER_OFF_DELTA 1.0
#define ORC_MIN_XY_RES 64
#define ORC_MAX_MULTRES 8
#define ORC_MIN_ROWS_PER_NODE 2
#define ORC_MIN_PROJ_ROWS 4
#define ORC_NHICD_K 10 /* XT_Init.c:645 */
#define ORC_PI 3.14159265358979323846
#define ORC_PI_2 1.57079632679489661923132169163975144
#define ORC_PI_4 0.785398163397448309615660845819875721

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) {
    fprintf(stderr, "orc: out of memory\n");
    exit(-1); /* allocate.c:42-45 convention */
  }
  return p;
}

/* ======================================================================= */
/* RNG                                                                     */
/* ======================================================================= */
/* randlib.c:119-141: minimal-standard generator, new = 16807*old mod 2^31-1,
 * returned as new/(2^31-1).  The reference keeps one global state seeded by
 * srandom2(761521) (XT_Engine.c:99); the oracle keeps it per problem. */
uint32_t orc_pm_next(uint32_t s) {
  return (uint32_t)(((uint64_t)s * 16807u) % 2147483647u);
}

void orc_random2_stream(uint32_t seed, int64_t n, double *out) {
  uint32_t s = seed;
  for (int64_t i = 0; i < n; i++) {
    s = orc_pm_next(s);
    out[i] = (double)s / 2147483647.0;
  }
}

/* ======================================================================= */
/* Colour-schedule permutation (no reference counterpart: the reference   */
/* shuffles with random2, XT_ICD_update.c:500-532).  Bit-identical on GPU. */
/* ======================================================================= */
static uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

uint32_t orc_perm_key(uint32_t seed, int32_t iter, int32_t frame, int32_t slice) {
  uint32_t k = seed ^ 0x5bd1e995U;
  k = mix32(k ^ ((uint32_t)iter * 0x9e3779b1U));
  k = mix32(k ^ ((uint32_t)frame * 0x85ebca77U));
  k = mix32(k ^ ((uint32_t)slice * 0xc2b2ae3dU));
  return k;
}

uint32_t orc_perm(uint32_t n, uint32_t N, uint32_t key) {
  uint32_t bits = 1;
  while ((1u << bits) < N) bits++;
  uint32_t h = (bits + 1) / 2;
  uint32_t mask = (1u << h) - 1u;
  do {
    uint32_t L = n >> h, R = n & mask;
    for (uint32_t r = 0; r < 4; r++) {
      uint32_t F = mix32(R ^ (key + r * 0x9e3779b9U)) & mask;
      uint32_t nl = R;
      R = L ^ F;
      L = nl;
    }
    n = (L << h) | R;
  } while (n >= N);
  return n;
}

/* inverse of orc_perm: the rounds undone in reverse order, cycle-walking
 * alike (a round maps (L, R) to (R, L ^ F(R))) */
uint32_t orc_perm_inv(uint32_t n, uint32_t N, uint32_t key) {
  uint32_t bits = 1;
  while ((1u << bits) < N) bits++;
  uint32_t h = (bits + 1) / 2;
  uint32_t mask = (1u << h) - 1u;
  do {
    uint32_t L = n >> h, R = n & mask;
    for (int r = 3; r >= 0; r--) {
      uint32_t F = mix32(L ^ (key + (uint32_t)r * 0x9e3779b9U)) & mask;
      uint32_t pl = R ^ F;
      R = L;
      L = pl;
    }
    n = (L << h) | R;
  } while (n >= N);
  return n;
}

/* ======================================================================= */
/* System model                                                            */
/* ======================================================================= */
/* XT_Profile.c:47-104 (calculateVoxelProfile), 108-124 (initializeBeamProfile)
 * and 129-176 (DetectorResponseProfile, H_r part; H_t is unused). */
void orc_detector_response(const float *angles, int32_t N_p, double delta_xy,
                           double delta_r, double OffsetR, double beam_width,
                           float *H_r) {
  float *prof = (float *)xcalloc(ORC_PROFILE_RES, sizeof(float));
  double beam[ORC_BEAM_RES];
  double bsum = 0;
  for (int i = 0; i < ORC_BEAM_RES; i++) {
    beam[i] = 0.54 - 0.46 * cos((2 * ORC_PI / ORC_BEAM_RES) * i);
    bsum += beam[i];
  }
  for (int i = 0; i < ORC_BEAM_RES; i++) beam[i] /= bsum;
  const double r0 = -beam_width / 2;
  const double step = beam_width / ORC_BEAM_RES;
  const double tconst = ORC_PROFILE_RES / (2 * delta_xy);
  (void)delta_r;
  for (int k = 0; k < N_p; k++) {
    /* voxel profile: trapezoid of the line integral through the voxel */
    double ang = angles[k];
    while (ang > ORC_PI_2) ang -= ORC_PI_2;
    while (ang < 0) ang += ORC_PI_2;
    double peak = (ang <= ORC_PI_4) ? delta_xy / cos(ang) : delta_xy / cos(ORC_PI_2 - ang);
    double c45 = cos(ORC_PI_4);
    double d1 = c45 * cos(ORC_PI_4 - ang);
    double d2 = c45 * fabs(cos(ORC_PI_4 + ang));
    double lc = 1 - d1, ln = 1 - d2, rn = 1 + d2, rc = 1 + d1;
    for (int j = 0; j < ORC_PROFILE_RES; j++) {
      double t = 2.0 * j / ORC_PROFILE_RES;
      double v;
      if (t <= lc || t >= rc) v = 0;
      else if (t > rn) v = peak * (rc - t) / (rc - rn);
      else if (t >= ln) v = peak;
      else v = peak * (t - lc) / (ln - lc);
      prof[j] = (float)v;
    }
    /* convolve with the normalised Hamming beam, sampled every OffsetR */
    for (int i = 0; i < ORC_NBINS; i++) {
      double centre = i * OffsetR;
      double rmin = centre - delta_xy;
      double s = 0;
      for (int q = 0; q < ORC_BEAM_RES; q++) {
        double r = r0 + q * step;
        int idx = (int)floor((r - rmin) * tconst);
        if (idx >= 0 && idx <= ORC_PROFILE_RES - 1) s += prof[idx] * beam[q];
      }
      H_r[(size_t)k * (ORC_NBINS + 1) + i] = (float)s;
    }
    H_r[(size_t)k * (ORC_NBINS + 1) + ORC_NBINS] = 0.0f;
  }
  free(prof);
}

/* XT_AMatrix.c:50-110 */
int32_t orc_amatrix_column(const float *H, double cosv, double sinv, int32_t row,
                           int32_t col, double x0, double y0, double delta_xy,
                           double R0, double RMax, double delta_r, int32_t N_r,
                           double OffsetR, double *values, int32_t *index) {
  double x = x0 + ((double)col + 0.5) * delta_xy;
  double y = y0 + ((double)row + 0.5) * delta_xy;
  double r = x * cosv - y * sinv;
  double rmin = r - delta_xy, rmax = r + delta_xy;
  if (rmax < R0 || rmin > RMax) return 0;
  int32_t lo = (int32_t)floor((rmin - R0) / delta_r);
  int32_t hi = (int32_t)floor((rmax - R0) / delta_r);
  if (hi >= N_r) hi = N_r - 1;
  if (lo < 0) lo = 0;
  int32_t n = 0;
  for (int32_t j = lo; j <= hi; j++) {
    double centre = R0 + ((double)j + 0.5) * delta_r;
    double dist = fabs(r - centre);
    int32_t bin = (int32_t)floor(dist / OffsetR);
    if (bin >= 0 && bin < ORC_NBINS) {
      double w1 = dist - bin * OffsetR;
      double w2 = OffsetR - w1;
      double f = (w2 * H[bin] + w1 * H[bin + 1]) / OffsetR;
      if (f > 0) {
        values[n] = f;
        index[n] = j;
        n++;
      }
    }
  }
  return n;
}

/* ======================================================================= */
/* qGGMRF (XT_ICD_update.c:83-129), p = 1.2, q = 2                         */
/* ======================================================================= */
double orc_qggmrf_value(double delta, double sq, double sqp, double c) {
  return (pow(fabs(delta), ORC_Q) / sq) / (c + pow(fabs(delta), ORC_Q - ORC_P) / sqp);
}

double orc_qggmrf_derivative(double delta, double sq, double sqp, double c) {
  double t1 = pow(fabs(delta), ORC_Q - ORC_P) / sqp;
  double t2 = pow(fabs(delta), ORC_Q - 1);
  double t3 = c + t1;
  double mag = (t2 / (t3 * sq)) * (ORC_Q - ((ORC_Q - ORC_P) * t1) / t3);
  return delta < 0 ? -mag : mag;
}

double orc_qggmrf_second_derivative(double sq, double c) { return ORC_Q / (sq * c); }

/* ======================================================================= */
/* Problem                                                                 */
/* ======================================================================= */
struct orc_problem {
  orc_options opt;
  /* sinogram */
  int32_t N_p, N_r, N_t;
  double delta_r, delta_t, R0, RMax, T0;
  double *cosv, *sinv;
  int32_t *view_off, *view_idx;
  double delta_recon;
  /* object */
  int32_t N_x, N_y, N_z, N_time, mult_xy, mult_z;
  double delta_xy, delta_z, x0, y0, OffsetR, radius_obj;
  /* prior */
  double sSq, sSqp, sTq, sTqp, c_s, c_t;
  double sfilt[27]; /* [dz][dy][dx] */
  double tfilt;
  double alpha;
  /* data */
  double T, delta, var_est, convg_thresh;
  int32_t remove_rings, quad_convex, stage_idx;
  int32_t updateProjOffset, updateVar, initICD, initMagUpMap;
  /* arrays */
  arr_t *obj, *e, *w, *y, *d;
  uint8_t *b, *mask;
  float *H_r, *mag;
  int32_t nblocks, prev_blocks;
  int32_t U; /* masked pixels */
  int32_t *xsel, *ysel, *nsel;   /* random lists [N_time][nblocks][N_xy] */
  int32_t *xnh, *ynh, *nnh;      /* NHICD lists */
  uint32_t rng;
  int64_t zero_count, update_count;
};

static size_t sino_idx(const orc_problem *p, int32_t v, int32_t r, int32_t t) {
  return ((size_t)v * p->N_r + r) * p->N_t + t;
}
static size_t obj_idx(const orc_problem *p, int32_t f, int32_t z, int32_t yy, int32_t xx) {
  return (((size_t)f * p->N_z + z) * p->N_y + yy) * p->N_x + xx;
}

static double prng(orc_problem *p) { /* random2() on the problem's state */
  p->rng = orc_pm_next(p->rng);
  return (double)p->rng / 2147483647.0;
}

/* XT_Init.c:69-96 (initSparseAnglesOfObject) and 100-131 (initRandomAngles). */
static int build_view_map(orc_problem *p, const float *times, const float *rt) {
  for (int i = 0; i < p->N_p - 1; i++)
    if (times[i + 1] < times[i]) return -1;
  for (int i = 0; i < p->N_time; i++)
    if (rt[i + 1] < rt[i]) return -1;
  p->view_off = (int32_t *)xcalloc(p->N_time + 1, sizeof(int32_t));
  p->view_idx = (int32_t *)xcalloc(p->N_p * (size_t)p->N_time + 1, sizeof(int32_t));
  int32_t start = 0, total = 0;
  for (int i = 0; i < p->N_time; i++) {
    p->view_off[i] = total;
    for (int j = start; j < p->N_p; j++) {
      double t = times[j];
      if (t >= rt[i] && t < rt[i + 1]) p->view_idx[total++] = j;
      if (t >= rt[i + 1]) {
        start = j;
        break;
      }
    }
  }
  p->view_off[p->N_time] = total;
  p->delta_recon = (double)total / p->N_time;
  return 0;
}

/* XT_Init.c:137-142 distance2node and 144-190 initFilter. */
static void build_filter(orc_problem *p) {
  double prior_const = (double)p->mult_xy * p->mult_xy * p->mult_xy * p->delta_recon;
  double sum = 0;
  p->tfilt = 1.0; /* 1/distance to a temporal neighbour */
  sum += 2.0 * p->tfilt;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++) {
        double d2 = (i - 1) * (i - 1) + (j - 1) * (j - 1) + (k - 1) * (k - 1);
        double v = d2 > 0 ? 1.0 / sqrt(d2) : 0.0;
        p->sfilt[i * 9 + j * 3 + k] = v;
        sum += v;
      }
  for (int i = 0; i < 27; i++) p->sfilt[i] = prior_const * p->sfilt[i] / sum;
  p->tfilt = prior_const * p->tfilt / sum;
}

/* XT_Init.c:389-398 (num_z_blocks) and 400-410 (prevnum_z_blocks). */
static void choose_blocks(orc_problem *p, int32_t prev_nz) {
  if (p->opt.schedule == ORC_SCHED_COLOUR) {
    p->nblocks = p->N_z;
    p->prev_blocks = prev_nz > 0 ? prev_nz : p->N_z;
    return;
  }
  int32_t nt = p->opt.num_threads > 0 ? p->opt.num_threads : 1;
  int32_t nb = nt / p->N_time;
  if (nb < 2) nb = 2;
  else if (nb > p->N_z) nb = p->N_z;
  p->nblocks = (nb / 2) * 2;
  int32_t pb = nt / p->N_time;
  if (pb < 2) pb = 2;
  else {
    if (p->initICD == 3 && pb > p->N_z / 2) pb = p->N_z / 2;
    else if (pb > p->N_z) pb = p->N_z;
  }
  p->prev_blocks = (pb / 2) * 2;
}

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
                             double var_est, const orc_options *opt) {
  orc_problem *p = (orc_problem *)xcalloc(1, sizeof(orc_problem));
  p->opt = *opt;
  /* initStatStructures, XT_Init.c:449-669 (one rank) */
  double Length_R = (double)vox_wid * proj_cols;
  double Length_T = (double)vox_wid * proj_rows;
  p->N_p = proj_num;
  p->N_r = proj_cols;
  p->N_t = proj_rows;
  p->N_time = recon_num;
  p->convg_thresh = convg_thresh;
  p->alpha = ORC_ALPHA;
  p->remove_rings = remove_rings;
  p->quad_convex = quad_convex;
  if (quad_convex == 0) {
    p->T = huber_T;
    p->delta = huber_delta;
  } else {
    p->T = ORC_ZINGER_OFF_T;
    p->delta = ORC_ZINGER_OFF_DELTA;
  }
  p->delta_r = Length_R / p->N_r;
  p->delta_t = Length_T / p->N_t;
  p->R0 = -(double)rot_center * p->delta_r;
  p->RMax = (p->N_r - (double)rot_center) * p->delta_r;
  p->T0 = -Length_T / 2.0;
  p->x0 = p->R0;
  p->y0 = -Length_R / 2.0;
  p->radius_obj = (double)vox_wid * proj_cols / 2;
  p->cosv = (double *)xcalloc(p->N_p, sizeof(double));
  p->sinv = (double *)xcalloc(p->N_p, sizeof(double));
  for (int i = 0; i < p->N_p; i++) {
    p->cosv[i] = cos(proj_angles[i]);
    p->sinv[i] = sin(proj_angles[i]);
  }
  if (build_view_map(p, proj_times, recon_times)) {
    orc_problem_free(p);
    return NULL;
  }
  /* initDynStructures, XT_Init.c:211-445 */
  p->stage_idx = stage_idx;
  p->mult_xy = mult_xy;
  p->mult_z = mult_z;
  if (stage_idx == 0) p->initICD = 0;
  else if (mult_z == prev_mult_z) p->initICD = 2;
  else if (prev_mult_z / mult_z == 2) p->initICD = 3;
  else p->initICD = 0;
  p->updateProjOffset = stage_idx == 0 ? 0 : (stage_idx == 1 ? 2 : 3);
  if (remove_rings == 0) p->updateProjOffset = 0;
  p->initMagUpMap = stage_idx > 0;
  p->updateVar = (stage_idx > 0 && quad_convex == 0);
  p->var_est = var_est;
  p->N_x = p->N_r / mult_xy;
  p->N_y = p->N_r / mult_xy;
  p->N_z = p->N_t / mult_z;
  p->delta_xy = mult_xy * p->delta_r;
  p->delta_z = mult_z * p->delta_t;
  p->OffsetR = (p->delta_xy / sqrt(2.0) + p->delta_r / 2.0) / ORC_NBINS;
  p->sSq = pow(sig_s * (double)mult_xy, ORC_Q);
  p->sSqp = pow(sig_s * (double)mult_xy, ORC_Q - ORC_P);
  p->sTq = pow(sig_t * p->delta_recon, ORC_Q);
  p->sTqp = pow(sig_t * p->delta_recon, ORC_Q - ORC_P);
  p->c_s = c_s;
  p->c_t = c_t;
  build_filter(p);
  choose_blocks(p, stage_idx == 0 ? 0 : p->N_t / prev_mult_z);

  size_t ns = (size_t)p->N_p * p->N_r * p->N_t;
  size_t no = (size_t)p->N_time * p->N_z * p->N_y * p->N_x;
  size_t nxy = (size_t)p->N_x * p->N_y;
  p->obj = (arr_t *)xcalloc(no, sizeof(arr_t));
  p->e = (arr_t *)xcalloc(ns, sizeof(arr_t));
  p->w = (arr_t *)xcalloc(ns, sizeof(arr_t));
  p->y = (arr_t *)xcalloc(ns, sizeof(arr_t));
  p->b = (uint8_t *)xcalloc(ns, 1);
  p->d = (arr_t *)xcalloc((size_t)p->N_r * p->N_t, sizeof(arr_t));
  p->H_r = (float *)xcalloc((size_t)p->N_p * (ORC_NBINS + 1), sizeof(float));
  p->mask = (uint8_t *)xcalloc(nxy, 1);
  p->mag = (float *)xcalloc((size_t)p->N_time * p->nblocks * nxy, sizeof(float));
  size_t nl = (size_t)p->N_time * p->nblocks * nxy;
  p->xsel = (int32_t *)xcalloc(nl, sizeof(int32_t));
  p->ysel = (int32_t *)xcalloc(nl, sizeof(int32_t));
  p->xnh = (int32_t *)xcalloc(nl, sizeof(int32_t));
  p->ynh = (int32_t *)xcalloc(nl, sizeof(int32_t));
  p->nsel = (int32_t *)xcalloc((size_t)p->N_time * p->nblocks, sizeof(int32_t));
  p->nnh = (int32_t *)xcalloc((size_t)p->N_time * p->nblocks, sizeof(int32_t));
  /* transpose [view][row][col] -> [view][col][row] (XT_Init.c:298-305) and
   * divide the weights by the variance estimate (XT_Init.c:327-330) */
  for (int i = 0; i < p->N_p; i++)
    for (int j = 0; j < p->N_r; j++)
      for (int k = 0; k < p->N_t; k++) {
        size_t src = (size_t)i * p->N_t * p->N_r + (size_t)k * p->N_r + j;
        size_t dst = sino_idx(p, i, j, k);
        p->y[dst] = (arr_t)projections[src];
        arr_t wv = (arr_t)weights[src];
        wv /= (arr_t)p->var_est;
        p->w[dst] = wv;
      }
  /* ProjSelect starts true (XT_Init.c:576-580) */
  memset(p->b, 1, ns);
  /* circular mask, XT_ICD_update.c:1140-1149 */
  int32_t U = 0;
  for (int j = 0; j < p->N_y; j++)
    for (int k = 0; k < p->N_x; k++) {
      double xx = p->x0 + ((double)k + 0.5) * p->delta_xy;
      double yy = p->y0 + ((double)j + 0.5) * p->delta_xy;
      uint8_t m = (xx * xx + yy * yy < p->radius_obj * p->radius_obj) ? 1 : 0;
      p->mask[(size_t)j * p->N_x + k] = m;
      U += m;
    }
  p->U = U;
  /* XT_ICD_update.c:1151 */
  orc_detector_response(proj_angles, p->N_p, p->delta_xy, p->delta_r, p->OffsetR,
                        p->delta_r, p->H_r);
  p->rng = 761521u; /* srandom2(761521), XT_Engine.c:99 */
  return p;
}

void orc_problem_free(orc_problem *p) {
  if (!p) return;
  free(p->cosv); free(p->sinv); free(p->view_off); free(p->view_idx);
  free(p->obj); free(p->e); free(p->w); free(p->y); free(p->b); free(p->d);
  free(p->H_r); free(p->mask); free(p->mag);
  free(p->xsel); free(p->ysel); free(p->xnh); free(p->ynh);
  free(p->nsel); free(p->nnh);
  free(p);
}

void orc_problem_info(const orc_problem *p, int32_t *ints, double *d) {
  int32_t iv[] = {p->N_p, p->N_r, p->N_t, p->N_x, p->N_y, p->N_z, p->N_time,
                  p->mult_xy, p->mult_z, p->nblocks, p->updateProjOffset,
                  p->updateVar, p->initICD, p->U};
  memcpy(ints, iv, sizeof(iv));
  double dv[] = {p->delta_r, p->delta_t, p->R0, p->RMax, p->T0, p->x0, p->y0,
                 p->delta_xy, p->delta_z, p->OffsetR, p->radius_obj, p->sSq,
                 p->sSqp, p->sTq, p->sTqp, p->c_s, p->c_t, p->tfilt, p->alpha,
                 p->T, p->delta, p->var_est, p->delta_recon};
  memcpy(d, dv, sizeof(dv));
  memcpy(d + sizeof(dv) / sizeof(double), p->sfilt, sizeof(p->sfilt));
}

void *orc_obj(orc_problem *p) { return p->obj; }
void *orc_err(orc_problem *p) { return p->e; }
void *orc_wgt(orc_problem *p) { return p->w; }
void *orc_proj(orc_problem *p) { return p->y; }
uint8_t *orc_sel(orc_problem *p) { return p->b; }
void *orc_off(orc_problem *p) { return p->d; }
float *orc_hr(orc_problem *p) { return p->H_r; }
uint8_t *orc_mask(orc_problem *p) { return p->mask; }
float *orc_mag(orc_problem *p) { return p->mag; }
const int32_t *orc_view_off(orc_problem *p) { return p->view_off; }
const int32_t *orc_view_idx(orc_problem *p) { return p->view_idx; }
double orc_var_est(const orc_problem *p) { return p->var_est; }
void orc_set_var_est(orc_problem *p, double v) { p->var_est = v; }
int64_t orc_zero_count(const orc_problem *p) { return p->zero_count; }
int64_t orc_update_count(const orc_problem *p) { return p->update_count; }
uint32_t orc_rng_state(const orc_problem *p) { return p->rng; }
void orc_set_rng_state(orc_problem *p, uint32_t s) { p->rng = s; }
void orc_set_select_counts(orc_problem *p, int32_t n) {
  for (int i = 0; i < p->N_time * p->nblocks; i++) p->nsel[i] = n;
}

static int32_t acol(const orc_problem *p, int32_t view, int32_t yy, int32_t xx,
                    double *vals, int32_t *idx) {
  return orc_amatrix_column(p->H_r + (size_t)view * (ORC_NBINS + 1), p->cosv[view],
                            p->sinv[view], yy, xx, p->x0, p->y0, p->delta_xy, p->R0,
                            p->RMax, p->delta_r, p->N_r, p->OffsetR, vals, idx);
}

static int32_t max_acol(const orc_problem *p) {
  return (int32_t)ceil(3 * p->delta_xy / p->delta_r) + 1;
}

/* ======================================================================= */
/* Forward projection and error sinogram: XT_ICD_update.c:655-716,         */
/* XT_ForwardProject.c:8-25                                                */
/* ======================================================================= */
void orc_forward_project(orc_problem *p, void *outv) {
  arr_t *out = (arr_t *)outv;
  size_t ns = (size_t)p->N_p * p->N_r * p->N_t;
  memset(out, 0, ns * sizeof(arr_t));
  int32_t mc = max_acol(p);
  /* frames own disjoint views, so frames can run concurrently */
#pragma omp parallel for schedule(dynamic, 1)
  for (int f = 0; f < p->N_time; f++) {
    double *vals = (double *)xcalloc(mc, sizeof(double));
    int32_t *idx = (int32_t *)xcalloc(mc, sizeof(int32_t));
    for (int j = 0; j < p->N_y; j++)
      for (int k = 0; k < p->N_x; k++)
        for (int q = p->view_off[f]; q < p->view_off[f + 1]; q++) {
          int32_t v = p->view_idx[q];
          int32_t n = acol(p, v, j, k, vals, idx);
          for (int s = 0; s < p->N_z; s++) {
            double pix = p->obj[obj_idx(p, f, s, j, k)];
            for (int m = 0; m < n; m++)
              for (int o = 0; o < p->mult_z; o++)
                out[sino_idx(p, v, idx[m], s * p->mult_z + o)] += pix * vals[m];
          }
        }
    free(vals);
    free(idx);
  }
}

static uint8_t select_bit(const orc_problem *p, double e, double w) {
  return fabs(e * sqrt(w)) < p->T ? 1 : 0;
}

void orc_init_error_sino(orc_problem *p) {
  size_t ns = (size_t)p->N_p * p->N_r * p->N_t;
  orc_forward_project(p, p->e);
#pragma omp parallel for
  for (int i = 0; i < p->N_p; i++)
    for (int j = 0; j < p->N_r; j++)
      for (int k = 0; k < p->N_t; k++) {
        size_t q = sino_idx(p, i, j, k);
        p->e[q] = p->y[q] - p->e[q] - p->d[(size_t)j * p->N_t + k];
        p->b[q] = select_bit(p, p->e[q], p->w[q]);
      }
  (void)ns;
}

/* ======================================================================= */
/* Cost: XT_ICD_update.c:186-319                                           */
/* ======================================================================= */
double orc_cost(orc_problem *p, double *fwd_out, double *prior_out) {
  double fwd = 0, prior = 0;
  const double T = p->T, dl = p->delta;
#pragma omp parallel for reduction(+ : fwd)
  for (int i = 0; i < p->N_p; i++)
    for (int j = 0; j < p->N_r; j++)
      for (int k = 0; k < p->N_t; k++) {
        size_t q = sino_idx(p, i, j, k);
        double u = p->e[q] * sqrt((double)p->w[q]);
        if (p->b[q]) fwd += u * u;
        else fwd += 2.0 * dl * T * fabs(u) + T * T * (1.0 - 2.0 * dl);
      }
  fwd /= 2.0;
  /* each unordered neighbour pair once: 13 forward spatial neighbours in
   * the same frame plus the next frame (XT_ICD_update.c:214-299) */
  static const int fw[13][3] = {{0, 0, 1},  {0, 1, -1}, {0, 1, 0},  {0, 1, 1},
                                {1, -1, 0}, {1, 0, 0},  {1, 1, 0},  {1, -1, -1},
                                {1, -1, 1}, {1, 0, -1}, {1, 1, -1}, {1, 1, 1},
                                {1, 0, 1}};
#pragma omp parallel for reduction(+ : prior)
  for (int f = 0; f < p->N_time; f++)
    for (int z = 0; z < p->N_z; z++)
      for (int j = 0; j < p->N_y; j++)
        for (int k = 0; k < p->N_x; k++) {
          double v = p->obj[obj_idx(p, f, z, j, k)];
          for (int n = 0; n < 13; n++) {
            int zz = z + fw[n][0], jj = j + fw[n][1], kk = k + fw[n][2];
            if (zz >= p->N_z || jj < 0 || jj >= p->N_y || kk < 0 || kk >= p->N_x) continue;
            double g = p->sfilt[(fw[n][0] + 1) * 9 + (fw[n][1] + 1) * 3 + (fw[n][2] + 1)];
            prior += g * orc_qggmrf_value(v - p->obj[obj_idx(p, f, zz, jj, kk)], p->sSq, p->sSqp, p->c_s);
          }
          if (f + 1 < p->N_time)
            prior += p->tfilt * orc_qggmrf_value(v - p->obj[obj_idx(p, f + 1, z, j, k)], p->sTq, p->sTqp, p->c_t);
        }
  fwd += (double)p->N_p * (double)p->N_r * (double)p->N_t * log(p->var_est) / 2;
  if (fwd_out) *fwd_out = fwd;
  if (prior_out) *prior_out = prior;
  return fwd + prior;
}

/* ======================================================================= */
/* Offsets: XT_ICD_update.c:753-927                                        */
/* ======================================================================= */
void orc_offset_update(orc_problem *p, int32_t type) {
  const int32_t NR = p->N_r, NT = p->N_t;
  double *num = (double *)xcalloc((size_t)NR * NT, sizeof(double));
  double *den = (double *)xcalloc((size_t)NR * NT, sizeof(double));
  const double Td = p->T * p->delta;
  /* per detector pixel: weights c = w (b true) or T*delta*sqrt(w)/|e|;
   * type 1/2 accumulate c*e (as c*sign(e)*|e|), type 3 accumulates c*(e+d) */
#pragma omp parallel for collapse(2)
  for (int i = 0; i < NR; i++)
    for (int j = 0; j < NT; j++) {
      double n = 0, dd = 0, off = p->d[(size_t)i * NT + j];
      for (int k = 0; k < p->N_p; k++) {
        size_t q = sino_idx(p, k, i, j);
        double e = p->e[q], w = p->w[q];
        double tmp = Td * sqrt(w);
        if (type == 3) {
          if (p->b[q]) {
            dd += w;
            n += (e + off) * w;
          } else {
            dd += tmp / fabs(e);
            n += (e + off) * tmp / fabs(e);
          }
        } else {
          if (p->b[q]) {
            n += e * w;
            dd += w;
          } else {
            double sg = (e > 0) - (e < 0);
            n += tmp * sg;
            dd += tmp / fabs(e);
          }
        }
      }
      num[(size_t)i * NT + j] = n;
      den[(size_t)i * NT + j] = dd;
    }
  double gamma = 0;
  if (type == 2) {
    double ns = 0, ds = 0;
    for (size_t q = 0; q < (size_t)NR * NT; q++) {
      ns += p->d[q] + num[q] / den[q];
      ds += 1.0 / den[q];
    }
    gamma = ns / ds;
  }
#pragma omp parallel for collapse(2)
  for (int i = 0; i < NR; i++)
    for (int j = 0; j < NT; j++) {
      size_t o = (size_t)i * NT + j;
      double delta_d;
      if (type == 3) {
        double x = num[o] / den[o]; /* constrained_quad_opt with M = 0 */
        delta_d = x - p->d[o];
        for (int k = 0; k < p->N_p; k++) {
          size_t q = sino_idx(p, k, i, j);
          p->e[q] += p->d[o] - (arr_t)x;
          p->b[q] = select_bit(p, p->e[q], p->w[q]);
        }
        p->d[o] = (arr_t)x;
        (void)delta_d;
      } else {
        double step = (num[o] - gamma) / den[o];
        p->d[o] = p->d[o] + step;
        for (int k = 0; k < p->N_p; k++) {
          size_t q = sino_idx(p, k, i, j);
          p->e[q] -= step;
          p->b[q] = select_bit(p, p->e[q], p->w[q]);
        }
      }
    }
  free(num);
  free(den);
}

/* ======================================================================= */
/* Variance: XT_ICD_update.c:720-751                                       */
/* ======================================================================= */
void orc_variance_update(orc_problem *p) {
  double acc = 0;
  const double old = p->var_est, Td = p->delta * p->T;
  size_t ns = (size_t)p->N_p * p->N_r * p->N_t;
#pragma omp parallel for reduction(+ : acc)
  for (size_t q = 0; q < ns; q++) {
    p->w[q] = p->w[q] * old;
    double e = p->e[q], w = p->w[q];
    if (p->b[q]) acc += e * e * w;
    else acc += fabs(e) * Td * sqrt(w * old);
  }
  p->var_est = acc / ((double)p->N_p * (double)p->N_r * (double)p->N_t);
  const double nv = p->var_est;
#pragma omp parallel for
  for (size_t q = 0; q < ns; q++) {
    p->w[q] /= nv;
    p->b[q] = select_bit(p, p->e[q], p->w[q]);
  }
}

/* ======================================================================= */
/* Voxel update                                                            */
/* ======================================================================= */
typedef struct {
  double s[27];   /* spatial neighbourhood [dz][dy][dx] */
  uint8_t sf[27]; /* in range */
  double t[2];
  uint8_t tf[2];
} nhood;

/* XT_VoxUpdate.c:168-233: gather the 26 + 2 neighbourhood of (f, z, y, x) */
static void gather(const orc_problem *p, int32_t f, int32_t z, int32_t yy, int32_t xx, nhood *nh) {
  if (f - 1 >= 0) { nh->t[0] = p->obj[obj_idx(p, f - 1, z, yy, xx)]; nh->tf[0] = 1; }
  else { nh->t[0] = 0; nh->tf[0] = 0; }
  if (f + 1 < p->N_time) { nh->t[1] = p->obj[obj_idx(p, f + 1, z, yy, xx)]; nh->tf[1] = 1; }
  else { nh->t[1] = 0; nh->tf[1] = 0; }
  for (int a = 0; a < 3; a++)
    for (int bq = 0; bq < 3; bq++)
      for (int c = 0; c < 3; c++) {
        int zz = z + a - 1, jj = yy + bq - 1, kk = xx + c - 1;
        int o = a * 9 + bq * 3 + c;
        if (zz >= 0 && zz < p->N_z && jj >= 0 && jj < p->N_y && kk >= 0 && kk < p->N_x) {
          nh->s[o] = p->obj[obj_idx(p, f, zz, jj, kk)];
          nh->sf[o] = 1;
        } else {
          nh->s[o] = 0;
          nh->sf[o] = 0;
        }
      }
  nh->s[13] = 0.0;
}

/* XT_ICD_update.c:134-183 (CE_FunctionalSubstitution) */
static double surrogate(const orc_problem *p, double V, double th1, double th2, const nhood *nh) {
  double s1 = 0, s2 = 0;
  for (int o = 0; o < 27; o++) {
    if (!nh->sf[o] || o == 13) continue;
    double d0 = V - nh->s[o];
    double bq = d0 != 0 ? orc_qggmrf_derivative(d0, p->sSq, p->sSqp, p->c_s) / d0
                        : orc_qggmrf_second_derivative(p->sSq, p->c_s);
    double tc = p->sfilt[o] * bq;
    s1 += tc * nh->s[o];
    s2 += tc;
  }
  for (int i = 0; i < 2; i++) {
    if (!nh->tf[i]) continue;
    double d0 = V - nh->t[i];
    double bq = d0 != 0 ? orc_qggmrf_derivative(d0, p->sTq, p->sTqp, p->c_t) / d0
                        : orc_qggmrf_second_derivative(p->sTq, p->c_t);
    double tc = p->tfilt * bq;
    s1 += tc * nh->t[i];
    s2 += tc;
  }
  double u = (s1 + th2 * V - th1) / (s2 + th2);
  return V + p->alpha * (u - V);
}

double orc_surrogate_update(orc_problem *p, int32_t f, int32_t z, int32_t yy,
                            int32_t xx, double th1, double th2) {
  nhood nh;
  gather(p, f, z, yy, xx, &nh);
  return surrogate(p, p->obj[obj_idx(p, f, z, yy, xx)], th1, th2, &nh);
}

/* XT_VoxUpdate.c:43-115.  ncol/vals/idx hold the A columns of the frame's
 * views (vals[q*mc + m]). */
static double voxel_update(orc_problem *p, int32_t f, int32_t z, int32_t yy, int32_t xx,
                           const int32_t *ncol, const double *vals, const int32_t *idx,
                           int32_t mc, const nhood *nh) {
  size_t oi = obj_idx(p, f, z, yy, xx);
  double V = p->obj[oi];
  double th1 = 0, th2 = 0;
  const double Td = p->T * p->delta;
  int32_t nv = p->view_off[f + 1] - p->view_off[f];
  for (int q = 0; q < nv; q++) {
    int32_t v = p->view_idx[p->view_off[f] + q];
    for (int m = 0; m < ncol[q]; m++) {
      double a = vals[q * mc + m];
      for (int o = 0; o < p->mult_z; o++) {
        size_t s = sino_idx(p, v, idx[q * mc + m], z * p->mult_z + o);
        double e = p->e[s], w = p->w[s];
        if (p->b[s]) {
          th2 += a * a * w;
          th1 += -(e * a * w);
        } else {
          double c = Td * sqrt(w) / fabs(e);
          th2 += a * a * c;
          th1 += -(e * a * c);
        }
      }
    }
  }
  double nvx = surrogate(p, V, th1, th2, nh);
  p->obj[oi] = (arr_t)nvx;
  double dx = (double)p->obj[oi] - V;
  for (int q = 0; q < nv; q++) {
    int32_t v = p->view_idx[p->view_off[f] + q];
    for (int m = 0; m < ncol[q]; m++) {
      double a = vals[q * mc + m];
      for (int o = 0; o < p->mult_z; o++) {
        size_t s = sino_idx(p, v, idx[q * mc + m], z * p->mult_z + o);
        p->e[s] -= a * dx;
        p->b[s] = select_bit(p, p->e[s], p->w[s]);
      }
    }
  }
  return p->obj[oi];
}

static void frame_columns(const orc_problem *p, int32_t f, int32_t yy, int32_t xx,
                          int32_t *ncol, double *vals, int32_t *idx, int32_t mc) {
  int32_t nv = p->view_off[f + 1] - p->view_off[f];
  for (int q = 0; q < nv; q++)
    ncol[q] = acol(p, p->view_idx[p->view_off[f] + q], yy, xx, vals + (size_t)q * mc,
                   idx + (size_t)q * mc);
}

double orc_voxel_update(orc_problem *p, int32_t f, int32_t z, int32_t yy, int32_t xx) {
  int32_t mc = max_acol(p);
  int32_t nv = p->view_off[f + 1] - p->view_off[f];
  int32_t *ncol = (int32_t *)xcalloc(nv, sizeof(int32_t));
  double *vals = (double *)xcalloc((size_t)nv * mc, sizeof(double));
  int32_t *idx = (int32_t *)xcalloc((size_t)nv * mc, sizeof(int32_t));
  frame_columns(p, f, yy, xx, ncol, vals, idx, mc);
  nhood nh;
  gather(p, f, z, yy, xx, &nh);
  double r = voxel_update(p, f, z, yy, xx, ncol, vals, idx, mc, &nh);
  free(ncol); free(vals); free(idx);
  return r;
}

/* XT_VoxUpdate.c:117-281 (updateVoxels_Atten).  Returns sum |x_new| of the
 * updated voxels; accumulates zero-skips.  mag points at the block's map. */
typedef struct {
  double vox_mag;
  int64_t zeros, updates;
} upd_stats;

static upd_stats update_lines(orc_problem *p, int32_t f, int32_t z0, int32_t z1,
                              int32_t n, const int32_t *xs, const int32_t *ys,
                              int32_t iter, float *mag) {
  upd_stats st = {0, 0, 0};
  int32_t mc = max_acol(p);
  int32_t nv = p->view_off[f + 1] - p->view_off[f];
  int32_t *ncol = (int32_t *)xcalloc(nv + 1, sizeof(int32_t));
  double *vals = (double *)xcalloc((size_t)(nv + 1) * mc, sizeof(double));
  int32_t *idx = (int32_t *)xcalloc((size_t)(nv + 1) * mc, sizeof(int32_t));
  for (int l = 0; l < n; l++) {
    int32_t xx = xs[l], yy = ys[l];
    size_t mi = (size_t)yy * p->N_x + xx;
    mag[mi] = 0;
    frame_columns(p, f, yy, xx, ncol, vals, idx, mc);
    for (int z = z0; z <= z1; z++) {
      if (p->mask[mi] != 1) continue;
      nhood nh;
      gather(p, f, z, yy, xx, &nh);
      double V = p->obj[obj_idx(p, f, z, yy, xx)];
      int skip = 0;
      if (V == 0.0 && iter > 1) { /* zero skipping, XT_VoxUpdate.c:236-259 */
        skip = 1;
        if (nh.t[0] > 0.0 || nh.t[1] > 0.0) skip = 0;
        for (int o = 0; o < 27; o++)
          if (nh.s[o] > 0.0) skip = 0;
      }
      if (!skip) {
        double nvx = voxel_update(p, f, z, yy, xx, ncol, vals, idx, mc, &nh);
        mag[mi] += (float)fabs(nvx - V);
        st.vox_mag += fabs(nvx);
        st.updates++;
      } else {
        st.zeros++;
      }
    }
  }
  free(ncol); free(vals); free(idx);
  return st;
}

double orc_update_voxels(orc_problem *p, int32_t f, int32_t z0, int32_t z1, int32_t n,
                         const int32_t *xs, const int32_t *ys, int32_t iter, int32_t block) {
  size_t nxy = (size_t)p->N_x * p->N_y;
  upd_stats st = update_lines(p, f, z0, z1, n, xs, ys, iter,
                              p->mag + ((size_t)f * p->nblocks + block) * nxy);
  p->zero_count += st.zeros;
  p->update_count += st.updates;
  return st.vox_mag;
}

/* ======================================================================= */
/* Voxel orders                                                            */
/* ======================================================================= */
/* XT_ICD_update.c:500-532 (randomly_select_x_y) */
static void random_select(orc_problem *p) {
  int32_t nxy = p->N_x * p->N_y;
  int32_t *counter = (int32_t *)xcalloc(nxy, sizeof(int32_t));
  for (int i = 0; i < p->N_time; i++)
    for (int blk = 0; blk < p->nblocks; blk++) {
      int32_t size = nxy;
      for (int q = 0; q < size; q++) counter[q] = q;
      size_t base = ((size_t)i * p->nblocks + blk) * nxy;
      int32_t cnt = 0;
      int32_t per = p->N_z / p->nblocks;
      for (int j = 0; j < nxy; j++) {
        int32_t index = (int32_t)floor(prng(p) * size);
        index = (index == size) ? size - 1 : index;
        int32_t col = counter[index] % p->N_x, row = counter[index] / p->N_x;
        if (per > 0 && p->mask[(size_t)row * p->N_x + col] == 1) {
          p->xsel[base + cnt] = col;
          p->ysel[base + cnt] = row;
          cnt++;
        }
        counter[index] = counter[size - 1];
        size--;
      }
      p->nsel[i * p->nblocks + blk] = cnt;
    }
  free(counter);
}

/* XT_NHICD.c:128-162 (ComputeVSC) */
static void vsc_filter(const orc_problem *p, const float *mag, float *out) {
  static const double hw[5][5] = {{0.0064, 0.0432, 0.0800, 0.0432, 0.0064},
                                   {0.0432, 0.2916, 0.5400, 0.2916, 0.0432},
                                   {0.0800, 0.5400, 1.0000, 0.5400, 0.0800},
                                   {0.0432, 0.2916, 0.5400, 0.2916, 0.0432},
                                   {0.0064, 0.0432, 0.0800, 0.0432, 0.0064}};
  for (int i = 0; i < p->N_y; i++)
    for (int j = 0; j < p->N_x; j++) {
      double acc = 0;
      for (int a = -2; a <= 2; a++)
        for (int c = -2; c <= 2; c++)
          if (i + a >= 0 && i + a < p->N_y && j + c >= 0 && j + c < p->N_x)
            acc += hw[a + 2][c + 2] * mag[(size_t)(i + a) * p->N_x + (j + c)];
      out[(size_t)i * p->N_x + j] = (float)acc;
    }
}

/* XT_NHICD.c:63-126: randomized select of the k-th smallest (1-based),
 * consuming the problem's random2 stream exactly as the reference. */
static uint32_t part(float *A, uint32_t lo, uint32_t hi) {
  double x = A[hi];
  uint32_t i = lo - 1;
  for (uint32_t j = lo; j < hi; j++)
    if (A[j] <= x) {
      i++;
      float t = A[i]; A[i] = A[j]; A[j] = t;
    }
  float t = A[i + 1]; A[i + 1] = A[hi]; A[hi] = t;
  return i + 1;
}

static double rselect(orc_problem *p, float *A, uint32_t lo, uint32_t hi, uint32_t k) {
  for (;;) {
    if (lo == hi) return A[lo];
    uint32_t rn = (uint32_t)floor(prng(p) * (hi - lo + 1));
    rn = (rn == hi - lo + 1) ? hi - lo : rn;
    uint32_t j = lo + rn;
    float t = A[hi]; A[hi] = A[j]; A[j] = t;
    uint32_t q = part(A, lo, hi);
    uint32_t m = q - lo + 1;
    if (k == m) return A[q];
    if (k < m) hi = q - 1;
    else { lo = q + 1; k -= m; }
  }
}

/* exact k-th smallest (1-based) without randomness: the colour schedule's
 * order statistic (same value as any correct selection). */
static int cmp_float(const void *a, const void *b) {
  float x = *(const float *)a, y = *(const float *)b;
  return (x > y) - (x < y);
}
static double dselect(float *A, uint32_t n, uint32_t k) {
  qsort(A, n, sizeof(float), cmp_float);
  return A[k - 1];
}

/* XT_NHICD.c:164-216 */
void orc_vsc_select(orc_problem *p) {
  int32_t nxy = p->N_x * p->N_y;
  float *filt = (float *)xcalloc(nxy, sizeof(float));
  float *copy = (float *)xcalloc(nxy, sizeof(float));
  for (int i = 0; i < p->N_time; i++)
    for (int blk = 0; blk < p->nblocks; blk++) {
      size_t cell = (size_t)i * p->nblocks + blk;
      int32_t want = p->nsel[cell] / ORC_NHICD_K;
      vsc_filter(p, p->mag + cell * nxy, filt);
      memcpy(copy, filt, nxy * sizeof(float));
      double thresh;
      if (p->opt.schedule == ORC_SCHED_REFERENCE)
        thresh = rselect(p, filt, 0, nxy - 1, nxy - want);
      else
        thresh = dselect(filt, nxy, nxy - want);
      int32_t cnt = 0;
      for (int j = 0; j < p->N_y; j++)
        for (int k = 0; k < p->N_x; k++)
          if (copy[(size_t)j * p->N_x + k] > thresh) {
            p->xnh[cell * nxy + cnt] = k;
            p->ynh[cell * nxy + cnt] = j;
            cnt++;
          }
      p->nnh[cell] = cnt;
    }
  free(filt);
  free(copy);
}

int32_t orc_nhicd_count(const orc_problem *p, int32_t f, int32_t blk, int32_t *xs, int32_t *ys) {
  size_t nxy = (size_t)p->N_x * p->N_y;
  size_t cell = (size_t)f * p->nblocks + blk;
  int32_t n = p->nnh[cell];
  if (xs) memcpy(xs, p->xnh + cell * nxy, n * sizeof(int32_t));
  if (ys) memcpy(ys, p->ynh + cell * nxy, n * sizeof(int32_t));
  return n;
}

/* colour schedule: the lines of cell (f, z) for one chunk, as the GPU walks
 * them: the unmasked pixels of permutation indices [n0, n1), in permutation
 * order (hicd_order 0) or in raster order (1). */
static int32_t colour_chunk(const orc_problem *p, int32_t iter, int32_t f, int32_t z,
                            int32_t n0, int32_t n1, int32_t *xs, int32_t *ys) {
  uint32_t N = (uint32_t)(p->N_x * p->N_y);
  uint32_t key = orc_perm_key(p->opt.order_seed, iter, f, z);
  int32_t cnt = 0;
  if (p->opt.hicd_order == 1) {
    for (uint32_t pix = 0; pix < N; pix++) {
      if (p->mask[pix] != 1) continue;
      int32_t n = (int32_t)orc_perm_inv(pix, N, key);
      if (n < n0 || n >= n1) continue;
      xs[cnt] = (int32_t)(pix % (uint32_t)p->N_x);
      ys[cnt] = (int32_t)(pix / (uint32_t)p->N_x);
      cnt++;
    }
    return cnt;
  }
  for (int32_t n = n0; n < n1; n++) {
    uint32_t pix = orc_perm((uint32_t)n, N, key);
    int32_t yy = (int32_t)(pix / (uint32_t)p->N_x), xx = (int32_t)(pix % (uint32_t)p->N_x);
    if (p->mask[pix] != 1) continue;
    xs[cnt] = xx;
    ys[cnt] = yy;
    cnt++;
  }
  return cnt;
}

int32_t orc_colour_chunk(const orc_problem *p, int32_t iter, int32_t f, int32_t z, int32_t n0,
                         int32_t n1, int32_t *xs, int32_t *ys) {
  return colour_chunk(p, iter, f, z, n0, n1, xs, ys);
}

/* ======================================================================= */
/* One ICD iteration: XT_ICD_update.c:936-1105                             */
/* ======================================================================= */
static void block_range(const orc_problem *p, int32_t blk, int32_t *z0, int32_t *z1) {
  int32_t per = p->N_z / p->nblocks;
  *z0 = blk * per;
  *z1 = (blk + 1) * per - 1;
  if (blk >= p->nblocks - 1) *z1 = p->N_z - 1;
}

/* cells of one colour phase: frame i with block (i even ? b : b+1) for
 * b = 0, 2, 4, ... (phase 0) or the opposite (phase 1). */
static int32_t phase_cells(const orc_problem *p, int32_t phase, int32_t *cf, int32_t *cb) {
  int32_t n = 0;
  for (int i = 0; i < p->N_time; i++)
    for (int blk = 0; blk < p->nblocks; blk += 2) {
      int32_t idx = ((i % 2 == 0) == (phase == 0)) ? blk : blk + 1;
      if (idx >= p->nblocks) continue;
      cf[n] = i;
      cb[n] = idx;
      n++;
    }
  return n;
}

static void maybe_offset(orc_problem *p) {
  if (p->updateProjOffset > 1) orc_offset_update(p, p->remove_rings);
}

/* run one colour phase; hicd: chunk j of K lists, else the NHICD lists */
static double run_phase(orc_problem *p, int32_t phase, int32_t iter, int32_t j, int hicd) {
  int32_t ncell_max = p->N_time * (p->nblocks / 2 + 1);
  int32_t *cf = (int32_t *)xcalloc(ncell_max, sizeof(int32_t));
  int32_t *cb = (int32_t *)xcalloc(ncell_max, sizeof(int32_t));
  int32_t nc = phase_cells(p, phase, cf, cb);
  size_t nxy = (size_t)p->N_x * p->N_y;
  double *vm = (double *)xcalloc(nc, sizeof(double));
  int64_t *zc = (int64_t *)xcalloc(nc, sizeof(int64_t));
  int64_t *uc = (int64_t *)xcalloc(nc, sizeof(int64_t));
  const int K = ORC_NHICD_K;
#pragma omp parallel for schedule(dynamic, 1)
  for (int c = 0; c < nc; c++) {
    int32_t f = cf[c], blk = cb[c], z0, z1;
    block_range(p, blk, &z0, &z1);
    size_t cell = (size_t)f * p->nblocks + blk;
    float *mag = p->mag + cell * nxy;
    upd_stats st;
    if (hicd && p->opt.schedule == ORC_SCHED_REFERENCE) {
      int32_t U = p->nsel[cell];
      int32_t s = j * (U / K), e = (j + 1) * (U / K) - 1;
      if (j == K - 1) e = U - 1;
      st = update_lines(p, f, z0, z1, e - s + 1, p->xsel + cell * nxy + s,
                        p->ysel + cell * nxy + s, iter, mag);
    } else if (hicd) {
      int32_t N = (int32_t)nxy;
      int32_t s = j * (N / K), e = (j == K - 1) ? N : (j + 1) * (N / K);
      int32_t *xs = (int32_t *)xcalloc(nxy, sizeof(int32_t));
      int32_t *ys = (int32_t *)xcalloc(nxy, sizeof(int32_t));
      int32_t n = colour_chunk(p, iter, f, z0, s, e, xs, ys);
      st = update_lines(p, f, z0, z1, n, xs, ys, iter, mag);
      free(xs);
      free(ys);
    } else {
      st = update_lines(p, f, z0, z1, p->nnh[cell], p->xnh + cell * nxy,
                        p->ynh + cell * nxy, iter, mag);
    }
    vm[c] = st.vox_mag;
    zc[c] = st.zeros;
    uc[c] = st.updates;
  }
  double tot = 0;
  for (int c = 0; c < nc; c++) {
    tot += vm[c];
    p->zero_count += zc[c];
    p->update_count += uc[c];
  }
  free(cf); free(cb); free(vm); free(zc); free(uc);
  return tot;
}

double orc_run_phase(orc_problem *p, int32_t phase, int32_t iter, int32_t chunk, int32_t hicd) {
  return run_phase(p, phase, iter, chunk, hicd);
}

int32_t orc_icd_iteration(orc_problem *p, int32_t iter, double *ratio) {
  const int K = ORC_NHICD_K;
  size_t nxy = (size_t)p->N_x * p->N_y;
  if (p->opt.schedule == ORC_SCHED_REFERENCE) {
    random_select(p);
  } else {
    for (int i = 0; i < p->N_time * p->nblocks; i++) p->nsel[i] = p->U;
  }
  double total_vox_mag = 0;
  for (int j = 0; j < K; j++) {
    total_vox_mag = 0; /* reset per chunk: quirk of XT_ICD_update.c:976 */
    total_vox_mag += run_phase(p, 0, iter, j, 1);
    maybe_offset(p);
    total_vox_mag += run_phase(p, 1, iter, j, 1);
    maybe_offset(p);
    orc_vsc_select(p);
    if (iter > 1 && p->opt.no_nhicd == 0) {
      run_phase(p, 0, iter, j, 0);
      maybe_offset(p);
      run_phase(p, 1, iter, j, 0);
      maybe_offset(p);
    }
  }
  if (p->updateVar == 1) orc_variance_update(p);
  /* XT_ICD_update.c:1066-1104 */
  double avg = 0;
  for (int i = 0; i < p->N_time; i++)
    for (int blk = 0; blk < p->nblocks; blk++) {
      size_t cell = (size_t)i * p->nblocks + blk;
      const float *mag = p->mag + cell * nxy;
      if (p->opt.schedule == ORC_SCHED_REFERENCE) {
        for (int l = 0; l < p->nsel[cell]; l++)
          avg += mag[(size_t)p->ysel[cell * nxy + l] * p->N_x + p->xsel[cell * nxy + l]];
      } else {
        for (size_t q = 0; q < nxy; q++)
          if (p->mask[q]) avg += mag[q];
      }
    }
  double pct = 100 * avg / total_vox_mag;
  if (ratio) *ratio = pct;
  return pct < p->convg_thresh ? 1 : 0;
}

/* ======================================================================= */
/* Multi-resolution initialisation: XT_ICD_update.c:323-651                */
/* ======================================================================= */
/* bilinear x2 along x then y of one plane (upsample_object_bilinear_2D,
 * XT_ICD_update.c:362-396); tmp sized Ny x 2Nx */
/* mid: how the x pass is stored.  The 2D upsample keeps it in a buffer of
 * the array type (Real_arr_t, XT_ICD_update.c:362-375): MID_ARR for the
 * object, MID_FLOAT for the float magnitude map; the 3D upsample keeps it in
 * double (Real_t buffers, XT_ICD_update.c:398-480): MID_DOUBLE. */
enum { MID_DOUBLE = 0, MID_ARR = 1, MID_FLOAT = 2 };
static double mid_round(double v, int mid) {
  return mid == MID_ARR ? (double)(ORC_ARR_T)v : mid == MID_FLOAT ? (double)(float)v : v;
}
static void up2d_plane(const double *in, int Ny, int Nx, double *tmp, double *out, int mid) {
  int W = 2 * Nx;
  for (int j = 0; j < Ny; j++) {
    const double *r = in + (size_t)j * Nx;
    double *t = tmp + (size_t)j * W;
    t[0] = mid_round(r[0], mid);
    t[1] = mid_round((3.0 * r[0] + r[1]) / 4.0, mid);
    t[W - 1] = mid_round(r[Nx - 1], mid);
    t[W - 2] = mid_round((r[Nx - 2] + 3.0 * r[Nx - 1]) / 4.0, mid);
    for (int k = 1; k < Nx - 1; k++) {
      t[2 * k] = mid_round((r[k - 1] + 3.0 * r[k]) / 4.0, mid);
      t[2 * k + 1] = mid_round((3.0 * r[k] + r[k + 1]) / 4.0, mid);
    }
  }
  for (int k = 0; k < W; k++) {
    out[k] = tmp[k];
    out[W + k] = (3.0 * tmp[k] + tmp[W + k]) / 4.0;
    out[(size_t)(2 * Ny - 1) * W + k] = tmp[(size_t)(Ny - 1) * W + k];
    out[(size_t)(2 * Ny - 2) * W + k] = (tmp[(size_t)(Ny - 2) * W + k] + 3.0 * tmp[(size_t)(Ny - 1) * W + k]) / 4.0;
  }
  for (int j = 1; j < Ny - 1; j++)
    for (int k = 0; k < W; k++) {
      out[(size_t)(2 * j) * W + k] = (tmp[(size_t)(j - 1) * W + k] + 3.0 * tmp[(size_t)j * W + k]) / 4.0;
      out[(size_t)(2 * j + 1) * W + k] = (3 * tmp[(size_t)j * W + k] + tmp[(size_t)(j + 1) * W + k]) / 4.0;
    }
}

/* z interpolation of upsampled planes (upsample_object_bilinear_3D,
 * XT_ICD_update.c:435-480): planes [Nz][P] -> [2Nz][P] */
static void up_z(const double *in, int Nz, size_t P, double *out) {
  for (size_t q = 0; q < P; q++) {
    out[q] = in[q];
    out[P + q] = (3.0 * in[q] + in[P + q]) / 4.0;
    out[(size_t)(2 * Nz - 1) * P + q] = in[(size_t)(Nz - 1) * P + q];
    out[(size_t)(2 * Nz - 2) * P + q] = (3.0 * in[(size_t)(Nz - 1) * P + q] + in[(size_t)(Nz - 2) * P + q]) / 4.0;
  }
  for (int s = 1; s < Nz - 1; s++)
    for (size_t q = 0; q < P; q++) {
      out[(size_t)(2 * s) * P + q] = (in[(size_t)(s - 1) * P + q] + 3.0 * in[(size_t)s * P + q]) / 4.0;
      out[(size_t)(2 * s + 1) * P + q] = (3.0 * in[(size_t)s * P + q] + in[(size_t)(s + 1) * P + q]) / 4.0;
    }
}

void orc_init_from_coarse(orc_problem *p, const void *prev_v, int32_t prev_nz,
                          const float *prev_mag, int32_t prev_blocks) {
  const arr_t *prev = (const arr_t *)prev_v;
  int Ny = p->N_y / 2, Nx = p->N_x / 2;
  size_t P2 = (size_t)p->N_y * p->N_x, P1 = (size_t)Ny * Nx;
  double *in = (double *)xcalloc(P1, sizeof(double));
  double *tmp = (double *)xcalloc((size_t)Ny * 2 * Nx, sizeof(double));
  double *up = (double *)xcalloc(P2 * prev_nz, sizeof(double));
  double *upz = (double *)xcalloc(P2 * 2 * prev_nz, sizeof(double));
  for (int f = 0; f < p->N_time; f++) {
    for (int s = 0; s < prev_nz; s++) {
      for (size_t q = 0; q < P1; q++) in[q] = prev[((size_t)f * prev_nz + s) * P1 + q];
      up2d_plane(in, Ny, Nx, tmp, up + (size_t)s * P2, p->initICD == 3 ? MID_DOUBLE : MID_ARR);
    }
    if (p->initICD == 3) {
      up_z(up, prev_nz, P2, upz);
      for (int s = 0; s < p->N_z; s++)
        for (size_t q = 0; q < P2; q++) p->obj[obj_idx(p, f, s, 0, 0) + q] = (arr_t)upz[(size_t)s * P2 + q];
    } else {
      for (int s = 0; s < p->N_z; s++)
        for (size_t q = 0; q < P2; q++) p->obj[obj_idx(p, f, s, 0, 0) + q] = (arr_t)up[(size_t)s * P2 + q];
    }
  }
  /* MagUpdateMap (XT_ICD_update.c:603-630): float arithmetic in 2D, double
   * buffers in 3D, as the reference */
  if (prev_mag && p->initMagUpMap) {
    if (prev_blocks == p->nblocks || prev_blocks * 2 == p->nblocks) {
      double *mup = (double *)xcalloc(P2 * prev_blocks, sizeof(double));
      double *mupz = (double *)xcalloc(P2 * 2 * prev_blocks, sizeof(double));
      for (int f = 0; f < p->N_time; f++) {
        for (int s = 0; s < prev_blocks; s++) {
          for (size_t q = 0; q < P1; q++) in[q] = prev_mag[((size_t)f * prev_blocks + s) * P1 + q];
          up2d_plane(in, Ny, Nx, tmp, mup + (size_t)s * P2,
                     prev_blocks == p->nblocks ? MID_FLOAT : MID_DOUBLE);
        }
        float *dst = p->mag + (size_t)f * p->nblocks * P2;
        if (prev_blocks == p->nblocks) {
          for (size_t q = 0; q < P2 * p->nblocks; q++) dst[q] = (float)mup[q];
        } else {
          up_z(mup, prev_blocks, P2, mupz);
          for (size_t q = 0; q < P2 * p->nblocks; q++) dst[q] = (float)mupz[q];
        }
      }
      free(mup);
      free(mupz);
    }
  }
  free(in); free(tmp); free(up); free(upz);
}

/* ======================================================================= */
/* ICD_BackProject loop: XT_ICD_update.c:1108-1244                         */
/* ======================================================================= */
int32_t orc_stage_run(orc_problem *p, int32_t max_iter, double *cost_out,
                      double *ratio_out, int32_t *iters) {
  if (max_iter <= 0) max_iter = p->opt.max_iterations > 0 ? p->opt.max_iterations : ORC_MAX_ITERS;
  orc_init_error_sino(p);
  double last = 0;
  if (p->opt.compute_cost) {
    last = orc_cost(p, NULL, NULL);
    if (cost_out) cost_out[0] = last;
  }
  int32_t it, status = 0;
  for (it = 1; it <= max_iter; it++) {
    double ratio = 0;
    int32_t flag = orc_icd_iteration(p, it, &ratio);
    if (ratio_out) ratio_out[it - 1] = ratio;
    if (p->opt.compute_cost) {
      double c = orc_cost(p, NULL, NULL);
      if (cost_out) cost_out[it] = c;
      if (p->opt.abort_on_cost_increase && c > last) {
        status = -1;
        break;
      }
      last = c;
    }
    if (flag != 0 && it > 1) break;
  }
  if (iters) *iters = it > max_iter ? max_iter : it;
  return status;
}

/* ======================================================================= */
/* reconstruct(): XT_Engine.c:81-203                                       */
/* ======================================================================= */
static int write_bin(const char *name, const void *data, size_t bytes) {
  FILE *fp = fopen(name, "wb");
  if (!fp) return -1;
  size_t w = fwrite(data, 1, bytes, fp);
  fclose(fp);
  return w == bytes ? 0 : -1;
}

int orc_reconstruct(void **object_out, const float *projections, const float *weights,
                    const float *proj_angles, const float *proj_times,
                    const float *recon_times, int32_t proj_rows, int32_t proj_cols,
                    int32_t proj_num, int32_t recon_num, float vox_wid, float rot_center,
                    float sig_s, float sig_t, float c_s, float c_t, float convg_thresh,
                    int32_t remove_rings, int32_t quad_convex, float huber_delta,
                    float huber_T, const orc_options *opt, double *final_cost,
                    int32_t *total_iters) {
  /* argument checks, XT_Engine.c:101-104 (one node) */
  if (proj_rows < ORC_MIN_PROJ_ROWS || proj_cols % 2 != 0 || proj_rows % 2 != 0 ||
      proj_rows < ORC_MIN_ROWS_PER_NODE)
    return -1;
  /* multi-resolution schedule, XT_Engine.c:111-141 */
  int32_t nres = (int32_t)(log(((float)proj_cols) / ORC_MIN_XY_RES) / log(2.0) + 1);
  if (nres > ORC_MAX_MULTRES) nres = ORC_MAX_MULTRES;
  if (nres < 2) nres = 2;
  int32_t mxy[ORC_MAX_MULTRES], mz[ORC_MAX_MULTRES];
  mxy[0] = 1;
  mz[0] = 1;
  for (int i = 1; i < nres; i++) {
    if (proj_cols % (mxy[i - 1] * 2) == 0) {
      mxy[i] = mxy[i - 1] * 2;
      if (proj_rows % (mz[i - 1] * 2) == 0 && proj_rows / (mz[i - 1] * 2) >= ORC_MIN_ROWS_PER_NODE)
        mz[i] = mz[i - 1] * 2;
      else
        mz[i] = mz[i - 1];
    } else {
      nres = i;
      break;
    }
  }
  int32_t sxy[ORC_MAX_MULTRES], sz[ORC_MAX_MULTRES];
  for (int i = 0; i < nres; i++) {
    sxy[i] = mxy[nres - 1 - i];
    sz[i] = mz[nres - 1 - i];
  }
  orc_problem *prev = NULL;
  double var = 1.0; /* VAR_PARAM_INIT */
  int32_t iters_all = 0;
  double cost = 0;
  for (int s = 0; s < nres; s++) {
    orc_problem *p = orc_problem_new(projections, weights, proj_angles, proj_times, recon_times,
                                     proj_rows, proj_cols, proj_num, recon_num, vox_wid, rot_center,
                                     sig_s, sig_t, c_s, c_t, convg_thresh, remove_rings, quad_convex,
                                     huber_delta, huber_T, sxy[s], sz[s], s, s ? sz[s - 1] : 0, var, opt);
    if (!p) {
      orc_problem_free(prev);
      return -1;
    }
    if (prev) {
      /* offsets continue from the previous stage (static ProjOffset) */
      memcpy(p->d, prev->d, (size_t)p->N_r * p->N_t * sizeof(arr_t));
      orc_init_from_coarse(p, prev->obj, prev->N_z, prev->mag, prev->nblocks);
      orc_problem_free(prev);
    }
    int32_t it = 0;
    int st = orc_stage_run(p, 0, NULL, NULL, &it);
    iters_all += it;
    if (st) {
      orc_problem_free(p);
      return -1;
    }
    var = p->var_est;
    if (final_cost) cost = orc_cost(p, NULL, NULL);
    if (opt->write_files) {
      char name[64];
      size_t plane = (size_t)p->N_z * p->N_y * p->N_x;
      float *buf = (float *)xcalloc(plane, sizeof(float));
      for (int f = 0; f < p->N_time; f++) {
        for (size_t q = 0; q < plane; q++) buf[q] = (float)p->obj[(size_t)f * plane + q];
        snprintf(name, sizeof(name), "object_time_%d.bin", f);
        write_bin(name, buf, plane * sizeof(float));
      }
      free(buf);
      size_t nd = (size_t)p->N_r * p->N_t;
      float *dbuf = (float *)xcalloc(nd, sizeof(float));
      for (size_t q = 0; q < nd; q++) dbuf[q] = (float)p->d[q];
      write_bin("proj_offset.bin", dbuf, nd * sizeof(float));
      free(dbuf);
      write_bin("variance_estimate.bin", &p->var_est, sizeof(double));
      write_bin("mag_update.bin", p->mag,
                (size_t)p->N_time * p->nblocks * p->N_y * p->N_x * sizeof(float));
      write_bin("status.bin", &s, sizeof(int32_t));
    }
    prev = p;
  }
  if (object_out) {
    size_t no = (size_t)prev->N_time * prev->N_z * prev->N_y * prev->N_x;
    arr_t *o = (arr_t *)xcalloc(no, sizeof(arr_t));
    memcpy(o, prev->obj, no * sizeof(arr_t));
    *object_out = o;
  }
  if (final_cost) *final_cost = cost;
  if (total_iters) *total_iters = iters_all;
  orc_problem_free(prev);
  return 0;
}

void orc_free(void *ptr) { free(ptr); }
