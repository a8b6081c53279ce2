// C ABI of libtimbir_hip: one multi-resolution stage resident on the GPU.
// Declarations and the reference functions each entry point replaces are in
// include/timbir_dev.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "timbir_dev.h"
#include "tb_common.h"

namespace tb {
__global__ void k_transpose_in(KP, const float *, const float *, double);
__global__ void k_to_ref_layout(KP, const float *, float *);
__global__ void k_fwd_interleave(KP, float4 *);
__global__ void k_fbp_filter(KP, float *);
__global__ void k_fbp_back(KP, const float *);
__global__ void k_forward(KP, const int *, const float4 *, float *, int);
__global__ void k_upsample(const float *, float *, int, int, int, int, int);
template <int NT>
__global__ void k_update_voxels(KP, const int2 *, int, int, int, int, int);
template <int VPL, int NB, int MZ, int G>
__global__ void k_icd_lds(KP, const int2 *, int, int, int, int);
template <int NB, int MZ, int VPL, int MAXTH>
__global__ void k_icd_mw(KP, const int2 *, int, int, int, int, int);
template <int NB, int MZ, int VPL, int MAXTH, int TF>
__global__ void k_icd_mwg(KP, const int2 *, int, int, int, int, int);
template <int NB, int MZ, int VPL, int MAXTH, int D, int C>
__global__ void k_icd_mwd(KP, const int2 *, int, int, int, int, int);
template <int MAXTH>
__global__ void k_icd_bat(KP, const int2 *, int, int, int, int, int);
template <int NBX, int R, int MAXTH>
__global__ void k_icd_split(KP, const int2 *, int, int, int, int, int, SplitBuf);
__global__ void k_split_reset(SplitBuf, int);
template <int NB, int MZ, int VPL, int MAXTH>
__global__ void k_icd_mw_m(const KP *, const int4 *, int, int, int, int);
template <int NB, int MZ, int VPL, int MAXTH, int TF>
__global__ void k_icd_mwg_m(const KP *, const int4 *, int, int, int, int);
template <int NBX, int R, int MAXTH>
__global__ void k_icd_wide_m(const KP *, const int4 *, int, int, int, int);
template <int NBX, int R, int PMAX, int MAXTH, bool HL>
__global__ void k_icd_wpass_m(const KP *, const int4 *, int, int, int, int);
template <int NBX, int R, int GP, int MAXTH, bool HL>
__global__ void k_icd_wloop_m(const KP *, const int4 *, int, int, int, int);
template <int NBX, int R, int MAXTH>
__global__ void k_icd_wide(KP, const int2 *, int, int, int, int, int);
template <int NBX, int R, int PMAX, int MAXTH, bool HL>
__global__ void k_icd_wpass(KP, const int2 *, int, int, int, int, int);
template <int NBX, int R, int GP, int MAXTH, bool HL>
__global__ void k_icd_wloop(KP, const int2 *, int, int, int, int, int);
template <int NB>
__global__ void k_build_columns(KP);
__global__ void k_offset_partial(KP, int, int, double *, double *);
__global__ void k_offset_finalize(KP, int, const double *, const double *, double *, double *);
__global__ void k_offset_gamma(KP, const double *, const double *, double *);
__global__ void k_offset_apply(KP, int, const double *, const double *, const double *);
__global__ void k_offset_store(KP, int, const double *, const double *, const double *);
__global__ void k_var_partial(KP, double, double *);
__global__ void k_var_apply(KP, double, double);
__global__ void k_cost_data(KP, double *);
__global__ void k_cost_prior(KP, double, double, double, double, double, double, const double *,
                             double, double *);
__global__ void k_vsc_filter(KP, float *);
__global__ void k_vsc_select(KP, const float *, int);
__global__ void k_mag_sum(KP, double *);
__global__ void k_fill_f64(double *, size_t, double);
__global__ void k_fill_f32(float *, size_t, float);
__global__ void k_fill_u64(unsigned long long *, size_t);
__global__ void k_build_order(KP, int);
}  // namespace tb

using namespace tb;

// ---------------------------------------------------------------------------
// errors and options
// ---------------------------------------------------------------------------
static thread_local std::string g_err;
static tb_options g_opt = {0, 0, kMaxIters, 0, 1, 0u, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

static int fail(const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return -1;
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail("%s:%d: %s: %s", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
  } while (0)

extern "C" const char *tb_last_error(void) { return g_err.c_str(); }
extern "C" void tb_get_options(tb_options *o) { *o = g_opt; }
extern "C" void tb_set_options(const tb_options *o) { g_opt = *o; }

// ---------------------------------------------------------------------------
// stage
// ---------------------------------------------------------------------------
struct tb_stage {
  // geometry and model (host)
  int N_p, N_r, N_t, N_x, N_y, N_z, N_time, mult_xy, mult_z, U;
  double delta_r, delta_t, R0, RMax, T0, x0, y0, delta_xy, delta_z, OffsetR, radius_obj;
  double sSq, sSqp, sTq, sTqp, c_s, c_t, sfilt[27], tfilt, alpha, T, delta, var_est;
  double convg_thresh, delta_recon;
  int remove_rings, quad_convex, stage_idx, updateProjOffset, updateVar, initICD, initMagUpMap;
  std::vector<int> view_off, view_idx, view_frame;
  std::vector<float> angles;
  int max_nv;
  // device
  int device = 0;
  hipStream_t stream = nullptr;
  float *e = nullptr, *w = nullptr, *y = nullptr, *d = nullptr, *obj = nullptr, *mag = nullptr;
  float *hr = nullptr, *filt = nullptr;
  double *cosv = nullptr, *sinv = nullptr, *sfilt_d = nullptr;
  int *view_off_d = nullptr, *view_idx_d = nullptr, *view_frame_d = nullptr;
  uint8_t *mask = nullptr;
  int *nh_list = nullptr, *nh_count = nullptr;
  int nh_cap = 0;
  int *nh_ord = nullptr, *nh_ord_count = nullptr, *ord = nullptr, *ord_off = nullptr;
  int ord_cap = 0, ord_iter = -1;
  float *dummy = nullptr;
  int *single = nullptr;
  float *colt = nullptr;  // column table of the multi-wavefront kernel (built on first use)
  // the split-cell kernel's exchange buffers (allocated on first use) and
  // the device's compute units
  unsigned long long *split_rec = nullptr;  // [slots][2][64][4] tagged partials
  unsigned *split_w = nullptr;              // err, abort
  hipStream_t abort_stream = nullptr;       // tb_stage_abort's writes
  int split_slots = 0, n_cus = 0;
  bool split_ran = false;
  int colt_nb = 0;
  size_t lds_max = 65536;
  int2 *cells[2] = {nullptr, nullptr};
  int ncells[2] = {0, 0};
  std::vector<int2> cells_h[2];  // the same lists on the host
  double *cell_voxmag = nullptr;
  unsigned long long *cell_upd = nullptr, *cell_zero = nullptr, *cell_ent = nullptr;
  double *off_pnum = nullptr, *off_pden = nullptr, *off_num = nullptr, *off_den = nullptr;
  double *off_gamma = nullptr;
  int off_chunks = 1, off_vpc = 1;
  double *red = nullptr, *red_h = nullptr;
  int red_n = 1024;
  // stats
  std::vector<hipEvent_t> ev;
  int ev_used = 0;
  double voxel_ms = 0, other_ms = 0;
  long long voxel_launches = 0, lds_launches = 0, global_launches = 0;
  long long fam_launches[TB_KERNEL_FAMILIES] = {0};
  long long colt_builds = 0;
  hipEvent_t o0 = nullptr, o1 = nullptr;

  KP kp() const {
    KP P;
    memset(&P, 0, sizeof(P));
    P.N_p = N_p; P.N_r = N_r; P.N_t = N_t; P.N_x = N_x; P.N_y = N_y; P.N_z = N_z;
    P.N_time = N_time; P.mz = mult_z; P.K = kNhicdK;
    P.R0 = R0; P.RMax = RMax; P.delta_r = delta_r; P.x0 = x0; P.y0 = y0;
    P.delta_xy = delta_xy; P.OffsetR = OffsetR;
    P.sSq = (float)sSq; P.sSqp = (float)sSqp; P.sTq = (float)sTq; P.sTqp = (float)sTqp;
    P.c_s = (float)c_s; P.c_t = (float)c_t; P.tfilt = (float)tfilt; P.alpha = (float)alpha;
    for (int i = 0; i < 27; i++) P.sfilt[i] = (float)sfilt[i];
    P.T = (float)T; P.Td = (float)(T * delta); P.T_d = T; P.Td_d = T * delta; P.dl_d = delta;
    P.e = e; P.w = w; P.y = y; P.d = d; P.obj = obj; P.mag = mag; P.hr = hr;
    P.cosv = cosv; P.sinv = sinv; P.view_off = view_off_d; P.view_idx = view_idx_d;
    P.mask = mask; P.nh_list = nh_list; P.nh_count = nh_count; P.nh_cap = nh_cap;
    P.nh_ord = nh_ord; P.nh_ord_count = nh_ord_count; P.ord = ord; P.ord_off = ord_off;
    P.ord_cap = ord_cap; P.dummy = dummy; P.single = single; P.nobj = (int)nobj();
    P.colt = colt; P.colt_nv = max_nv; P.colt_rec = colt_nb ? col_rec(colt_nb) : 0;
    P.u_inv = 1.0 / delta_r; P.u_r0 = R0 / delta_r; P.u_dxy = delta_xy / delta_r;
    P.u_max = (RMax - R0) / delta_r; P.u_kk = delta_r / OffsetR; P.mxy = mult_xy;
    {
      const double t2 = T * T;  // T is a float parameter: T^2 is exact in double
      P.T2h = (float)t2;
      P.T2l = (float)(t2 - (double)P.T2h);
    }
    P.cell_voxmag = cell_voxmag; P.cell_upd = cell_upd; P.cell_zero = cell_zero;
    P.cell_ent = cell_ent; P.order_seed = g_opt.order_seed; P.hicd_order = g_opt.hicd_order;
    return P;
  }
  size_t nsino() const { return (size_t)N_p * N_r * N_t; }
  size_t nobj() const { return (size_t)N_time * N_z * N_y * N_x; }
  size_t nxy() const { return (size_t)N_x * N_y; }
};

static int grid_for(size_t n, int block = 256) {
  size_t g = (n + block - 1) / block;
  return (int)std::min<size_t>(std::max<size_t>(g, 1), 65535 * 8);
}

// ---------------------------------------------------------------------------
// host restatements of the stage set-up
// ---------------------------------------------------------------------------
// XT_Profile.c:47-176: trapezoid voxel profile convolved with the normalised
// Hamming beam, sampled every OffsetR, one row of 65 values per view.  Views
// are independent; they are split over host threads (each row is computed
// by the same sequential code whichever thread runs it).
static void detector_response(const float *angles, int N_p, double delta_xy, double OffsetR,
                              double beam_width, float *H_r) {
  double beam[kBeamRes], bsum = 0;
  for (int i = 0; i < kBeamRes; i++) {
    beam[i] = 0.54 - 0.46 * cos((2 * kPi / kBeamRes) * i);
    bsum += beam[i];
  }
  for (int i = 0; i < kBeamRes; i++) beam[i] /= bsum;
  const double r0 = -beam_width / 2, step = beam_width / kBeamRes;
  const double tconst = kProfileRes / (2 * delta_xy);
  auto rows = [&](int k0, int k1) {
    std::vector<float> prof(kProfileRes);
    for (int k = k0; k < k1; k++) {
      double ang = angles[k];
      while (ang > kPi2) ang -= kPi2;
      while (ang < 0) ang += kPi2;
      const double peak = (ang <= kPi4) ? delta_xy / cos(ang) : delta_xy / cos(kPi2 - ang);
      const double c45 = cos(kPi4);
      const double d1 = c45 * cos(kPi4 - ang), d2 = c45 * fabs(cos(kPi4 + ang));
      const double lc = 1 - d1, ln = 1 - d2, rn = 1 + d2, rc = 1 + d1;
      for (int j = 0; j < kProfileRes; j++) {
        const double t = 2.0 * j / kProfileRes;
        double v;
        if (t <= lc || t >= rc) v = 0;
        else if (t > rn) v = peak * (rc - t) / (rc - rn);
        else if (t >= ln) v = peak;
        else v = peak * (t - lc) / (ln - lc);
        prof[j] = (float)v;
      }
      for (int i = 0; i < kRespBins; i++) {
        const double rmin = i * OffsetR - delta_xy;
        double s = 0;
        for (int q = 0; q < kBeamRes; q++) {
          const double r = r0 + q * step;
          const int idx = (int)floor((r - rmin) * tconst);
          if (idx >= 0 && idx <= kProfileRes - 1) s += prof[idx] * beam[q];
        }
        H_r[(size_t)k * kHrStride + i] = (float)s;
      }
      H_r[(size_t)k * kHrStride + kRespBins] = 0.0f;
    }
  };
  const int hw = (int)std::thread::hardware_concurrency();
  const int nt = std::max(1, std::min({hw > 0 ? hw : 1, 16, (N_p + 31) / 32}));
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; t++)
    pool.emplace_back(rows, (int)((long long)N_p * t / nt), (int)((long long)N_p * (t + 1) / nt));
  rows(0, (int)((long long)N_p / nt));
  for (auto &th : pool) th.join();
}

extern "C" void tb_stage_destroy(tb_stage *st) {
  if (!st) return;
  if (st->stream) (void)hipStreamSynchronize(st->stream);
  if (st->abort_stream) (void)hipStreamDestroy(st->abort_stream);
  void *bufs[] = {st->e, st->w, st->y, st->d, st->obj, st->mag, st->hr, st->filt, st->cosv,
                  st->sinv, st->sfilt_d, st->view_off_d, st->view_idx_d, st->view_frame_d,
                  st->mask, st->nh_list, st->nh_count, st->nh_ord, st->nh_ord_count, st->ord,
                  st->ord_off, st->dummy, st->single, st->colt, st->cells[0], st->cells[1],
                  st->cell_voxmag, st->cell_upd, st->cell_zero, st->cell_ent, st->off_pnum,
                  st->off_pden, st->off_num, st->off_den, st->off_gamma, st->red,
                  st->split_rec, st->split_w};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (st->red_h) (void)hipHostFree(st->red_h);
  for (auto &e : st->ev) (void)hipEventDestroy(e);
  if (st->o0) (void)hipEventDestroy(st->o0);
  if (st->o1) (void)hipEventDestroy(st->o1);
  if (st->stream) (void)hipStreamDestroy(st->stream);
  delete st;
}

template <class T>
static hipError_t dalloc(T **p, size_t n) {
  return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
}

static int stage_create_impl(tb_stage *st, const float *projections, const float *weights,
                             const float *proj_angles, const float *proj_times,
                             const float *recon_times, int proj_rows, int proj_cols,
                             int proj_num, int recon_num, float vox_wid, float rot_center,
                             float sig_s, float sig_t, float c_s, float c_t, float convg_thresh,
                             int remove_rings, int quad_convex, float huber_delta, float huber_T,
                             int mult_xy, int mult_z, int stage_idx, int prev_mult_z,
                             double var_est) {
  if (proj_rows <= 0 || proj_cols <= 0 || proj_num <= 0 || recon_num <= 0)
    return fail("tb_stage_create: non-positive size");
  if (mult_xy <= 0 || mult_z <= 0 || proj_cols % mult_xy || proj_rows % mult_z)
    return fail("tb_stage_create: mult_xy=%d mult_z=%d do not divide %d x %d", mult_xy, mult_z,
                proj_cols, proj_rows);
  // initStatStructures (XT_Init.c:449-669), one process
  const double Length_R = (double)vox_wid * proj_cols, Length_T = (double)vox_wid * proj_rows;
  st->N_p = proj_num; st->N_r = proj_cols; st->N_t = proj_rows; st->N_time = recon_num;
  st->convg_thresh = convg_thresh; st->alpha = kAlpha;
  st->remove_rings = remove_rings; st->quad_convex = quad_convex;
  st->T = quad_convex == 0 ? (double)huber_T : kZingerOffT;
  st->delta = quad_convex == 0 ? (double)huber_delta : kZingerOffDelta;
  st->delta_r = Length_R / st->N_r;
  st->delta_t = Length_T / st->N_t;
  st->R0 = -(double)rot_center * st->delta_r;
  st->RMax = (st->N_r - (double)rot_center) * st->delta_r;
  st->T0 = -Length_T / 2.0;
  st->x0 = st->R0;
  st->y0 = -Length_R / 2.0;
  st->radius_obj = (double)vox_wid * proj_cols / 2;
  // view -> frame map (XT_Init.c:69-131): half-open [rt[i], rt[i+1])
  for (int i = 0; i < proj_num - 1; i++)
    if (proj_times[i + 1] < proj_times[i])
      return fail("Projection acquisition time decreased for increasing index.");
  for (int i = 0; i < recon_num; i++)
    if (recon_times[i + 1] < recon_times[i])
      return fail("Reconstruction times decreased for increasing index.");
  st->view_off.assign(recon_num + 1, 0);
  st->view_frame.assign(proj_num, -1);
  st->view_idx.clear();
  int start = 0;
  for (int i = 0; i < recon_num; i++) {
    st->view_off[i] = (int)st->view_idx.size();
    for (int j = start; j < proj_num; j++) {
      const double t = proj_times[j];
      if (t >= recon_times[i] && t < recon_times[i + 1]) {
        st->view_idx.push_back(j);
        st->view_frame[j] = i;
      }
      if (t >= recon_times[i + 1]) {
        start = j;
        break;
      }
    }
  }
  st->view_off[recon_num] = (int)st->view_idx.size();
  st->delta_recon = (double)st->view_idx.size() / recon_num;
  st->max_nv = 0;
  for (int i = 0; i < recon_num; i++)
    st->max_nv = std::max(st->max_nv, st->view_off[i + 1] - st->view_off[i]);
  // initDynStructures (XT_Init.c:211-445)
  st->stage_idx = stage_idx; st->mult_xy = mult_xy; st->mult_z = mult_z;
  if (stage_idx == 0) st->initICD = 0;
  else if (mult_z == prev_mult_z) st->initICD = 2;
  else if (prev_mult_z == 2 * mult_z) st->initICD = 3;
  else return fail("Multi-resolution scaling is not supported");
  st->updateProjOffset = stage_idx == 0 ? 0 : (stage_idx == 1 ? 2 : 3);
  if (remove_rings == 0) st->updateProjOffset = 0;
  st->initMagUpMap = stage_idx > 0;
  st->updateVar = stage_idx > 0 && quad_convex == 0;
  st->var_est = var_est;
  st->N_x = st->N_r / mult_xy; st->N_y = st->N_r / mult_xy; st->N_z = st->N_t / mult_z;
  // kernels index the object and a sinogram with 32-bit words and pack a
  // pixel as (y << 16) | x
  if (st->N_x >= 32768 || st->nobj() + 128 > (size_t)INT32_MAX ||
      st->nsino() > (size_t)INT32_MAX)
    return fail("tb_stage_create: problem too large for 32-bit indexing");
  st->delta_xy = mult_xy * st->delta_r;
  st->delta_z = mult_z * st->delta_t;
  st->OffsetR = (st->delta_xy / sqrt(2.0) + st->delta_r / 2.0) / kRespBins;
  st->sSq = pow(sig_s * (double)mult_xy, kMrfQ);
  st->sSqp = pow(sig_s * (double)mult_xy, kMrfQ - kMrfP);
  st->sTq = pow(sig_t * st->delta_recon, kMrfQ);
  st->sTqp = pow(sig_t * st->delta_recon, kMrfQ - kMrfP);
  st->c_s = c_s; st->c_t = c_t;
  // initFilter (XT_Init.c:144-190): inverse-distance weights normalised so
  // the 26 spatial and 2 temporal weights sum to mult_xy^3 * delta_recon
  {
    const double prior_const = (double)mult_xy * mult_xy * mult_xy * st->delta_recon;
    double sum = 2.0;
    for (int i = 0; i < 27; i++) {
      const int a = i / 9 - 1, b = (i / 3) % 3 - 1, c = i % 3 - 1;
      const double d2 = a * a + b * b + c * c;
      st->sfilt[i] = d2 > 0 ? 1.0 / sqrt(d2) : 0.0;
      sum += st->sfilt[i];
    }
    for (int i = 0; i < 27; i++) st->sfilt[i] = prior_const * st->sfilt[i] / sum;
    st->tfilt = prior_const * 1.0 / sum;
  }
  // mask (XT_ICD_update.c:1140-1149)
  std::vector<uint8_t> mask_h(st->nxy());
  st->U = 0;
  for (int j = 0; j < st->N_y; j++)
    for (int k = 0; k < st->N_x; k++) {
      const double x = st->x0 + ((double)k + 0.5) * st->delta_xy;
      const double yv = st->y0 + ((double)j + 0.5) * st->delta_xy;
      const uint8_t m = (x * x + yv * yv < st->radius_obj * st->radius_obj) ? 1 : 0;
      mask_h[(size_t)j * st->N_x + k] = m;
      st->U += m;
    }
  std::vector<float> hr_h((size_t)st->N_p * kHrStride);
  detector_response(proj_angles, st->N_p, st->delta_xy, st->OffsetR, st->delta_r, hr_h.data());
  std::vector<double> cs(st->N_p), sn(st->N_p);
  for (int i = 0; i < st->N_p; i++) {  // calculateSinCos (XT_Init.c:193-204)
    cs[i] = cos((double)proj_angles[i]);
    sn[i] = sin((double)proj_angles[i]);
  }
  st->angles.assign(proj_angles, proj_angles + proj_num);

  // device buffers, on the device current at creation
  HIPCHK(hipGetDevice(&st->device));
  HIPCHK(hipStreamCreateWithFlags(&st->stream, hipStreamNonBlocking));
  const size_t ns = st->nsino(), no = st->nobj(), nxy = st->nxy();
  const int ncell = st->N_time * st->N_z;
  HIPCHK(dalloc(&st->e, ns));
  HIPCHK(dalloc(&st->w, ns));
  HIPCHK(dalloc(&st->y, ns));
  HIPCHK(dalloc(&st->d, (size_t)st->N_r * st->N_t));
  // past the object: 64 words that take the stores of lanes with no voxel,
  // then a NaN word that stands for a neighbour outside the object
  HIPCHK(dalloc(&st->obj, no + 128));
  HIPCHK(dalloc(&st->mag, no + 64));
  HIPCHK(dalloc(&st->filt, no));
  HIPCHK(dalloc(&st->hr, hr_h.size()));
  HIPCHK(dalloc(&st->cosv, st->N_p));
  HIPCHK(dalloc(&st->sinv, st->N_p));
  HIPCHK(dalloc(&st->sfilt_d, 27));
  HIPCHK(dalloc(&st->view_off_d, st->view_off.size()));
  HIPCHK(dalloc(&st->view_idx_d, st->view_idx.size()));
  HIPCHK(dalloc(&st->view_frame_d, st->view_frame.size()));
  HIPCHK(dalloc(&st->mask, nxy));
  st->nh_cap = std::max(1, st->U / kNhicdK);
  HIPCHK(dalloc(&st->nh_list, (size_t)ncell * st->nh_cap));
  HIPCHK(dalloc(&st->nh_count, ncell));
  HIPCHK(dalloc(&st->nh_ord, (size_t)ncell * st->nh_cap));
  HIPCHK(dalloc(&st->nh_ord_count, ncell));
  st->ord_cap = std::max(1, st->U);
  HIPCHK(dalloc(&st->ord, (size_t)ncell * st->ord_cap));
  HIPCHK(dalloc(&st->ord_off, (size_t)ncell * (kNhicdK + 1)));
  HIPCHK(dalloc(&st->dummy, (size_t)ncell * 64));
  HIPCHK(dalloc(&st->single, 64));
  {
    int dev = 0, lds = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    st->lds_max = (size_t)lds;
  }
  HIPCHK(dalloc(&st->cell_voxmag, ncell));
  HIPCHK(dalloc(&st->cell_upd, ncell));
  HIPCHK(dalloc(&st->cell_zero, ncell));
  HIPCHK(dalloc(&st->cell_ent, ncell));
  const int npix = st->N_r * st->N_t;
  // 32 views per partial: the partial sums (npix x N_p / 32 of them) and
  // the finalize pass over them, which adds them in chunk order, are both
  // a few rounds of eight loads in flight
  st->off_vpc = std::max(1, std::min(st->N_p, 32));
  st->off_chunks = (st->N_p + st->off_vpc - 1) / st->off_vpc;
  HIPCHK(dalloc(&st->off_pnum, (size_t)npix * st->off_chunks));
  HIPCHK(dalloc(&st->off_pden, (size_t)npix * st->off_chunks));
  HIPCHK(dalloc(&st->off_num, npix));
  HIPCHK(dalloc(&st->off_den, npix));
  HIPCHK(dalloc(&st->off_gamma, 1));
  HIPCHK(dalloc(&st->red, st->red_n));
  HIPCHK(hipHostMalloc((void **)&st->red_h, st->red_n * sizeof(double), hipHostMallocDefault));
  // cells of each colour: (f + z) & 1
  for (int c = 0; c < 2; c++) {
    std::vector<int2> cl;
    for (int f = 0; f < st->N_time; f++)
      for (int z = 0; z < st->N_z; z++)
        if (((f + z) & 1) == c) cl.push_back(make_int2(f, z));
    st->ncells[c] = (int)cl.size();
    HIPCHK(dalloc(&st->cells[c], cl.size()));
    if (!cl.empty())
      HIPCHK(hipMemcpyAsync(st->cells[c], cl.data(), cl.size() * sizeof(int2),
                            hipMemcpyHostToDevice, st->stream));
    st->cells_h[c] = cl;
  }
  HIPCHK(hipMemcpyAsync(st->hr, hr_h.data(), hr_h.size() * sizeof(float), hipMemcpyHostToDevice,
                        st->stream));
  HIPCHK(hipMemcpyAsync(st->cosv, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice,
                        st->stream));
  HIPCHK(hipMemcpyAsync(st->sinv, sn.data(), sn.size() * sizeof(double), hipMemcpyHostToDevice,
                        st->stream));
  HIPCHK(hipMemcpyAsync(st->sfilt_d, st->sfilt, sizeof(st->sfilt), hipMemcpyHostToDevice,
                        st->stream));
  HIPCHK(hipMemcpyAsync(st->view_off_d, st->view_off.data(), st->view_off.size() * sizeof(int),
                        hipMemcpyHostToDevice, st->stream));
  if (!st->view_idx.empty())
    HIPCHK(hipMemcpyAsync(st->view_idx_d, st->view_idx.data(), st->view_idx.size() * sizeof(int),
                          hipMemcpyHostToDevice, st->stream));
  HIPCHK(hipMemcpyAsync(st->view_frame_d, st->view_frame.data(),
                        st->view_frame.size() * sizeof(int), hipMemcpyHostToDevice, st->stream));
  HIPCHK(hipMemcpyAsync(st->mask, mask_h.data(), nxy, hipMemcpyHostToDevice, st->stream));
  // raw scan -> device layout
  float *raw_p = nullptr, *raw_w = nullptr;
  HIPCHK(dalloc(&raw_p, ns));
  HIPCHK(dalloc(&raw_w, ns));
  HIPCHK(hipMemcpyAsync(raw_p, projections, ns * sizeof(float), hipMemcpyHostToDevice,
                        st->stream));
  HIPCHK(hipMemcpyAsync(raw_w, weights, ns * sizeof(float), hipMemcpyHostToDevice, st->stream));
  KP P = st->kp();
  hipLaunchKernelGGL(k_transpose_in, dim3(grid_for(ns)), dim3(256), 0, st->stream, P, raw_p,
                     raw_w, var_est);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(no)), dim3(256), 0, st->stream, st->obj, no, 0.f);
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(no)), dim3(256), 0, st->stream, st->mag, no, 0.f);
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(128), 0, st->stream, st->obj + no, (size_t)128,
                     __builtin_nanf(""));
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(npix)), dim3(256), 0, st->stream, st->d,
                     (size_t)npix, 0.f);
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(ncell)), dim3(256), 0, st->stream,
                     st->cell_voxmag, (size_t)ncell, 0.0);
  // no NHICD lines until the first selection (all-zero bits = int 0)
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(ncell)), dim3(256), 0, st->stream,
                     reinterpret_cast<float *>(st->nh_count), (size_t)ncell, 0.f);
  hipLaunchKernelGGL(k_fill_f32, dim3(grid_for(ncell)), dim3(256), 0, st->stream,
                     reinterpret_cast<float *>(st->nh_ord_count), (size_t)ncell, 0.f);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ncell)), dim3(256), 0, st->stream, st->cell_upd,
                     (size_t)ncell);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ncell)), dim3(256), 0, st->stream, st->cell_zero,
                     (size_t)ncell);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ncell)), dim3(256), 0, st->stream, st->cell_ent,
                     (size_t)ncell);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st->stream));
  HIPCHK(hipFree(raw_p));
  HIPCHK(hipFree(raw_w));
  st->ev.resize(256);
  for (auto &e : st->ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventCreate(&st->o0));
  HIPCHK(hipEventCreate(&st->o1));
  return 0;
}

extern "C" int tb_stage_create(tb_stage **out, const float *projections, const float *weights,
                               const float *proj_angles, const float *proj_times,
                               const float *recon_times, int32_t proj_rows, int32_t proj_cols,
                               int32_t proj_num, int32_t recon_num, float vox_wid,
                               float rot_center, float sig_s, float sig_t, float c_s, float c_t,
                               float convg_thresh, int32_t remove_rings, int32_t quad_convex,
                               float huber_delta, float huber_T, int32_t mult_xy,
                               int32_t mult_z, int32_t stage_idx, int32_t prev_mult_z,
                               double var_est) {
  *out = nullptr;
  tb_stage *st = new tb_stage();
  int rc = stage_create_impl(st, projections, weights, proj_angles, proj_times, recon_times,
                             proj_rows, proj_cols, proj_num, recon_num, vox_wid, rot_center,
                             sig_s, sig_t, c_s, c_t, convg_thresh, remove_rings, quad_convex,
                             huber_delta, huber_T, mult_xy, mult_z, stage_idx, prev_mult_z,
                             var_est);
  if (rc) {
    std::string keep = g_err;
    tb_stage_destroy(st);
    g_err = keep;
    return -1;
  }
  *out = st;
  return 0;
}

extern "C" int tb_stage_info(const tb_stage *st, int32_t *ints, double *d) {
  int32_t iv[] = {st->N_p, st->N_r, st->N_t, st->N_x, st->N_y, st->N_z, st->N_time,
                  st->mult_xy, st->mult_z, st->N_z, st->updateProjOffset, st->updateVar,
                  st->initICD, st->U};
  memcpy(ints, iv, sizeof(iv));
  double dv[] = {st->delta_r, st->delta_t, st->R0, st->RMax, st->T0, st->x0, st->y0,
                 st->delta_xy, st->delta_z, st->OffsetR, st->radius_obj, st->sSq, st->sSqp,
                 st->sTq, st->sTqp, st->c_s, st->c_t, st->tfilt, st->alpha, st->T, st->delta,
                 st->var_est, st->delta_recon};
  memcpy(d, dv, sizeof(dv));
  memcpy(d + sizeof(dv) / sizeof(double), st->sfilt, sizeof(st->sfilt));
  return 0;
}

// ---------------------------------------------------------------------------
// transfers
// ---------------------------------------------------------------------------
static int h2d(tb_stage *st, void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  return 0;
}
static int d2h(tb_stage *st, void *dst, const void *src, size_t bytes) {
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  return 0;
}

extern "C" int tb_set_object(tb_stage *st, const float *o) {
  return h2d(st, st->obj, o, st->nobj() * sizeof(float));
}
extern "C" int tb_get_object(tb_stage *st, float *o) {
  return d2h(st, o, st->obj, st->nobj() * sizeof(float));
}
extern "C" int tb_set_offset(tb_stage *st, const float *d) {
  return h2d(st, st->d, d, (size_t)st->N_r * st->N_t * sizeof(float));
}
extern "C" int tb_get_offset(tb_stage *st, float *d) {
  return d2h(st, d, st->d, (size_t)st->N_r * st->N_t * sizeof(float));
}
extern "C" int tb_set_magmap(tb_stage *st, const float *m) {
  return h2d(st, st->mag, m, st->nobj() * sizeof(float));
}
extern "C" int tb_get_magmap(tb_stage *st, float *m) {
  return d2h(st, m, st->mag, st->nobj() * sizeof(float));
}
extern "C" int tb_get_detector_response(tb_stage *st, float *h) {
  return d2h(st, h, st->hr, (size_t)st->N_p * kHrStride * sizeof(float));
}
extern "C" double tb_get_var_est(const tb_stage *st) { return st->var_est; }

static int sino_to_host(tb_stage *st, const float *src, float *dst) {
  float *tmp = nullptr;
  HIPCHK(dalloc(&tmp, st->nsino()));
  KP P = st->kp();
  hipLaunchKernelGGL(k_to_ref_layout, dim3(grid_for(st->nsino())), dim3(256), 0, st->stream, P,
                     src, tmp);
  HIPCHK(hipGetLastError());
  int rc = d2h(st, dst, tmp, st->nsino() * sizeof(float));
  (void)hipFree(tmp);
  return rc;
}
extern "C" int tb_get_error_sino(tb_stage *st, float *e) { return sino_to_host(st, st->e, e); }
extern "C" int tb_get_weights(tb_stage *st, float *w) { return sino_to_host(st, st->w, w); }

// initObject with initICD 2 (x-y) or 3 (x-y-z), XT_ICD_update.c:323-651:
// bilinear upsampling by 2; 2D intermediate rows are rounded to float as the
// reference's Real_arr_t buffers are, 3D intermediates stay double.
static void up2d(const float *in, int Ny, int Nx, bool float_mid, std::vector<double> &tmp,
                 double *out) {
  const int W = 2 * Nx;
  tmp.assign((size_t)Ny * W, 0.0);
  auto rnd = [&](double v) { return float_mid ? (double)(float)v : v; };
  for (int j = 0; j < Ny; j++) {
    const float *r = in + (size_t)j * Nx;
    double *t = tmp.data() + (size_t)j * W;
    t[0] = rnd(r[0]);
    t[1] = rnd((3.0 * r[0] + r[1]) / 4.0);
    t[W - 1] = rnd(r[Nx - 1]);
    t[W - 2] = rnd((r[Nx - 2] + 3.0 * r[Nx - 1]) / 4.0);
    for (int k = 1; k < Nx - 1; k++) {
      t[2 * k] = rnd((r[k - 1] + 3.0 * r[k]) / 4.0);
      t[2 * k + 1] = rnd((3.0 * r[k] + r[k + 1]) / 4.0);
    }
  }
  const double *T = tmp.data();
  for (int k = 0; k < W; k++) {
    out[k] = T[k];
    out[W + k] = (3.0 * T[k] + T[W + k]) / 4.0;
    out[(size_t)(2 * Ny - 1) * W + k] = T[(size_t)(Ny - 1) * W + k];
    out[(size_t)(2 * Ny - 2) * W + k] =
        (T[(size_t)(Ny - 2) * W + k] + 3.0 * T[(size_t)(Ny - 1) * W + k]) / 4.0;
  }
  for (int j = 1; j < Ny - 1; j++)
    for (int k = 0; k < W; k++) {
      out[(size_t)(2 * j) * W + k] = (T[(size_t)(j - 1) * W + k] + 3.0 * T[(size_t)j * W + k]) / 4.0;
      out[(size_t)(2 * j + 1) * W + k] =
          (3 * T[(size_t)j * W + k] + T[(size_t)(j + 1) * W + k]) / 4.0;
    }
}

static void upz(const double *in, int Nz, size_t P, double *out) {
  for (size_t q = 0; q < P; q++) {
    out[q] = in[q];
    out[P + q] = (3.0 * in[q] + in[P + q]) / 4.0;
    out[(size_t)(2 * Nz - 1) * P + q] = in[(size_t)(Nz - 1) * P + q];
    out[(size_t)(2 * Nz - 2) * P + q] =
        (3.0 * in[(size_t)(Nz - 1) * P + q] + in[(size_t)(Nz - 2) * P + q]) / 4.0;
  }
  for (int s = 1; s < Nz - 1; s++)
    for (size_t q = 0; q < P; q++) {
      out[(size_t)(2 * s) * P + q] = (in[(size_t)(s - 1) * P + q] + 3.0 * in[(size_t)s * P + q]) / 4.0;
      out[(size_t)(2 * s + 1) * P + q] =
          (3.0 * in[(size_t)s * P + q] + in[(size_t)(s + 1) * P + q]) / 4.0;
    }
}

static void upsample(const float *prev, int prev_nz, int Ny, int Nx, int N_time, int out_nz,
                     bool three_d, bool float_mid, std::vector<float> &out) {
  const size_t P2 = (size_t)4 * Ny * Nx, P1 = (size_t)Ny * Nx;
  out.assign((size_t)N_time * out_nz * P2, 0.f);
  std::vector<double> tmp, up(P2 * prev_nz), uz(P2 * 2 * prev_nz);
  for (int f = 0; f < N_time; f++) {
    for (int s = 0; s < prev_nz; s++)
      up2d(prev + ((size_t)f * prev_nz + s) * P1, Ny, Nx, float_mid && !three_d, tmp,
           up.data() + (size_t)s * P2);
    const double *src = up.data();
    if (three_d) {
      upz(up.data(), prev_nz, P2, uz.data());
      src = uz.data();
    }
    for (size_t q = 0; q < (size_t)out_nz * P2; q++)
      out[(size_t)f * out_nz * P2 + q] = (float)src[q];
  }
}

extern "C" int tb_init_from_coarse(tb_stage *st, const float *prev_obj, int32_t prev_nz,
                                   const float *prev_mag) {
  if (st->initICD != 2 && st->initICD != 3) return fail("tb_init_from_coarse: stage 0");
  const bool three = st->initICD == 3;
  if ((three && prev_nz * 2 != st->N_z) || (!three && prev_nz != st->N_z))
    return fail("tb_init_from_coarse: previous N_z %d incompatible with %d", prev_nz, st->N_z);
  std::vector<float> o;
  upsample(prev_obj, prev_nz, st->N_y / 2, st->N_x / 2, st->N_time, st->N_z, three, true, o);
  if (tb_set_object(st, o.data())) return -1;
  if (prev_mag && st->initMagUpMap) {
    // one z-block per slice in both stages, so the block counts follow N_z
    upsample(prev_mag, prev_nz, st->N_y / 2, st->N_x / 2, st->N_time, st->N_z, three, true, o);
    if (tb_set_magmap(st, o.data())) return -1;
  }
  return 0;
}

// The same handoff from the previous stage's device buffers: the object and
// the update-magnitude map are upsampled in HBM, no host copy.
extern "C" int tb_init_from_stage(tb_stage *st, tb_stage *prev) {
  if (!prev) return fail("tb_init_from_stage: no previous stage");
  if (st->initICD != 2 && st->initICD != 3) return fail("tb_init_from_stage: stage 0");
  const bool three = st->initICD == 3;
  if (prev->N_time != st->N_time || 2 * prev->N_x != st->N_x || 2 * prev->N_y != st->N_y ||
      (three ? 2 * prev->N_z : prev->N_z) != st->N_z)
    return fail("tb_init_from_stage: previous stage %dx%dx%d does not upsample to %dx%dx%d",
                prev->N_x, prev->N_y, prev->N_z, st->N_x, st->N_y, st->N_z);
  HIPCHK(hipStreamSynchronize(prev->stream));
  hipLaunchKernelGGL(k_upsample, dim3(grid_for(st->nobj())), dim3(256), 0, st->stream, prev->obj,
                     st->obj, st->N_time, prev->N_z, prev->N_y, prev->N_x, three ? 1 : 0);
  if (st->initMagUpMap)
    hipLaunchKernelGGL(k_upsample, dim3(grid_for(st->nobj())), dim3(256), 0, st->stream,
                       prev->mag, st->mag, st->N_time, prev->N_z, prev->N_y, prev->N_x,
                       three ? 1 : 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st->stream));
  return 0;
}

// ---------------------------------------------------------------------------
// timing helpers
// ---------------------------------------------------------------------------
static hipEvent_t *ev_pair(tb_stage *st) {
  if (st->ev_used + 2 > (int)st->ev.size()) return nullptr;
  hipEvent_t *p = &st->ev[st->ev_used];
  st->ev_used += 2;
  return p;
}

static int harvest_events(tb_stage *st) {
  if (st->split_ran) {
    // a cross-workgroup wait of any split-cell launch since the last check
    // that timed out (the word only accumulates on the device; it is
    // cleared here, after it has been read)
    unsigned err = 0;
    HIPCHK(hipMemcpyAsync(&err, st->split_w, sizeof(err),
                          hipMemcpyDeviceToHost, st->stream));
    HIPCHK(hipStreamSynchronize(st->stream));
    st->split_ran = false;
    if (err) {
      hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, st->stream,
                         reinterpret_cast<float *>(st->split_w), (size_t)1, 0.0f);
      HIPCHK(hipGetLastError());
      HIPCHK(hipStreamSynchronize(st->stream));
      return fail("k_icd_split: a cross-workgroup wait timed out (results invalid)");
    }
  }
  if (st->ev_used == 0) return 0;
  HIPCHK(hipEventSynchronize(st->ev[st->ev_used - 1]));
  for (int i = 0; i < st->ev_used; i += 2) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, st->ev[i], st->ev[i + 1]));
    st->voxel_ms += ms;
  }
  st->ev_used = 0;
  return 0;
}

// ---------------------------------------------------------------------------
// operations
// ---------------------------------------------------------------------------
// The slice-interleaved object copy the forward projector reads: one record
// of kFwdZ floats per (frame, group of kFwdZ slices, pixel).
static size_t fwd_records(const tb_stage *st) {
  const int ngroups = (st->N_z + kFwdZ - 1) / kFwdZ;
  return (size_t)st->N_time * ngroups * st->N_x * st->N_y;
}

// Forward projection into `out` (mode 0: A x, mode 1: the error sinogram):
// the object is copied slice-interleaved into objt (2 float4 per record;
// a temporary buffer when objt is null), then k_forward runs a block per
// (view, tile of kFwdT bins) x group of kFwdZ slices.  Returns after the
// stream has finished.
static int run_forward(tb_stage *st, float *out, int mode, float4 *objt = nullptr) {
  KP P = st->kp();
  const int ngroups = (st->N_z + kFwdZ - 1) / kFwdZ;
  const int ntile = (st->N_r + kFwdT - 1) / kFwdT;
  const size_t nrec = fwd_records(st);
  const bool own = objt == nullptr;
  if (own) HIPCHK(dalloc(&objt, 2 * nrec));
  hipLaunchKernelGGL(k_fwd_interleave, dim3(grid_for(nrec)), dim3(256), 0, st->stream, P, objt);
  hipLaunchKernelGGL(k_forward, dim3(ntile * st->N_p, ngroups), dim3(kFwdT), 0, st->stream, P,
                     st->view_frame_d, objt, out, mode);
  hipError_t err = hipGetLastError();
  if (err == hipSuccess) err = hipStreamSynchronize(st->stream);
  if (own) (void)hipFree(objt);
  HIPCHK(err);
  return 0;
}

extern "C" int tb_init_error_sino(tb_stage *st) {
  if (run_forward(st, st->e, 1)) return -1;
  return 0;
}

// Scratch of tb_init_error_sino_static: the interleaved object copy, in a
// file-scope array so that the call allocates nothing.  Every GPU the
// library runs on holds its own copy for the life of the process; host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name in host code is not a device address).
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) float4 scratch[kStaticScratchBytes / sizeof(float4)];
static std::mutex g_scratch_mu;
static float4 *g_scratch_dev[kMaxDevices];

static int static_scratch(int dev, float4 **out) {
  if (dev < 0 || dev >= kMaxDevices) return fail("static scratch: device %d out of range", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail("static scratch: no address on device %d", dev);
    g_scratch_dev[dev] = static_cast<float4 *>(p);
  }
  *out = g_scratch_dev[dev];
  return 0;
}

extern "C" int tb_init_error_sino_static(tb_stage *st) {
  const size_t need = 2 * fwd_records(st) * sizeof(float4);
  if (need > kStaticScratchBytes)
    return fail("tb_init_error_sino_static: the interleaved object takes %zu bytes "
                "(%d frames x %d slices x %d x %d), more than the %zu-byte static scratch; "
                "use tb_init_error_sino",
                need, st->N_time, st->N_z, st->N_y, st->N_x, kStaticScratchBytes);
  int cur = 0;
  HIPCHK(hipGetDevice(&cur));
  if (cur != st->device) HIPCHK(hipSetDevice(st->device));
  float4 *buf = nullptr;
  int rc = static_scratch(st->device, &buf);
  if (!rc) rc = run_forward(st, st->e, 1, buf);
  if (cur != st->device) {
    const std::string keep = g_err;
    if (hipSetDevice(cur) != hipSuccess && !rc) rc = fail("tb_init_error_sino_static: restoring device %d", cur);
    else if (rc) g_err = keep;
  }
  return rc;
}

extern "C" int tb_init_fbp(tb_stage *st) {
  if ((size_t)st->N_r * sizeof(double) > 65536)
    return fail("tb_init_fbp: %d detector bins exceed the filter's LDS row", st->N_r);
  KP P = st->kp();
  float *q = nullptr;
  HIPCHK(dalloc(&q, (size_t)st->N_z * st->N_p * st->N_r));
  hipLaunchKernelGGL(k_fbp_filter, dim3(st->N_p, st->N_z), dim3(256),
                     st->N_r * sizeof(double), st->stream, P, q);
  hipLaunchKernelGGL(k_fbp_back, dim3(grid_for(st->nobj())), dim3(256), 0, st->stream, P, q);
  hipError_t err = hipGetLastError();
  if (err == hipSuccess) err = hipStreamSynchronize(st->stream);
  (void)hipFree(q);
  HIPCHK(err);
  return 0;
}

extern "C" int tb_forward_project(tb_stage *st, float *out) {
  float *tmp = nullptr;
  HIPCHK(dalloc(&tmp, st->nsino()));
  if (run_forward(st, tmp, 0)) {
    (void)hipFree(tmp);
    return -1;
  }
  int rc = sino_to_host(st, tmp, out);
  (void)hipFree(tmp);
  return rc;
}

static int reduce_partials(tb_stage *st, int nblocks, double *out) {
  HIPCHK(hipMemcpyAsync(st->red_h, st->red, nblocks * sizeof(double), hipMemcpyDeviceToHost,
                        st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  double s = 0;
  for (int i = 0; i < nblocks; i++) s += st->red_h[i];
  *out = s;
  return 0;
}

extern "C" int tb_cost(tb_stage *st, double *total, double *forward, double *prior) {
  KP P = st->kp();
  const int nb = std::min(st->red_n, grid_for(st->nsino()));
  hipLaunchKernelGGL(k_cost_data, dim3(nb), dim3(256), 0, st->stream, P, st->red);
  HIPCHK(hipGetLastError());
  double fwd, pri;
  if (reduce_partials(st, nb, &fwd)) return -1;
  fwd /= 2.0;
  const int nb2 = std::min(st->red_n, grid_for(st->nobj()));
  hipLaunchKernelGGL(k_cost_prior, dim3(nb2), dim3(256), 0, st->stream, P, st->sSq, st->sSqp,
                     st->c_s, st->sTq, st->sTqp, st->c_t, st->sfilt_d, st->tfilt, st->red);
  HIPCHK(hipGetLastError());
  if (reduce_partials(st, nb2, &pri)) return -1;
  fwd += (double)st->N_p * (double)st->N_r * (double)st->N_t * log(st->var_est) / 2;
  if (forward) *forward = fwd;
  if (prior) *prior = pri;
  if (total) *total = fwd + pri;
  return 0;
}

static int offset_update_async(tb_stage *st, int type) {
  if (type < 1 || type > 3) return fail("offset type %d", type);
  KP P = st->kp();
  const int npix = st->N_r * st->N_t;
  hipLaunchKernelGGL(k_offset_partial, dim3(grid_for(npix, 64), st->off_chunks), dim3(64), 0,
                     st->stream, P, type, st->off_vpc, st->off_pnum, st->off_pden);
  hipLaunchKernelGGL(k_offset_finalize, dim3(grid_for(npix)), dim3(256), 0, st->stream, P,
                     st->off_chunks, st->off_pnum, st->off_pden, st->off_num, st->off_den);
  if (type == 2)
    hipLaunchKernelGGL(k_offset_gamma, dim3(1), dim3(256), 0, st->stream, P, st->off_num,
                       st->off_den, st->off_gamma);
  hipLaunchKernelGGL(k_offset_apply, dim3(grid_for(st->nsino())), dim3(256), 0, st->stream, P,
                     type, st->off_num, st->off_den, st->off_gamma);
  hipLaunchKernelGGL(k_offset_store, dim3(grid_for(npix)), dim3(256), 0, st->stream, P, type,
                     st->off_num, st->off_den, st->off_gamma);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int tb_offset_update(tb_stage *st, int32_t type) {
  if (offset_update_async(st, type)) return -1;
  HIPCHK(hipStreamSynchronize(st->stream));
  return 0;
}

extern "C" int tb_variance_update(tb_stage *st) {
  KP P = st->kp();
  const int nb = std::min(st->red_n, grid_for(st->nsino()));
  hipLaunchKernelGGL(k_var_partial, dim3(nb), dim3(256), 0, st->stream, P, st->var_est, st->red);
  HIPCHK(hipGetLastError());
  double acc;
  if (reduce_partials(st, nb, &acc)) return -1;
  const double vnew = acc / ((double)st->N_p * (double)st->N_r * (double)st->N_t);
  hipLaunchKernelGGL(k_var_apply, dim3(grid_for(st->nsino())), dim3(256), 0, st->stream, P,
                     st->var_est, vnew);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st->stream));
  st->var_est = vnew;
  return 0;
}

// the HICD voxel order of iteration `iter` for every cell, built once per
// iteration
static int ensure_order(tb_stage *st, int iter) {
  if (st->ord_iter == iter) return 0;
  KP P = st->kp();
  hipLaunchKernelGGL(k_build_order, dim3(st->N_time * st->N_z), dim3(256), 0, st->stream, P, iter);
  HIPCHK(hipGetLastError());
  st->ord_iter = iter;
  return 0;
}

// The A-matrix column table of the multi-wavefront kernel: it depends on the
// stage's geometry only, so it is built once, by a kernel over every
// (frame, pixel, view slot).  Returns 0 when the table is ready, 1 when it
// would not fit (the caller then uses another kernel), -1 on a HIP error.
static int ensure_columns(tb_stage *st, int nb) {
  if (st->colt && st->colt_nb == nb && !g_opt.no_column_table) return 0;
  if (st->colt) {
    HIPCHK(hipStreamSynchronize(st->stream));
    HIPCHK(hipFree(st->colt));
    st->colt = nullptr;
    st->colt_nb = 0;
  }
  if (g_opt.no_column_table) return 1;
  const int rec = col_rec(nb);
  const size_t n = (size_t)st->N_time * st->nxy() * st->max_nv * rec;
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b));
  if (n * sizeof(float) > free_b / 4) return 1;
  if (dalloc(&st->colt, n) != hipSuccess) {
    (void)hipGetLastError();
    st->colt = nullptr;
    return 1;
  }
  st->colt_nb = nb;
  st->colt_builds++;
  KP P = st->kp();
  const size_t recs = n / rec;
  const int grid = (int)std::min<size_t>((recs + 255) / 256, 65536);
  if (nb == 17)
    hipLaunchKernelGGL(k_build_columns<17>, dim3(grid), dim3(256), 0, st->stream, P);
  else if (nb == 9) hipLaunchKernelGGL(k_build_columns<9>, dim3(grid), dim3(256), 0, st->stream, P);
  else if (nb == 5)
    hipLaunchKernelGGL(k_build_columns<5>, dim3(grid), dim3(256), 0, st->stream, P);
  else hipLaunchKernelGGL(k_build_columns<3>, dim3(grid), dim3(256), 0, st->stream, P);
  HIPCHK(hipGetLastError());
  return 0;
}

static int autotune_pick(tb_stage *st, int iter, int *kernel, int *vpl);

// The split-cell kernel (k_icd_split) for a colour of `ncells` cells: the
// column width and rows per lane of the stage, S segments per cell (as many
// as the compute units allow, each of at least 4 views, at most 64), a
// cooperative launch so that every workgroup is resident.  Sets *launched;
// returns 0 also when the kernel does not serve the stage.
constexpr int kSplitMaxCells = 32;
constexpr int kSplitSegMax = 64;
// Split-cell launches on a device, from any stage and stream, run one after
// another: each waits for the previous one's completion event, so two
// cooperative grids are never resident together and the launch path's
// set-up calls are made by one host thread at a time.  The events live as
// long as the process.
static std::mutex g_split_mu;
static hipEvent_t g_split_ev[kMaxDevices];
static int launch_split(tb_stage *st, const int2 *cells, int ncells, int iter, int chunk, int seq,
                        int count_mag, bool *launched) {
  const int w = 2 * st->mult_xy + 1;
  const int nbx = w <= 3 ? 3 : w <= 5 ? 5 : w <= 9 ? 9 : w <= 17 ? 17 : 0;
  const int mz = st->mult_z;
  if (!nbx || st->N_r < nbx || (nbx == 3 && mz != 1) || ncells <= 0) return 0;
  if (g_opt.split_segments == 1) return 0;
  std::lock_guard<std::mutex> guard(g_split_mu);
  const int R = (nbx > 3 && mz % 2 == 0) ? 2 : 1;
  const int ngr = mz / R;
  using split_fn = void (*)(KP, const int2 *, int, int, int, int, int, SplitBuf);
  split_fn fn = nbx == 17 ? (R == 2 ? k_icd_split<17, 2, 8> : k_icd_split<17, 1, 8>)
              : nbx == 9  ? (R == 2 ? k_icd_split<9, 2, 8> : k_icd_split<9, 1, 8>)
              : nbx == 5  ? (R == 2 ? k_icd_split<5, 2, 8> : k_icd_split<5, 1, 8>)
                          : k_icd_split<3, 1, 8>;
  if (!st->n_cus) {
    int dev = 0, n = 0, coop = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    HIPCHK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    st->n_cus = coop ? n : -1;
  }
  if (st->n_cus <= 0) return 0;
  // segments: views of at least 4 each, lanes of at most 8 theta waves
  int S = std::min(kSplitSegMax, std::max(1, st->n_cus / ncells));
  if (g_opt.split_segments > 0) S = std::min(S, g_opt.split_segments);
  S = std::min(S, std::max(1, st->max_nv / 4));
  int vseg = (st->max_nv + S - 1) / S;
  while (vseg * ngr > 8 * 64 && S < kSplitSegMax) {
    S++;
    vseg = (st->max_nv + S - 1) / S;
  }
  if (vseg * ngr > 8 * 64) return 0;
  int nth = (vseg * ngr + 63) / 64;
  size_t lds = (size_t)vseg * kHrStride * sizeof(float);
  if (lds + 4096 > st->lds_max) return 0;
  HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  // every workgroup resident at once
  int per_cu = 0;
  HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * (nth + 1), lds));
  while (S >= 2 && (long long)ncells * S > (long long)per_cu * st->n_cus) {
    S--;
    vseg = (st->max_nv + S - 1) / S;
    if (vseg * ngr > 8 * 64) return 0;
    const int nth2 = (vseg * ngr + 63) / 64;
    const size_t lds2 = (size_t)vseg * kHrStride * sizeof(float);
    if (nth2 != nth || lds2 != lds) {
      nth = nth2;
      lds = lds2;
      if (lds + 4096 > st->lds_max) return 0;
      HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * (nth + 1), lds));
    }
  }
  if (S < 2) return 0;
  // exchange buffers, one slot per cell of the larger colour
  const int slots = std::max(st->ncells[0], st->ncells[1]);
  if (st->split_slots < slots) {
    if (st->split_rec) (void)hipFree(st->split_rec);
    st->split_rec = nullptr;
    st->split_slots = 0;
    HIPCHK(dalloc(&st->split_rec, (size_t)slots * 2 * kSplitSegMax * 4));
    if (!st->split_w) {
      HIPCHK(dalloc(&st->split_w, 2));
      hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, st->stream,
                         reinterpret_cast<float *>(st->split_w), (size_t)2, 0.0f);
    }
    st->split_slots = slots;
  }
  if (ensure_columns(st, nbx) < 0) return -1;
  KP P = st->kp();
  if (seq == 0 && ensure_order(st, iter)) return -1;
  SplitBuf B;
  B.rec = st->split_rec;
  B.err = st->split_w;
  B.abort = st->split_w + 1;
  B.S = S;
  B.smax = kSplitSegMax;
  // polls per wait: 2^22 (about a second) unless the option asks fewer
  B.spin_limit = g_opt.split_spin_limit > 0 ? (unsigned)g_opt.split_spin_limit - 1u : 1u << 22;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail("launch_split: device %d", dev);
  if (g_split_ev[dev]) HIPCHK(hipStreamWaitEvent(st->stream, g_split_ev[dev], 0));
  else HIPCHK(hipEventCreateWithFlags(&g_split_ev[dev], hipEventDisableTiming));
  hipLaunchKernelGGL(k_split_reset, dim3(std::min(64, st->split_slots)), dim3(256), 0, st->stream,
                     B, st->split_slots);
  HIPCHK(hipGetLastError());
  void *args[] = {&P, &cells, &iter, &chunk, &seq, &count_mag, const_cast<int *>(&ngr), &B};
  HIPCHK(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(fn), dim3(ncells * S),
                                    dim3(64 * (nth + 1)), args, (unsigned)lds, st->stream));
  HIPCHK(hipEventRecord(g_split_ev[dev], st->stream));
  st->global_launches++;
  st->fam_launches[TB_K_SPLIT]++;
  st->split_ran = true;
  *launched = true;
  return 0;
}

// The voxel-update kernel launch_update picks for a stage and phase, for
// callers that launch it themselves (the multi-problem batch):
//   kind 0: (KP, cells, iter, chunk, seq, count_mag, extra), block `block`
//           threads, `lds` bytes of dynamic LDS, family TB_K_*;
//   kind 1: the one-wavefront LDS kernel (KP, cells, iter, chunk, seq,
//           count_mag);
//   kind 2: the one-voxel kernel (KP, cells, iter, chunk, seq, pix,
//           count_mag).
struct VoxPlan {
  const void *fn = nullptr;
  int kind = 0, block = 0, extra = 0, family = 0;
  size_t lds = 0;
  bool global = true;
};

static int launch_update_as(tb_stage *st, const int2 *cells, int ncells, int iter, int chunk,
                            int seq, int pix, int count_mag, VoxPlan *plan);

// With tb_options.autotune and kernel 0, the kernel and views per lane the
// autotuner measured fastest for the stage's shape (timed on first use)
static int launch_update(tb_stage *st, const int2 *cells, int ncells, int iter, int chunk,
                         int seq, int pix, int count_mag, VoxPlan *plan = nullptr) {
  if (ncells == 0) return 0;
  if (!g_opt.autotune || g_opt.kernel != 0 || seq == 2)
    return launch_update_as(st, cells, ncells, iter, chunk, seq, pix, count_mag, plan);
  int k = 0, v = 0;
  if (autotune_pick(st, iter, &k, &v)) return -1;
  const tb_options saved = g_opt;
  g_opt.kernel = k;
  g_opt.views_per_lane = v;
  const int rc = launch_update_as(st, cells, ncells, iter, chunk, seq, pix, count_mag, plan);
  g_opt = saved;
  return rc;
}

static int launch_update_as(tb_stage *st, const int2 *cells, int ncells, int iter, int chunk,
                            int seq, int pix, int count_mag, VoxPlan *plan) {
  if (ncells == 0) return 0;
  KP P = st->kp();
  hipEvent_t *evp = nullptr;
  if (!plan) {
    evp = ev_pair(st);
    if (!evp) {
      if (harvest_events(st)) return -1;
      evp = ev_pair(st);
    }
    HIPCHK(hipEventRecord(evp[0], st->stream));
  }
  // kernel 0 (auto): the multi-wavefront LDS kernel, else the
  // multi-wavefront kernel with e in global memory, else the one-wavefront
  // LDS kernel, else the one-voxel kernel; 1: one-voxel only; 2:
  // one-wavefront where it applies; 3: multi-wavefront LDS where it applies;
  // 4: multi-wavefront global where it applies.  The LDS kernels need the
  // cell's (e, w) block in LDS and a view's column within 3 bins x 1 row
  // (finest stage) or 5 bins x 2 rows.
  using mw_fn = void (*)(KP, const int2 *, int, int, int, int, int);
  int nb = 0;
  if (2 * st->mult_xy + 1 <= 3 && st->mult_z == 1) nb = 3;
  else if (2 * st->mult_xy + 1 <= 5 && st->mult_z == 2) nb = 5;
  if (st->N_r < nb) nb = 0;
  bool launched = false;
  bool deep = false;     // the launch below is k_icd_mwd
  bool batched = false;  // the launch below is k_icd_bat
  auto run_mw = [&](mw_fn fn, int colnb, int nth, int vpw, size_t lds, bool global) -> int {
    // the column table may be allocated now: take the parameters again
    const int ec = ensure_columns(st, colnb);
    if (ec != 0) return ec;
    P = st->kp();
    if (!P.colt || P.colt_rec != col_rec(colnb) || P.colt_nv != st->max_nv)
      return fail("launch_update: column table missing");
    if (seq == 0 && ensure_order(st, iter)) return -1;
    if (lds)
      HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int fam = global ? (batched ? TB_K_BAT : deep ? TB_K_MWD : TB_K_MWG) : TB_K_MW;
    launched = true;
    if (plan) {
      plan->fn = reinterpret_cast<const void *>(fn);
      plan->kind = 0;
      plan->block = 64 * (nth + 1);
      plan->extra = vpw;
      plan->family = fam;
      plan->lds = lds;
      plan->global = global;
      return 0;
    }
    hipLaunchKernelGGL(fn, dim3(ncells), dim3(64 * (nth + 1)), lds, st->stream, P, cells, iter,
                       chunk, seq, count_mag, vpw);
    if (global) st->global_launches++;
    else st->lds_launches++;
    st->fam_launches[fam]++;
    return 0;
  };
  // the same for the wide-column and pass kernels
  auto run_wide = [&](mw_fn fn, int nth, int extra, size_t lds, int fam) {
    launched = true;
    if (plan) {
      plan->fn = reinterpret_cast<const void *>(fn);
      plan->kind = 0;
      plan->block = 64 * (nth + 1);
      plan->extra = extra;
      plan->family = fam;
      plan->lds = lds;
      plan->global = true;
      return;
    }
    hipLaunchKernelGGL(fn, dim3(ncells), dim3(64 * (nth + 1)), lds, st->stream, P, cells, iter,
                       chunk, seq, count_mag, extra);
    st->global_launches++;
    st->fam_launches[fam]++;
  };
  // FP32 theta sums (theta_fp 1 or 2) exist in the global-e kernel only
  const int tf = g_opt.theta_fp;
  if (tf < 0 || tf > 2) return fail("theta_fp %d: 0, 1 or 2", tf);
  if (nb && tf == 0 && (g_opt.kernel == 0 || g_opt.kernel == 3)) {
    // one wavefront per SIMD where the views allow: at most 3 theta waves
    // (one view per lane, else two), else one view per lane and more waves
    const int vw = 64 / (nb * st->mult_z);
    const int nv = st->max_nv;
    int vpl = 1, nth = (nv + vw - 1) / vw;
    const bool two = g_opt.views_per_lane == 2 ||
                     (g_opt.views_per_lane == 0 && nth > 3 && (nv + 2 * vw - 1) / (2 * vw) <= 3);
    if (two) {
      vpl = 2;
      nth = (nv + 2 * vw - 1) / (2 * vw);
    }
    const int vpw = (nv + nth - 1) / nth;
    mw_fn fn = nullptr;
    if (nb == 3) {
      if (vpl == 2) fn = nth <= 2 ? k_icd_mw<3, 1, 2, 2> : nth <= 3 ? k_icd_mw<3, 1, 2, 3> : nullptr;
      else fn = nth <= 4 ? k_icd_mw<3, 1, 1, 4> : nth <= 8 ? k_icd_mw<3, 1, 1, 8>
              : nth <= 15 ? k_icd_mw<3, 1, 1, 15> : nullptr;
    } else {
      if (vpl == 2) fn = nth <= 3 ? k_icd_mw<5, 2, 2, 3> : nullptr;
      else fn = nth <= 8 ? k_icd_mw<5, 2, 1, 8> : nth <= 15 ? k_icd_mw<5, 2, 1, 15> : nullptr;
    }
    const size_t lds = icd_mw_lds_bytes(nv, st->N_r, st->mult_z, nth * vpl * 64);
    if (fn && lds + 1024 <= st->lds_max && run_mw(fn, nb, nth, vpw, lds, false) < 0) return -1;
  }
  if (!launched && tf == 0 && g_opt.kernel == 9 && 2 * st->mult_xy + 1 <= 3 && st->mult_z == 1 &&
      st->N_r >= 3) {
    // the batched finest-stage kernel: one view per lane, 3-bin columns
    const int cap = 64 / 3;
    const int nth = (st->max_nv + cap - 1) / cap;
    const int vpw = (st->max_nv + nth - 1) / nth;
    mw_fn fn = nth <= 8 ? k_icd_bat<8> : nth <= 15 ? k_icd_bat<15> : nullptr;
    batched = fn != nullptr;
    if (fn && run_mw(fn, 3, nth, vpw, 0, true) < 0) return -1;
  }
  if (!launched && tf == 0 && g_opt.kernel == 8) {
    // the deep-pipeline global-e kernel: the lane layout of k_icd_mwg with
    // `la` voxels' (e, w) rows in flight (tb_options.row_lookahead)
    const int w = 2 * st->mult_xy + 1;
    const int gnb = w <= 3 ? 3 : w <= 5 ? 5 : 0;
    const int mz = st->mult_z;
    const int la = g_opt.row_lookahead == 0 ? 3 : g_opt.row_lookahead;
    if (la != 1 && la != 3 && la != 7) return fail("row_lookahead %d: 0, 1, 3 or 7", la);
    auto pick = [&](int vpl, int nth) -> mw_fn {
      if (gnb == 3 && mz == 1) {
        if (la == 1) return vpl == 1 && nth <= 8 ? k_icd_mwd<3, 1, 1, 8, 1, 2> : nullptr;
        if (la == 7) return vpl == 1 && nth <= 8 ? k_icd_mwd<3, 1, 1, 8, 7, 1> : nullptr;
        if (vpl == 1) return nth <= 8 ? k_icd_mwd<3, 1, 1, 8, 3, 4> : k_icd_mwd<3, 1, 1, 15, 3, 4>;
      } else if (gnb == 5 && mz == 2 && la == 3) {
        if (vpl == 1) return nth <= 8 ? k_icd_mwd<5, 2, 1, 8, 3, 4> : k_icd_mwd<5, 2, 1, 15, 3, 4>;
      }
      return nullptr;
    };
    mw_fn fn = nullptr;
    int nth = 0, vpw = 0;
    if (gnb && st->N_r >= gnb) {
      const int cap = 64 / gnb;  // one view per lane
      nth = (st->max_nv + cap - 1) / cap;
      vpw = (st->max_nv + nth - 1) / nth;
      if (nth <= 15) fn = pick(1, nth);
    }
    deep = fn != nullptr;
    if (fn && run_mw(fn, gnb, nth, vpw, 0, true) < 0) return -1;
  }
  if (!launched && (g_opt.kernel == 0 || g_opt.kernel == 4)) {
    // lane = (view, bin) with the slice's mult_z rows: NB bins, VW = 64 / NB
    // views per lane layer, VPL layers.  The fewest views per lane that fit
    // in 15 theta waves: these waves are latency-bound, and more of them
    // sharing a SIMD hide each other's waits (measured on the C2 coarse
    // stages: two or four views per lane in fewer waves ran 10-40% slower)
    const int w = 2 * st->mult_xy + 1;
    const int gnb = w <= 3 ? 3 : w <= 5 ? 5 : w <= 9 ? 9 : 0;
    const int mz = st->mult_z;
    // instantiations: (NB, MZ, VPL) -> the widest MAXTH that keeps the
    // registers unspilled (15, or 11 with four views per lane)
    auto pick = [&](int vpl, int nth) -> mw_fn {
      if (tf != 0) {
        // the FP32 theta variants: 3- and 5-bin columns
        if (gnb == 3 && mz == 1) {
          if (vpl == 1) return nth <= 8 ? (tf == 1 ? k_icd_mwg<3, 1, 1, 8, 1> : k_icd_mwg<3, 1, 1, 8, 2>)
                                        : (tf == 1 ? k_icd_mwg<3, 1, 1, 15, 1> : k_icd_mwg<3, 1, 1, 15, 2>);
          if (vpl == 2) return tf == 1 ? k_icd_mwg<3, 1, 2, 15, 1> : k_icd_mwg<3, 1, 2, 15, 2>;
          if (vpl == 4 && nth <= 11) return tf == 1 ? k_icd_mwg<3, 1, 4, 11, 1> : k_icd_mwg<3, 1, 4, 11, 2>;
        } else if (gnb == 5 && mz == 2) {
          if (vpl == 1) return nth <= 8 ? (tf == 1 ? k_icd_mwg<5, 2, 1, 8, 1> : k_icd_mwg<5, 2, 1, 8, 2>)
                                        : (tf == 1 ? k_icd_mwg<5, 2, 1, 15, 1> : k_icd_mwg<5, 2, 1, 15, 2>);
          if (vpl == 2) return tf == 1 ? k_icd_mwg<5, 2, 2, 15, 1> : k_icd_mwg<5, 2, 2, 15, 2>;
          if (vpl == 4 && nth <= 11) return tf == 1 ? k_icd_mwg<5, 2, 4, 11, 1> : k_icd_mwg<5, 2, 4, 11, 2>;
        }
        return nullptr;
      }
      if (gnb == 3 && mz == 1) {
        if (vpl == 1) return nth <= 8 ? k_icd_mwg<3, 1, 1, 8, 0> : k_icd_mwg<3, 1, 1, 15, 0>;
        if (vpl == 2) return k_icd_mwg<3, 1, 2, 15, 0>;
        if (vpl == 4 && nth <= 11) return k_icd_mwg<3, 1, 4, 11, 0>;
      } else if (gnb == 5 && mz == 2) {
        if (vpl == 1) return nth <= 8 ? k_icd_mwg<5, 2, 1, 8, 0> : k_icd_mwg<5, 2, 1, 15, 0>;
        if (vpl == 2) return k_icd_mwg<5, 2, 2, 15, 0>;
        if (vpl == 4 && nth <= 11) return k_icd_mwg<5, 2, 4, 11, 0>;
      } else if (gnb == 9 && mz == 4) {
        if (vpl == 1) return k_icd_mwg<9, 4, 1, 15, 0>;
        if (vpl == 2) return k_icd_mwg<9, 4, 2, 15, 0>;
      } else if (gnb == 9 && mz == 2) {
        if (vpl == 1) return k_icd_mwg<9, 2, 1, 15, 0>;
        if (vpl == 2) return k_icd_mwg<9, 2, 2, 15, 0>;
        if (vpl == 4 && nth <= 11) return k_icd_mwg<9, 2, 4, 11, 0>;
      }
      return nullptr;
    };
    mw_fn fn = nullptr;
    int nth = 0, vpw = 0;
    for (int vpl = 1; vpl <= 4 && !fn && gnb && st->N_r >= gnb; vpl *= 2) {
      const int cap = (64 / gnb) * vpl;
      nth = (st->max_nv + cap - 1) / cap;
      vpw = (st->max_nv + nth - 1) / nth;
      if (nth > 15) continue;
      fn = pick(vpl, nth);
    }
    if (fn && run_mw(fn, gnb, nth, vpw, 0, true) < 0) return -1;
  }
  if (!launched && tf != 0)
    return fail("theta_fp %d: FP32 theta sums exist in the multi-wavefront global-e kernel "
                "with 3- or 5-bin columns only, which does not serve this stage", tf);
  // few cells per colour (single-frame scans' coarse stages): the
  // wide-column update with each cell's views split over S workgroups of one
  // cooperative launch (kernel 10 asks for it alone; never in plan mode,
  // whose callers may capture the launch in a graph)
  if (!launched && !plan && tf == 0 &&
      (g_opt.kernel == 10 || (g_opt.kernel == 0 && ncells <= kSplitMaxCells))) {
    const int rc = launch_split(st, cells, ncells, iter, chunk, seq, count_mag, &launched);
    if (rc < 0) return -1;
  }
  // kernel 6 asks for the pass form alone
  const bool wpass_first = g_opt.kernel == 6;
  if (!launched && (g_opt.kernel == 0 || g_opt.kernel == 5) && 2 * st->mult_xy + 1 <= 17 &&
      st->N_r >= 17 && !wpass_first) {
    // columns of up to 17 bins (the coarsest stages): a lane per (view, R
    // detector rows), the fewest rows per lane that fit in 8 theta waves
    const int mz = st->mult_z;
    for (int R = 1; R <= 2 && !launched; R++) {
      if (mz % R != 0 || st->max_nv * (mz / R) > 8 * 64) continue;
      const int nth = (st->max_nv * (mz / R) + 63) / 64;
      const size_t lds = (size_t)st->max_nv * kHrStride * sizeof(float);
      if (lds + 2048 > st->lds_max) break;
      mw_fn fn = R == 1 ? k_icd_wide<17, 1, 8> : k_icd_wide<17, 2, 8>;
      // A values from the stage's 17-bin column table when it fits in
      // memory (the kernel evaluates them otherwise)
      if (ensure_columns(st, 17) < 0) return -1;
      P = st->kp();
      if (seq == 0 && ensure_order(st, iter)) return -1;
      HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      run_wide(fn, nth, mz / R, lds, TB_K_WIDE);
    }
  }
  if (!launched && (g_opt.kernel == 0 || g_opt.kernel == 5 || g_opt.kernel == 6) &&
      (st->mult_z % 2 == 0 || st->mult_z == 1)) {
    // 11 theta waves own blocks of views and walk their (view, bin, R rows)
    // entries in passes (R = 2, or 1 for single-row slices); the views'
    // detector responses go to LDS when they fit
    const int R = st->mult_z % 2 == 0 ? 2 : 1;
    const int w = 2 * st->mult_xy + 1;
    const int nbx = R == 1 ? (w <= 3 ? 3 : 0) : w <= 5 ? 5 : w <= 9 ? 9 : w <= 17 ? 17 : 0;
    constexpr int nth = 11;
    const int vpw = (st->max_nv + nth - 1) / nth;
    const int passes = (vpw * nbx * (st->mult_z / R) + 63) / 64;
    // per-view coefficients, per-wave A values, then the detector responses
    // when they fit
    const size_t lds_c = (size_t)st->max_nv * 32 + (size_t)nth * vpw * (nbx + 1) * sizeof(float);
    const size_t lds_h = lds_c + (size_t)st->max_nv * kHrStride * sizeof(float);
    const bool hl = lds_h + 4096 <= st->lds_max;
    const size_t lds = hl ? lds_h : lds_c;
    mw_fn fn = nullptr;
    if (nbx == 17)
      fn = passes <= 16 ? (hl ? k_icd_wpass<17, 2, 16, 11, true> : k_icd_wpass<17, 2, 16, 11, false>)
         : passes <= 24 ? (hl ? k_icd_wpass<17, 2, 24, 11, true> : k_icd_wpass<17, 2, 24, 11, false>)
                        : nullptr;
    else if (nbx == 9 && passes <= 16)
      fn = hl ? k_icd_wpass<9, 2, 16, 11, true> : k_icd_wpass<9, 2, 16, 11, false>;
    else if (nbx == 5 && passes <= 16)
      fn = hl ? k_icd_wpass<5, 2, 16, 11, true> : k_icd_wpass<5, 2, 16, 11, false>;
    else if (nbx == 3 && passes <= 16)
      fn = hl ? k_icd_wpass<3, 1, 16, 11, true> : k_icd_wpass<3, 1, 16, 11, false>;
    if (fn && st->N_r >= nbx && lds + 4096 <= st->lds_max) {
      // A values from the stage's nbx-bin column table when it fits in
      // memory (the kernel evaluates them otherwise)
      if (ensure_columns(st, nbx) < 0) return -1;
      P = st->kp();
      if (seq == 0 && ensure_order(st, iter)) return -1;
      HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      run_wide(fn, nth, vpw, lds, TB_K_WPASS);
    }
  }
  if (!launched && (g_opt.kernel == 0 || g_opt.kernel == 5 || g_opt.kernel == 7) &&
      (st->mult_z % 2 == 0 || st->mult_z == 1)) {
    // more passes than k_icd_wpass holds (the coarse stages of scans with
    // many views per frame): 15 theta waves walk their entries in groups
    // of 8 passes and read e again for the write
    const int R = st->mult_z % 2 == 0 ? 2 : 1;
    const int w = 2 * st->mult_xy + 1;
    const int nbx = R == 1 ? (w <= 3 ? 3 : 0) : w <= 5 ? 5 : w <= 9 ? 9 : w <= 17 ? 17 : 0;
    const int nth = std::min(15, std::max(1, st->max_nv));
    const int vpw = (st->max_nv + nth - 1) / nth;
    const size_t lds = (size_t)st->max_nv * 32 + (size_t)nth * vpw * (nbx + 1) * sizeof(float);
    mw_fn fn = nbx == 17 ? k_icd_wloop<17, 2, 8, 15, false>
             : nbx == 9  ? k_icd_wloop<9, 2, 8, 15, false>
             : nbx == 5  ? k_icd_wloop<5, 2, 8, 15, false>
             : nbx == 3  ? k_icd_wloop<3, 1, 8, 15, false> : nullptr;
    if (fn && st->N_r >= nbx && lds + 4096 <= st->lds_max) {
      // A values from the stage's nbx-bin column table when it fits in
      // memory (the kernel evaluates them otherwise)
      if (ensure_columns(st, nbx) < 0) return -1;
      P = st->kp();
      if (seq == 0 && ensure_order(st, iter)) return -1;
      HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(fn),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      run_wide(fn, nth, vpw, lds, TB_K_WLOOP);
    }
  }

  if (!launched && nb && g_opt.kernel != 1) {
    using lds_fn = void (*)(KP, const int2 *, int, int, int, int);
    lds_fn fast = nullptr;
    const int vpl = (st->max_nv + 63) / 64;
    if (nb == 3) fast = vpl == 1 ? k_icd_lds<1, 3, 1, 4> : vpl == 2 ? k_icd_lds<2, 3, 1, 4>
                      : vpl <= 4 ? k_icd_lds<4, 3, 1, 2> : nullptr;
    else fast = vpl == 1 ? k_icd_lds<1, 5, 2, 4> : vpl == 2 ? k_icd_lds<2, 5, 2, 2> : nullptr;
    const size_t lds = icd_lds_bytes(st->max_nv, st->N_r, st->mult_z, nb);
    if (fast && lds <= st->lds_max) {
      if (seq == 0 && ensure_order(st, iter)) return -1;
      HIPCHK(hipFuncSetAttribute(reinterpret_cast<const void *>(fast),
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      launched = true;
      if (plan) {
        plan->fn = reinterpret_cast<const void *>(fast);
        plan->kind = 1;
        plan->block = 64;
        plan->family = TB_K_LDS;
        plan->lds = lds;
        plan->global = false;
        return 0;
      }
      hipLaunchKernelGGL(fast, dim3(ncells), dim3(64), lds, st->stream, P, cells, iter, chunk,
                         seq, count_mag);
      st->lds_launches++;
      st->fam_launches[TB_K_LDS]++;
    }
  }
  if (plan) {
    if (!launched) {
      plan->kind = 2;
      plan->block = st->max_nv <= 64 ? 64 : st->max_nv <= 128 ? 128 : 256;
      plan->fn = st->max_nv <= 64    ? reinterpret_cast<const void *>(k_update_voxels<64>)
                 : st->max_nv <= 128 ? reinterpret_cast<const void *>(k_update_voxels<128>)
                                     : reinterpret_cast<const void *>(k_update_voxels<256>);
      plan->family = TB_K_ONE_VOXEL;
      plan->lds = 0;
      plan->global = false;
    }
    return 0;
  }
  if (!launched) st->fam_launches[TB_K_ONE_VOXEL]++;
  if (launched) {
    // counted above
  } else if (st->max_nv <= 64)
    hipLaunchKernelGGL(k_update_voxels<64>, dim3(ncells), dim3(64), 0, st->stream, P, cells,
                       iter, chunk, seq, pix, count_mag);
  else if (st->max_nv <= 128)
    hipLaunchKernelGGL(k_update_voxels<128>, dim3(ncells), dim3(128), 0, st->stream, P, cells,
                       iter, chunk, seq, pix, count_mag);
  else
    hipLaunchKernelGGL(k_update_voxels<256>, dim3(ncells), dim3(256), 0, st->stream, P, cells,
                       iter, chunk, seq, pix, count_mag);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(evp[1], st->stream));
  st->voxel_launches++;
  return 0;
}

extern "C" int tb_voxel_update(tb_stage *st, int32_t f, int32_t z, int32_t y, int32_t x,
                               float *value) {
  if (f < 0 || f >= st->N_time || z < 0 || z >= st->N_z || y < 0 || y >= st->N_y || x < 0 ||
      x >= st->N_x)
    return fail("tb_voxel_update: voxel out of range");
  int2 c = make_int2(f, z);
  int2 *dc = nullptr;
  HIPCHK(dalloc(&dc, 1));
  HIPCHK(hipMemcpyAsync(dc, &c, sizeof(c), hipMemcpyHostToDevice, st->stream));
  const int pix = y * st->N_x + x;
  const int packed = (y << 16) | x;
  HIPCHK(hipMemcpyAsync(st->single, &packed, sizeof(int), hipMemcpyHostToDevice, st->stream));
  if (launch_update(st, dc, 1, 1, 0, 2, pix, 0)) return -1;
  HIPCHK(hipMemcpyAsync(value, st->obj + ((size_t)f * st->N_z + z) * st->nxy() + pix,
                        sizeof(float), hipMemcpyDeviceToHost, st->stream));
  HIPCHK(hipStreamSynchronize(st->stream));
  (void)hipFree(dc);
  return harvest_events(st);
}

extern "C" int tb_update_colour(tb_stage *st, int32_t colour, int32_t iter, int32_t chunk,
                                int32_t nhicd) {
  if (colour < 0 || colour > 1 || chunk < 0 || chunk >= kNhicdK)
    return fail("tb_update_colour: bad colour/chunk");
  if (launch_update(st, st->cells[colour], st->ncells[colour], iter, chunk, nhicd ? 1 : 0, 0,
                    nhicd ? 0 : 1))
    return -1;
  HIPCHK(hipStreamSynchronize(st->stream));
  return harvest_events(st);
}

static int vsc_async(tb_stage *st) {
  KP P = st->kp();
  hipLaunchKernelGGL(k_vsc_filter, dim3((st->N_x + 15) / 16, (st->N_y + 15) / 16, st->N_time * st->N_z),
                     dim3(256), 0, st->stream, P, st->filt);
  // NHICDSelectNum = floor(UpdateSelectNum / NHICD_Iterations), XT_NHICD.c:182
  hipLaunchKernelGGL(k_vsc_select, dim3(st->N_time * st->N_z), dim3(1024), 0, st->stream, P,
                     st->filt, st->U / kNhicdK);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int tb_vsc_select(tb_stage *st) {
  if (vsc_async(st)) return -1;
  HIPCHK(hipStreamSynchronize(st->stream));
  return 0;
}

extern "C" int tb_nhicd_list(tb_stage *st, int32_t f, int32_t z, int32_t *xs, int32_t *ys) {
  const int cell = f * st->N_z + z;
  int n = 0;
  if (d2h(st, &n, st->nh_count + cell, sizeof(int))) return -1;
  std::vector<int> l(std::max(n, 1));
  if (n && d2h(st, l.data(), st->nh_list + (size_t)cell * st->nh_cap, n * sizeof(int))) return -1;
  for (int i = 0; i < n; i++) {
    if (xs) xs[i] = l[i] % st->N_x;
    if (ys) ys[i] = l[i] / st->N_x;
  }
  return n;
}

static int maybe_offset(tb_stage *st) {
  if (st->updateProjOffset > 1) return offset_update_async(st, st->remove_rings);
  return 0;
}

// convergence rule (XT_ICD_update.c:1066-1104): the masked update
// magnitudes of the iteration against |x| of the last chunk's voxels
static int convergence_check(tb_stage *st, int32_t *converged, double *ratio) {
  const int ncell = st->N_time * st->N_z;
  KP P = st->kp();
  const int nb = std::min(st->red_n, grid_for(st->nobj()));
  hipLaunchKernelGGL(k_mag_sum, dim3(nb), dim3(256), 0, st->stream, P, st->red);
  HIPCHK(hipGetLastError());
  double avg;
  if (reduce_partials(st, nb, &avg)) return -1;
  std::vector<double> vm(ncell);
  if (d2h(st, vm.data(), st->cell_voxmag, ncell * sizeof(double))) return -1;
  double vox_mag = 0;
  for (double v : vm) vox_mag += v;
  if (harvest_events(st)) return -1;
  const double pct = 100 * avg / vox_mag;
  if (ratio) *ratio = pct;
  if (converged) *converged = pct < st->convg_thresh ? 1 : 0;
  return 0;
}

extern "C" int tb_icd_iteration(tb_stage *st, int32_t iter, int32_t *converged, double *ratio) {
  const int ncell = st->N_time * st->N_z;
  for (int j = 0; j < kNhicdK; j++) {
    // total_vox_mag restarts every chunk (XT_ICD_update.c:976)
    hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(ncell)), dim3(256), 0, st->stream,
                       st->cell_voxmag, (size_t)ncell, 0.0);
    if (launch_update(st, st->cells[0], st->ncells[0], iter, j, 0, 0, 1)) return -1;
    if (maybe_offset(st)) return -1;
    if (launch_update(st, st->cells[1], st->ncells[1], iter, j, 0, 0, 1)) return -1;
    if (maybe_offset(st)) return -1;
    if (vsc_async(st)) return -1;
    if (iter > 1 && g_opt.no_nhicd == 0) {
      if (launch_update(st, st->cells[0], st->ncells[0], iter, j, 1, 0, 0)) return -1;
      if (maybe_offset(st)) return -1;
      if (launch_update(st, st->cells[1], st->ncells[1], iter, j, 1, 0, 0)) return -1;
      if (maybe_offset(st)) return -1;
    }
  }
  if (st->updateVar == 1 && tb_variance_update(st)) return -1;
  return convergence_check(st, converged, ratio);
}

extern "C" int tb_stage_run(tb_stage *st, int32_t max_iter, double *costs, int32_t *iters) {
  if (max_iter <= 0) max_iter = g_opt.max_iterations > 0 ? g_opt.max_iterations : kMaxIters;
  if (tb_init_error_sino(st)) return -1;
  double last = 0;
  const bool want_cost = g_opt.compute_cost || costs;
  if (want_cost) {
    if (tb_cost(st, &last, nullptr, nullptr)) return -1;
    if (costs) costs[0] = last;
  }
  int it;
  for (it = 1; it <= max_iter; it++) {
    int conv = 0;
    double ratio = 0;
    if (tb_icd_iteration(st, it, &conv, &ratio)) return -1;
    if (want_cost) {
      double c;
      if (tb_cost(st, &c, nullptr, nullptr)) return -1;
      if (costs) costs[it] = c;
      if (g_opt.compute_cost && g_opt.abort_on_cost_increase && c > last)
        return fail("Cost value increased.");
      last = c;
    }
    if (conv && it > 1) break;
  }
  if (iters) *iters = it > max_iter ? max_iter : it;
  return 0;
}

// ---------------------------------------------------------------------------
// Batches of independent problems advanced in lockstep.  Each member is a
// stage of its own scan; tb_batch_iteration runs one ICD iteration
// (tb_icd_iteration's sequence) of every member at once on the batch's
// stream: each voxel-update phase is one launch per kernel configuration
// over the cells of all members that use it (the kernels' multi-problem
// entry points read the member's parameters from a device table), with the
// members' offset, selection and fill kernels between them.  A member's
// kernels, arguments and order are those of its own tb_icd_iteration, so
// its result is bit for bit the same.  From the second iteration on the
// launches of an iteration are captured once in a HIP graph and replayed
// (the kernels depend on the iteration only through iter > 1; the voxel
// order of each iteration is built before the replay).
// ---------------------------------------------------------------------------
namespace {
const void *multi_of(const void *fn) {
#define TB_FN(...) reinterpret_cast<const void *>(__VA_ARGS__)
  static const std::pair<const void *, const void *> tab[] = {
      {TB_FN(k_icd_mw<3, 1, 1, 4>), TB_FN(k_icd_mw_m<3, 1, 1, 4>)},
      {TB_FN(k_icd_mw<3, 1, 2, 2>), TB_FN(k_icd_mw_m<3, 1, 2, 2>)},
      {TB_FN(k_icd_mw<3, 1, 2, 3>), TB_FN(k_icd_mw_m<3, 1, 2, 3>)},
      {TB_FN(k_icd_mw<3, 1, 1, 8>), TB_FN(k_icd_mw_m<3, 1, 1, 8>)},
      {TB_FN(k_icd_mw<3, 1, 1, 15>), TB_FN(k_icd_mw_m<3, 1, 1, 15>)},
      {TB_FN(k_icd_mw<5, 2, 1, 8>), TB_FN(k_icd_mw_m<5, 2, 1, 8>)},
      {TB_FN(k_icd_mw<5, 2, 1, 15>), TB_FN(k_icd_mw_m<5, 2, 1, 15>)},
      {TB_FN(k_icd_mw<5, 2, 2, 3>), TB_FN(k_icd_mw_m<5, 2, 2, 3>)},
      {TB_FN(k_icd_wide<17, 1, 8>), TB_FN(k_icd_wide_m<17, 1, 8>)},
      {TB_FN(k_icd_wide<17, 2, 8>), TB_FN(k_icd_wide_m<17, 2, 8>)},
      {TB_FN(k_icd_wloop<17, 2, 8, 15, false>), TB_FN(k_icd_wloop_m<17, 2, 8, 15, false>)},
      {TB_FN(k_icd_wloop<9, 2, 8, 15, false>), TB_FN(k_icd_wloop_m<9, 2, 8, 15, false>)},
      {TB_FN(k_icd_wloop<5, 2, 8, 15, false>), TB_FN(k_icd_wloop_m<5, 2, 8, 15, false>)},
      {TB_FN(k_icd_wloop<3, 1, 8, 15, false>), TB_FN(k_icd_wloop_m<3, 1, 8, 15, false>)},
      {TB_FN(k_icd_wpass<17, 2, 16, 11, false>), TB_FN(k_icd_wpass_m<17, 2, 16, 11, false>)},
      {TB_FN(k_icd_wpass<17, 2, 24, 11, false>), TB_FN(k_icd_wpass_m<17, 2, 24, 11, false>)},
      {TB_FN(k_icd_wpass<9, 2, 16, 11, false>), TB_FN(k_icd_wpass_m<9, 2, 16, 11, false>)},
      {TB_FN(k_icd_wpass<5, 2, 16, 11, false>), TB_FN(k_icd_wpass_m<5, 2, 16, 11, false>)},
      {TB_FN(k_icd_wpass<17, 2, 16, 11, true>), TB_FN(k_icd_wpass_m<17, 2, 16, 11, true>)},
      {TB_FN(k_icd_wpass<17, 2, 24, 11, true>), TB_FN(k_icd_wpass_m<17, 2, 24, 11, true>)},
      {TB_FN(k_icd_wpass<9, 2, 16, 11, true>), TB_FN(k_icd_wpass_m<9, 2, 16, 11, true>)},
      {TB_FN(k_icd_wpass<5, 2, 16, 11, true>), TB_FN(k_icd_wpass_m<5, 2, 16, 11, true>)},
      {TB_FN(k_icd_wpass<3, 1, 16, 11, false>), TB_FN(k_icd_wpass_m<3, 1, 16, 11, false>)},
      {TB_FN(k_icd_wpass<3, 1, 16, 11, true>), TB_FN(k_icd_wpass_m<3, 1, 16, 11, true>)},
      {TB_FN(k_icd_mwg<3, 1, 1, 8, 0>), TB_FN(k_icd_mwg_m<3, 1, 1, 8, 0>)},
      {TB_FN(k_icd_mwg<3, 1, 1, 15, 0>), TB_FN(k_icd_mwg_m<3, 1, 1, 15, 0>)},
      {TB_FN(k_icd_mwg<5, 2, 1, 8, 0>), TB_FN(k_icd_mwg_m<5, 2, 1, 8, 0>)},
      {TB_FN(k_icd_mwg<5, 2, 1, 15, 0>), TB_FN(k_icd_mwg_m<5, 2, 1, 15, 0>)},
      {TB_FN(k_icd_mwg<9, 4, 1, 15, 0>), TB_FN(k_icd_mwg_m<9, 4, 1, 15, 0>)},
      {TB_FN(k_icd_mwg<9, 4, 2, 15, 0>), TB_FN(k_icd_mwg_m<9, 4, 2, 15, 0>)},
      {TB_FN(k_icd_mwg<3, 1, 2, 15, 0>), TB_FN(k_icd_mwg_m<3, 1, 2, 15, 0>)},
      {TB_FN(k_icd_mwg<3, 1, 4, 11, 0>), TB_FN(k_icd_mwg_m<3, 1, 4, 11, 0>)},
      {TB_FN(k_icd_mwg<5, 2, 2, 15, 0>), TB_FN(k_icd_mwg_m<5, 2, 2, 15, 0>)},
      {TB_FN(k_icd_mwg<5, 2, 4, 11, 0>), TB_FN(k_icd_mwg_m<5, 2, 4, 11, 0>)},
      {TB_FN(k_icd_mwg<9, 2, 1, 15, 0>), TB_FN(k_icd_mwg_m<9, 2, 1, 15, 0>)},
      {TB_FN(k_icd_mwg<9, 2, 2, 15, 0>), TB_FN(k_icd_mwg_m<9, 2, 2, 15, 0>)},
      {TB_FN(k_icd_mwg<9, 2, 4, 11, 0>), TB_FN(k_icd_mwg_m<9, 2, 4, 11, 0>)},
  };
#undef TB_FN
  for (const auto &e : tab)
    if (e.first == fn) return e.second;
  return nullptr;
}
}  // namespace

struct tb_batch {
  hipStream_t stream = nullptr;
  int device = 0;
  int flags = 0;
  KP *tab = nullptr;  // device: one parameter block per member
  size_t tab_cap = 0;
  int4 *cells = nullptr;  // device: the multi launches' cell lists
  size_t cells_cap = 0;
  std::vector<KP> tab_h;
  std::vector<int4> cells_h;
  // launch groups: a multi-problem kernel, its block and LDS, and the
  // range of its cells per colour; solo members launch their own kernel
  struct Group {
    const void *multi;
    int block, family;
    size_t lds;
    bool global;
    std::vector<int> members;
    size_t off[2], n[2];
  };
  std::vector<Group> groups;
  std::vector<int> solo;
  std::vector<VoxPlan> plans;
  std::vector<tb_stage *> members;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool graph_ok = false;
  tb_batch_stats stats{};
  hipEvent_t ev[2] = {nullptr, nullptr};
};

static void batch_drop_graph(tb_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_ok = false;
}

extern "C" int tb_batch_create(tb_batch **out, int32_t flags) {
  *out = nullptr;
  auto *b = new tb_batch();
  b->flags = flags;
  if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess ||
      hipEventCreate(&b->ev[0]) != hipSuccess || hipEventCreate(&b->ev[1]) != hipSuccess) {
    tb_batch_destroy(b);
    return fail("tb_batch_create: HIP stream or events");
  }
  *out = b;
  return 0;
}

extern "C" void tb_batch_destroy(tb_batch *b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  if (b->tab) (void)hipFree(b->tab);
  if (b->cells) (void)hipFree(b->cells);
  for (hipEvent_t e : b->ev)
    if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

extern "C" int tb_batch_get_stats(const tb_batch *b, tb_batch_stats *out) {
  *out = b->stats;
  return 0;
}

// the voxel-update phase (colour c) of every member
static int batch_phase(tb_batch *b, int c, int iter, int chunk, int seq, int count_mag) {
  using mm_fn = void (*)(const KP *, const int4 *, int, int, int, int);
  using mw_fn = void (*)(KP, const int2 *, int, int, int, int, int);
  using lds_fn = void (*)(KP, const int2 *, int, int, int, int);
  using one_fn = void (*)(KP, const int2 *, int, int, int, int, int);
  for (const auto &g : b->groups) {
    if (!g.n[c]) continue;
    hipLaunchKernelGGL(reinterpret_cast<mm_fn>(const_cast<void *>(g.multi)), dim3((unsigned)g.n[c]),
                       dim3(g.block), g.lds, b->stream, b->tab, b->cells + g.off[c], iter, chunk,
                       seq, count_mag);
  }
  for (int i : b->solo) {
    tb_stage *st = b->members[i];
    const VoxPlan &pl = b->plans[i];
    if (!st->ncells[c] || !pl.fn) continue;
    const KP P = st->kp();
    if (pl.kind == 0)
      hipLaunchKernelGGL(reinterpret_cast<mw_fn>(const_cast<void *>(pl.fn)), dim3(st->ncells[c]),
                         dim3(pl.block), pl.lds, b->stream, P, st->cells[c], iter, chunk, seq,
                         count_mag, pl.extra);
    else if (pl.kind == 1)
      hipLaunchKernelGGL(reinterpret_cast<lds_fn>(const_cast<void *>(pl.fn)), dim3(st->ncells[c]),
                         dim3(pl.block), pl.lds, b->stream, P, st->cells[c], iter, chunk, seq,
                         count_mag);
    else
      hipLaunchKernelGGL(reinterpret_cast<one_fn>(const_cast<void *>(pl.fn)), dim3(st->ncells[c]),
                         dim3(pl.block), 0, b->stream, P, st->cells[c], iter, chunk, seq, 0,
                         count_mag);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// tb_icd_iteration's launches for every member, up to the variance update
static int batch_body(tb_batch *b, int iter) {
  const int n = (int)b->members.size();
  for (int j = 0; j < kNhicdK; j++) {
    for (tb_stage *st : b->members) {
      const int ncell = st->N_time * st->N_z;
      hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(ncell)), dim3(256), 0, b->stream,
                         st->cell_voxmag, (size_t)ncell, 0.0);
    }
    if (batch_phase(b, 0, iter, j, 0, 1)) return -1;
    for (tb_stage *st : b->members)
      if (maybe_offset(st)) return -1;
    if (batch_phase(b, 1, iter, j, 0, 1)) return -1;
    for (tb_stage *st : b->members)
      if (maybe_offset(st)) return -1;
    for (tb_stage *st : b->members)
      if (vsc_async(st)) return -1;
    if (iter > 1 && g_opt.no_nhicd == 0) {
      if (batch_phase(b, 0, iter, j, 1, 0)) return -1;
      for (tb_stage *st : b->members)
        if (maybe_offset(st)) return -1;
      if (batch_phase(b, 1, iter, j, 1, 0)) return -1;
      for (tb_stage *st : b->members)
        if (maybe_offset(st)) return -1;
    }
  }
  (void)n;
  return 0;
}

// plans, launch groups, parameter table and cell lists of the members;
// returns 1 when they differ from those the graph was captured with
static int batch_prepare(tb_batch *b, tb_stage *const *sts, int n, int iter, bool *changed) {
  std::vector<VoxPlan> plans(n);
  for (int i = 0; i < n; i++) {
    tb_stage *st = sts[i];
    const int c = st->ncells[0] ? 0 : 1;
    if (!st->ncells[c]) continue;
    if (launch_update(st, st->cells[c], st->ncells[c], iter, 0, 0, 0, 1, &plans[i])) return -1;
  }
  std::vector<tb_batch::Group> groups;
  std::vector<int> solo;
  for (int i = 0; i < n; i++) {
    const VoxPlan &pl = plans[i];
    if (!pl.fn) continue;
    const void *m = pl.kind == 0 ? multi_of(pl.fn) : nullptr;
    if (!m) {
      solo.push_back(i);
      continue;
    }
    tb_batch::Group *g = nullptr;
    for (auto &h : groups)
      if (h.multi == m && h.block == pl.block) g = &h;
    if (!g) {
      groups.push_back({m, pl.block, pl.family, 0, pl.global, {}, {0, 0}, {0, 0}});
      g = &groups.back();
    }
    g->lds = std::max(g->lds, pl.lds);
    g->members.push_back(i);
  }
  std::vector<int4> cells;
  for (auto &g : groups)
    for (int c = 0; c < 2; c++) {
      g.off[c] = cells.size();
      for (int i : g.members) {
        tb_stage *st = sts[i];
        for (const int2 &q : st->cells_h[c]) cells.push_back(make_int4(q.x, q.y, i, plans[i].extra));
      }
      g.n[c] = cells.size() - g.off[c];
    }
  std::vector<KP> tab(n);
  for (int i = 0; i < n; i++) tab[i] = sts[i]->kp();
  *changed = b->members.size() != (size_t)n ||
             !std::equal(b->members.begin(), b->members.end(), sts) ||
             b->cells_h.size() != cells.size() ||
             memcmp(b->cells_h.data(), cells.data(), cells.size() * sizeof(int4)) != 0 ||
             b->tab_h.size() != tab.size() ||
             memcmp(b->tab_h.data(), tab.data(), tab.size() * sizeof(KP)) != 0 ||
             b->groups.size() != groups.size();
  for (size_t k = 0; !*changed && k < groups.size(); k++)
    *changed = groups[k].multi != b->groups[k].multi || groups[k].block != b->groups[k].block ||
               groups[k].lds != b->groups[k].lds;
  for (int i = 0; !*changed && i < n; i++)
    *changed = plans[i].fn != b->plans[i].fn || plans[i].extra != b->plans[i].extra;
  if (*changed) {
    if (tab.size() > b->tab_cap) {
      if (b->tab) (void)hipFree(b->tab);
      b->tab = nullptr;
      b->tab_cap = 0;
      HIPCHK(dalloc(&b->tab, tab.size()));
      b->tab_cap = tab.size();
    }
    if (cells.size() > b->cells_cap) {
      if (b->cells) (void)hipFree(b->cells);
      b->cells = nullptr;
      b->cells_cap = 0;
      HIPCHK(dalloc(&b->cells, std::max<size_t>(cells.size(), 1)));
      b->cells_cap = std::max<size_t>(cells.size(), 1);
    }
    HIPCHK(hipMemcpyAsync(b->tab, tab.data(), tab.size() * sizeof(KP), hipMemcpyHostToDevice,
                          b->stream));
    if (!cells.empty())
      HIPCHK(hipMemcpyAsync(b->cells, cells.data(), cells.size() * sizeof(int4),
                            hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));  // the host copies stay valid until here
    b->tab_h = std::move(tab);
    b->cells_h = std::move(cells);
    b->groups = std::move(groups);
    b->solo = std::move(solo);
    b->plans = std::move(plans);
    b->members.assign(sts, sts + n);
    for (const auto &g : b->groups)
      if (g.lds)
        HIPCHK(hipFuncSetAttribute(g.multi, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)g.lds));
  }
  return 0;
}

// the members' voxel-launch counters for one iteration
static void batch_count(tb_batch *b, int iter) {
  const int per = (iter > 1 && g_opt.no_nhicd == 0) ? 2 : 1;
  for (size_t i = 0; i < b->members.size(); i++) {
    tb_stage *st = b->members[i];
    const VoxPlan &pl = b->plans[i];
    if (!pl.fn) continue;
    const int k = per * kNhicdK * ((st->ncells[0] ? 1 : 0) + (st->ncells[1] ? 1 : 0));
    st->voxel_launches += k;
    st->fam_launches[pl.family] += k;
    if (pl.global) st->global_launches += k;
    else st->lds_launches += k;
  }
  for (const auto &g : b->groups) b->stats.launches += per * kNhicdK * ((g.n[0] ? 1 : 0) + (g.n[1] ? 1 : 0));
  for (int i : b->solo) {
    tb_stage *st = b->members[i];
    b->stats.launches += per * kNhicdK * ((st->ncells[0] ? 1 : 0) + (st->ncells[1] ? 1 : 0));
  }
}

static int batch_iteration(tb_batch *b, tb_stage *const *sts, int n, int iter,
                           int32_t *converged, double *ratio) {
  bool changed = false;
  if (batch_prepare(b, sts, n, iter, &changed)) return -1;
  if (changed) batch_drop_graph(b);
  HIPCHK(hipEventRecord(b->ev[0], b->stream));
  const bool use_graph = (b->flags & TB_BATCH_GRAPH) && iter > 1;
  if (use_graph && !b->graph_ok) {
    batch_drop_graph(b);
    HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    const int rc = batch_body(b, iter);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(b->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return -1;
    }
    HIPCHK(ce);
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_ok = true;
    b->stats.captures++;
  }
  if (use_graph) {
    HIPCHK(hipGraphLaunch(b->exec, b->stream));
    b->stats.graph_replays++;
  } else if (batch_body(b, iter)) {
    return -1;
  }
  HIPCHK(hipEventRecord(b->ev[1], b->stream));
  batch_count(b, iter);
  for (int i = 0; i < n; i++) {
    tb_stage *st = sts[i];
    if (st->updateVar == 1 && tb_variance_update(st)) return -1;
    if (convergence_check(st, converged ? converged + i : nullptr, ratio ? ratio + i : nullptr))
      return -1;
  }
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
  b->stats.ms += ms;
  b->stats.iterations++;
  b->stats.groups = (int64_t)b->groups.size();
  b->stats.solo = (int64_t)b->solo.size();
  return 0;
}

extern "C" int tb_batch_iteration(tb_batch *b, tb_stage *const *stages, int32_t n, int32_t iter,
                                  int32_t *converged, double *ratio) {
  if (!b || n < 0 || (n && !stages)) return fail("tb_batch_iteration: bad arguments");
  if (n == 0) return 0;
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev != b->device) return fail("tb_batch_iteration: the batch belongs to device %d", b->device);
  // every member's work goes to the batch's stream for the call
  std::vector<hipStream_t> own(n);
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < i; k++)
      if (stages[k] == stages[i]) return fail("tb_batch_iteration: stage %d given twice", i);
    HIPCHK(hipStreamSynchronize(stages[i]->stream));
    own[i] = stages[i]->stream;
    stages[i]->stream = b->stream;
  }
  const int rc = batch_iteration(b, stages, n, iter, converged, ratio);
  if (rc) (void)hipStreamSynchronize(b->stream);
  for (int i = 0; i < n; i++) stages[i]->stream = own[i];
  return rc;
}

// ---------------------------------------------------------------------------
// Autotuner (no counterpart in the reference): for a stage shape (mult_xy,
// mult_z, views per frame, bins, image, slices, frames) it times one HICD
// colour phase with every kernel configuration that applies (kernel options
// 2-9, one or two views per lane for the multi-wavefront LDS kernel) and
// keeps the fastest.  Each trial runs on the stage's own buffers and is
// undone: the error sinogram, object, update map and per-cell counters are
// copied aside before and back after it, and the stage's launch counters
// and kernel time are restored, so tuning leaves the stage as it found it.
// Every configuration runs the same sweep (the GPU tests hold each to the
// oracle), but their sums round differently: with tuning on, which one ran
// depends on the timings, so results are reproducible only with the same
// cache (tb_autotune_put fills it from a saved one).
// ---------------------------------------------------------------------------
namespace {
struct TuneKey {
  int v[8];
  bool operator<(const TuneKey &o) const { return memcmp(v, o.v, sizeof(v)) < 0; }
};
std::mutex g_tune_mu;
std::map<TuneKey, tb_tune_entry> g_tune;

TuneKey tune_key(const tb_stage *st) {
  return {{st->mult_xy, st->mult_z, st->max_nv, st->N_r, st->N_x, st->N_y, st->N_z, st->N_time}};
}
}  // namespace

static int autotune_run(tb_stage *st, int iter, tb_tune_entry *best) {
  if (!st->ncells[0]) {
    best->kernel = 0;
    best->views_per_lane = 0;
    best->ms = 0;
    return 0;
  }
  if (harvest_events(st)) return -1;
  // state the trials change, copied aside
  const int ncell = st->N_time * st->N_z;
  const size_t ns = st->nsino(), no = st->nobj();
  float *e0 = nullptr, *x0 = nullptr, *m0 = nullptr;
  double *c0 = nullptr;
  unsigned long long *u0 = nullptr;
  auto release = [&]() {
    for (void *p : {(void *)e0, (void *)x0, (void *)m0, (void *)c0, (void *)u0})
      if (p) (void)hipFree(p);
  };
  if (dalloc(&e0, ns) != hipSuccess || dalloc(&x0, no) != hipSuccess ||
      dalloc(&m0, no) != hipSuccess || dalloc(&c0, ncell) != hipSuccess ||
      dalloc(&u0, 3 * (size_t)ncell) != hipSuccess) {
    release();
    return fail("autotune: no memory for the trial copies");
  }
  auto copy = [&](bool save) -> hipError_t {
    auto cp = [&](void *a, void *b, size_t n) {
      return hipMemcpyAsync(save ? a : b, save ? b : a, n, hipMemcpyDeviceToDevice, st->stream);
    };
    hipError_t r = cp(e0, st->e, ns * sizeof(float));
    if (r == hipSuccess) r = cp(x0, st->obj, no * sizeof(float));
    if (r == hipSuccess) r = cp(m0, st->mag, no * sizeof(float));
    if (r == hipSuccess) r = cp(c0, st->cell_voxmag, ncell * sizeof(double));
    if (r == hipSuccess) r = cp(u0, st->cell_upd, ncell * 8);
    if (r == hipSuccess) r = cp(u0 + ncell, st->cell_zero, ncell * 8);
    if (r == hipSuccess) r = cp(u0 + 2 * (size_t)ncell, st->cell_ent, ncell * 8);
    return r;
  };
  const double vms = st->voxel_ms;
  const long long vl = st->voxel_launches, ll = st->lds_launches, gl = st->global_launches;
  long long fam[TB_KERNEL_FAMILIES];
  memcpy(fam, st->fam_launches, sizeof(fam));
  const tb_options saved = g_opt;
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = 0;
  best->ms = -1;
  std::vector<const void *> seen;
  if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess ||
      copy(true) != hipSuccess)
    rc = fail("autotune: events or copies");
  const int kernels[] = {3, 4, 5, 6, 7, 8, 9, 10, 2};
  for (int k : kernels) {
    for (int v = (k == 3 ? 1 : 0); v <= (k == 3 ? 2 : 0) && !rc; v++) {
      g_opt = saved;
      g_opt.autotune = 0;
      g_opt.kernel = k;
      g_opt.views_per_lane = v;
      VoxPlan pl;
      if (launch_update_as(st, st->cells[0], st->ncells[0], iter, 0, 0, 0, 1, &pl)) {
        rc = -1;
        break;
      }
      if (!pl.fn || pl.kind != 0 ||
          std::find(seen.begin(), seen.end(), pl.fn) != seen.end())
        continue;
      seen.push_back(pl.fn);
      float ms = 0;
      if (hipEventRecord(ev[0], st->stream) != hipSuccess ||
          launch_update_as(st, st->cells[0], st->ncells[0], iter, 0, 0, 0, 1, nullptr) ||
          hipEventRecord(ev[1], st->stream) != hipSuccess || harvest_events(st) ||
          hipEventSynchronize(ev[1]) != hipSuccess ||
          hipEventElapsedTime(&ms, ev[0], ev[1]) != hipSuccess || copy(false) != hipSuccess) {
        rc = rc ? rc : fail("autotune: trial of kernel %d failed", k);
        break;
      }
      if (best->ms < 0 || ms < best->ms) {
        best->kernel = k;
        best->views_per_lane = v;
        best->ms = ms;
      }
    }
  }
  g_opt = saved;
  if (hipStreamSynchronize(st->stream) != hipSuccess && !rc) rc = fail("autotune: sync");
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  release();
  st->voxel_ms = vms;
  st->voxel_launches = vl;
  st->lds_launches = ll;
  st->global_launches = gl;
  memcpy(st->fam_launches, fam, sizeof(fam));
  if (!rc && best->ms < 0) {
    best->kernel = 0;  // nothing applied: the default choice
    best->views_per_lane = 0;
    best->ms = 0;
  }
  return rc;
}

static int autotune_pick(tb_stage *st, int iter, int *kernel, int *vpl) {
  const TuneKey key = tune_key(st);
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      *kernel = it->second.kernel;
      *vpl = it->second.views_per_lane;
      return 0;
    }
  }
  tb_tune_entry e;
  memset(&e, 0, sizeof(e));
  memcpy(e.shape, key.v, sizeof(key.v));
  if (autotune_run(st, iter, &e)) return -1;
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[key] = e;
  }
  *kernel = e.kernel;
  *vpl = e.views_per_lane;
  return 0;
}

extern "C" int tb_autotune(tb_stage *st, tb_tune_entry *out) {
  int k = 0, v = 0;
  if (autotune_pick(st, st->ord_iter > 0 ? st->ord_iter : 1, &k, &v)) return -1;
  std::lock_guard<std::mutex> g(g_tune_mu);
  if (out) *out = g_tune[tune_key(st)];
  return 0;
}

extern "C" int tb_autotune_entries(tb_tune_entry *out, int32_t cap) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  int n = 0;
  for (const auto &kv : g_tune) {
    if (out && n < cap) out[n] = kv.second;
    n++;
  }
  return n;
}

extern "C" int tb_autotune_put(const tb_tune_entry *e) {
  if (e->kernel < 0 || e->kernel > 10 || e->views_per_lane < 0 || e->views_per_lane > 2)
    return fail("tb_autotune_put: kernel %d / views per lane %d", e->kernel, e->views_per_lane);
  TuneKey k;
  memcpy(k.v, e->shape, sizeof(k.v));
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune[k] = *e;
  return 0;
}

extern "C" void tb_autotune_clear(void) {
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune.clear();
}

extern "C" int tb_get_stats(tb_stage *st, tb_stats *out) {
  if (harvest_events(st)) return -1;
  const int ncell = st->N_time * st->N_z;
  std::vector<unsigned long long> u(ncell), z(ncell), e(ncell);
  if (d2h(st, u.data(), st->cell_upd, ncell * 8) || d2h(st, z.data(), st->cell_zero, ncell * 8) ||
      d2h(st, e.data(), st->cell_ent, ncell * 8))
    return -1;
  memset(out, 0, sizeof(*out));
  for (int i = 0; i < ncell; i++) {
    out->updates += (int64_t)u[i];
    out->zero_skips += (int64_t)z[i];
    out->entries += (int64_t)e[i];
  }
  out->voxel_ms = st->voxel_ms;
  out->other_ms = st->other_ms;
  out->voxel_launches = st->voxel_launches;
  out->lds_launches = st->lds_launches;
  out->global_launches = st->global_launches;
  for (int k = 0; k < TB_KERNEL_FAMILIES; k++) out->family_launches[k] = st->fam_launches[k];
  out->column_builds = st->colt_builds;
  out->column_bins = st->colt ? st->colt_nb : 0;
  out->alg_bytes = 12.0 * (double)out->entries;
  return 0;
}

extern "C" int tb_reset_stats(tb_stage *st) {
  if (harvest_events(st)) return -1;
  const int ncell = st->N_time * st->N_z;
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ncell)), dim3(256), 0, st->stream, st->cell_upd,
                     (size_t)ncell);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ncell)), dim3(256), 0, st->stream, st->cell_zero,
                     (size_t)ncell);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(ncell)), dim3(256), 0, st->stream, st->cell_ent,
                     (size_t)ncell);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st->stream));
  st->voxel_ms = 0;
  st->other_ms = 0;
  st->voxel_launches = 0;
  st->lds_launches = 0;
  st->global_launches = 0;
  for (auto &k : st->fam_launches) k = 0;
  return 0;
}

extern "C" int tb_stage_object_device(tb_stage *st, void **ptr, uint64_t *bytes) {
  if (!st || !ptr) return fail("tb_stage_object_device: null argument");
  *ptr = st->obj;
  if (bytes) *bytes = (uint64_t)st->nobj() * sizeof(float);
  return 0;
}

extern "C" int tb_synchronize(tb_stage *st) {
  HIPCHK(hipStreamSynchronize(st->stream));
  return 0;
}

extern "C" int tb_stage_abort(tb_stage *st) {
  if (!st) return fail("tb_stage_abort: no stage");
  std::lock_guard<std::mutex> guard(g_split_mu);
  if (!st->split_w) return 0;  // no split-cell launch yet
  if (!st->abort_stream) HIPCHK(hipStreamCreateWithFlags(&st->abort_stream, hipStreamNonBlocking));
  static const unsigned one = 1u;
  HIPCHK(hipMemcpyAsync(st->split_w + 1, &one, sizeof(one), hipMemcpyHostToDevice, st->abort_stream));
  HIPCHK(hipStreamSynchronize(st->abort_stream));
  return 0;
}
