// Device helpers shared by the kernels of libtimbir_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include "tb_common.h"

namespace tb {

__device__ __forceinline__ size_t sidx(const KP &P, int z, int v, int r, int m) {
  return (((size_t)z * P.N_p + v) * P.N_r + r) * P.mz + m;
}

// A-matrix column of voxel (row, col) in one view: calcAMatrixColumnforAngle
// (XT_AMatrix.c:50-110) in double with contraction off, so every bin and
// value equals the reference's.  fn(bin, value) is called for each nonzero.
template <class F>
__device__ __forceinline__ void for_acol(const KP &P, const float *__restrict__ H, double c,
                                         double s, int row, int col, F &&fn) {
#pragma clang fp contract(off)
  double x = P.x0 + ((double)col + 0.5) * P.delta_xy;
  double y = P.y0 + ((double)row + 0.5) * P.delta_xy;
  double r = x * c - y * s;
  double rmin = r - P.delta_xy, rmax = r + P.delta_xy;
  if (rmax < P.R0 || rmin > P.RMax) return;
  int lo = (int)floor((rmin - P.R0) / P.delta_r);
  int hi = (int)floor((rmax - P.R0) / P.delta_r);
  if (hi >= P.N_r) hi = P.N_r - 1;
  if (lo < 0) lo = 0;
  for (int j = lo; j <= hi; j++) {
    double centre = P.R0 + ((double)j + 0.5) * P.delta_r;
    double dist = fabs(r - centre);
    int bin = (int)floor(dist / P.OffsetR);
    if (bin >= 0 && bin < kRespBins) {
      double w1 = dist - bin * P.OffsetR;
      double w2 = P.OffsetR - w1;
      double f1 = (w2 * (double)H[bin] + w1 * (double)H[bin + 1]) / P.OffsetR;
      if (f1 > 0) fn(j, f1);
    }
  }
}

// Value of bin `j` of the column above, or 0 when j is outside it.
__device__ __forceinline__ double acol_bin(const KP &P, const float *__restrict__ H, double c,
                                           double s, int row, int col, int j) {
#pragma clang fp contract(off)
  double x = P.x0 + ((double)col + 0.5) * P.delta_xy;
  double y = P.y0 + ((double)row + 0.5) * P.delta_xy;
  double r = x * c - y * s;
  double rmin = r - P.delta_xy, rmax = r + P.delta_xy;
  if (rmax < P.R0 || rmin > P.RMax) return 0.0;
  int lo = (int)floor((rmin - P.R0) / P.delta_r);
  int hi = (int)floor((rmax - P.R0) / P.delta_r);
  if (hi >= P.N_r) hi = P.N_r - 1;
  if (lo < 0) lo = 0;
  if (j < lo || j > hi) return 0.0;
  double centre = P.R0 + ((double)j + 0.5) * P.delta_r;
  double dist = fabs(r - centre);
  int bin = (int)floor(dist / P.OffsetR);
  if (bin < 0 || bin >= kRespBins) return 0.0;
  double w1 = dist - bin * P.OffsetR;
  double w2 = P.OffsetR - w1;
  double f1 = (w2 * (double)H[bin] + w1 * (double)H[bin + 1]) / P.OffsetR;
  return f1 > 0 ? f1 : 0.0;
}

// ProjSelect b = |e sqrt(w)| < T evaluated as the reference does, in double
// (XT_VoxUpdate.c:108-111); the float test decides all but a thin band.
__device__ __forceinline__ bool selected(const KP &P, float e, float w) {
  float u2 = e * e * w;
  float T2 = P.T * P.T;
  if (u2 < T2 * 0.99999f) return true;
  if (u2 > T2 * 1.00001f) return false;
  return fabs((double)e * sqrt((double)w)) < P.T_d;
}

// Neumaier compensated add.
__device__ __forceinline__ void cadd(float &s, float &c, float x) {
  float t = s + x;
  if (fabsf(s) >= fabsf(x)) c += (s - t) + x;
  else c += (x - t) + s;
  s = t;
}

// (s, c) pair sum over a wavefront with error-free TwoSum at each level.
__device__ __forceinline__ void wave_csum(float &s, float &c) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    float so = __shfl_xor(s, off);
    float co = __shfl_xor(c, off);
    float t = s + so;
    float bp = t - s;
    float err = (s - (t - bp)) + (so - bp);
    s = t;
    c = c + co + err;
  }
}

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// qGGMRF rho'(d)/d for p = 1.2, q = 2 (XT_ICD_update.c:98-129): with q = 2
// |d|^(q-2) = 1 and the d = 0 limit q/(sigma^q c) is the same expression.
__device__ __forceinline__ float qggmrf_b(float d, float sq, float sqp, float c) {
  float a = fabsf(d);
  float t1 = powf(a, (float)(kMrfQ - kMrfP)) / sqp;
  float t3 = c + t1;
  return ((float)kMrfQ - (float)(kMrfQ - kMrfP) * t1 / t3) / (t3 * sq);
}

// Wavefront sums: four DPP levels inside each row of 16 lanes (quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror) leave every
// lane holding its row's sum; then gfx950's permlane16 / permlane32 swaps
// pair rows (0,1), (2,3) and then the two halves.  A swap of a register
// with itself leaves one copy with one row of the pair and the other copy
// with the other row, so their sum is the pair's total, in every lane.  The
// total ends in every lane ((r0 + r1) + (r2 + r3)).  Every lane must be
// active.
template <int CTRL>
__device__ __forceinline__ int dpp_all(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(dpp_all<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)dpp_all<CTRL>((int)(unsigned)b);
  const unsigned hi = (unsigned)dpp_all<CTRL>((int)(unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <bool HALVES>
__device__ __forceinline__ float swap_sum_f(float v) {
  const unsigned u = __float_as_uint(v);
  const auto r = HALVES ? __builtin_amdgcn_permlane32_swap(u, u, false, false)
                        : __builtin_amdgcn_permlane16_swap(u, u, false, false);
  return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
template <bool HALVES>
__device__ __forceinline__ double swap_sum_d(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned l = (unsigned)b, h = (unsigned)(b >> 32);
  const auto lo = HALVES ? __builtin_amdgcn_permlane32_swap(l, l, false, false)
                         : __builtin_amdgcn_permlane16_swap(l, l, false, false);
  const auto hi = HALVES ? __builtin_amdgcn_permlane32_swap(h, h, false, false)
                         : __builtin_amdgcn_permlane16_swap(h, h, false, false);
  const double a = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  const double c = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
  return a + c;
}
__device__ __forceinline__ float wave_total(float v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  v += dpp_f<0x140>(v);  // row_mirror: row sums in every lane of a row
  v = swap_sum_f<false>(v);
  return swap_sum_f<true>(v);
}
__device__ __forceinline__ double wave_total_d(double v) {
  v += dpp_d<0xB1>(v);
  v += dpp_d<0x4E>(v);
  v += dpp_d<0x141>(v);
  v += dpp_d<0x140>(v);
  v = swap_sum_d<false>(v);
  return swap_sum_d<true>(v);
}
// A double and a float sum taken level by level, so that each level's lane
// moves fill the other tree's wait states.
__device__ __forceinline__ void wave_total_df(double &d, float &f) {
  d += dpp_d<0xB1>(d);
  f += dpp_f<0xB1>(f);
  d += dpp_d<0x4E>(d);
  f += dpp_f<0x4E>(f);
  d += dpp_d<0x141>(d);
  f += dpp_f<0x141>(f);
  d += dpp_d<0x140>(d);
  f += dpp_f<0x140>(f);
  d = swap_sum_d<false>(d);
  f = swap_sum_f<false>(f);
  d = swap_sum_d<true>(d);
  f = swap_sum_f<true>(f);
}
// Two float sums in plain FP32, level by level as wave_total_df.
__device__ __forceinline__ void wave_total_ff(float &a, float &b) {
  a += dpp_f<0xB1>(a);
  b += dpp_f<0xB1>(b);
  a += dpp_f<0x4E>(a);
  b += dpp_f<0x4E>(b);
  a += dpp_f<0x141>(a);
  b += dpp_f<0x141>(b);
  a += dpp_f<0x140>(a);
  b += dpp_f<0x140>(b);
  a = swap_sum_f<false>(a);
  b = swap_sum_f<false>(b);
  a = swap_sum_f<true>(a);
  b = swap_sum_f<true>(b);
}

// Error-free TwoSum of a and b: a + b = s + err exactly (Knuth).
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &err) {
#pragma clang fp contract(off)
  s = a + b;
  const float bp = s - a;
  err = (a - (s - bp)) + (b - bp);
}
// One level of a compensated pair tree: (s, c) += (so, co).
__device__ __forceinline__ void cc_level(float &s, float &c, float so, float co) {
  float t, err;
  two_sum(s, so, t, err);
  s = t;
  c = c + co + err;
}
template <int CTRL>
__device__ __forceinline__ void cc_dpp(float &s, float &c) {
  cc_level(s, c, dpp_f<CTRL>(s), dpp_f<CTRL>(c));
}
template <bool HALVES>
__device__ __forceinline__ void cc_swap(float &s, float &c) {
  const unsigned us = __float_as_uint(s), uc = __float_as_uint(c);
  const auto rs = HALVES ? __builtin_amdgcn_permlane32_swap(us, us, false, false)
                         : __builtin_amdgcn_permlane16_swap(us, us, false, false);
  const auto rc = HALVES ? __builtin_amdgcn_permlane32_swap(uc, uc, false, false)
                         : __builtin_amdgcn_permlane16_swap(uc, uc, false, false);
  s = __uint_as_float(rs[0]);
  c = __uint_as_float(rc[0]);
  cc_level(s, c, __uint_as_float(rs[1]), __uint_as_float(rc[1]));
}
// Two compensated FP32 sums (s1, c1) and (s2, c2) over the wavefront: the
// same DPP / permlane tree as wave_total, each level a TwoSum whose rounding
// error joins the compensation term; the totals end in every lane.
__device__ __forceinline__ void wave_total_cc(float &s1, float &c1, float &s2, float &c2) {
  cc_dpp<0xB1>(s1, c1);
  cc_dpp<0xB1>(s2, c2);
  cc_dpp<0x4E>(s1, c1);
  cc_dpp<0x4E>(s2, c2);
  cc_dpp<0x141>(s1, c1);
  cc_dpp<0x141>(s2, c2);
  cc_dpp<0x140>(s1, c1);
  cc_dpp<0x140>(s2, c2);
  cc_swap<false>(s1, c1);
  cc_swap<false>(s2, c2);
  cc_swap<true>(s1, c1);
  cc_swap<true>(s2, c2);
}

// Two float sums over lanes 0..31 only (rows 0 and 1), taken level by level;
// the totals end in every lane of those rows.
__device__ __forceinline__ void lo32_total_ff(float &a, float &b) {
  a += dpp_f<0xB1>(a);
  b += dpp_f<0xB1>(b);
  a += dpp_f<0x4E>(a);
  b += dpp_f<0x4E>(b);
  a += dpp_f<0x141>(a);
  b += dpp_f<0x141>(b);
  a += dpp_f<0x140>(a);
  b += dpp_f<0x140>(b);
  a = swap_sum_f<false>(a);
  b = swap_sum_f<false>(b);
}

}  // namespace tb
