// Shared host/device definitions of libtimbir_hip.
#pragma once
#include <stdint.h>
#include <stddef.h>
#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#endif

namespace tb {

// XT_Constants.h:62-102 of the reference
constexpr double kMrfP = 1.2;
constexpr double kMrfQ = 2.0;
constexpr int kProfileRes = 1536;
constexpr int kBeamRes = 512;
constexpr int kRespBins = 64;  // DETECTOR_RESPONSE_BINS; rows hold 65 values
constexpr int kHrStride = kRespBins + 1;
constexpr int kFwdZ = 8;    // slices one forward-projection thread accumulates
constexpr int kFwdT = 256;  // detector bins per forward-projection block
constexpr double kAlpha = 1.5;
constexpr int kMaxIters = 1000;
constexpr double kZingerOffT = 100000.0;
constexpr double kZingerOffDelta = 1.0;
constexpr int kMinXYRes = 64;
constexpr int kMaxMultres = 8;
constexpr int kMinRowsPerNode = 2;
constexpr int kMinProjRows = 4;
constexpr int kNhicdK = 10;  // NHICD_Iterations, XT_Init.c:645
constexpr double kPi = 3.14159265358979323846;
constexpr double kPi2 = 1.57079632679489661923132169163975144;
constexpr double kPi4 = 0.785398163397448309615660845819875721;

// Everything a kernel needs, passed by value.  Device sinogram layout is
// [N_z][N_p][N_r][mz] (detector row t = z*mz + m): a cell (frame, slice)
// reads one contiguous [views][N_r][mz] block.
struct KP {
  int N_p, N_r, N_t, N_x, N_y, N_z, N_time, mz, K;
  double R0, RMax, delta_r, x0, y0, delta_xy, OffsetR;
  // detector coordinate in bin units u = (r - R0) / delta_r: 1/delta_r,
  // R0/delta_r, delta_xy/delta_r, (RMax - R0)/delta_r; delta_r/OffsetR
  double u_inv, u_r0, u_dxy, u_max;
  double u_kk;
  int mxy;           // delta_xy / delta_r (the stage's x-y multiplier)
  float T2h, T2l;    // T^2 as an unevaluated float sum (exact)
  float sSq, sSqp, sTq, sTqp, c_s, c_t, tfilt, alpha;
  float sfilt[27];
  float T, Td;       // generalized Huber threshold and T*delta
  double T_d, Td_d;  // same in double
  double dl_d;       // generalized Huber delta
  float *e, *w, *y, *d, *obj, *mag;
  const float *hr;
  const double *cosv, *sinv;
  const int *view_off, *view_idx;
  const uint8_t *mask;
  int *nh_list, *nh_count;  // NHICD lines, raster order (all selected)
  int nh_cap;
  int *nh_ord, *nh_ord_count;  // the same lines inside the mask, (y << 16) | x
  int *ord, *ord_off;          // per-iteration voxel order ((y << 16) | x), chunk starts
  int ord_cap;
  float *dummy;                // per-lane scratch word for masked-off stores
  const int *single;           // one-pixel sequence (single voxel update)
  int nobj;                    // object words; obj and mag hold 64 more
  // A-matrix column table of the LDS kernels: per (frame, pixel, view slot)
  // the values at bins w0 .. w0 + nb - 1 and w0 (as int bits, -1 outside
  // the detector), colt_rec floats per record, colt_nv view slots per frame
  const float *colt;
  int colt_rec, colt_nv;
  double *cell_voxmag;
  unsigned long long *cell_upd, *cell_zero, *cell_ent;
  uint32_t order_seed;
  int hicd_order;  // 0: a chunk's pixels in permutation order, 1: in raster order
};

// LDS bytes of the LDS-resident voxel-update kernel for a cell of nv views:
// the (e, w) block with an odd row stride, one dummy column window per lane,
// then the views' detector-response rows.
inline size_t icd_lds_bytes(int nv, int N_r, int mz, int nb) {
  const size_t ew = (size_t)nv * (N_r * mz + 1) + (size_t)64 * nb * mz;
  return ew * 2 * sizeof(float) + (size_t)nv * kHrStride * sizeof(float);
}

// Dynamic LDS bytes of the multi-wavefront kernel for a cell of nv views
// and nlanes theta-wave lane slots (theta waves x views per lane x 64): the
// (e, w) block with an odd row stride and one dummy word pair per lane slot.
inline size_t icd_mw_lds_bytes(int nv, int N_r, int mz, int nlanes) {
  const size_t ew = (size_t)nv * (N_r * mz + 1) + (size_t)nlanes;
  return ew * 2 * sizeof(float);
}

// Colour-schedule voxel order: keyed 4-round Feistel bijection on
// [0, 2^(2h)) with cycle walking onto [0, N).  No reference counterpart
// (the reference shuffles with random2, XT_ICD_update.c:500-532).
#if defined(__HIPCC__) || defined(__HIP__)
#define TB_HD __host__ __device__ __forceinline__
#else
#define TB_HD inline
#endif

TB_HD uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

TB_HD uint32_t perm_key(uint32_t seed, int iter, int frame, int slice) {
  uint32_t k = seed ^ 0x5bd1e995U;
  k = mix32(k ^ ((uint32_t)iter * 0x9e3779b1U));
  k = mix32(k ^ ((uint32_t)frame * 0x85ebca77U));
  k = mix32(k ^ ((uint32_t)slice * 0xc2b2ae3dU));
  return k;
}

TB_HD uint32_t perm(uint32_t n, uint32_t N, uint32_t key) {
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

// The inverse of perm: the Feistel rounds backwards, cycle-walking alike.
TB_HD uint32_t perm_inv(uint32_t n, uint32_t N, uint32_t key) {
  uint32_t bits = 1;
  while ((1u << bits) < N) bits++;
  uint32_t h = (bits + 1) / 2;
  uint32_t mask = (1u << h) - 1u;
  do {
    uint32_t L = n >> h, R = n & mask;
    for (int r = 3; r >= 0; r--) {
      // a round maps (L, R) to (R, L ^ F(R)); undo it from (L, R) = (R0, L0 ^ F(R0))
      uint32_t F = mix32(L ^ (key + (uint32_t)r * 0x9e3779b9U)) & mask;
      uint32_t pl = R ^ F;
      R = L;
      L = pl;
    }
    n = (L << h) | R;
  } while (n >= N);
  return n;
}

// Floats per column-table record for columns of nb bins: nb values and w0,
// rounded up to whole 16-byte words.
TB_HD constexpr int col_rec(int nb) { return (nb + 4) / 4 * 4; }

// Cross-workgroup exchange of the split-cell voxel kernel (k_icd_split):
// per cell slot and voxel parity, one record per workgroup of its partial
// theta sums, each 64-bit word tagged with the voxel's sequence number; a
// word set when a wait timed out and one the host may set to end every wait.
struct SplitBuf {
  unsigned long long *rec;  // [slots][2][smax][4]: theta1 high, theta1 low, theta2, unused
  unsigned *err;            // set when a wait timed out or was aborted
  const unsigned *abort;    // nonzero: end every wait
  int S, smax;
  unsigned spin_limit;
};

}  // namespace tb
