// HIP kernels of the MBIR_4D hot path for gfx950 (CDNA4).
//
// Cells: (frame f, slice z).  Voxels of different cells touch disjoint
// error-sinogram entries (different frames own different views, different
// slices own different detector rows: XT_VoxUpdate.c:62-65), and cells of
// one colour (f + z) & 1 are never prior neighbours, so one colour's cells
// are updated concurrently without atomics; inside a cell voxels are
// updated one after another (Gauss-Seidel), as ICD requires.
#include <hip/hip_runtime.h>
#include <math.h>
#include "tb_common.h"
#include "tb_devutil.h"

namespace tb {

// ---------------------------------------------------------------------------
// Input transpose: [view][row][col] -> [z][view][col][m], weights / var_est
// (XT_Init.c:298-305, 327-330).
// ---------------------------------------------------------------------------
__global__ void k_transpose_in(KP P, const float *__restrict__ proj, const float *__restrict__ wgt,
                               double var_est) {
  size_t n = (size_t)P.N_p * P.N_t * P.N_r;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    int col = (int)(i % P.N_r);
    int row = (int)((i / P.N_r) % P.N_t);
    int v = (int)(i / ((size_t)P.N_r * P.N_t));
    size_t o = sidx(P, row / P.mz, v, col, row % P.mz);
    P.y[o] = proj[i];
    P.w[o] = (float)((double)wgt[i] / var_est);
  }
}

// Device layout -> reference layout [view][col][row] for outputs.
__global__ void k_to_ref_layout(KP P, const float *__restrict__ src, float *__restrict__ dst) {
  size_t n = (size_t)P.N_p * P.N_r * P.N_t;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    int t = (int)(i % P.N_t);
    int r = (int)((i / P.N_t) % P.N_r);
    int v = (int)(i / ((size_t)P.N_t * P.N_r));
    dst[i] = src[sidx(P, t / P.mz, v, r, t % P.mz)];
  }
}

// ---------------------------------------------------------------------------
// initObject's bilinear x2 upsampling (XT_ICD_update.c:323-651) from a coarser
// stage's array, on the device: in is [N_time][prev_nz][Ny][Nx], out is
// [N_time][out_nz][2 Ny][2 Nx] with out_nz = prev_nz (x-y, initICD 2) or
// 2 prev_nz (x-y-z, initICD 3).  Every output value is evaluated with the
// double expressions of the host restatement (x pass, y pass, z pass, edge
// rows first), contraction off; in the x-y form the x pass is rounded to
// float, as the reference's Real_arr_t rows are.  Bit for bit the host path.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double up_x(const float *r, int Nx, int k2, bool fmid) {
#pragma clang fp contract(off)
  const int W = 2 * Nx;
  double v;
  if (k2 == W - 1) v = r[Nx - 1];
  else if (k2 == W - 2) v = (r[Nx - 2] + 3.0 * r[Nx - 1]) / 4.0;
  else if (k2 == 0) v = r[0];
  else if (k2 == 1) v = (3.0 * r[0] + r[1]) / 4.0;
  else {
    const int k = k2 >> 1;
    v = (k2 & 1) ? (3.0 * r[k] + r[k + 1]) / 4.0 : (r[k - 1] + 3.0 * r[k]) / 4.0;
  }
  return fmid ? (double)(float)v : v;
}

__device__ __forceinline__ double up_xy(const float *plane, int Ny, int Nx, int j2, int k2,
                                        bool fmid) {
#pragma clang fp contract(off)
  auto T = [&](int j) { return up_x(plane + (size_t)j * Nx, Nx, k2, fmid); };
  const int H = 2 * Ny;
  if (j2 == 0) return T(0);
  if (j2 == 1) return (3.0 * T(0) + T(1)) / 4.0;
  if (j2 == H - 1) return T(Ny - 1);
  if (j2 == H - 2) return (T(Ny - 2) + 3.0 * T(Ny - 1)) / 4.0;
  const int j = j2 >> 1;
  return (j2 & 1) ? (3.0 * T(j) + T(j + 1)) / 4.0 : (T(j - 1) + 3.0 * T(j)) / 4.0;
}

__global__ void k_upsample(const float *__restrict__ in, float *__restrict__ out, int N_time,
                           int prev_nz, int Ny, int Nx, int three) {
#pragma clang fp contract(off)
  const int W = 2 * Nx, H = 2 * Ny, out_nz = three ? 2 * prev_nz : prev_nz;
  const size_t plane_in = (size_t)Ny * Nx, plane_out = (size_t)H * W;
  const size_t n = (size_t)N_time * out_nz * plane_out;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int x2 = (int)(i % W);
    const int y2 = (int)((i / W) % H);
    const int z2 = (int)((i / plane_out) % out_nz);
    const int f = (int)(i / (plane_out * out_nz));
    const float *base = in + (size_t)f * prev_nz * plane_in;
    double v;
    if (!three) {
      v = up_xy(base + (size_t)z2 * plane_in, Ny, Nx, y2, x2, true);
    } else {
      auto U = [&](int s) { return up_xy(base + (size_t)s * plane_in, Ny, Nx, y2, x2, false); };
      const int Z = 2 * prev_nz;
      if (z2 == 0) v = U(0);
      else if (z2 == 1) v = (3.0 * U(0) + U(1)) / 4.0;
      else if (z2 == Z - 1) v = U(prev_nz - 1);
      else if (z2 == Z - 2) v = (3.0 * U(prev_nz - 1) + U(prev_nz - 2)) / 4.0;
      else {
        const int s2 = z2 >> 1;
        v = (z2 & 1) ? (3.0 * U(s2) + U(s2 + 1)) / 4.0 : (U(s2 - 1) + 3.0 * U(s2)) / 4.0;
      }
    }
    out[i] = (float)v;
  }
}

// ---------------------------------------------------------------------------
// Forward projection, gather form.  Thread per (bin r, view v, group of
// kFwdZ slices); block per kFwdT bins of one view.  Each entry sums the
// voxels of the frame that owns v whose column covers r, in raster order,
// accumulating exactly as forward_project_voxel_AMat1D
// (XT_ForwardProject.c:8-25) does: float += double(pixel) * double(A), so
// the result is bit-identical to the reference's.  The A entry depends only
// on (view, pixel, bin): each is evaluated once and applied to the group's
// slices, read as one 32-byte record from the slice-interleaved copy that
// k_fwd_interleave writes.  Each thread visits only the rows and columns
// whose centres project into reach of its bin, so the threads of a view do
// similar work at every angle.  Every detector row of a slice receives the
// same value.  mode 0: out = A x; mode 1: e = y - A x - d.  Blocks are
// numbered view-major inside each XCD's share (block b runs on XCD b mod
// 8), so one XCD's L2 holds the objects of few frames.
// ---------------------------------------------------------------------------

// objt[f][g][row][col][q] = obj[f][g * kFwdZ + q][row][col], zero past N_z
__global__ void k_fwd_interleave(KP P, float4 *__restrict__ objt) {
  const int ngroups = (P.N_z + kFwdZ - 1) / kFwdZ;
  const size_t Nxy = (size_t)P.N_x * P.N_y;
  const size_t n = (size_t)P.N_time * ngroups * Nxy;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t pix = i % Nxy, fg = i / Nxy;
    const int g = (int)(fg % ngroups), f = (int)(fg / ngroups);
    float v[kFwdZ];
#pragma unroll
    for (int q = 0; q < kFwdZ; q++) {
      const int z = g * kFwdZ + q;
      v[q] = z < P.N_z ? P.obj[((size_t)f * P.N_z + z) * Nxy + pix] : 0.0f;
    }
    objt[2 * i] = make_float4(v[0], v[1], v[2], v[3]);
    objt[2 * i + 1] = make_float4(v[4], v[5], v[6], v[7]);
  }
}

// first and last index of the cells (centres at o + (k + 0.5) h, k in
// [0, n)) whose centre t satisfies t * g in [lo, hi], padded by one cell
__device__ __forceinline__ void fwd_span(double lo, double hi, double g, double o, double h,
                                         int n, int &k0, int &k1) {
  double ta = lo / g, tb = hi / g;
  if (ta > tb) { double t = ta; ta = tb; tb = t; }
  double ka = (ta - o) / h - 0.5, kb = (tb - o) / h - 0.5;
  // clamp before converting so huge ranges cannot overflow an int
  ka = fmin(fmax(ka, -4.0), (double)n + 4.0);
  kb = fmin(fmax(kb, -4.0), (double)n + 4.0);
  k0 = max(0, (int)floor(ka) - 1);
  k1 = min(n - 1, (int)ceil(kb) + 1);
}

__global__ __launch_bounds__(kFwdT) void k_forward(KP P, const int *__restrict__ view_frame,
                                                   const float4 *__restrict__ objt,
                                                   float *__restrict__ out, int mode) {
#pragma clang fp contract(off)
  __shared__ float s_H[kHrStride];
  const int ntile = (P.N_r + kFwdT - 1) / kFwdT;
  const int nb = ntile * P.N_p;
  int b = blockIdx.x;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  const int v = b / ntile;
  const int r = (b % ntile) * kFwdT + (int)threadIdx.x;
  const int g = blockIdx.y, ngroups = gridDim.y;
  const int z0 = g * kFwdZ;
  const int nz = min(kFwdZ, P.N_z - z0);
  const int f = view_frame[v];
  float acc[kFwdZ];
#pragma unroll
  for (int q = 0; q < kFwdZ; q++) acc[q] = 0.0f;
  if (f >= 0) {
    for (int i = threadIdx.x; i < kHrStride; i += blockDim.x)
      s_H[i] = P.hr[(size_t)v * kHrStride + i];
    __syncthreads();
  }
  if (f >= 0 && r < P.N_r) {
    const double c = P.cosv[v], s = P.sinv[v];
    const float4 *ob = objt + (size_t)(f * ngroups + g) * P.N_x * P.N_y * 2;
    // voxels whose projected centre x c - y s lies in [lo, hi) can cover bin r
    const double lo = P.R0 + r * P.delta_r - P.delta_xy - 1e-9;
    const double hi = P.R0 + (r + 1) * P.delta_r + P.delta_xy + 1e-9;
    int y_first = 0, y_last = P.N_y - 1;
    if (fabs(s) > 1e-6) {
      // y s in (x c - hi, x c - lo] for some column centre x
      const double xa = (P.x0 + 0.5 * P.delta_xy) * c;
      const double xb = (P.x0 + (P.N_x - 0.5) * P.delta_xy) * c;
      fwd_span(fmin(xa, xb) - hi, fmax(xa, xb) - lo, s, P.y0, P.delta_xy, P.N_y, y_first,
               y_last);
    }
    for (int row = y_first; row <= y_last; row++) {
      const double y = P.y0 + ((double)row + 0.5) * P.delta_xy;
      int k0 = 0, k1 = P.N_x - 1;
      // x c in [lo + y s, hi + y s)
      if (fabs(c) > 1e-6) fwd_span(lo + y * s, hi + y * s, c, P.x0, P.delta_xy, P.N_x, k0, k1);
      for (int col = k0; col <= k1; col++) {
        const float4 *px = ob + ((size_t)row * P.N_x + col) * 2;
        const float4 p0 = px[0], p1 = px[1];
        const double a = acol_bin(P, s_H, c, s, row, col, r);
        if (a != 0.0) {
          const float pv[kFwdZ] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w};
#pragma unroll
          for (int q = 0; q < kFwdZ; q++) acc[q] = (float)((double)acc[q] + (double)pv[q] * a);
        }
      }
    }
  }
  if (r >= P.N_r) return;
#pragma unroll
  for (int q = 0; q < kFwdZ; q++) {
    if (q >= nz) break;
    const int z = z0 + q;
    for (int m = 0; m < P.mz; m++) {
      size_t o = sidx(P, z, v, r, m);
      if (mode == 0) out[o] = acc[q];
      else out[o] = (P.y[o] - acc[q]) - P.d[(size_t)r * P.N_t + z * P.mz + m];
    }
  }
}

// ---------------------------------------------------------------------------
// Filtered back-projection (an initial object; the reference starts every
// reconstruction from zero, OBJECT_INIT_VAL in XT_Constants.h).  Parallel
// beam, ray offset r = x cos(theta) - y sin(theta) as the A matrix uses.
//   k_fbp_filter: q[z][v][r] = ramp * p, p the slice's measurements y - d
//     averaged over its mz detector rows; the Ram-Lak kernel sampled at the
//     bin pitch h (Kak and Slaney, eq. 3.61): 1/(4h^2) at 0, -1/(n pi h)^2
//     at odd n, 0 at even n, times h.  Linear convolution (zero outside the
//     detector), FP64 sums in a fixed order.
//   k_fbp_back: x[f][z][pixel] = (pi / n_f) sum over the n_f views of frame
//     f of q at the pixel's ray, linearly interpolated between bin centres;
//     zero outside the reconstruction mask.
// ---------------------------------------------------------------------------
__global__ void k_fbp_filter(KP P, float *__restrict__ q) {
  extern __shared__ double s_p[];  // one detector row
  const int v = blockIdx.x, z = blockIdx.y;
  for (int r = threadIdx.x; r < P.N_r; r += blockDim.x) {
    double acc = 0.0;
    for (int m = 0; m < P.mz; m++) {
      const size_t o = sidx(P, z, v, r, m);
      acc += (double)P.y[o] - (double)P.d[(size_t)r * P.N_t + z * P.mz + m];
    }
    s_p[r] = acc / P.mz;
  }
  __syncthreads();
  const double h = P.delta_r;
  const double k_odd = -1.0 / (kPi * kPi);
  for (int r = threadIdx.x; r < P.N_r; r += blockDim.x) {
    double acc = 0.25 * s_p[r];
    // odd offsets n = r - k
    for (int k = (r & 1) ^ 1; k < P.N_r; k += 2) {
      const double n = (double)(r - k);
      acc += k_odd * s_p[k] / (n * n);
    }
    q[((size_t)z * P.N_p + v) * P.N_r + r] = (float)(acc / h);
  }
}

__global__ void k_fbp_back(KP P, const float *__restrict__ q) {
  const size_t Nxy = (size_t)P.N_x * P.N_y;
  const size_t n = (size_t)P.N_time * P.N_z * Nxy;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const size_t pix = i % Nxy;
    const int z = (int)((i / Nxy) % P.N_z), f = (int)(i / (Nxy * P.N_z));
    const int row = (int)(pix / P.N_x), col = (int)(pix % P.N_x);
    float out = 0.0f;
    const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;
    if (P.mask[pix] && nv > 0) {
      const double x = P.x0 + ((double)col + 0.5) * P.delta_xy;
      const double y = P.y0 + ((double)row + 0.5) * P.delta_xy;
      double acc = 0.0;
      for (int k = 0; k < nv; k++) {
        const int v = P.view_idx[vb + k];
        const double u = (x * P.cosv[v] - y * P.sinv[v] - P.R0) / P.delta_r - 0.5;
        const double j = floor(u);
        const double t = u - j;
        const int j0 = (int)j;
        const float *qv = q + ((size_t)z * P.N_p + v) * P.N_r;
        const double q0 = (j0 >= 0 && j0 < P.N_r) ? (double)qv[j0] : 0.0;
        const double q1 = (j0 + 1 >= 0 && j0 + 1 < P.N_r) ? (double)qv[j0 + 1] : 0.0;
        acc += (1.0 - t) * q0 + t * q1;
      }
      out = (float)(acc * kPi / nv);
    }
    P.obj[i] = out;
  }
}

// ---------------------------------------------------------------------------
// Voxel update: one workgroup per cell of one colour, voxels in sequence.
// Per voxel (compute_voxel_update_Atten, XT_VoxUpdate.c:43-115):
//   threads stride over the frame's views, build the A column on the fly and
//   accumulate theta1 = -sum A e w~, theta2 = sum A^2 w~ (Neumaier FP32);
//   lanes 0..28 of wave 0 gather the 26 spatial + 2 temporal neighbours and
//   form the qGGMRF surrogate sums (CE_FunctionalSubstitution :134-183);
//   every thread solves x <- V + alpha (u - V) and subtracts A dx from its own
//   views' entries.  A thread only ever touches its own views' entries, so
//   the error-sinogram read-modify-write is thread-private.
// seq: 0 HICD chunk of the keyed permutation, 1 NHICD list, 2 one pixel.
// ---------------------------------------------------------------------------
template <int NT>
__global__ __launch_bounds__(NT) void k_update_voxels(KP P, const int2 *__restrict__ cells,
                                                      int iter, int chunk, int seq, int pix1,
                                                      int count_mag) {
  constexpr int NW = NT / 64;
  __shared__ float s_part[NW][4];
  __shared__ float s_prior[2];
  __shared__ float s_V;
  __shared__ int s_skip;

  const int2 cl = cells[blockIdx.x];
  const int f = cl.x, z = cl.y;
  const int cell = f * P.N_z + z;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int Nxy = P.N_x * P.N_y;
  const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;
  float *magc = P.mag + (size_t)cell * Nxy;
  float *objc = P.obj + (size_t)cell * Nxy;

  int n0 = 0, n1 = 1;
  uint32_t key = 0;
  if (seq == 0) {
    n0 = chunk * (Nxy / P.K);
    n1 = (chunk == P.K - 1) ? Nxy : (chunk + 1) * (Nxy / P.K);
    key = perm_key(P.order_seed, iter, f, z);
  } else if (seq == 1) {
    n1 = P.nh_count[cell];
  }

  double voxmag = 0.0;
  unsigned long long nupd = 0, nzero = 0, nent = 0;

  // neighbour slot of this lane (wave 0, lanes 0..28)
  int ndz = 0, ndy = 0, ndx = 0, ndt = 0;
  float ng = 0.0f;
  bool nspatial = false;
  if (wid == 0 && lane < 29) {
    if (lane < 27) {
      ndz = lane / 9 - 1;
      ndy = (lane / 3) % 3 - 1;
      ndx = lane % 3 - 1;
      nspatial = lane != 13;
      ng = P.sfilt[lane];
    } else {
      ndt = lane == 27 ? -1 : 1;
      ng = P.tfilt;
    }
  }

  // a HICD chunk in raster order (hicd_order 1): every pixel in turn, those
  // outside the chunk's permutation-index range skipped (uniformly)
  const bool raster = seq == 0 && P.hicd_order == 1;
  for (int n = raster ? 0 : n0; n < (raster ? Nxy : n1); n++) {
    int pix;
    if (raster) {
      pix = n;
      const int q = (int)perm_inv((uint32_t)pix, (uint32_t)Nxy, key);
      if (q < n0 || q >= n1) continue;
    } else if (seq == 0) {
      pix = (int)perm((uint32_t)n, (uint32_t)Nxy, key);
    } else if (seq == 1) {
      pix = P.nh_list[(size_t)cell * P.nh_cap + n];
    } else {
      pix = pix1;
    }
    const int yy = pix / P.N_x, xx = pix - yy * P.N_x;
    if (P.mask[pix] != 1) {
      if (seq == 1 && tid == 0) magc[pix] = 0.0f;  // XT_VoxUpdate.c:156
      continue;
    }

    // neighbourhood (XT_VoxUpdate.c:168-233)
    float nbv = 0.0f;
    bool nbok = false;
    if (wid == 0 && lane < 29) {
      if (lane < 27) {
        int zz = z + ndz, jj = yy + ndy, kk = xx + ndx;
        if (zz >= 0 && zz < P.N_z && jj >= 0 && jj < P.N_y && kk >= 0 && kk < P.N_x) {
          const float *q = &P.obj[(((size_t)f * P.N_z + zz) * P.N_y + jj) * P.N_x + kk];
          // in-plane neighbours may have been written earlier in this launch
          // by another lane: read them coherently from L2, not from the CU's L1
          nbv = (ndz == 0) ? __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : *q;
          nbok = true;
        }
      } else {
        int ff = f + ndt;
        if (ff >= 0 && ff < P.N_time) {
          nbv = P.obj[((size_t)ff * P.N_z + z) * Nxy + pix];
          nbok = true;
        }
      }
    }

    // theta1 / theta2 over the frame's views
    float t1s = 0.f, t1c = 0.f, t2s = 0.f, t2c = 0.f;
    for (int q = tid; q < nv; q += NT) {
      const int v = P.view_idx[vb + q];
      const float *H = P.hr + (size_t)v * kHrStride;
      for_acol(P, H, P.cosv[v], P.sinv[v], yy, xx, [&](int r, double ad) {
        const float a = (float)ad;
        for (int m = 0; m < P.mz; m++) {
          const size_t s = sidx(P, z, v, r, m);
          const float e = P.e[s], w = P.w[s];
          const float cw = selected(P, e, w) ? w : P.Td * sqrtf(w) / fabsf(e);
          cadd(t2s, t2c, a * a * cw);
          cadd(t1s, t1c, -(e * a * cw));
          nent++;
        }
      });
    }
    wave_csum(t1s, t1c);
    wave_csum(t2s, t2c);

    if (wid == 0) {
      const float V = __shfl(nbv, 13);
      // zero skipping (XT_VoxUpdate.c:236-259)
      const bool pos = (lane < 29) && (lane != 13) && (nbv > 0.0f);
      const bool anypos = __any(pos);
      float pb = 0.f, pbx = 0.f;
      if (lane < 29 && nbok && lane != 13) {
        const float dlt = V - nbv;
        const float b = (lane < 27) ? qggmrf_b(dlt, P.sSq, P.sSqp, P.c_s)
                                    : qggmrf_b(dlt, P.sTq, P.sTqp, P.c_t);
        pb = ng * b;
        pbx = ng * b * (nbv - V);
      }
      pb = wave_sum(pb);
      pbx = wave_sum(pbx);
      if (lane == 0) {
        s_prior[0] = pb;
        s_prior[1] = pbx;
        s_V = V;
        s_skip = (V == 0.0f && iter > 1 && !anypos) ? 1 : 0;
      }
      (void)nspatial;
    }
    if (lane == 0) {
      s_part[wid][0] = t1s;
      s_part[wid][1] = t1c;
      s_part[wid][2] = t2s;
      s_part[wid][3] = t2c;
    }
    __syncthreads();

    if (!s_skip) {
      float T1 = 0.f, T1c = 0.f, T2 = 0.f, T2c = 0.f;
#pragma unroll
      for (int w = 0; w < NW; w++) {
        cadd(T1, T1c, s_part[w][0]);
        T1c += s_part[w][1];
        cadd(T2, T2c, s_part[w][2]);
        T2c += s_part[w][3];
      }
      const float th1 = T1 + T1c, th2 = T2 + T2c;
      const float V = s_V;
      // x = V + alpha (u - V), u - V = (sum g b (x_j - V) - theta1) / (sum g b + theta2)
      const float step = P.alpha * (s_prior[1] - th1) / (s_prior[0] + th2);
      const float xnew = V + step;
      const float dx = xnew - V;
      if (tid == 0) {
        __hip_atomic_store(&objc[pix], xnew, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        magc[pix] = fabsf(dx);
        if (count_mag) voxmag += fabs((double)xnew);
        nupd++;
      }
      if (dx != 0.0f) {
        for (int q = tid; q < nv; q += NT) {
          const int v = P.view_idx[vb + q];
          const float *H = P.hr + (size_t)v * kHrStride;
          for_acol(P, H, P.cosv[v], P.sinv[v], yy, xx, [&](int r, double ad) {
            const float a = (float)ad;
            for (int m = 0; m < P.mz; m++) {
              const size_t s = sidx(P, z, v, r, m);
              P.e[s] = P.e[s] - a * dx;
            }
          });
        }
      }
    } else {
      if (tid == 0) {
        magc[pix] = 0.0f;
        nzero++;
      }
    }
    __syncthreads();
  }

  // per-cell accounting
  nent = (unsigned long long)wave_sum_d((double)nent);
  __shared__ unsigned long long s_ent[NW];
  if (lane == 0) s_ent[wid] = nent;
  __syncthreads();
  if (tid == 0) {
    unsigned long long te = 0;
    for (int w = 0; w < NW; w++) te += s_ent[w];
    if (seq == 0 || seq == 1) {
      P.cell_voxmag[cell] += voxmag;
      P.cell_upd[cell] += nupd;
      P.cell_zero[cell] += nzero;
      P.cell_ent[cell] += te;
    }
  }
}

template __global__ void k_update_voxels<64>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_update_voxels<128>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_update_voxels<256>(KP, const int2 *, int, int, int, int, int);

// ---------------------------------------------------------------------------
// Offset passes (XT_ICD_update.c:753-917).  Per detector pixel (r, t) the
// weights c = w (b) or T delta sqrt(w)/|e| (not b) are reduced over all views
// in double: num = sum c e (type 1/2, with c*sign(e)*|e| for not-b) or
// sum c (e + d) (type 3), den = sum c.  Partial sums over view chunks, then
// a fixed-order finalize.
// ---------------------------------------------------------------------------
__global__ void k_offset_partial(KP P, int type, int views_per_chunk, double *__restrict__ pnum,
                                 double *__restrict__ pden) {
  const int npix = P.N_r * P.N_t;
  const int chunk = blockIdx.y;
  for (int px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
    // px enumerates (z, r, m) so consecutive threads touch consecutive floats
    const int m = px % P.mz;
    const int r = (px / P.mz) % P.N_r;
    const int z = px / (P.mz * P.N_r);
    const int t = z * P.mz + m;
    const double dof = P.d[(size_t)r * P.N_t + t];
    double num = 0, den = 0;
    const int v0 = chunk * views_per_chunk, v1 = min(P.N_p, v0 + views_per_chunk);
    // eight views' (e, w) loaded at once, then added in view order
    for (int vb = v0; vb < v1; vb += 8) {
      float ev[8], wv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const size_t s = sidx(P, z, min(vb + u, v1 - 1), r, m);
        ev[u] = P.e[s];
        wv[u] = P.w[s];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        if (vb + u >= v1) break;
        const double e = ev[u], w = wv[u];
        const double tmp = P.Td_d * sqrt(w);
        const bool b = fabs(e * sqrt(w)) < P.T_d;
        if (type == 3) {
          if (b) { den += w; num += (e + dof) * w; }
          else { den += tmp / fabs(e); num += (e + dof) * tmp / fabs(e); }
        } else {
          if (b) { num += e * w; den += w; }
          else { double sg = (e > 0) - (e < 0); num += tmp * sg; den += tmp / fabs(e); }
        }
      }
    }
    pnum[(size_t)chunk * npix + px] = num;
    pden[(size_t)chunk * npix + px] = den;
  }
}

__global__ void k_offset_finalize(KP P, int nchunks, const double *__restrict__ pnum,
                                  const double *__restrict__ pden, double *__restrict__ num,
                                  double *__restrict__ den) {
  const int npix = P.N_r * P.N_t;
  for (int px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
    double a = 0, b = 0;
    // eight partials loaded at once, then added in chunk order
    for (int cb = 0; cb < nchunks; cb += 8) {
      double pa[8], pb[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int c = min(cb + u, nchunks - 1);
        pa[u] = pnum[(size_t)c * npix + px];
        pb[u] = pden[(size_t)c * npix + px];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        if (cb + u >= nchunks) break;
        a += pa[u];
        b += pb[u];
      }
    }
    num[px] = a;
    den[px] = b;
  }
}

// type 2: gamma = sum(d + num/den) / sum(1/den), one workgroup
__global__ void k_offset_gamma(KP P, const double *__restrict__ num, const double *__restrict__ den,
                               double *__restrict__ gamma) {
  const int npix = P.N_r * P.N_t;
  double a = 0, b = 0;
  for (int px = threadIdx.x; px < npix; px += blockDim.x) {
    const int m = px % P.mz, r = (px / P.mz) % P.N_r, z = px / (P.mz * P.N_r);
    const double dof = P.d[(size_t)r * P.N_t + z * P.mz + m];
    a += dof + num[px] / den[px];
    b += 1.0 / den[px];
  }
  __shared__ double sa[256], sb[256];
  sa[threadIdx.x] = a;
  sb[threadIdx.x] = b;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sa[threadIdx.x] += sa[threadIdx.x + o];
      sb[threadIdx.x] += sb[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *gamma = sa[0] / sb[0];
}

// apply: d and e for every view
__global__ void k_offset_apply(KP P, int type, const double *__restrict__ num,
                               const double *__restrict__ den, const double *__restrict__ gamma) {
  const int npix = P.N_r * P.N_t;
  const size_t n = (size_t)npix * P.N_p;
  const double g = (type == 2) ? *gamma : 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int m = (int)(i % P.mz);
    const int r = (int)((i / P.mz) % P.N_r);
    const int v = (int)((i / ((size_t)P.mz * P.N_r)) % P.N_p);
    const int z = (int)(i / ((size_t)P.mz * P.N_r * P.N_p));
    const int px = (z * P.N_r + r) * P.mz + m;
    const size_t dofi = (size_t)r * P.N_t + z * P.mz + m;
    const float dold = P.d[dofi];
    float &e = P.e[i];
    if (type == 3) {
      const float x = (float)(num[px] / den[px]);
      e += dold - x;
    } else {
      const double step = (num[px] - g) / den[px];
      e = (float)((double)e - step);
    }
  }
}

// new offsets after k_offset_apply has consumed the old ones
__global__ void k_offset_store(KP P, int type, const double *__restrict__ num,
                               const double *__restrict__ den, const double *__restrict__ gamma) {
  const int npix = P.N_r * P.N_t;
  const double g = (type == 2) ? *gamma : 0.0;
  for (int px = blockIdx.x * blockDim.x + threadIdx.x; px < npix; px += gridDim.x * blockDim.x) {
    const int m = px % P.mz, r = (px / P.mz) % P.N_r, z = px / (P.mz * P.N_r);
    const size_t dofi = (size_t)r * P.N_t + z * P.mz + m;
    if (type == 3) P.d[dofi] = (float)(num[px] / den[px]);
    else P.d[dofi] = (float)((double)P.d[dofi] + (num[px] - g) / den[px]);
  }
}

// ---------------------------------------------------------------------------
// Variance (XT_ICD_update.c:720-751) and cost (:186-319): FP64 partials per
// workgroup, finished on the host in a fixed order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum_d(double v, double *sh) {
  v = wave_sum_d(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += sh[i];
  __syncthreads();
  return t;
}

__global__ void k_var_partial(KP P, double var_old, double *__restrict__ part) {
  __shared__ double sh[16];
  const size_t n = (size_t)P.N_p * P.N_r * P.N_t;
  double acc = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float wr = (float)((double)P.w[i] * var_old);
    const double e = P.e[i], w = wr;
    if (fabs(e * sqrt((double)P.w[i])) < P.T_d) acc += e * e * w;
    else acc += fabs(e) * P.Td_d * sqrt(w * var_old);
  }
  double t = block_sum_d(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_var_apply(KP P, double var_old, double var_new) {
  const size_t n = (size_t)P.N_p * P.N_r * P.N_t;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float wr = (float)((double)P.w[i] * var_old);
    P.w[i] = (float)((double)wr / var_new);
  }
}

__global__ void k_cost_data(KP P, double *__restrict__ part) {
  __shared__ double sh[16];
  const size_t n = (size_t)P.N_p * P.N_r * P.N_t;
  double acc = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double u = (double)P.e[i] * sqrt((double)P.w[i]);
    if (fabs(u) < P.T_d) acc += u * u;
    else acc += 2.0 * P.dl_d * P.T_d * fabs(u) + P.T_d * P.T_d * (1.0 - 2.0 * P.dl_d);
  }
  double t = block_sum_d(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__device__ __forceinline__ double qggmrf_value_d(double d, double sq, double sqp, double c) {
  double a = fabs(d);
  return (a * a / sq) / (c + pow(a, kMrfQ - kMrfP) / sqp);
}

__global__ void k_cost_prior(KP P, double sSq, double sSqp, double cs, double sTq, double sTqp,
                             double ct, const double *__restrict__ sfilt, double tfilt,
                             double *__restrict__ part) {
  __shared__ double sh[16];
  const int Nxy = P.N_x * P.N_y;
  const size_t n = (size_t)P.N_time * P.N_z * Nxy;
  // the 13 forward neighbours in (dz, dy, dx) (XT_ICD_update.c:223-292)
  const int fw[13][3] = {{0, 0, 1},  {0, 1, -1}, {0, 1, 0},  {0, 1, 1},  {1, -1, 0},
                         {1, 0, 0},  {1, 1, 0},  {1, -1, -1}, {1, -1, 1}, {1, 0, -1},
                         {1, 1, -1}, {1, 1, 1},  {1, 0, 1}};
  double acc = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int k = (int)(i % P.N_x);
    const int j = (int)((i / P.N_x) % P.N_y);
    const int z = (int)((i / Nxy) % P.N_z);
    const int f = (int)(i / ((size_t)Nxy * P.N_z));
    const double v = P.obj[i];
    for (int q = 0; q < 13; q++) {
      const int zz = z + fw[q][0], jj = j + fw[q][1], kk = k + fw[q][2];
      if (zz >= P.N_z || jj < 0 || jj >= P.N_y || kk < 0 || kk >= P.N_x) continue;
      const double g = sfilt[(fw[q][0] + 1) * 9 + (fw[q][1] + 1) * 3 + (fw[q][2] + 1)];
      const double o = P.obj[(((size_t)f * P.N_z + zz) * P.N_y + jj) * P.N_x + kk];
      acc += g * qggmrf_value_d(v - o, sSq, sSqp, cs);
    }
    if (f + 1 < P.N_time)
      acc += tfilt * qggmrf_value_d(v - (double)P.obj[i + (size_t)P.N_z * Nxy], sTq, sTqp, ct);
  }
  double t = block_sum_d(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// 5x5 Hamming-weighted sum of the update-magnitude map (ComputeVSC,
// XT_NHICD.c:128-162): one 16 x 16 tile of one plane (frame, slice) per
// workgroup, staged in LDS with its 2-pixel halo.  Taps are added in the
// reference's row-major order in double with contraction off, so every
// filtered value, and with it the NHICD selection, equals the CPU
// restatement's.
__global__ __launch_bounds__(256) void k_vsc_filter(KP P, float *__restrict__ filt) {
#pragma clang fp contract(off)
  const double tab[5][5] = {{0.0064, 0.0432, 0.0800, 0.0432, 0.0064},
                            {0.0432, 0.2916, 0.5400, 0.2916, 0.0432},
                            {0.0800, 0.5400, 1.0000, 0.5400, 0.0800},
                            {0.0432, 0.2916, 0.5400, 0.2916, 0.0432},
                            {0.0064, 0.0432, 0.0800, 0.0432, 0.0064}};
  __shared__ float tile[20][21];
  const int Nxy = P.N_x * P.N_y;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int x0 = blockIdx.x * 16, y0 = blockIdx.y * 16;
  const float *m = P.mag + (size_t)blockIdx.z * Nxy;
  for (int t = threadIdx.x; t < 400; t += 256) {
    const int r = t / 20, c = t - r * 20;
    const int y = y0 + r - 2, x = x0 + c - 2;
    const bool in = (y >= 0) & (y < P.N_y) & (x >= 0) & (x < P.N_x);
    tile[r][c] = in ? m[y * P.N_x + x] : 0.0f;
  }
  __syncthreads();
  const int j = y0 + ty, k = x0 + tx;
  if (j >= P.N_y || k >= P.N_x) return;
  double acc = 0;
#pragma unroll
  for (int a = -2; a <= 2; a++)
#pragma unroll
    for (int c = -2; c <= 2; c++)
      if (j + a >= 0 && j + a < P.N_y && k + c >= 0 && k + c < P.N_x)
        acc += tab[a + 2][c + 2] * (double)tile[ty + 2 + a][tx + 2 + c];
  filt[(size_t)blockIdx.z * Nxy + j * P.N_x + k] = (float)acc;
}

// NHICD line selection of one cell per workgroup of 1024 threads
// (VSC_based_Voxel_Line_Select, XT_NHICD.c:161-216): the threshold is the
// (N_xy - want)-th smallest filtered magnitude, found exactly by a 4-pass
// 8-bit radix select over the float bits (the values are >= 0, so their bit
// patterns order like the values), and the lines above it are listed in
// raster order.  The reference finds the same order statistic with a
// randomized quickselect (XT_NHICD.c:63-128).
constexpr int kVscThreads = 1024;
__global__ __launch_bounds__(kVscThreads) void k_vsc_select(KP P, const float *__restrict__ filt,
                                                            int want) {
  constexpr int NW = kVscThreads / 64;
  const int cell = blockIdx.x;
  const int Nxy = P.N_x * P.N_y;
  const uint32_t *vals = reinterpret_cast<const uint32_t *>(filt + (size_t)cell * Nxy);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __shared__ uint32_t hist[256];
  __shared__ uint32_t s_prefix, s_k;
  __shared__ int s_cnt[2][NW];
  if (threadIdx.x == 0) {
    s_prefix = 0;
    s_k = (uint32_t)(Nxy - want);  // 1-based rank of the threshold
  }
  __syncthreads();
  uint32_t thresh_bits = 0xffffffffu;  // want <= 0: nothing above
  if (want > 0) {
    for (int pass = 0; pass < 4; pass++) {
      const int shift = 24 - 8 * pass;
      const uint32_t hmask = pass == 0 ? 0u : (0xffffffffu << (32 - 8 * pass));
      if (threadIdx.x < 256) hist[threadIdx.x] = 0;
      __syncthreads();
      const uint32_t pre = s_prefix;
      // eight independent loads in flight per thread
      for (int i0 = threadIdx.x; i0 < Nxy; i0 += 8 * kVscThreads) {
        uint32_t b[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int i = i0 + u * kVscThreads;
          b[u] = i < Nxy ? vals[i] : 0u;
        }
#pragma unroll
        for (int u = 0; u < 8; u++)
          if ((i0 + u * kVscThreads < Nxy) && (b[u] & hmask) == pre)
            atomicAdd(&hist[(b[u] >> shift) & 0xff], 1u);
      }
      __syncthreads();
      if (wid == 0) {
        // inclusive scan of the 256 counts, 4 per lane, then the first
        // bucket whose running count reaches k
        const uint32_t c0 = hist[4 * lane], c1 = hist[4 * lane + 1], c2 = hist[4 * lane + 2],
                       c3 = hist[4 * lane + 3];
        const uint32_t own = c0 + c1 + c2 + c3;
        uint32_t incl = own;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
          const uint32_t v = __shfl_up(incl, o);
          if (lane >= o) incl += v;
        }
        const uint32_t excl = incl - own;
        const uint32_t k = s_k;
        const bool here = (excl < k) & (incl >= k);
        const unsigned long long bal = __ballot(here);
        const int owner = bal ? __ffsll((long long)bal) - 1 : 63;
        if (lane == owner) {
          uint32_t acc = excl;
          int q = 0;
          const uint32_t cs[4] = {c0, c1, c2, c3};
          for (; q < 3; q++) {
            if (acc + cs[q] >= k) break;
            acc += cs[q];
          }
          s_k = k - acc;
          s_prefix = pre | ((uint32_t)(4 * lane + q) << shift);
        }
      }
      __syncthreads();
    }
    thresh_bits = s_prefix;
  }
  const float thresh = __uint_as_float(thresh_bits);
  // compaction in raster order: wave w owns the contiguous range
  // [w * per, (w + 1) * per) and walks it 64 pixels at a time (coalesced,
  // ballot offsets); one block-wide prefix of the wave counts in between
  const int per = (Nxy + NW - 1) / NW;
  const int w0 = min(wid * per, Nxy), w1 = min(w0 + per, Nxy);
  int nsel = 0, nin = 0;
  if (want > 0)
    for (int i = w0 + lane; i - lane < w1; i += 64) {
      const bool f = i < w1 && __uint_as_float(vals[i]) > thresh;
      const bool in = f && P.mask[i] == 1;
      nsel += __popcll(__ballot(f));
      nin += __popcll(__ballot(in));
    }
  if (lane == 0) {
    s_cnt[0][wid] = nsel;
    s_cnt[1][wid] = nin;
  }
  __syncthreads();
  int ps = 0, pi = 0, ts = 0, ti = 0;
  for (int w = 0; w < NW; w++) {
    if (w < wid) {
      ps += s_cnt[0][w];
      pi += s_cnt[1][w];
    }
    ts += s_cnt[0][w];
    ti += s_cnt[1][w];
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  if (want > 0)
    for (int i = w0 + lane; i - lane < w1; i += 64) {
      const bool f = i < w1 && __uint_as_float(vals[i]) > thresh;
      const bool in = f && P.mask[i] == 1;
      const unsigned long long bf = __ballot(f), bi = __ballot(in);
      if (f) {
        const int pos = ps + __popcll(bf & below);
        if (pos < P.nh_cap) P.nh_list[(size_t)cell * P.nh_cap + pos] = i;
      }
      if (in) {
        const int pos = pi + __popcll(bi & below);
        if (pos < P.nh_cap)
          P.nh_ord[(size_t)cell * P.nh_cap + pos] = ((i / P.N_x) << 16) | (i % P.N_x);
      }
      // a selected line outside the mask only has its map entry cleared
      // (XT_VoxUpdate.c:156)
      if (f && !in) P.mag[(size_t)cell * Nxy + i] = 0.0f;
      ps += __popcll(bf);
      pi += __popcll(bi);
    }
  if (threadIdx.x == 0) {
    P.nh_count[cell] = min(ts, P.nh_cap);
    P.nh_ord_count[cell] = min(ti, P.nh_cap);
  }
}

// sum of the magnitude map over masked pixels (convergence numerator,
// XT_ICD_update.c:1068-1072)
__global__ void k_mag_sum(KP P, double *__restrict__ part) {
  __shared__ double sh[16];
  const int Nxy = P.N_x * P.N_y;
  const size_t n = (size_t)P.N_time * P.N_z * Nxy;
  double acc = 0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    if (P.mask[i % Nxy]) acc += P.mag[i];
  double t = block_sum_d(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_fill_f64(double *p, size_t n, double v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_f32(float *p, size_t n, float v) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

}  // namespace tb
