// ICD voxel-update kernels for gfx950 (updateVoxels_Atten and
// compute_voxel_update_Atten, XT_VoxUpdate.c:43-289), one workgroup per cell.
//
// A cell is (frame f, slice z).  Every error-sinogram entry it touches lies
// in the block [views of f][N_r][rows of z], and no other cell of the same
// colour (f + z) & 1 touches it, so cells of one colour run side by side with
// no atomics.  Inside a cell the voxels are updated one after another, and
// each kernel keeps that sequence exact by letting only one wavefront read or
// write a given e entry: the operations of one wavefront to one address, in
// LDS or in global memory, take effect in issue order.
//
//   k_icd_lds    one wavefront per cell, lane = view, the (e, w) block in LDS
//   k_icd_mw     theta waves (views split between them) + a prior wave, the
//                (e, w) block in LDS, the A values from the column table
//   k_icd_mwg    the same roles with e and w in global memory, for blocks
//                that do not fit in LDS (coarse stages, large finest stages)
//   k_icd_wide   columns of up to 17 bins (coarsest stages): lane = (view,
//                detector rows), A values evaluated in the lane
//   k_icd_wpass  many views per cell: waves own blocks of views and walk
//                their entries in coalesced passes
//   k_icd_wloop  k_icd_wpass with any number of passes, in groups, e read
//                again for the write
//
// In the multi-wavefront kernels every wave publishes its partial theta sums
// in LDS, all waves meet at one barrier per voxel, and each evaluates the
// same step from the same sums, so all hold the identical update.  The prior
// wave loads the 28 neighbour values one group of voxels ahead and patches
// in-flight copies when a neighbour is updated in between.
#include <hip/hip_runtime.h>

#include <type_traits>
#include <math.h>
#include "tb_common.h"
#include "tb_devutil.h"

namespace tb {

// Per-iteration voxel order of every cell, and where each of the K HICD
// chunks starts.  Chunk j holds the unmasked pixels whose index in the
// cell's keyed permutation of the N_x*N_y pixels lies in [j per, (j + 1)
// per) (the last chunk to the end), as XT_ICD_update.c:983-986 splits the
// reference's random list; pixels are (y << 16) | x.  With hicd_order 0 a
// chunk's pixels are listed in permutation order, with 1 in raster order
// (consecutive voxels then share the cache lines of their sinogram rows,
// column records and neighbours).  One workgroup per cell.
__global__ __launch_bounds__(256) void k_build_order(KP P, int iter) {
  const int cell = blockIdx.x;
  const int f = cell / P.N_z, z = cell - f * P.N_z;
  const int Nxy = P.N_x * P.N_y;
  const uint32_t key = perm_key(P.order_seed, iter, f, z);
  int *ord = P.ord + (size_t)cell * P.ord_cap;
  int *off = P.ord_off + (size_t)cell * (P.K + 1);
  const int per = Nxy / P.K;
  __shared__ int s_wcnt[4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int base = 0;
  // appends this thread's pixel (when ok) after `base`, in thread order
  auto append = [&](bool ok, int pix) -> int {
    const unsigned long long bal = __ballot(ok);
    const int inwave = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s_wcnt[wid] = __popcll(bal);
    __syncthreads();
    int pos = base;
    for (int w = 0; w < wid; w++) pos += s_wcnt[w];
    pos += inwave;
    if (ok && pos < P.ord_cap) ord[pos] = ((pix / P.N_x) << 16) | (pix % P.N_x);
    return pos;
  };
  if (P.hicd_order == 1) {
    for (int j = 0; j < P.K; j++) {
      if (threadIdx.x == 0) off[j] = min(base, P.ord_cap);
      for (int start = 0; start < Nxy; start += blockDim.x) {
        const int pix = start + threadIdx.x;
        bool ok = false;
        if (pix < Nxy && P.mask[pix] == 1) {
          const int n = (int)perm_inv((uint32_t)pix, (uint32_t)Nxy, key);
          ok = (per > 0 ? min(n / per, P.K - 1) : 0) == j;
        }
        append(ok, pix);
        for (int w = 0; w < 4; w++) base += s_wcnt[w];
        __syncthreads();
      }
    }
  } else {
    for (int start = 0; start < Nxy; start += blockDim.x) {
      const int n = start + threadIdx.x;
      int pix = 0;
      bool ok = false;
      if (n < Nxy) {
        pix = (int)perm((uint32_t)n, (uint32_t)Nxy, key);
        ok = P.mask[pix] == 1;
      }
      const int pos = append(ok, pix);
      // chunk j starts at permutation index j * per; the count of masked
      // pixels before index n is pos (before this thread's own pixel)
      if (n < Nxy && per > 0 && n % per == 0 && n / per < P.K) off[n / per] = pos;
      for (int w = 0; w < 4; w++) base += s_wcnt[w];
      __syncthreads();
    }
  }
  if (threadIdx.x == 0) off[P.K] = min(base, P.ord_cap);
}

__device__ __forceinline__ void wait_vm_all() {
  __builtin_amdgcn_sched_barrier(0);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), expcnt and lgkmcnt untouched
  __builtin_amdgcn_sched_barrier(0);
}

// Generalized-Huber selection b = |e sqrt(w)| < T (XT_VoxUpdate.c:108-111)
// decided exactly: e^2 w is carried as q + (error terms) with FMA residuals
// and compared with T^2 = T2h + T2l, so the outcome is the exact one (the
// reference's double evaluation differs from it only within 1e-16 of T).
__device__ __forceinline__ bool huber_sel(const KP &P, float e, float w) {
  const float p = e * e;
  const float pe = fmaf(e, e, -p);
  const float q = p * w;
  const float qe = fmaf(p, w, -q);
  const float corr = fmaf(pe, w, qe);
  return (q - P.T2h) + (corr - P.T2l) < 0.0f;
}

constexpr int kNoPix = 0x7fffffff;  // empty slot (never equals a packed pixel)

// One wavefront per cell, lane = view (VPL views per lane).  The workgroup
// copies the cell's (e, w) block into LDS (64 views x 256 bins x (e, w) is
// 128 KB of the 160 KB), runs the cell's whole voxel sequence against it and
// copies e back at the end.  The 28 neighbour values of each voxel stay in
// global memory and are loaded one group of G voxels ahead.  The loop is laid
// out so that its one wait for global memory, s_waitcnt vmcnt(0), comes a
// whole group after the loads it covers:
//   fill group g (column footprints, neighbour loads, next group's pixels)
//   solve group g-1 from registers and LDS
//   wait; store group g-1's voxel values; group g becomes current.
template <int VPL, int NB, int MZ, int G>
__global__ __launch_bounds__(64) void k_icd_lds(KP P, const int2 *__restrict__ cells, int iter,
                                                int chunk, int seq, int count_mag) {
  extern __shared__ float2 s_ew[];  // [nv][RS] (e, w), lane windows; then H rows
  const int2 cl = cells[blockIdx.x];
  const int f = cl.x, z = cl.y;
  const int cell = f * P.N_z + z;
  const int lane = threadIdx.x;
  const int Nxy = P.N_x * P.N_y;
  const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;
  const int RW = P.N_r * MZ;  // entries of one view
  const int RS = RW + 1;      // LDS row stride, odd to spread banks
  const int rows = nv * RS;
  const int ldum = rows + lane * NB * MZ;  // the lane's own dummy window
  float *s_h = reinterpret_cast<float *>(s_ew + rows + 64 * NB * MZ);

  // the cell's sequence
  const int *list;
  int n0 = 0, cnt = 0;
  if (seq == 0) {
    list = P.ord + (size_t)cell * P.ord_cap;
    n0 = P.ord_off[(size_t)cell * (P.K + 1) + chunk];
    cnt = P.ord_off[(size_t)cell * (P.K + 1) + chunk + 1] - n0;
  } else if (seq == 1) {
    list = P.nh_ord + (size_t)cell * P.nh_cap;
    cnt = P.nh_ord_count[cell];
  } else {
    list = P.single;
    cnt = 1;
  }
  if (cnt <= 0) return;

  // stage the cell's (e, w) block and the frame's detector-response rows
  for (int q = 0; q < nv; q++) {
    const size_t g = sidx(P, z, P.view_idx[vb + q], 0, 0);
    for (int t = lane; t < RW; t += 64) s_ew[q * RS + t] = make_float2(P.e[g + t], P.w[g + t]);
    if (lane == 0) s_ew[q * RS + RW] = make_float2(0.0f, 0.0f);
  }
  for (int t = rows + lane; t < rows + 64 * NB * MZ; t += 64) s_ew[t] = make_float2(0.0f, 0.0f);
  for (int i = lane; i < nv * kHrStride; i += 64) {
    const int q = i / kHrStride, b = i - q * kHrStride;
    s_h[i] = P.hr[(size_t)P.view_idx[vb + q] * kHrStride + b];
  }
  __syncthreads();

  // the lane's views: detector coordinate in bins u = xx*Au - yy*Bu + Cu
  int qv[VPL];
  bool vok[VPL];
  double Au[VPL], Bu[VPL], Cu[VPL];
#pragma unroll
  for (int i = 0; i < VPL; i++) {
    const int q = lane + 64 * i;
    vok[i] = q < nv;
    qv[i] = vok[i] ? q : 0;
    const int v = P.view_idx[vb + qv[i]];
    const double c = P.cosv[v], sn = P.sinv[v];
    Au[i] = P.delta_xy * c * P.u_inv;
    Bu[i] = P.delta_xy * sn * P.u_inv;
    Cu[i] = ((P.x0 + 0.5 * P.delta_xy) * c - (P.y0 + 0.5 * P.delta_xy) * sn - P.R0) * P.u_inv;
  }
  const int m = P.mxy;
  // the lane's neighbour (lanes 0..26 spatial [dz][dy][dx], 27/28 temporal)
  int ndz = 0, ndy = 0, ndx = 0, ndt = 0;
  float ng = 0.0f, nsq = 1.0f, nisqp = 1.0f, nc = 1.0f;
  const bool nlane = lane < 29;
  if (nlane) {
    if (lane < 27) {
      ndz = lane / 9 - 1;
      ndy = (lane / 3) % 3 - 1;
      ndx = lane % 3 - 1;
      ng = P.sfilt[lane];
      nsq = P.sSq, nisqp = 1.0f / P.sSqp, nc = P.c_s;
    } else {
      ndt = lane == 27 ? -1 : 1;
      ng = P.tfilt;
      nsq = P.sTq, nisqp = 1.0f / P.sTqp, nc = P.c_t;
    }
  }
  const bool inplane = lane < 27 && ndz == 0 && lane != 13;
  const int noff = ndy * 65536 + ndx;  // packed-coordinate offset of the neighbour
  const int zz = z + ndz, ff = f + ndt;
  const bool zok = nlane && zz >= 0 && zz < P.N_z && ff >= 0 && ff < P.N_time;
  const int zc = min(max(zz, 0), P.N_z - 1), fc = min(max(ff, 0), P.N_time - 1);
  const int nplane = (fc * P.N_z + zc) * Nxy;
  const int cbase = cell * Nxy;
  const float two_c = 2.0f * nc;

  // current group (A) and the group in flight (B)
  // Neighbour values that do not exist (outside the object, lanes 29..63)
  // are NaN: their loads read the NaN word the object buffer carries past
  // its end, and a patch not yet applied is NaN too.  Keeping validity in the
  // values keeps per-lane booleans out of scalar registers.
  int apix[G], abase[G][VPL], anid[G];
  int bpix[G], bbase[G][VPL], bnid[G];
  float aa[G][VPL][NB], ba[G][VPL][NB];
  float anb[G], bnb[G], bnv[G];

  auto fill = [&](int k, int pix) {
    bpix[k] = pix;
    const bool valid = pix != kNoPix;
    const int yy = valid ? (pix >> 16) : 0, xx = valid ? (pix & 0xffff) : 0;
    const double xd = (double)xx, yd = (double)yy;
#pragma unroll
    for (int i = 0; i < VPL; i++) {
      // calcAMatrixColumnforAngle (XT_AMatrix.c:50-110) in bin units: bins
      // floor(u) - m .. floor(u) + m, value H at |u - j - 0.5| delta_r/OffsetR
      const double u = fma(xd, Au[i], fma(-yd, Bu[i], Cu[i]));
      // bitwise tests throughout: short-circuit && would branch around the
      // LDS reads below and serialise their latency
      const bool inr = valid & vok[i] & (u + m >= 0.0) & (u - m <= P.u_max);
      const int lo = (int)floor(u) - m;
      const int loc = max(lo, 0), hic = min(lo + 2 * m, P.N_r - 1);
      const int w0 = min(loc, P.N_r - NB);  // window of NB bins inside the row
      bbase[k][i] = inr ? qv[i] * RS + w0 * MZ : ldum;
      const double d = u - ((double)w0 + 0.5);
      const float *H = s_h + qv[i] * kHrStride;
#pragma unroll
      for (int ub = 0; ub < NB; ub++) {
        const double t = fabs(d - (double)ub) * P.u_kk;
        const int bin = (int)t;
        const bool ok = inr & (w0 + ub >= loc) & (w0 + ub <= hic) & (bin < kRespBins);
        const int bb = ok ? bin : 0;
        const float fr = (float)(t - (double)bin);
        const float h0 = H[bb], h1 = H[bb + 1];
        // a multiply rather than a select, so the reads are not moved under a
        // branch
        ba[k][i][ub] = fmaf(fr, h1 - h0, h0) * (ok ? 1.0f : 0.0f);
      }
    }
    // neighbour value (the NaN word when it does not exist)
    const int jj = yy + ndy, kk = xx + ndx;
    const bool nok = valid & zok & (jj >= 0) & (jj < P.N_y) & (kk >= 0) & (kk < P.N_x);
    bnb[k] = P.obj[(unsigned)(nok ? nplane + jj * P.N_x + kk : P.nobj + 64)];
    bnid[k] = (valid & inplane) ? pix + noff : -2;
    bnv[k] = __int_as_float(0x7fc00000);
  };

  double voxmag = 0.0;
  unsigned long long nupd = 0, nzero = 0;
  int nent = 0;

  const int ngroups = (cnt + G - 1) / G;
  // pixels of group 0
  int raw = list[n0 + min(lane, cnt - 1)];
  wait_vm_all();
  int pnext[G];
#pragma unroll
  for (int k = 0; k < G; k++) pnext[k] = k < cnt ? __builtin_amdgcn_readlane(raw, k) : kNoPix;
#pragma unroll
  for (int k = 0; k < G; k++) {
    apix[k] = kNoPix;
    anid[k] = -2;
    anb[k] = __int_as_float(0x7fc00000);
#pragma unroll
    for (int i = 0; i < VPL; i++) {
      abase[k][i] = ldum;
#pragma unroll
      for (int ub = 0; ub < NB; ub++) aa[k][i][ub] = 0.0f;
    }
  }

  for (int g = 0; g <= ngroups; g++) {
    // group g in flight
#pragma unroll
    for (int k = 0; k < G; k++) fill(k, pnext[k]);
    raw = list[n0 + min((g + 1) * G + lane, cnt - 1)];

    // solve group g - 1 (empty slots run as no-ops: no bins, dx = 0)
    float rx = 0.0f, rm = 0.0f;
    int ridx = P.nobj + lane;
#pragma unroll
    for (int s = 0; s < G; s++) {
      const int pix = apix[s];
      const bool valid = pix != kNoPix;
      // theta1 / theta2 (XT_VoxUpdate.c:54-83) from the LDS block
      float2 ew[VPL][NB][MZ];
#pragma unroll
      for (int i = 0; i < VPL; i++)
#pragma unroll
        for (int ub = 0; ub < NB; ub++)
#pragma unroll
          for (int mm = 0; mm < MZ; mm++) ew[i][ub][mm] = s_ew[abase[s][i] + ub * MZ + mm];
      float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
      for (int i = 0; i < VPL; i++)
#pragma unroll
        for (int ub = 0; ub < NB; ub++) {
          const float a = aa[s][i][ub];
#pragma unroll
          for (int mm = 0; mm < MZ; mm++) {
            const float e = ew[i][ub][mm].x, w = ew[i][ub][mm].y;
            const float cw_out =
                P.Td * __builtin_amdgcn_sqrtf(w) * __builtin_amdgcn_rcpf(fabsf(e));
            const float cw = huber_sel(P, e, w) ? w : cw_out;
            const float ac = a * cw;  // cw is finite: e = 0 is always selected
            s2 = fmaf(a, ac, s2);
            s1 = fmaf(-e, ac, s1);
            nent += a > 0.0f ? 1 : 0;
          }
        }
      // prior surrogate (XT_ICD_update.c:134-183): qGGMRF b(d) for p = 1.2,
      // q = 2, (2 - 0.8 t1/t3) / (t3 sigma^q) = (2c + 1.2 t1) / (t3^2 sigma^q)
      const float nbv = anb[s];
      const bool vld = (nbv == nbv) & (lane != 13);
      const float V = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nbv), 13));
      const bool anypos = __ballot((lane != 13) & (nbv > 0.0f)) != 0ull;
      const float dlt = vld ? V - nbv : 0.0f;
      const float t1 = __builtin_amdgcn_exp2f(0.8f * __builtin_amdgcn_logf(fabsf(dlt))) * nisqp;
      const float t3 = nc + t1;
      const float b = fmaf(1.2f, t1, two_c) * __builtin_amdgcn_rcpf(t3 * t3 * nsq);
      const float gb = (vld ? ng : 0.0f) * b;
      // sums over the wavefront: theta1 in double (it cancels), the rest in float
      const float pb = wave_total(gb);
      const float pbx = wave_total(-gb * dlt);
      const float th1 = (float)wave_total_d((double)s1);
      const float th2 = wave_total(s2);
      // zero skipping (XT_VoxUpdate.c:236-259)
      const bool skip = (V == 0.0f) && iter > 1 && !anypos;
      const float step = P.alpha * (pbx - th1) * __builtin_amdgcn_rcpf(pb + th2);
      const float xn = (skip || !valid) ? V : V + step;
      // an empty slot's V is the NaN of a missing neighbour: its dx must be
      // 0 so that the unchanged e values it writes back stay finite
      const float dx = valid ? xn - V : 0.0f;
      // e -= A dx in LDS (entries outside the column and empty slots write
      // their unchanged value back)
#pragma unroll
      for (int i = 0; i < VPL; i++)
#pragma unroll
        for (int ub = 0; ub < NB; ub++)
#pragma unroll
          for (int mm = 0; mm < MZ; mm++)
            s_ew[abase[s][i] + ub * MZ + mm].x = fmaf(-aa[s][i][ub], dx, ew[i][ub][mm].x);
      // in-plane neighbour values of the voxels still to come
#pragma unroll
      for (int o = s + 1; o < G; o++) anb[o] = anid[o] == pix ? xn : anb[o];
#pragma unroll
      for (int o = 0; o < G; o++) bnv[o] = bnid[o] == pix ? xn : bnv[o];
      // this voxel's value and map entry, stored by lane s after the wait
      const int lin0 = cbase + (valid ? pix >> 16 : 0) * P.N_x + (valid ? pix & 0xffff : 0);
      const int lin = valid ? lin0 : P.nobj + s;
      rx = lane == s ? xn : rx;
      rm = lane == s ? (skip ? 0.0f : fabsf(dx)) : rm;
      ridx = lane == s ? lin : ridx;
      voxmag += (valid && count_mag && !skip) ? fabs((double)xn) : 0.0;
      nupd += (valid && !skip) ? 1 : 0;
      nzero += (valid && skip) ? 1 : 0;
    }

    wait_vm_all();
    P.obj[(unsigned)ridx] = rx;
    P.mag[(unsigned)ridx] = rm;

    // group g becomes current
#pragma unroll
    for (int k = 0; k < G; k++) {
      apix[k] = bpix[k];
      anid[k] = bnid[k];
      anb[k] = bnv[k] == bnv[k] ? bnv[k] : bnb[k];
#pragma unroll
      for (int i = 0; i < VPL; i++) {
        abase[k][i] = bbase[k][i];
#pragma unroll
        for (int ub = 0; ub < NB; ub++) aa[k][i][ub] = ba[k][i][ub];
      }
    }
#pragma unroll
    for (int k = 0; k < G; k++)
      pnext[k] = (g + 1) * G + k < cnt ? __builtin_amdgcn_readlane(raw, k) : kNoPix;
  }

  // e back to global memory
  __syncthreads();
  for (int q = 0; q < nv; q++) {
    const size_t gq = sidx(P, z, P.view_idx[vb + q], 0, 0);
    for (int t = lane; t < RW; t += 64) P.e[gq + t] = s_ew[q * RS + t].x;
  }
  const double ne = wave_sum_d((double)nent);
  if (lane == 0 && seq != 2) {
    P.cell_voxmag[cell] += voxmag;
    P.cell_upd[cell] += nupd;
    P.cell_zero[cell] += nzero;
    P.cell_ent[cell] += (unsigned long long)ne;
  }
}

// ---------------------------------------------------------------------------
// Multi-wavefront form: one workgroup per cell, split by role.
//
// Waves 0..nth-1 ("theta waves") each own vpw views of the cell, one A-matrix
// entry (view, bin, row) per lane for each of VPL views, and the LDS rows of
// those views: no other wave reads or writes them, so the in-order LDS
// argument above still gives the sequential sweep.  Where the views allow,
// the workgroup has at most four waves, one per SIMD (two views per lane
// for 64 views of 3 bins: 2 theta waves).  The last wave ("prior wave")
// owns the object side:
// neighbour prefetch one group ahead (with patching, as above), the qGGMRF
// prior, and the object / map stores.  Per voxel every wave publishes its
// partial sums in LDS and meets the others at one barrier; then every wave
// evaluates the same closed-form step from the same sums in the same order,
// so all waves hold the identical dx.
// ---------------------------------------------------------------------------
struct StepIn {
  double t1[2][16];  // per theta wave, double-buffered by voxel parity
  float t2[2][16];
  float pr[2][4];    // pb, pbx, V, skip
  unsigned long long nent[16];
};

// How the theta sums are formed (tb_options.theta_fp): in every mode each
// lane's products are FP32.  TF 0: theta1 summed in FP64 over lanes and
// waves, theta2 in FP32.  TF 1: both in FP32 with compensated (TwoSum)
// accumulation over lanes; the waves' compensated totals s + c are exact in
// FP64 and added there.  TF 2: both in plain FP32 throughout.
template <int MAXTH, int TF = 0>
__device__ __forceinline__ float solve_step(const KP &P, const StepIn &S, int par, int nth,
                                            float *V_out, bool *skip_out) {
#pragma clang fp contract(off)
  // every slot is read at once (one LDS latency) and added as a fixed
  // pairwise tree over the MAXTH slots; slots of absent waves hold zero
  using T1T = typename std::conditional<TF == 2, float, double>::type;
  T1T a1[MAXTH];
  float a2[MAXTH];
#pragma unroll
  for (int w = 0; w < MAXTH; w++) {
    a1[w] = (T1T)S.t1[par][w];
    a2[w] = S.t2[par][w];
  }
#pragma unroll
  for (int m = MAXTH; m > 1; m = (m + 1) / 2) {
    const int h = (m + 1) / 2;
#pragma unroll
    for (int w = 0; w + h < m; w++) {
      a1[w] = a1[w] + a1[w + h];
      a2[w] = a2[w] + a2[w + h];
    }
  }
  const double T1 = (double)a1[0];
  const float T2 = a2[0];
  (void)nth;
  const float pb = S.pr[par][0], pbx = S.pr[par][1], V = S.pr[par][2];
  const bool skip = S.pr[par][3] != 0.0f;
  const float th1 = (float)T1;
  const float num = P.alpha * (pbx - th1);
  const float step = num * __builtin_amdgcn_rcpf(pb + T2);
  *V_out = V;
  *skip_out = skip;
  return skip ? V : V + step;
}

// A-matrix column of a voxel in one view (calcAMatrixColumnforAngle,
// XT_AMatrix.c:50-110) in bin units: u = (r - R0) / delta_r of the voxel
// centre is xx * Au - yy * Bu + Cu; the column covers bins
// floor(u) - m .. floor(u) + m (m = delta_xy / delta_r) clipped to the
// detector, and its value at bin j is the lerp of H_r at |u - (j + 0.5)| in
// units of OffsetR.  Contraction is off, so every caller (the column table,
// the kernels that evaluate columns on the fly) gets the same bits.
__device__ __forceinline__ void col_coef(const KP &P, int v, double &Au, double &Bu,
                                         double &Cu) {
#pragma clang fp contract(off)
  const double cs = P.cosv[v], sn = P.sinv[v];
  Au = P.delta_xy * cs * P.u_inv;
  Bu = P.delta_xy * sn * P.u_inv;
  Cu = ((P.x0 + 0.5 * P.delta_xy) * cs - (P.y0 + 0.5 * P.delta_xy) * sn - P.R0) * P.u_inv;
}
__device__ __forceinline__ double col_u(int xx, int yy, double Au, double Bu, double Cu) {
  return fma((double)xx, Au, fma(-(double)yy, Bu, Cu));
}
// Window start w0 = min(first bin, N_r - nb) of a column of at most nb bins,
// or -1 when the voxel misses the detector; loc / hic are the clipped first
// and last bins.
__device__ __forceinline__ int col_window(const KP &P, double u, int nb, int &loc, int &hic) {
  const int m = P.mxy;
  const bool on = (u + m >= 0.0) & (u - m <= P.u_max);
  const int lo = (int)floor(u) - m;
  loc = max(lo, 0);
  hic = min(lo + 2 * m, P.N_r - 1);
  return on ? min(loc, P.N_r - nb) : -1;
}
__device__ __forceinline__ float col_value(const KP &P, const float *H, double u, int j, int loc,
                                           int hic) {
#pragma clang fp contract(off)
  const double t = fabs(u - ((double)j + 0.5)) * P.u_kk;
  const int bin = (int)t;
  const bool ok = (j >= loc) & (j <= hic) & (bin < kRespBins);
  const int bb = ok ? bin : 0;
  const float fr = (float)(t - (double)bin);
  const float h0 = H[bb], h1 = H[bb + 1];
  return fmaf(fr, h1 - h0, h0) * (ok ? 1.0f : 0.0f);
}

// Column table of the multi-wavefront kernel: for every (frame, pixel, view
// slot) the A-matrix values at bins w0 .. w0 + NB - 1 of that view and w0,
// or w0 = -1 when the voxel misses the detector in that view or the slot
// holds no view.  Records of col_rec(NB) floats (NB + 1 rounded up to a
// multiple of 4: 4, 8, 12 for NB = 3, 5, 9).
template <int NB>
__global__ __launch_bounds__(256) void k_build_columns(KP P) {
  constexpr int REC = col_rec(NB);
  const int Nxy = P.N_x * P.N_y;
  const size_t nrec = (size_t)P.N_time * Nxy * P.colt_nv;
  float *out = const_cast<float *>(P.colt);
  for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < nrec;
       idx += (size_t)gridDim.x * blockDim.x) {
    const int q = (int)(idx % P.colt_nv);
    const size_t fp = idx / P.colt_nv;
    const int lin = (int)(fp % Nxy), f = (int)(fp / Nxy);
    const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;
    float a[NB];
    int w0 = -1;
#pragma unroll
    for (int ub = 0; ub < NB; ub++) a[ub] = 0.0f;
    if (q < nv) {
      const int v = P.view_idx[vb + q];
      double Au, Bu, Cu;
      col_coef(P, v, Au, Bu, Cu);
      const int yy = lin / P.N_x, xx = lin - yy * P.N_x;
      const double u = col_u(xx, yy, Au, Bu, Cu);
      int loc, hic;
      w0 = col_window(P, u, NB, loc, hic);
      if (w0 >= 0) {
        const float *H = P.hr + (size_t)v * kHrStride;
#pragma unroll
        for (int ub = 0; ub < NB; ub++) a[ub] = col_value(P, H, u, w0 + ub, loc, hic);
      }
    }
    float *r = out + idx * REC;
#pragma unroll
    for (int ub = 0; ub < NB; ub++) r[ub] = a[ub];
    r[NB] = __int_as_float(w0);
#pragma unroll
    for (int k = NB + 1; k < REC; k++) r[k] = 0.0f;
  }
}

template __global__ void k_build_columns<3>(KP);
template __global__ void k_build_columns<5>(KP);
template __global__ void k_build_columns<9>(KP);
template __global__ void k_build_columns<17>(KP);

// The prior wave of the multi-wavefront kernels: the object side of the
// voxel sequence (neighbour prefetch one group ahead with in-flight patching,
// the qGGMRF prior surrogate of XT_ICD_update.c:134-183, zero-skip,
// XT_VoxUpdate.c:236-259, the object and update-map stores), meeting the
// theta waves at one barrier per voxel.
// PAD: the theta waves run whole groups of G voxels, the last one padded
// with empty slots; the prior wave meets them at every slot, publishing a
// skipped step (dx = 0) for an empty one.
template <int MAXTH, bool PAD = false>
__device__ __forceinline__ void prior_wave(const KP &P, StepIn &S, const int *list, int n0,
                                           int cnt, int f, int z, int iter, int seq,
                                           int count_mag, int nth, int lane) {
  constexpr int G = 8;
  const int cell = f * P.N_z + z;
  const int Nxy = P.N_x * P.N_y;

  int ndz = 0, ndy = 0, ndx = 0, ndt = 0;
  float ng = 0.0f, nsq = 1.0f, nisqp = 1.0f, nc = 1.0f;
  const bool nlane = lane < 29;
  if (nlane) {
    if (lane < 27) {
      ndz = lane / 9 - 1;
      ndy = (lane / 3) % 3 - 1;
      ndx = lane % 3 - 1;
      ng = P.sfilt[lane];
      nsq = P.sSq, nisqp = 1.0f / P.sSqp, nc = P.c_s;
    } else {
      ndt = lane == 27 ? -1 : 1;
      ng = P.tfilt;
      nsq = P.sTq, nisqp = 1.0f / P.sTqp, nc = P.c_t;
    }
  }
  const bool inplane = lane < 27 && ndz == 0 && lane != 13;
  const int noff = ndy * 65536 + ndx;
  const int zz = z + ndz, ff = f + ndt;
  const bool zok = nlane && zz >= 0 && zz < P.N_z && ff >= 0 && ff < P.N_time;
  const int zc = min(max(zz, 0), P.N_z - 1), fc = min(max(ff, 0), P.N_time - 1);
  const int nplane = (fc * P.N_z + zc) * Nxy;
  const int cbase = cell * Nxy;
  const float two_c = 2.0f * nc;

  int apix[G], anid[G], bpix[G], bnid[G];
  float anb[G], bnb[G], bnv[G];
  // packed pixels of the pending voxels, one per lane (A in lanes 0..G-1,
  // B in lanes G..2G-1): one vector test per update finds the rare case of
  // an in-plane neighbour among them
  int pend = kNoPix;
  auto fill = [&](int k, int pix) {
    bpix[k] = pix;
    const bool valid = pix != kNoPix;
    const int yy = valid ? (pix >> 16) : 0, xx = valid ? (pix & 0xffff) : 0;
    const int jj = yy + ndy, kk = xx + ndx;
    const bool nok = valid & zok & (jj >= 0) & (jj < P.N_y) & (kk >= 0) & (kk < P.N_x);
    bnb[k] = P.obj[(unsigned)(nok ? nplane + jj * P.N_x + kk : P.nobj + 64)];
    bnid[k] = (valid & inplane) ? pix + noff : -2;
    bnv[k] = __int_as_float(0x7fc00000);
  };
  double vmacc = 0.0;  // lane k: magnitudes of the voxels at group position k
  unsigned long long nupd = 0, nzero = 0;
  const int ngroups = (cnt + G - 1) / G;
  int raw = list[n0 + min(lane, cnt - 1)];
  wait_vm_all();
  int pnext[G];
#pragma unroll
  for (int k = 0; k < G; k++) pnext[k] = k < cnt ? __builtin_amdgcn_readlane(raw, k) : kNoPix;
#pragma unroll
  for (int k = 0; k < G; k++) {
    apix[k] = kNoPix;
    anid[k] = -2;
    anb[k] = __int_as_float(0x7fc00000);
  }
  for (int g = 0; g <= ngroups; g++) {
#pragma unroll
    for (int k = 0; k < G; k++) fill(k, pnext[k]);
    raw = list[n0 + min((g + 1) * G + lane, cnt - 1)];
#pragma unroll
    for (int k = 0; k < G; k++) pend = lane == G + k ? bpix[k] : pend;
    float rx = 0.0f, rm = 0.0f;
    int ridx = P.nobj + lane;
    int rsk = -1;  // lane k: voxel k of the group skipped (1), updated (0), absent (-1)
#pragma unroll
    for (int s = 0; s < G; s++) {
      const int pix = apix[s];
      const int n = (g - 1) * G + s;
      const int par = n & 1;
      if (pix == kNoPix) {  // only past the end of the sequence
        if (PAD && g > 0) {
          if (lane == 0) {
            S.pr[par][0] = 1.0f;
            S.pr[par][1] = 0.0f;
            S.pr[par][2] = 0.0f;
            S.pr[par][3] = 1.0f;
          }
          __syncthreads();
        }
        continue;
      }
      const float nbv = anb[s];
      const bool vld = (nbv == nbv) & (lane != 13);
      const float V = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nbv), 13));
      const bool anypos = __ballot((lane != 13) & (nbv > 0.0f)) != 0ull;
      const float dlt = vld ? V - nbv : 0.0f;
      const float t1 =
          __builtin_amdgcn_exp2f(0.8f * __builtin_amdgcn_logf(fabsf(dlt))) * nisqp;
      const float t3 = nc + t1;
      const float b = fmaf(1.2f, t1, two_c) * __builtin_amdgcn_rcpf(t3 * t3 * nsq);
      const float gb = (vld ? ng : 0.0f) * b;
      float pb = gb, pbx = -gb * dlt;
      lo32_total_ff(pb, pbx);  // neighbour lanes are 0..28
      const bool skip = (V == 0.0f) && iter > 1 && !anypos;
      if (lane == 0) {
        S.pr[par][0] = pb;
        S.pr[par][1] = pbx;
        S.pr[par][2] = V;
        S.pr[par][3] = skip ? 1.0f : 0.0f;
      }
      __syncthreads();
      float Vs;
      bool sk;
      const float xn = solve_step<MAXTH>(P, S, par, nth, &Vs, &sk);
      const float dx = xn - V;
      // (dy + 1) * 65536 + (dx + 1) with |dy|, |dx| <= 1 for a neighbour
      const unsigned dd = (unsigned)(pend - pix + 65537);
      const bool adj = (pend != kNoPix) & (dd != 65537u) & ((dd & 0xffffu) <= 2u) &
                       (dd <= 131074u);
      if (__ballot(adj) != 0ull) {
#pragma unroll
        for (int o = s + 1; o < G; o++) anb[o] = anid[o] == pix ? xn : anb[o];
#pragma unroll
        for (int o = 0; o < G; o++) bnv[o] = bnid[o] == pix ? xn : bnv[o];
      }
      const int lin = cbase + (pix >> 16) * P.N_x + (pix & 0xffff);
      rx = lane == s ? xn : rx;
      rm = lane == s ? (skip ? 0.0f : fabsf(dx)) : rm;
      ridx = lane == s ? lin : ridx;
      rsk = lane == s ? (skip ? 1 : 0) : rsk;
    }
    // the group's update counts and |x| sums (updateVoxels_Atten's return
    // value, XT_VoxUpdate.c:263-281), lane by lane
    vmacc += (count_mag && rsk == 0) ? fabs((double)rx) : 0.0;
    nupd += (unsigned long long)__popcll(__ballot(rsk == 0));
    nzero += (unsigned long long)__popcll(__ballot(rsk == 1));
    wait_vm_all();
    P.obj[(unsigned)ridx] = rx;
    P.mag[(unsigned)ridx] = rm;
#pragma unroll
    for (int k = 0; k < G; k++) {
      apix[k] = bpix[k];
      anid[k] = bnid[k];
      anb[k] = bnv[k] == bnv[k] ? bnv[k] : bnb[k];
    }
    // B moves to lanes 0..G-1
    pend = __shfl(pend, (lane + G) & 63);
    pend = lane < G ? pend : kNoPix;
#pragma unroll
    for (int k = 0; k < G; k++)
      pnext[k] = (g + 1) * G + k < cnt ? __builtin_amdgcn_readlane(raw, k) : kNoPix;
  }
  const double voxmag = wave_sum_d(vmacc);
  if (lane == 0 && seq != 2) {
    P.cell_voxmag[cell] += voxmag;
    P.cell_upd[cell] += nupd;
    P.cell_zero[cell] += nzero;
  }
}

template <int NB, int MZ, int VPL, int MAXTH>
__device__ __forceinline__ void k_icd_mw_body(const KP &P, const int2 cl, int iter,
                                             int chunk, int seq, int count_mag, int vpw) {
  constexpr int EPV = NB * MZ;  // entries of one view
  constexpr int VW = 64 / EPV;  // views of one lane layer
  constexpr int G = 8;          // voxels per prefetch group
  constexpr int REC = col_rec(NB);
  extern __shared__ float2 s_ew[];  // [nv][RS] (e, w), then one dummy pair per lane and layer
  __shared__ StepIn S;
  const int f = cl.x, z = cl.y;
  const int cell = f * P.N_z + z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nth = (int)(blockDim.x >> 6) - 1;
  const int Nxy = P.N_x * P.N_y;
  const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;
  const int RW = P.N_r * MZ;
  const int RS = RW + 1;
  const int rows = nv * RS;

  const int *list;
  int n0 = 0, cnt = 0;
  if (seq == 0) {
    list = P.ord + (size_t)cell * P.ord_cap;
    n0 = P.ord_off[(size_t)cell * (P.K + 1) + chunk];
    cnt = P.ord_off[(size_t)cell * (P.K + 1) + chunk + 1] - n0;
  } else if (seq == 1) {
    list = P.nh_ord + (size_t)cell * P.nh_cap;
    cnt = P.nh_ord_count[cell];
  } else {
    list = P.single;
    cnt = 1;
  }
  if (cnt <= 0) return;

  // stage the (e, w) block (whole rows per wave)
  const int nwaves = nth + 1;
  for (int q = wave; q < nv; q += nwaves) {
    const size_t g = sidx(P, z, P.view_idx[vb + q], 0, 0);
    for (int t = lane; t < RW; t += 64) s_ew[q * RS + t] = make_float2(P.e[g + t], P.w[g + t]);
    if (lane == 0) s_ew[q * RS + RW] = make_float2(0.0f, 0.0f);
  }
  for (int t = threadIdx.x; t < nth * VPL * 64; t += blockDim.x)
    s_ew[rows + t] = make_float2(0.f, 0.f);
  if (threadIdx.x < 32) {
    (&S.t1[0][0])[threadIdx.x] = 0.0;
    (&S.t2[0][0])[threadIdx.x] = 0.0f;
  }
  __syncthreads();

  if (wave < nth) {
    // ---------------- theta wave ----------------
    // Lane layer k holds the entry (view q_k, bin w0 + ub, row mm) of the
    // current voxel; its A value and w0 come from the column table,
    // prefetched one group of G voxels ahead.  The next voxel's (e, w) is
    // read before this voxel's e is written and corrected afterwards by
    // a(this voxel, that bin) * dx: the same fma the lane owning that entry
    // stores, so no bits change.
    const int lq = lane / EPV, ent = lane - lq * EPV;
    const int ub = ent / MZ, mm = ent - ub * MZ;
    bool active[VPL];
    int qv[VPL], dummy[VPL];
#pragma unroll
    for (int k = 0; k < VPL; k++) {
      const int qq = k * VW + lq, q = wave * vpw + qq;
      active[k] = (lq < VW) & (qq < vpw) & (q < nv);
      qv[k] = active[k] ? q : 0;
      dummy[k] = rows + (wave * VPL + k) * 64 + lane;
    }
    const float4 *colt = reinterpret_cast<const float4 *>(P.colt);
    const size_t fbase = (size_t)f * Nxy;
    struct Rec {
      float a[NB];
      int w0;
    };
    auto load = [&](int pix, int k) -> Rec {
      const bool valid = pix != kNoPix;
      const int lin = valid ? (pix >> 16) * P.N_x + (pix & 0xffff) : 0;
      const size_t r = ((fbase + lin) * P.colt_nv + qv[k]) * (REC / 4);
      Rec o;
      const float4 r0 = colt[r];
      if (REC == 4) {
        o.a[0] = r0.x;
        o.a[1] = r0.y;
        o.a[2] = r0.z;
        o.w0 = __float_as_int(r0.w);
      } else {
        const float4 r1 = colt[r + 1];
        const float v[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
#pragma unroll
        for (int t = 0; t < NB; t++) o.a[t] = v[t];
        o.w0 = __float_as_int(v[NB]);
      }
      // no arithmetic on the loaded registers here: any would make the
      // compiler wait for the load at once; validity is applied at use
      return o;
    };
    auto pick = [&](const Rec &r, int j) -> float {
      float v = 0.0f;
#pragma unroll
      for (int t = 0; t < NB; t++) v = j == t ? r.a[t] : v;
      return v;
    };
    // w0 of a record as this lane uses it: -1 for an empty slot or a lane
    // without a view
    auto w0_of = [&](const Rec &r, int pix, int k) -> int {
      return (active[k] & (pix != kNoPix)) ? r.w0 : -1;
    };
    auto entry_at = [&](int w0, int k) -> int {
      return w0 >= 0 ? qv[k] * RS + (w0 + ub) * MZ + mm : dummy[k];
    };

    int apix[G], bpix[G];
    Rec arec[G][VPL], brec[G][VPL];
    const int ngroups = (cnt + G - 1) / G;
    int raw = list[n0 + min(lane, cnt - 1)];
    int pnext[G];
#pragma unroll
    for (int j = 0; j < G; j++) pnext[j] = j < cnt ? __builtin_amdgcn_readlane(raw, j) : kNoPix;
#pragma unroll
    for (int j = 0; j < G; j++) {
      apix[j] = kNoPix;
#pragma unroll
      for (int k = 0; k < VPL; k++) {
        arec[j][k].w0 = -1;
#pragma unroll
        for (int t = 0; t < NB; t++) arec[j][k].a[t] = 0.0f;
      }
    }
    // current voxel's entries: A value, (e, w) with e already corrected for
    // the previous voxel's update, LDS slot, window start and record
    float ca[VPL], ce[VPL], cwv[VPL];
    int cat[VPL], cw0[VPL];
    Rec crec[VPL];
#pragma unroll
    for (int k = 0; k < VPL; k++) {
      ca[k] = 0.0f;
      ce[k] = 0.0f;
      cwv[k] = 0.0f;
      cat[k] = dummy[k];
      cw0[k] = -1;
      crec[k] = arec[0][k];
    }
    unsigned nent = 0;  // entries with a > 0 (at most 2 x 65535 x 3 per lane)
    for (int g = 0; g <= ngroups; g++) {
#pragma unroll
      for (int j = 0; j < G; j++) {
        bpix[j] = pnext[j];
#pragma unroll
        for (int k = 0; k < VPL; k++) brec[j][k] = load(pnext[j], k);
      }
      raw = list[n0 + min((g + 1) * G + lane, cnt - 1)];
#pragma unroll
      for (int s = 0; s < G; s++) {
        const int pix = apix[s];
        if (pix == kNoPix) continue;  // only past the end of the sequence
        const int n = (g - 1) * G + s;
        const int par = n & 1;
        // this voxel's partial sums (XT_VoxUpdate.c:61-92)
        double t1;
        float t2;
        // generalized-Huber weights (XT_VoxUpdate.c:66-80): w inside the
        // threshold, T delta sqrt(w) / |e| outside; the outer weight is
        // evaluated only when some lane of the wave needs it
        float cwk[VPL];
        bool sel[VPL], allsel = true;
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          sel[k] = huber_sel(P, ce[k], cwv[k]);
          cwk[k] = cwv[k];
          allsel &= sel[k];
        }
        if (__builtin_expect(__ballot(!allsel) != 0ull, 0)) {
#pragma unroll
          for (int k = 0; k < VPL; k++) {
            const float cw_out =
                P.Td * __builtin_amdgcn_sqrtf(cwv[k]) * __builtin_amdgcn_rcpf(fabsf(ce[k]));
            cwk[k] = sel[k] ? cwv[k] : cw_out;
          }
        }
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          const float cw = cwk[k];
          const float ac = ca[k] * cw;
          const double p1 = (double)(-ce[k] * ac);
          const float p2 = ca[k] * ac;
          t1 = k == 0 ? p1 : t1 + p1;
          t2 = k == 0 ? p2 : t2 + p2;
          nent += ca[k] > 0.0f ? 1u : 0u;
        }
        // the next voxel's entries and their (e, w), read before this
        // voxel's e is written; independent of the sums, so they fill the
        // reduction's latency
        const int npix = s + 1 < G ? apix[s + 1] : bpix[0];
        int nw0[VPL], nat[VPL];
        float2 ewn[VPL];
        float cn_prev[VPL], na[VPL];
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          const Rec &nr = s + 1 < G ? arec[s + 1][k] : brec[0][k];
          nw0[k] = w0_of(nr, npix, k);
          nat[k] = entry_at(nw0[k], k);
          ewn[k] = s_ew[nat[k]];
          const int d = nw0[k] + ub - cw0[k];  // the next voxel's bin in this voxel's window
          const float pd = pick(crec[k], d), pu = pick(nr, ub);
          cn_prev[k] = ((nw0[k] >= 0) & (cw0[k] >= 0)) ? pd : 0.0f;
          na[k] = nw0[k] >= 0 ? pu : 0.0f;
          // computed here, not sunk past the barrier onto the critical path
          asm volatile("" ::"v"(cn_prev[k]), "v"(na[k]), "v"(nat[k]));
        }
        wave_total_df(t1, t2);
        if (lane == 0) {
          S.t1[par][wave] = t1;
          S.t2[par][wave] = t2;
        }
        __syncthreads();
        float V;
        bool skip;
        const float xn = solve_step<MAXTH>(P, S, par, nth, &V, &skip);
        const float dx = xn - V;
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          s_ew[cat[k]].x = fmaf(-ca[k], dx, ce[k]);
          ce[k] = fmaf(-cn_prev[k], dx, ewn[k].x);
          cwv[k] = ewn[k].y;
          crec[k] = s + 1 < G ? arec[s + 1][k] : brec[0][k];
          cw0[k] = nw0[k];
          ca[k] = na[k];
          cat[k] = nat[k];
        }
      }
#pragma unroll
      for (int j = 0; j < G; j++) {
        apix[j] = bpix[j];
#pragma unroll
        for (int k = 0; k < VPL; k++) arec[j][k] = brec[j][k];
      }
#pragma unroll
      for (int j = 0; j < G; j++)
        pnext[j] = (g + 1) * G + j < cnt ? __builtin_amdgcn_readlane(raw, j) : kNoPix;
      if (g == 0) {
        // the first voxel's entries
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          crec[k] = arec[0][k];
          cw0[k] = w0_of(crec[k], apix[0], k);
          ca[k] = cw0[k] >= 0 ? pick(crec[k], ub) : 0.0f;
          cat[k] = entry_at(cw0[k], k);
          const float2 ew0 = s_ew[cat[k]];
          ce[k] = ew0.x;
          cwv[k] = ew0.y;
        }
      }
    }
    const double ne = wave_sum_d((double)nent);
    if (lane == 0) S.nent[wave] = (unsigned long long)ne;
  } else {
    // ---------------- prior wave ----------------
    prior_wave<MAXTH>(P, S, list, n0, cnt, f, z, iter, seq, count_mag, nth, lane);
  }

  // e back to global memory; entry count
  __syncthreads();
  for (int q = wave; q < nv; q += nwaves) {
    const size_t gq = sidx(P, z, P.view_idx[vb + q], 0, 0);
    for (int t = lane; t < RW; t += 64) P.e[gq + t] = s_ew[q * RS + t].x;
  }
  if (threadIdx.x == 0 && seq != 2) {
    unsigned long long ne = 0;
    for (int w = 0; w < nth; w++) ne += S.nent[w];
    P.cell_ent[cell] += ne;
  }
}

// one problem: the stage's cell list
template <int NB, int MZ, int VPL, int MAXTH>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_mw(KP P, const int2 *__restrict__ cells,
    int iter, int chunk, int seq, int count_mag, int vpw) {
  k_icd_mw_body<NB, MZ, VPL, MAXTH>(P, cells[blockIdx.x], iter, chunk, seq, count_mag, vpw);
}
// many problems: cells (frame, slice, problem, last argument), parameters
// from the problems' table
template <int NB, int MZ, int VPL, int MAXTH>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_mw_m(const KP *__restrict__ tab,
    const int4 *__restrict__ cells, int iter, int chunk, int seq, int count_mag) {
  const int4 c = cells[blockIdx.x];
  k_icd_mw_body<NB, MZ, VPL, MAXTH>(tab[c.z], make_int2(c.x, c.y), iter, chunk, seq, count_mag, c.w);
}


// views per wave: VW * VPL = 21 * VPL (3 bins x 1 row) or 6 * VPL (5 bins x 2 rows)
template __global__ void k_icd_mw<3, 1, 1, 4>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mw_m<3, 1, 1, 4>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mw<3, 1, 2, 2>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mw_m<3, 1, 2, 2>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mw<3, 1, 2, 3>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mw_m<3, 1, 2, 3>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mw<3, 1, 1, 8>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mw_m<3, 1, 1, 8>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mw<3, 1, 1, 15>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mw_m<3, 1, 1, 15>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mw<5, 2, 1, 8>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mw_m<5, 2, 1, 8>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mw<5, 2, 1, 15>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mw_m<5, 2, 1, 15>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mw<5, 2, 2, 3>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mw_m<5, 2, 2, 3>(const KP *, const int4 *, int, int, int, int);

// ---------------------------------------------------------------------------
// Multi-wavefront voxel update with the error sinogram left in global memory
// (L2-resident), for cells whose [views][N_r][rows] block does not fit in
// LDS: the coarse multires stages (5 or 9 bins x 2 or 4 rows per view) and
// large finest stages.  The roles are those of k_icd_mw: theta waves own
// whole views of the cell and are the only waves that read or write their e
// rows, so the sequential sweep follows from program order (one wavefront's
// global loads and stores to one address take effect in issue order); the
// prior wave is the same.  A lane holds one (view, bin) of the voxel's column
// with all MZ detector rows of the slice, which are contiguous in the
// [N_z][N_p][N_r][mz] layout, for VPL views.  Per voxel n a lane:
//   issues the column-table entry (a, w0) of voxel n + 2 and the (e, w) rows
//   of voxel n + 1 (addressed by the w0 that arrived during voxel n - 1),
//   reduces voxel n's theta sums, meets the other waves, solves,
//   waits once for everything it issued, stores voxel n's e, and corrects
//   voxel n + 1's e, read before that store, by a(voxel n, that bin) * dx:
//   the same fma the lane owning the bin stores, so no bits change.
// ---------------------------------------------------------------------------
template <int MZ>
__device__ __forceinline__ void ld_rows(const float *p, float (&o)[MZ]) {
  if constexpr (MZ == 1) {
    o[0] = p[0];
  } else if constexpr (MZ == 2) {
    const float2 v = *reinterpret_cast<const float2 *>(p);
    o[0] = v.x;
    o[1] = v.y;
  } else {
#pragma unroll
    for (int i = 0; i < MZ; i += 4) {
      const float4 v = *reinterpret_cast<const float4 *>(p + i);
      o[i] = v.x;
      o[i + 1] = v.y;
      o[i + 2] = v.z;
      o[i + 3] = v.w;
    }
  }
}

template <int MZ>
__device__ __forceinline__ void st_rows(float *p, const float (&o)[MZ]) {
  if constexpr (MZ == 1) {
    p[0] = o[0];
  } else if constexpr (MZ == 2) {
    *reinterpret_cast<float2 *>(p) = make_float2(o[0], o[1]);
  } else {
#pragma unroll
    for (int i = 0; i < MZ; i += 4)
      *reinterpret_cast<float4 *>(p + i) = make_float4(o[i], o[i + 1], o[i + 2], o[i + 3]);
  }
}

template <int NB, int MZ, int VPL, int MAXTH, int TF>
__device__ __forceinline__ void k_icd_mwg_body(const KP &P, const int2 cl, int iter,
                                              int chunk, int seq, int count_mag, int vpw) {
  constexpr int G = 8;             // pixels read per list load
  constexpr int VW = 64 / NB;      // views of one lane layer
  constexpr int REC = col_rec(NB);
  __shared__ StepIn S;
  const int f = cl.x, z = cl.y;
  const int cell = f * P.N_z + z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nth = (int)(blockDim.x >> 6) - 1;
  const int Nxy = P.N_x * P.N_y;
  const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;

  const int *list;
  int n0 = 0, cnt = 0;
  if (seq == 0) {
    list = P.ord + (size_t)cell * P.ord_cap;
    n0 = P.ord_off[(size_t)cell * (P.K + 1) + chunk];
    cnt = P.ord_off[(size_t)cell * (P.K + 1) + chunk + 1] - n0;
  } else if (seq == 1) {
    list = P.nh_ord + (size_t)cell * P.nh_cap;
    cnt = P.nh_ord_count[cell];
  } else {
    list = P.single;
    cnt = 1;
  }
  if (cnt <= 0) return;
  if (threadIdx.x < 32) {
    (&S.t1[0][0])[threadIdx.x] = 0.0;
    (&S.t2[0][0])[threadIdx.x] = 0.0f;
  }
  __syncthreads();

  if (wave < nth) {
    // ---------------- theta wave ----------------
    const int lq = lane / NB, ub = lane - lq * NB;
    bool act[VPL];
    int qv[VPL];
    size_t ebase[VPL];
#pragma unroll
    for (int k = 0; k < VPL; k++) {
      const int qq = k * VW + lq, q = wave * vpw + qq;
      act[k] = (lq < VW) & (qq < vpw) & (q < nv);
      qv[k] = act[k] ? q : 0;
      ebase[k] = sidx(P, z, nv > 0 ? P.view_idx[vb + qv[k]] : 0, ub, 0);
    }
    const float *colt = P.colt;
    const size_t fbase = (size_t)f * Nxy;
    auto rec_at = [&](int pix, int k) -> size_t {
      const int lin = pix != kNoPix ? (pix >> 16) * P.N_x + (pix & 0xffff) : 0;
      return ((fbase + lin) * (size_t)P.colt_nv + qv[k]) * REC;
    };
    // The shuffled (HICD) voxels of 3-bin columns, one view per lane, take
    // their A value and window from the view's geometry in the lane, with
    // the column table's own arithmetic (the same bits): their records
    // would be 2 KB reads at random places of the table.  Raster-ordered
    // lines read the table, and so do lanes of several views, for which the
    // arithmetic costs more than the reads.
    const bool calc = NB == 3 && VPL == 1 && seq == 0;
    double cAu[VPL], cBu[VPL], cCu[VPL];
    const float *cH[VPL];
#pragma unroll
    for (int k = 0; k < VPL; k++) {
      const int v = nv > 0 ? P.view_idx[vb + qv[k]] : 0;
      col_coef(P, v, cAu[k], cBu[k], cCu[k]);
      cH[k] = P.hr + (size_t)v * kHrStride;
    }
    auto col_at = [&](int pix, int k, float &a, int &w0) {
      if (calc) {
        const int yy = pix != kNoPix ? pix >> 16 : 0, xx = pix != kNoPix ? pix & 0xffff : 0;
        const double u = col_u(xx, yy, cAu[k], cBu[k], cCu[k]);
        int loc, hic;
        w0 = col_window(P, u, NB, loc, hic);
        a = w0 >= 0 ? col_value(P, cH[k], u, w0 + ub, loc, hic) : 0.0f;
      } else {
        const size_t r = rec_at(pix, k);
        a = colt[r + ub];
        w0 = __float_as_int(colt[r + NB]);
      }
    };
    const int *lst = list + n0;
    auto pix_at = [&](int raw, int k, int n) -> int {
      return n < cnt ? __builtin_amdgcn_readlane(raw, k) : kNoPix;
    };

    // voxel n ("c"): a, w0, (e corrected, w) rows and their address;
    // voxel n + 1 ("n"): a, w0 arrived, (e, w) rows in flight
    float ca[VPL], na[VPL], ce[VPL][MZ], cwv[VPL][MZ], ne[VPL][MZ], nwv[VPL][MZ];
    int cw0[VPL], nw0[VPL];
    size_t cea[VPL], nea[VPL];
    int rawc = lst[min(lane, cnt - 1)];
    int rawn = lst[min(G + lane, cnt - 1)];
    wait_vm_all();
    {
      const int p0 = pix_at(rawc, 0, 0), p1 = pix_at(rawc, 1, 1);
      float a0[VPL], a1[VPL];
      int w00[VPL], w01[VPL];
#pragma unroll
      for (int k = 0; k < VPL; k++) {
        col_at(p0, k, a0[k], w00[k]);
        col_at(p1, k, a1[k], w01[k]);
      }
      wait_vm_all();
#pragma unroll
      for (int k = 0; k < VPL; k++) {
        cw0[k] = act[k] ? w00[k] : -1;
        ca[k] = cw0[k] >= 0 ? a0[k] : 0.0f;
        const bool v1 = act[k] & (p1 != kNoPix);
        nw0[k] = v1 ? w01[k] : -1;
        na[k] = nw0[k] >= 0 ? a1[k] : 0.0f;
        cea[k] = ebase[k] + (size_t)max(cw0[k], 0) * MZ;
        ld_rows<MZ>(P.e + cea[k], ce[k]);
        ld_rows<MZ>(P.w + cea[k], cwv[k]);
      }
      wait_vm_all();
    }
    int rawnn = rawn;
    unsigned long long nent = 0;
    const int ngroups = (cnt + G - 1) / G;
    for (int g = 0; g < ngroups; g++) {
#pragma unroll
      for (int s = 0; s < G; s++) {
        const int n = g * G + s;
        if (n >= cnt) break;
        const int par = n & 1;
        if (s == 0) rawnn = lst[min((g + 2) * G + lane, cnt - 1)];
        // issue: column entries of voxel n + 2, (e, w) rows of voxel n + 1
        const int p2 = s + 2 < G ? pix_at(rawc, s + 2, n + 2) : pix_at(rawn, s + 2 - G, n + 2);
        float a2[VPL];
        int w2[VPL];
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          col_at(p2, k, a2[k], w2[k]);
          nea[k] = ebase[k] + (size_t)max(nw0[k], 0) * MZ;
          ld_rows<MZ>(P.e + nea[k], ne[k]);
          ld_rows<MZ>(P.w + nea[k], nwv[k]);
        }
        // voxel n's theta sums (XT_VoxUpdate.c:61-92)
        // generalized-Huber weights (XT_VoxUpdate.c:66-80); the outer
        // weight only when some lane of the wave needs it
        float cwk[VPL][MZ];
        bool sel[VPL][MZ], allsel = true;
#pragma unroll
        for (int k = 0; k < VPL; k++)
#pragma unroll
          for (int m = 0; m < MZ; m++) {
            sel[k][m] = huber_sel(P, ce[k][m], cwv[k][m]);
            cwk[k][m] = cwv[k][m];
            allsel &= sel[k][m];
          }
        if (__builtin_expect(__ballot(!allsel) != 0ull, 0)) {
#pragma unroll
          for (int k = 0; k < VPL; k++)
#pragma unroll
            for (int m = 0; m < MZ; m++) {
              const float cw_out = P.Td * __builtin_amdgcn_sqrtf(cwv[k][m]) *
                                   __builtin_amdgcn_rcpf(fabsf(ce[k][m]));
              cwk[k][m] = sel[k][m] ? cwv[k][m] : cw_out;
            }
        }
        double t1;
        float t2;
        if constexpr (TF == 0) {
#pragma unroll
          for (int k = 0; k < VPL; k++) {
#pragma unroll
            for (int m = 0; m < MZ; m++) {
              const float e = ce[k][m];
              const float cw = cwk[k][m];
              const float ac = ca[k] * cw;
              const double p1 = (double)(-e * ac);
              const float p2 = ca[k] * ac;
              t1 = (k == 0 && m == 0) ? p1 : t1 + p1;
              t2 = (k == 0 && m == 0) ? p2 : t2 + p2;
            }
            nent += ca[k] > 0.0f ? MZ : 0;
          }
          wave_total_df(t1, t2);
        } else {
          // FP32 sums: compensated (TF 1) or plain (TF 2)
          float s1 = 0.f, c1 = 0.f, s2 = 0.f, c2 = 0.f;
#pragma unroll
          for (int k = 0; k < VPL; k++) {
#pragma unroll
            for (int m = 0; m < MZ; m++) {
              const float e = ce[k][m];
              const float cw = cwk[k][m];
              const float ac = ca[k] * cw;
              const float p1 = -e * ac;
              const float p2 = ca[k] * ac;
              if constexpr (TF == 1) {
                cadd(s1, c1, p1);
                cadd(s2, c2, p2);
              } else {
                s1 += p1;
                s2 += p2;
              }
            }
            nent += ca[k] > 0.0f ? MZ : 0;
          }
          if constexpr (TF == 1) {
            wave_total_cc(s1, c1, s2, c2);
            t1 = (double)s1 + (double)c1;
            t2 = s2 + c2;
          } else {
            wave_total_ff(s1, s2);
            t1 = (double)s1;
            t2 = s2;
          }
        }
        if (lane == 0) {
          S.t1[par][wave] = t1;
          S.t2[par][wave] = t2;
        }
        // a of this voxel at the bin each lane holds for voxel n + 1
        float cn[VPL];
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          const int d = nw0[k] + ub - cw0[k];
          const bool ov = (nw0[k] >= 0) & (cw0[k] >= 0) & (d >= 0) & (d < NB);
          const float av = __shfl(ca[k], lq * NB + (ov ? d : 0));
          cn[k] = ov ? av : 0.0f;
        }
        __syncthreads();
        float V;
        bool skip;
        const float xn = solve_step<MAXTH, TF>(P, S, par, nth, &V, &skip);
        const float dx = xn - V;
        wait_vm_all();
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          if (ca[k] != 0.0f) {
            float o[MZ];
#pragma unroll
            for (int m = 0; m < MZ; m++) o[m] = fmaf(-ca[k], dx, ce[k][m]);
            st_rows<MZ>(P.e + cea[k], o);
          }
#pragma unroll
          for (int m = 0; m < MZ; m++) {
            ce[k][m] = fmaf(-cn[k], dx, ne[k][m]);
            cwv[k][m] = nwv[k][m];
          }
          ca[k] = na[k];
          cw0[k] = nw0[k];
          cea[k] = nea[k];
          const bool v2 = act[k] & (p2 != kNoPix);
          nw0[k] = v2 ? w2[k] : -1;
          na[k] = nw0[k] >= 0 ? a2[k] : 0.0f;
        }
      }
      rawc = rawn;
      rawn = rawnn;
    }
    const double nsum = wave_sum_d((double)nent);
    if (lane == 0) S.nent[wave] = (unsigned long long)nsum;
  } else {
    // ---------------- prior wave ----------------
    prior_wave<MAXTH>(P, S, list, n0, cnt, f, z, iter, seq, count_mag, nth, lane);
  }
  __syncthreads();
  if (threadIdx.x == 0 && seq != 2) {
    unsigned long long ne = 0;
    for (int w = 0; w < nth; w++) ne += S.nent[w];
    P.cell_ent[cell] += ne;
  }
}

// one problem: the stage's cell list
template <int NB, int MZ, int VPL, int MAXTH, int TF>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_mwg(KP P, const int2 *__restrict__ cells,
    int iter, int chunk, int seq, int count_mag, int vpw) {
  k_icd_mwg_body<NB, MZ, VPL, MAXTH, TF>(P, cells[blockIdx.x], iter, chunk, seq, count_mag, vpw);
}
// many problems: cells (frame, slice, problem, last argument), parameters
// from the problems' table
template <int NB, int MZ, int VPL, int MAXTH, int TF>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_mwg_m(const KP *__restrict__ tab,
    const int4 *__restrict__ cells, int iter, int chunk, int seq, int count_mag) {
  const int4 c = cells[blockIdx.x];
  k_icd_mwg_body<NB, MZ, VPL, MAXTH, TF>(tab[c.z], make_int2(c.x, c.y), iter, chunk, seq, count_mag, c.w);
}


// ---------------------------------------------------------------------------
// k_icd_mwd: the global-e multi-wavefront update (k_icd_mwg's lane layout,
// prior wave and exchange, theta1 in FP64) with a deeper pipeline, for
// cells whose (e, w) rows miss the L2 and come from the memory-side cache
// or HBM.  At voxel n a lane has the (e, w) rows of voxels n + 1 .. n + D
// in flight and the column-table entries of voxels n + D + 1 .. n + D + C:
//   issues the column entry of voxel n + D + C and the (e, w) rows of voxel
//   n + D (addressed by the column entry that arrived C voxels ago),
//   replays on voxel n's entries the updates of voxels n - D .. n - 1 that
//   its rows were read before (the owner lanes' own fmaf(-a, dx, e), in
//   the same order, so the bits are those of the sequential sweep),
//   reduces voxel n's theta sums, takes voxel n's A value at the bins it
//   holds for voxels n + 1 .. n + D (their replay coefficients), meets the
//   other waves, solves and stores voxel n's e.
// Only the waits of the compiler's own counters remain in the loop, so a
// row waits for its data D voxels after it was issued.  D + 1 and C divide
// the group of G voxels, so the unrolled group rotates the register rings
// into place without copies of rows still in flight.
// ---------------------------------------------------------------------------
template <int NB, int MZ, int VPL, int MAXTH, int D, int C>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_mwd(KP P, const int2 *__restrict__ cells,
                                                            int iter, int chunk, int seq,
                                                            int count_mag, int vpw) {
  constexpr int G = 8;  // pixels read per list load
  static_assert(D >= 1 && C >= 1 && D + C <= G && G % (D + 1) == 0 && G % C == 0,
                "look-ahead rings must rotate within one group");
  constexpr int VW = 64 / NB;  // views of one lane layer
  constexpr int REC = col_rec(NB);
  __shared__ StepIn S;
  const int2 cl = cells[blockIdx.x];
  const int f = cl.x, z = cl.y;
  const int cell = f * P.N_z + z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nth = (int)(blockDim.x >> 6) - 1;
  const int Nxy = P.N_x * P.N_y;
  const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;

  const int *list;
  int n0 = 0, cnt = 0;
  if (seq == 0) {
    list = P.ord + (size_t)cell * P.ord_cap;
    n0 = P.ord_off[(size_t)cell * (P.K + 1) + chunk];
    cnt = P.ord_off[(size_t)cell * (P.K + 1) + chunk + 1] - n0;
  } else if (seq == 1) {
    list = P.nh_ord + (size_t)cell * P.nh_cap;
    cnt = P.nh_ord_count[cell];
  } else {
    list = P.single;
    cnt = 1;
  }
  if (cnt <= 0) return;
  if (threadIdx.x < 32) {
    (&S.t1[0][0])[threadIdx.x] = 0.0;
    (&S.t2[0][0])[threadIdx.x] = 0.0f;
  }
  __syncthreads();

  if (wave < nth) {
    // ---------------- theta wave ----------------
    const int lq = lane / NB, ub = lane - lq * NB;
    bool act[VPL];
    int qv[VPL];
    size_t ebase[VPL];
#pragma unroll
    for (int k = 0; k < VPL; k++) {
      const int qq = k * VW + lq, q = wave * vpw + qq;
      act[k] = (lq < VW) & (qq < vpw) & (q < nv);
      qv[k] = act[k] ? q : 0;
      ebase[k] = sidx(P, z, nv > 0 ? P.view_idx[vb + qv[k]] : 0, ub, 0);
    }
    const float *colt = P.colt;
    const size_t fbase = (size_t)f * Nxy;
    auto rec_at = [&](int pix, int k) -> size_t {
      const int lin = pix != kNoPix ? (pix >> 16) * P.N_x + (pix & 0xffff) : 0;
      return ((fbase + lin) * (size_t)P.colt_nv + qv[k]) * REC;
    };
    const int *lst = list + n0;
    auto pix_at = [&](int raw, int k, int n) -> int {
      return n < cnt ? __builtin_amdgcn_readlane(raw, k) : kNoPix;
    };

    // column entries in flight: ring slot j holds voxel n + D + j
    float ra[C][VPL];
    int rw[C][VPL];
    bool rv[C];
    // voxels n .. n + D: slot i holds voxel n + i (a, w0, the (e, w) rows
    // as read, and the coefficients of the updates to replay on them:
    // rc[i][t] belongs to voxel n + i - D + t)
    float qa[D + 1][VPL], qe[D + 1][VPL][MZ], qw[D + 1][VPL][MZ], rc[D + 1][D][VPL];
    int q0[D + 1][VPL];
    float dxs[D];  // dx of voxels n - D .. n - 1
    int rawc = lst[min(lane, cnt - 1)];
    int rawn = lst[min(G + lane, cnt - 1)];
    // prologue: the column entries of voxels 0 .. D + C - 1, then the rows
    // of voxels 0 .. D - 1
    {
      float a0[D + C][VPL];
      int w0[D + C][VPL];
      bool v0[D + C];
#pragma unroll
      for (int j = 0; j < D + C; j++) {
        const int p = pix_at(rawc, j, j);
        v0[j] = p != kNoPix;
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          const size_t r = rec_at(p, k);
          a0[j][k] = colt[r + ub];
          w0[j][k] = __float_as_int(colt[r + NB]);
        }
      }
#pragma unroll
      for (int i = 0; i < D; i++)
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          const int w = (act[k] & v0[i]) ? w0[i][k] : -1;
          q0[i][k] = w;
          qa[i][k] = w >= 0 ? a0[i][k] : 0.0f;
          ld_rows<MZ>(P.e + ebase[k] + (size_t)max(w, 0) * MZ, qe[i][k]);
          ld_rows<MZ>(P.w + ebase[k] + (size_t)max(w, 0) * MZ, qw[i][k]);
#pragma unroll
          for (int t = 0; t < D; t++) rc[i][t][k] = 0.0f;
        }
#pragma unroll
      for (int j = 0; j < C; j++) {
        rv[j] = v0[D + j];
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          ra[j][k] = a0[D + j][k];
          rw[j][k] = w0[D + j][k];
        }
      }
#pragma unroll
      for (int t = 0; t < D; t++) dxs[t] = 0.0f;
    }
    int rawnn = rawn;
    unsigned long long nent = 0;
    const int ngroups = (cnt + G - 1) / G;
    for (int g = 0; g < ngroups; g++) {
#pragma unroll
      for (int s = 0; s < G; s++) {
        const int n = g * G + s;  // past cnt: an empty slot, A = 0
        const int par = n & 1;
        if (s == 0) rawnn = lst[min((g + 2) * G + lane, cnt - 1)];
        // issue: the column entry of voxel n + D + C, the rows of voxel n + D
        {
          constexpr int so0 = D + C;
          const int so = s + so0;
          const int p = so < G ? pix_at(rawc, so, n + so0) : pix_at(rawn, so - G, n + so0);
          float na[VPL];
          int nw[VPL];
#pragma unroll
          for (int k = 0; k < VPL; k++) {
            const size_t r = rec_at(p, k);
            na[k] = colt[r + ub];
            nw[k] = __float_as_int(colt[r + NB]);
          }
#pragma unroll
          for (int k = 0; k < VPL; k++) {
            const int w = (act[k] & rv[0]) ? rw[0][k] : -1;
            q0[D][k] = w;
            qa[D][k] = w >= 0 ? ra[0][k] : 0.0f;
            ld_rows<MZ>(P.e + ebase[k] + (size_t)max(w, 0) * MZ, qe[D][k]);
            ld_rows<MZ>(P.w + ebase[k] + (size_t)max(w, 0) * MZ, qw[D][k]);
#pragma unroll
            for (int t = 0; t < D; t++) rc[D][t][k] = 0.0f;
          }
#pragma unroll
          for (int j = 0; j + 1 < C; j++) {
            rv[j] = rv[j + 1];
#pragma unroll
            for (int k = 0; k < VPL; k++) {
              ra[j][k] = ra[j + 1][k];
              rw[j][k] = rw[j + 1][k];
            }
          }
          rv[C - 1] = p != kNoPix;
#pragma unroll
          for (int k = 0; k < VPL; k++) {
            ra[C - 1][k] = na[k];
            rw[C - 1][k] = nw[k];
          }
        }
        // voxel n's rows brought up to date: the updates of voxels n - D ..
        // n - 1, in order
        float ce[VPL][MZ];
#pragma unroll
        for (int k = 0; k < VPL; k++)
#pragma unroll
          for (int m = 0; m < MZ; m++) {
            float e = qe[0][k][m];
#pragma unroll
            for (int t = 0; t < D; t++) e = fmaf(-rc[0][t][k], dxs[t], e);
            ce[k][m] = e;
          }
        // voxel n's theta sums (XT_VoxUpdate.c:61-92) with the
        // generalized-Huber weights (XT_VoxUpdate.c:66-80)
        float cwk[VPL][MZ];
        bool sel[VPL][MZ], allsel = true;
#pragma unroll
        for (int k = 0; k < VPL; k++)
#pragma unroll
          for (int m = 0; m < MZ; m++) {
            sel[k][m] = huber_sel(P, ce[k][m], qw[0][k][m]);
            cwk[k][m] = qw[0][k][m];
            allsel &= sel[k][m];
          }
        if (__builtin_expect(__ballot(!allsel) != 0ull, 0)) {
#pragma unroll
          for (int k = 0; k < VPL; k++)
#pragma unroll
            for (int m = 0; m < MZ; m++) {
              const float cw_out = P.Td * __builtin_amdgcn_sqrtf(qw[0][k][m]) *
                                   __builtin_amdgcn_rcpf(fabsf(ce[k][m]));
              cwk[k][m] = sel[k][m] ? qw[0][k][m] : cw_out;
            }
        }
        double t1;
        float t2;
#pragma unroll
        for (int k = 0; k < VPL; k++) {
#pragma unroll
          for (int m = 0; m < MZ; m++) {
            const float ac = qa[0][k] * cwk[k][m];
            const double p1 = (double)(-ce[k][m] * ac);
            const float p2 = qa[0][k] * ac;
            t1 = (k == 0 && m == 0) ? p1 : t1 + p1;
            t2 = (k == 0 && m == 0) ? p2 : t2 + p2;
          }
          nent += qa[0][k] > 0.0f ? MZ : 0;
        }
        wave_total_df(t1, t2);
        if (lane == 0) {
          S.t1[par][wave] = t1;
          S.t2[par][wave] = t2;
        }
        // voxel n's A value at the bins this lane holds for voxels n + 1 ..
        // n + D: the coefficient of voxel n's update on them
#pragma unroll
        for (int i = 1; i <= D; i++)
#pragma unroll
          for (int k = 0; k < VPL; k++) {
            const int d = q0[i][k] + ub - q0[0][k];
            const bool ov = (q0[i][k] >= 0) & (q0[0][k] >= 0) & (d >= 0) & (d < NB);
            const float av = __shfl(qa[0][k], lq * NB + (ov ? d : 0));
            rc[i][D - i][k] = ov ? av : 0.0f;
          }
        __syncthreads();
        float V;
        bool skip;
        const float xn = solve_step<MAXTH, 0>(P, S, par, nth, &V, &skip);
        const float dx = xn - V;
#pragma unroll
        for (int k = 0; k < VPL; k++) {
          if (qa[0][k] != 0.0f) {
            float o[MZ];
#pragma unroll
            for (int m = 0; m < MZ; m++) o[m] = fmaf(-qa[0][k], dx, ce[k][m]);
            st_rows<MZ>(P.e + ebase[k] + (size_t)q0[0][k] * MZ, o);
          }
        }
        // rotate the rings: voxel n + 1 becomes the current one
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
          for (int k = 0; k < VPL; k++) {
            qa[i][k] = qa[i + 1][k];
            q0[i][k] = q0[i + 1][k];
#pragma unroll
            for (int m = 0; m < MZ; m++) {
              qe[i][k][m] = qe[i + 1][k][m];
              qw[i][k][m] = qw[i + 1][k][m];
            }
#pragma unroll
            for (int t = 0; t < D; t++) rc[i][t][k] = rc[i + 1][t][k];
          }
#pragma unroll
        for (int t = 0; t + 1 < D; t++) dxs[t] = dxs[t + 1];
        dxs[D - 1] = dx;
      }
      rawc = rawn;
      rawn = rawnn;
    }
    const double nsum = wave_sum_d((double)nent);
    if (lane == 0) S.nent[wave] = (unsigned long long)nsum;
  } else {
    // ---------------- prior wave ----------------
    prior_wave<MAXTH, true>(P, S, list, n0, cnt, f, z, iter, seq, count_mag, nth, lane);
  }
  __syncthreads();
  if (threadIdx.x == 0 && seq != 2) {
    unsigned long long ne = 0;
    for (int w = 0; w < nth; w++) ne += S.nent[w];
    P.cell_ent[cell] += ne;
  }
}

// ---------------------------------------------------------------------------
// k_icd_bat: the finest stage's voxel update (3-bin columns, one detector
// row per slice, one view per lane) in rounds of K voxels of each cell's
// sequence, with three barriers per round instead of one per voxel.
//
// The sweep is the sequential one of XT_VoxUpdate.c:43-115, in the same
// order and with the same result up to the order of the theta sums.  A
// round takes the error sinogram as it finds it (e0).  Voxel i's theta1 is
//   theta1_i = -sum_j a_ij wt(e_j) e_j,  e_j = e0_j - sum_{m<i} a_mj dx_m
// (wt the generalized-Huber weight).  An entry that no earlier voxel of the
// round touches keeps e0; one that stays in the quadratic branch (wt = w)
// changes theta1 by dx_m a_mj a_ij w_j and leaves theta2; so with
//   G_mi = sum over those entries of a_mj a_ij w_j
// theta1_i = theta1_i(e0) + sum_{m<i} dx_m G_mi and theta2_i = theta2_i(e0)
// hold exactly whenever every entry voxel i shares with an earlier voxel of
// the round is in the quadratic branch both at e0 and after their updates.
//
// Round: the theta lanes (view, bin offset) read each voxel's column-table
// record whole (one round ahead), its (e, w) entry, and every other voxel's
// A value at that entry, and fold the K theta1 (FP64), K theta2 and
// K(K-1)/2 G partial sums so that each lane ends with one total; the waves'
// totals meet in LDS (barrier 1).  The prior wave, which has meanwhile
// evaluated the K voxels' qGGMRF surrogates (two voxels per pass), solves
// the K voxels in order from those sums, re-evaluating the surrogate of a
// voxel whose in-plane neighbour the round has just updated, and publishes
// the steps (barrier 2).  Each theta lane then checks its shared entries
// with those steps (barrier 3): from the first voxel with a shared entry
// outside the quadratic branch on either side, the round continues one
// voxel at a time with theta taken from the entries' current values (two
// barriers each), the prior wave re-solving it.  Finally every theta lane
// writes the final e of its entry if no later voxel of the round has a
// nonzero A value there: e0 minus the round's updates of that entry in
// order, fmaf(-a, dx, e) each, the sequential sweep's arithmetic.
// ---------------------------------------------------------------------------
constexpr int kBat = 4;    // voxels per round
constexpr int kBatF = 16;  // FP32 sums per round: theta2 x 4, G x 6, padding

struct BatIn {
  double t1[16][kBat];          // per theta wave: theta1(e0) of the round's voxels
  float f[16][kBatF];           // per theta wave: theta2 of each voxel, then G_mi
  alignas(16) float dx[kBat];   // the round's steps
  int bad[16];                  // per theta wave: voxels whose theta was not exact
  unsigned long long nent[16];
};

// Folding sums: many values per lane become one total per lane.  Level L
// pairs the values (2t, 2t + 1) across lanes that differ in one bit (L = 0:
// lane bit 5, permlane32 swap; 1: bit 4, permlane16 swap; 2: bit 3, DPP
// row_ror:8; 3: bit 2, DPP row_half_mirror (lane 7 - l within 8); 4: bit 1
// and 5: bit 0, DPP quad perms): the lane with the bit clear keeps the sum
// of value 2t, its partner that of value 2t + 1.  Once one value is left
// the remaining levels add it up over the lanes that still differ.
template <int L>
__device__ __forceinline__ float fold_level(float a, float b) {
  if constexpr (L <= 1) {
    const auto r = L == 0 ? __builtin_amdgcn_permlane32_swap(__float_as_uint(a), __float_as_uint(b),
                                                             false, false)
                          : __builtin_amdgcn_permlane16_swap(__float_as_uint(a), __float_as_uint(b),
                                                             false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
  } else {
    constexpr int CTRL = L == 2 ? 0x128 : L == 3 ? 0x141 : L == 4 ? 0x4E : 0xB1;
    const bool hi = (threadIdx.x >> (5 - L)) & 1;
    const float send = hi ? a : b, keep = hi ? b : a;
    return keep + dpp_f<CTRL>(send);
  }
}
template <int L>
__device__ __forceinline__ float fold_all(float v) {
  if constexpr (L <= 1) return swap_sum_f<L == 0>(v);
  else return v + dpp_f<L == 2 ? 0x128 : L == 3 ? 0x141 : L == 4 ? 0x4E : 0xB1>(v);
}
template <int L>
__device__ __forceinline__ double fold_level_d(double a, double b) {
  static_assert(L <= 1, "the FP64 fold pairs across rows only");
  const unsigned long long ua = (unsigned long long)__double_as_longlong(a);
  const unsigned long long ub = (unsigned long long)__double_as_longlong(b);
  const unsigned al = (unsigned)ua, ah = (unsigned)(ua >> 32);
  const unsigned bl = (unsigned)ub, bh = (unsigned)(ub >> 32);
  const auto lo = L == 0 ? __builtin_amdgcn_permlane32_swap(al, bl, false, false)
                         : __builtin_amdgcn_permlane16_swap(al, bl, false, false);
  const auto hi = L == 0 ? __builtin_amdgcn_permlane32_swap(ah, bh, false, false)
                         : __builtin_amdgcn_permlane16_swap(ah, bh, false, false);
  const double x = __longlong_as_double((long long)(((unsigned long long)hi[0] << 32) | lo[0]));
  const double y = __longlong_as_double((long long)(((unsigned long long)hi[1] << 32) | lo[1]));
  return x + y;
}
// 16 FP32 values: lane l ends with the total of value
// (l5 | l4 << 1 | l3 << 2 | l2 << 3) (lane bits), equal in the 4 lanes of a
// quad
__device__ __forceinline__ float fold16(const float (&v)[16]) {
  float u[8], w[4], x[2];
#pragma unroll
  for (int t = 0; t < 8; t++) u[t] = fold_level<0>(v[2 * t], v[2 * t + 1]);
#pragma unroll
  for (int t = 0; t < 4; t++) w[t] = fold_level<1>(u[2 * t], u[2 * t + 1]);
#pragma unroll
  for (int t = 0; t < 2; t++) x[t] = fold_level<2>(w[2 * t], w[2 * t + 1]);
  float y = fold_level<3>(x[0], x[1]);
  y = fold_all<4>(y);
  return fold_all<5>(y);
}
__device__ __forceinline__ int fold16_index(int lane) {
  return ((lane >> 5) & 1) | ((lane >> 3) & 2) | ((lane >> 1) & 4) | ((lane << 1) & 8);
}
// 4 FP64 values: lane l ends with the total of value (l5 | l4 << 1), equal
// in the 16 lanes of a row
__device__ __forceinline__ double fold4d(const double (&v)[4]) {
  const double u0 = fold_level_d<0>(v[0], v[1]), u1 = fold_level_d<0>(v[2], v[3]);
  double y = fold_level_d<1>(u0, u1);
  y += dpp_d<0xB1>(y);
  y += dpp_d<0x4E>(y);
  y += dpp_d<0x141>(y);
  y += dpp_d<0x140>(y);
  return y;
}
__device__ __forceinline__ int fold4d_index(int lane) { return ((lane >> 5) & 1) | ((lane >> 3) & 2); }

// The qGGMRF surrogate (XT_ICD_update.c:134-183) of two voxels at once: the
// lanes of each half-wave hold the 29 neighbour values of one voxel (lane
// 13 + 32h the voxel itself, NaN for a neighbour outside the volume) and
// the per-neighbour constants; pb and pbx end in every lane of the half.
struct PriorLane {
  float ng, nsq, nisqp, nc, two_c;
  int ln;
};
__device__ __forceinline__ void prior_pair(const PriorLane &C, float nbv, float V, float &pb,
                                           float &pbx) {
  const bool vld = (nbv == nbv) & (C.ln != 13);
  const float dlt = vld ? V - nbv : 0.0f;
  const float t1 = __builtin_amdgcn_exp2f(0.8f * __builtin_amdgcn_logf(fabsf(dlt))) * C.nisqp;
  const float t3 = C.nc + t1;
  const float b = fmaf(1.2f, t1, C.two_c) * __builtin_amdgcn_rcpf(t3 * t3 * C.nsq);
  const float gb = (vld ? C.ng : 0.0f) * b;
  pb = gb;
  pbx = -gb * dlt;
  lo32_total_ff(pb, pbx);
}

template <int MAXTH>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_bat(KP P, const int2 *__restrict__ cells,
                                                            int iter, int chunk, int seq,
                                                            int count_mag, int vpw) {
  constexpr int NB = 3, G = 8, K = kBat;
  static_assert(G % K == 0 && K == 4, "whole rounds of 4 voxels per group");
  constexpr int VW = 64 / NB;  // views of one lane layer
  __shared__ BatIn S;
  const int2 cl = cells[blockIdx.x];
  const int f = cl.x, z = cl.y;
  const int cell = f * P.N_z + z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nth = (int)(blockDim.x >> 6) - 1;
  const int Nxy = P.N_x * P.N_y;
  const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;

  const int *list;
  int n0 = 0, cnt = 0;
  if (seq == 0) {
    list = P.ord + (size_t)cell * P.ord_cap;
    n0 = P.ord_off[(size_t)cell * (P.K + 1) + chunk];
    cnt = P.ord_off[(size_t)cell * (P.K + 1) + chunk + 1] - n0;
  } else if (seq == 1) {
    list = P.nh_ord + (size_t)cell * P.nh_cap;
    cnt = P.nh_ord_count[cell];
  } else {
    list = P.single;
    cnt = 1;
  }
  if (cnt <= 0) return;
  for (int t = threadIdx.x; t < 16 * kBat; t += blockDim.x) (&S.t1[0][0])[t] = 0.0;
  for (int t = threadIdx.x; t < 16 * kBatF; t += blockDim.x) (&S.f[0][0])[t] = 0.0f;
  if (threadIdx.x < 16) S.bad[threadIdx.x] = 0;
  __syncthreads();
  const int ngroups = (cnt + G - 1) / G;
  // the first voxel of the round whose theta was not exact (K: none), the
  // same in every wave
  auto first_bad = [&]() -> int {
    int m = 0;
#pragma unroll
    for (int w = 0; w < MAXTH; w++) m |= S.bad[w];
    return m ? __builtin_ctz((unsigned)m) : K;
  };
  // a theta total over the waves: a fixed pairwise tree over MAXTH slots
  // (absent waves hold zero), slot `col` of each wave's row
  auto sum_t1 = [&](int col) -> double {
    double v[MAXTH];
#pragma unroll
    for (int w = 0; w < MAXTH; w++) v[w] = S.t1[w][col];
#pragma unroll
    for (int m = MAXTH; m > 1; m = (m + 1) / 2) {
      const int h = (m + 1) / 2;
#pragma unroll
      for (int w = 0; w + h < m; w++) v[w] = v[w] + v[w + h];
    }
    return v[0];
  };
  auto sum_f = [&](int col) -> float {
    float v[MAXTH];
#pragma unroll
    for (int w = 0; w < MAXTH; w++) v[w] = S.f[w][col];
#pragma unroll
    for (int m = MAXTH; m > 1; m = (m + 1) / 2) {
      const int h = (m + 1) / 2;
#pragma unroll
      for (int w = 0; w + h < m; w++) v[w] = v[w] + v[w + h];
    }
    return v[0];
  };

  if (wave < nth) {
    // ---------------- theta wave ----------------
    const int lq = lane / NB, ub = lane - lq * NB;
    const int qq = lq, q = wave * vpw + qq;
    const bool act = (lq < VW) & (qq < vpw) & (q < nv);
    const int qv = act ? q : 0;
    const size_t ebase = sidx(P, z, nv > 0 ? P.view_idx[vb + qv] : 0, ub, 0);
    const float4 *colt = reinterpret_cast<const float4 *>(P.colt);
    const size_t fbase = (size_t)f * Nxy;
    auto rec_at = [&](int pix) -> size_t {
      const int lin = pix != kNoPix ? (pix >> 16) * P.N_x + (pix & 0xffff) : 0;
      return (fbase + lin) * (size_t)P.colt_nv + qv;
    };
    const int *lst = list + n0;
    auto pix_at = [&](int raw, int k, int n) -> int {
      return n < cnt ? __builtin_amdgcn_readlane(raw, k) : kNoPix;
    };
    // A value of a record at window offset d (0 outside the window)
    auto rec_a = [](const float4 &r, int d) -> float {
      const float v = d == 0 ? r.x : d == 1 ? r.y : r.z;
      return (unsigned)d < (unsigned)NB ? v : 0.0f;
    };
    // theta terms of one entry (XT_VoxUpdate.c:61-92) with the
    // generalized-Huber weight (XT_VoxUpdate.c:66-80)
    auto terms = [&](float a, float e, float w, bool sel, double &p1, float &p2) {
      const float cw = sel ? w : P.Td * __builtin_amdgcn_sqrtf(w) * __builtin_amdgcn_rcpf(fabsf(e));
      const float ac = a * cw;
      p1 = (double)(-e * ac);
      p2 = a * ac;
    };

    // column records of the current round and of the next one
    float4 rc[K], rn[K];
    bool vc[K], vn[K];
    int rawc = lst[min(lane, cnt - 1)];
    int rawn = lst[min(G + lane, cnt - 1)];
#pragma unroll
    for (int i = 0; i < K; i++) {
      const int p = pix_at(rawc, i, i);
      vc[i] = p != kNoPix;
      rc[i] = colt[rec_at(p)];
    }
    int rawnn = rawn;
    unsigned long long nent = 0;
    for (int g = 0; g < ngroups; g++) {
#pragma unroll
      for (int bb = 0; bb < G / K; bb++) {
        const int s0 = bb * K;  // the round's first voxel in the group
        if (bb == 0) rawnn = lst[min((g + 2) * G + lane, cnt - 1)];
        // this round's entries: window start, own A value, (e0, w)
        int w0[K];
        float a[K], e[K], wv[K];
        size_t ea[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
          w0[i] = (act & vc[i]) ? __float_as_int(rc[i].w) : -1;
          a[i] = w0[i] >= 0 ? (ub == 0 ? rc[i].x : ub == 1 ? rc[i].y : rc[i].z) : 0.0f;
          ea[i] = ebase + (size_t)max(w0[i], 0);
          e[i] = P.e[ea[i]];
          wv[i] = P.w[ea[i]];
        }
        // the next round's records
#pragma unroll
        for (int i = 0; i < K; i++) {
          const int sn = s0 + K + i;
          const int p = sn < G ? pix_at(rawc, sn, g * G + sn) : pix_at(rawn, sn - G, g * G + sn);
          vn[i] = p != kNoPix;
          rn[i] = colt[rec_at(p)];
        }
        // A value of voxel m at the entry this lane holds for voxel i
        float cmi[K][K];
        bool later[K], shared[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
          later[i] = false;
          shared[i] = false;
#pragma unroll
          for (int m = 0; m < K; m++) {
            if (m == i) continue;
            const int d = w0[i] + ub - w0[m];
            const float c = (w0[i] >= 0) & (w0[m] >= 0) ? rec_a(rc[m], d) : 0.0f;
            cmi[m][i] = c;
            if (m > i) later[i] |= c != 0.0f;
            else shared[i] |= c != 0.0f;
          }
        }
        // theta sums of e0 and the pair sums G_mi
        double t1[K];
        float fv[kBatF];
        bool sel0[K];
#pragma unroll
        for (int i = 0; i < K; i++) {
          sel0[i] = huber_sel(P, e[i], wv[i]);
          terms(a[i], e[i], wv[i], sel0[i], t1[i], fv[i]);
          const float aq = sel0[i] ? a[i] * wv[i] : 0.0f;
#pragma unroll
          for (int m = 0; m < i; m++) fv[K + i * (i - 1) / 2 + m] = cmi[m][i] * aq;
          nent += a[i] > 0.0f ? 1 : 0;
        }
#pragma unroll
        for (int t = K + K * (K - 1) / 2; t < kBatF; t++) fv[t] = 0.0f;
        const double t1s = fold4d(t1);
        const float fs = fold16(fv);
        if ((lane & 15) == 0) S.t1[wave][fold4d_index(lane)] = t1s;
        if ((lane & 3) == 0) S.f[wave][fold16_index(lane)] = fs;
        __syncthreads();  // 1: the round's sums are in LDS
        __syncthreads();  // 2: the prior wave's steps are in LDS
        float dx[K];
        {
          const float4 d4 = *reinterpret_cast<const float4 *>(S.dx);
          dx[0] = d4.x;
          dx[1] = d4.y;
          dx[2] = d4.z;
          dx[3] = d4.w;
        }
        // each shared entry in the quadratic branch before and after the
        // earlier voxels' updates?
        int badm = 0;
#pragma unroll
        for (int i = 1; i < K; i++) {
          float v = e[i];
#pragma unroll
          for (int m = 0; m < i; m++) v = fmaf(-cmi[m][i], dx[m], v);
          const bool bad = shared[i] & (!sel0[i] | !huber_sel(P, v, wv[i]));
          badm |= __ballot(bad) != 0ull ? 1 << i : 0;
        }
        if (lane == 0) S.bad[wave] = badm;
        __syncthreads();  // 3: the waves' checks are in LDS
        const int ib = first_bad();
        // from voxel ib on, one voxel at a time from the entries' current
        // values
#pragma unroll
        for (int i = 1; i < K; i++) {
          if (i < ib) continue;
          {
            const float4 d4 = *reinterpret_cast<const float4 *>(S.dx);
            dx[0] = d4.x;
            dx[1] = d4.y;
            dx[2] = d4.z;
            dx[3] = d4.w;
          }
          float v = e[i];
#pragma unroll
          for (int m = 0; m < i; m++) v = fmaf(-cmi[m][i], dx[m], v);
          double p1;
          float p2;
          terms(a[i], v, wv[i], huber_sel(P, v, wv[i]), p1, p2);
          wave_total_df(p1, p2);
          if (lane == 0) {
            S.t1[wave][0] = p1;
            S.f[wave][0] = p2;
          }
          __syncthreads();  // the voxel's sums are in LDS
          __syncthreads();  // its step is in LDS
        }
        {
          const float4 d4 = *reinterpret_cast<const float4 *>(S.dx);
          dx[0] = d4.x;
          dx[1] = d4.y;
          dx[2] = d4.z;
          dx[3] = d4.w;
        }
        // the final e of each entry, written by the last voxel of the round
        // with a nonzero A value there
#pragma unroll
        for (int i = 0; i < K; i++) {
          if (a[i] != 0.0f && !later[i]) {
            float v = e[i];
#pragma unroll
            for (int m = 0; m < i; m++) v = fmaf(-cmi[m][i], dx[m], v);
            P.e[ea[i]] = fmaf(-a[i], dx[i], v);
          }
        }
#pragma unroll
        for (int i = 0; i < K; i++) {
          rc[i] = rn[i];
          vc[i] = vn[i];
        }
      }
      rawc = rawn;
      rawn = rawnn;
    }
    const double nsum = wave_sum_d((double)nent);
    if (lane == 0) S.nent[wave] = (unsigned long long)nsum;
  } else {
    // ---------------- prior wave ----------------
    const int ln = lane & 31;
    PriorLane C;
    C.ln = ln;
    int ndz = 0, ndy = 0, ndx = 0, ndt = 0;
    C.ng = 0.0f;
    C.nsq = 1.0f;
    C.nisqp = 1.0f;
    C.nc = 1.0f;
    if (ln < 29) {
      if (ln < 27) {
        ndz = ln / 9 - 1;
        ndy = (ln / 3) % 3 - 1;
        ndx = ln % 3 - 1;
        C.ng = P.sfilt[ln];
        C.nsq = P.sSq, C.nisqp = 1.0f / P.sSqp, C.nc = P.c_s;
      } else {
        ndt = ln == 27 ? -1 : 1;
        C.ng = P.tfilt;
        C.nsq = P.sTq, C.nisqp = 1.0f / P.sTqp, C.nc = P.c_t;
      }
    }
    C.two_c = 2.0f * C.nc;
    // the neighbour this lane loads (lanes 0..28; a swap gives the upper
    // half-wave another voxel's values)
    const bool nlane = lane < 29;
    const bool inplane = lane < 27 && ndz == 0 && lane != 13;
    const int noff = ndy * 65536 + ndx;
    const int zz = z + ndz, ff = f + ndt;
    const bool zok = nlane && zz >= 0 && zz < P.N_z && ff >= 0 && ff < P.N_time;
    const int zc = min(max(zz, 0), P.N_z - 1), fc = min(max(ff, 0), P.N_time - 1);
    const int nplane = (fc * P.N_z + zc) * Nxy;
    const int cbase = cell * Nxy;
    const float qnan = __int_as_float(0x7fc00000);

    int apix[G], anid[G], bpix[G], bnid[G];
    float anb[G], bnb[G], bnv[G];
    int pend = kNoPix;
    auto fill = [&](int k, int pix) {
      bpix[k] = pix;
      const bool valid = pix != kNoPix;
      const int yy = valid ? (pix >> 16) : 0, xx = valid ? (pix & 0xffff) : 0;
      const int jj = yy + ndy, kk = xx + ndx;
      const bool nok = valid & zok & (jj >= 0) & (jj < P.N_y) & (kk >= 0) & (kk < P.N_x);
      bnb[k] = P.obj[(unsigned)(nok ? nplane + jj * P.N_x + kk : P.nobj + 64)];
      bnid[k] = (valid & inplane) ? pix + noff : -2;
      bnv[k] = qnan;
    };
    // one voxel's surrogate from the lower half-wave's lanes
    auto prior_one = [&](float nbv, float &pb, float &pbx, float &V, bool &skip) {
      V = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nbv), 13));
      const bool anypos = (__ballot((ln != 13) & (nbv > 0.0f)) & 0xffffffffull) != 0ull;
      float qb, qbx;
      prior_pair(C, nbv, V, qb, qbx);
      pb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qb), 0));
      pbx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qbx), 0));
      skip = (V == 0.0f) && iter > 1 && !anypos;
    };
    double vmacc = 0.0;
    unsigned long long nupd = 0, nzero = 0;
    int raw = list[n0 + min(lane, cnt - 1)];
    wait_vm_all();
    int pnext[G];
#pragma unroll
    for (int k = 0; k < G; k++) pnext[k] = k < cnt ? __builtin_amdgcn_readlane(raw, k) : kNoPix;
#pragma unroll
    for (int k = 0; k < G; k++) {
      apix[k] = kNoPix;
      anid[k] = -2;
      anb[k] = qnan;
    }
    for (int g = 0; g <= ngroups; g++) {
#pragma unroll
      for (int k = 0; k < G; k++) fill(k, pnext[k]);
      raw = list[n0 + min((g + 1) * G + lane, cnt - 1)];
#pragma unroll
      for (int k = 0; k < G; k++) pend = lane == G + k ? bpix[k] : pend;
      float rx = 0.0f, rm = 0.0f;
      int ridx = P.nobj + lane;
      int rsk = -1;  // lane k: voxel k of the group skipped (1), updated (0), absent (-1)
      // voxel s's step from its surrogate and theta sums; the object and
      // update-map values and the neighbour patches of the voxels after it
      auto step_voxel = [&](int s, float pb, float pbx, float V, bool skip, double T1,
                            float T2, bool *redo, int s0) -> float {
        const int pix = apix[s];
        const float th1 = (float)T1;
        const float num = P.alpha * (pbx - th1);
        const float step = num * __builtin_amdgcn_rcpf(pb + T2);
        const float xn = skip ? V : V + step;
        const float dx = xn - V;
        // (dy + 1) * 65536 + (dx + 1) with |dy|, |dx| <= 1 for a neighbour
        const unsigned dd = (unsigned)(pend - pix + 65537);
        const bool adj = (pend != kNoPix) & (dd != 65537u) & ((dd & 0xffffu) <= 2u) &
                         (dd <= 131074u);
        if (__ballot(adj) != 0ull) {
#pragma unroll
          for (int o = 0; o < G; o++) {
            if (o <= s) continue;
            const bool hit = anid[o] == pix;
            anb[o] = hit ? xn : anb[o];
            if (o < s0 + K) redo[o - s0] |= __ballot(hit) != 0ull;
          }
#pragma unroll
          for (int o = 0; o < G; o++) bnv[o] = bnid[o] == pix ? xn : bnv[o];
        }
        const int lin = cbase + (pix >> 16) * P.N_x + (pix & 0xffff);
        rx = lane == s ? xn : rx;
        rm = lane == s ? (skip ? 0.0f : fabsf(dx)) : rm;
        ridx = lane == s ? lin : ridx;
        rsk = lane == s ? (skip ? 1 : 0) : rsk;
        return dx;
      };
      if (g > 0) {
#pragma unroll
        for (int bb = 0; bb < G / K; bb++) {
          const int s0 = bb * K;
          // the round's surrogates, two voxels per pass
          float pb[K], pbx[K], V[K];
          bool skip[K], redo[K];
#pragma unroll
          for (int i = 0; i < K; i += 2) {
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(anb[s0 + i]),
                                                             __float_as_uint(anb[s0 + i + 1]),
                                                             false, false);
            const float nbv = __uint_as_float(sw[0]);  // lanes 32..63: voxel i + 1's lanes 0..31
            const float Va = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nbv), 13));
            const float Vb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nbv), 45));
            const unsigned long long pos = __ballot((ln != 13) & (nbv > 0.0f));
            float qb, qbx;
            prior_pair(C, nbv, lane < 32 ? Va : Vb, qb, qbx);
            pb[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qb), 0));
            pbx[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qbx), 0));
            pb[i + 1] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qb), 32));
            pbx[i + 1] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qbx), 32));
            V[i] = Va;
            V[i + 1] = Vb;
            skip[i] = (Va == 0.0f) && iter > 1 && (pos & 0xffffffffull) == 0ull;
            skip[i + 1] = (Vb == 0.0f) && iter > 1 && (pos >> 32) == 0ull;
          }
#pragma unroll
          for (int i = 0; i < K; i++) redo[i] = false;
          __syncthreads();  // 1: the theta waves' sums are in LDS
          // lane t < 16: FP32 total t (theta2 of voxel t, then G); lane i < 4
          // also theta1 of voxel i
          const float fa = sum_f(lane & 15);
          const double da = sum_t1(lane & 3);
          float T2[K], Gp[K * (K - 1) / 2], dxs[K];
          double T1[K];
#pragma unroll
          for (int i = 0; i < K; i++) {
            T2[i] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fa), i));
            const unsigned long long bits = (unsigned long long)__double_as_longlong(da);
            const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)bits, i);
            const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(bits >> 32), i);
            T1[i] = __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
          }
#pragma unroll
          for (int t = 0; t < K * (K - 1) / 2; t++)
            Gp[t] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(fa), K + t));
          // the round's voxels in order (XT_VoxUpdate.c:43-115)
#pragma unroll
          for (int i = 0; i < K; i++) {
            const int s = s0 + i;
            dxs[i] = 0.0f;
            if (apix[s] == kNoPix) continue;  // only past the end of the sequence
            if (redo[i]) prior_one(anb[s], pb[i], pbx[i], V[i], skip[i]);
            double t = T1[i];
#pragma unroll
            for (int m = 0; m < i; m++) t += (double)dxs[m] * (double)Gp[i * (i - 1) / 2 + m];
            dxs[i] = step_voxel(s, pb[i], pbx[i], V[i], skip[i], t, T2[i], redo, s0);
          }
          if (lane == 0) *reinterpret_cast<float4 *>(S.dx) = make_float4(dxs[0], dxs[1], dxs[2], dxs[3]);
          __syncthreads();  // 2: the steps are in LDS
          __syncthreads();  // 3: the theta waves' checks are in LDS
          const int ib = first_bad();
          // from voxel ib on, one voxel at a time: theta from the entries'
          // current values, the surrogate from the neighbours' current values
#pragma unroll
          for (int i = 1; i < K; i++) {
            if (i < ib) continue;
            const int s = s0 + i;
            __syncthreads();  // the voxel's sums are in LDS
            const double T1i = sum_t1(0);
            const float T2i = sum_f(0);
            if (apix[s] != kNoPix) {
              float pbi, pbxi, Vi;
              bool ski;
              prior_one(anb[s], pbi, pbxi, Vi, ski);
              const float dx = step_voxel(s, pbi, pbxi, Vi, ski, T1i, T2i, redo, s0);
              if (lane == 0) S.dx[i] = dx;
            }
            __syncthreads();  // its step is in LDS
          }
        }
      }
      // the group's update counts and |x| sums (updateVoxels_Atten's return
      // value, XT_VoxUpdate.c:263-281), lane by lane
      vmacc += (count_mag && rsk == 0) ? fabs((double)rx) : 0.0;
      nupd += (unsigned long long)__popcll(__ballot(rsk == 0));
      nzero += (unsigned long long)__popcll(__ballot(rsk == 1));
      wait_vm_all();
      P.obj[(unsigned)ridx] = rx;
      P.mag[(unsigned)ridx] = rm;
#pragma unroll
      for (int k = 0; k < G; k++) {
        apix[k] = bpix[k];
        anid[k] = bnid[k];
        anb[k] = bnv[k] == bnv[k] ? bnv[k] : bnb[k];
      }
      // B moves to lanes 0..G-1
      pend = __shfl(pend, (lane + G) & 63);
      pend = lane < G ? pend : kNoPix;
#pragma unroll
      for (int k = 0; k < G; k++)
        pnext[k] = (g + 1) * G + k < cnt ? __builtin_amdgcn_readlane(raw, k) : kNoPix;
    }
    const double voxmag = wave_sum_d(vmacc);
    if (lane == 0 && seq != 2) {
      P.cell_voxmag[cell] += voxmag;
      P.cell_upd[cell] += nupd;
      P.cell_zero[cell] += nzero;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && seq != 2) {
    unsigned long long ne = 0;
    for (int w = 0; w < nth; w++) ne += S.nent[w];
    P.cell_ent[cell] += ne;
  }
}

// ---------------------------------------------------------------------------
// Multi-wavefront voxel update for wide columns: the coarsest multires
// stages, whose columns (up to NBX bins x the slice's mz rows, e.g. 17 bins x
// 8 rows at mult 8) exceed the column table's 9 bins.  A lane owns one view
// of the cell and R of the slice's detector rows, for every bin: it computes
// the voxel's column in that view (col_u / col_window / col_value, the
// arithmetic of the column table), reads the NBX x R (e, w) entries of its
// window, adds its theta terms, and after the exchange writes its own
// entries back, so each entry is read and written by one lane only and the
// sequential sweep follows from program order.  The prior wave, the
// exchange and the step are those of k_icd_mw.  The views' detector
// responses sit in LDS.
// ---------------------------------------------------------------------------
template <int R>
__device__ __forceinline__ void ld_r(const float *p, float (&o)[R]) {
  if constexpr (R == 1) {
    o[0] = p[0];
  } else {
    const float2 v = *reinterpret_cast<const float2 *>(p);
    o[0] = v.x;
    o[1] = v.y;
  }
}

template <int NBX, int R, int MAXTH>
__device__ __forceinline__ void k_icd_wide_body(const KP &P, const int2 cl, int iter,
                                               int chunk, int seq, int count_mag, int ngr) {
  __shared__ StepIn S;
  extern __shared__ float s_H[];  // [views of the frame][kHrStride]
  const int f = cl.x, z = cl.y;
  const int cell = f * P.N_z + z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nth = (int)(blockDim.x >> 6) - 1;
  const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;

  const int *list;
  int n0 = 0, cnt = 0;
  if (seq == 0) {
    list = P.ord + (size_t)cell * P.ord_cap;
    n0 = P.ord_off[(size_t)cell * (P.K + 1) + chunk];
    cnt = P.ord_off[(size_t)cell * (P.K + 1) + chunk + 1] - n0;
  } else if (seq == 1) {
    list = P.nh_ord + (size_t)cell * P.nh_cap;
    cnt = P.nh_ord_count[cell];
  } else {
    list = P.single;
    cnt = 1;
  }
  if (cnt <= 0 || nv <= 0) return;
  for (int i = threadIdx.x; i < nv * kHrStride; i += blockDim.x)
    s_H[i] = P.hr[(size_t)P.view_idx[vb + i / kHrStride] * kHrStride + i % kHrStride];
  if (threadIdx.x < 32) {
    (&S.t1[0][0])[threadIdx.x] = 0.0;
    (&S.t2[0][0])[threadIdx.x] = 0.0f;
  }
  __syncthreads();

  if (wave < nth) {
    // ---------------- theta wave ----------------
    const int t = wave * 64 + lane;
    const int q = t / ngr, g = t - q * ngr;
    const bool act = q < nv;
    const int v = P.view_idx[vb + (act ? q : 0)];
    double Au, Bu, Cu;
    col_coef(P, v, Au, Bu, Cu);
    const float *H = s_H + (act ? q : 0) * kHrStride;
    const float *eb = P.e + sidx(P, z, v, 0, g * R);
    const float *wb = P.w + sidx(P, z, v, 0, g * R);
    float *ebw = P.e + sidx(P, z, v, 0, g * R);
    const int *lst = list + n0;
    unsigned long long nent = 0;
    int pix_next = lst[0];
    for (int n = 0; n < cnt; n++) {
      const int par = n & 1;
      const int pix = pix_next;
      pix_next = lst[min(n + 1, cnt - 1)];
      const int yy = pix >> 16, xx = pix & 0xffff;
      const double u = col_u(xx, yy, Au, Bu, Cu);
      int loc, hic;
      const int w0 = act ? col_window(P, u, NBX, loc, hic) : -1;
      const int ws = max(w0, 0);
      float e[NBX][R], w[NBX][R], a[NBX];
#pragma unroll
      for (int j = 0; j < NBX; j++) {
        ld_r<R>(eb + (size_t)(ws + j) * P.mz, e[j]);
        ld_r<R>(wb + (size_t)(ws + j) * P.mz, w[j]);
      }
      if (P.colt && P.colt_rec == col_rec(NBX)) {
        // the stage's NBX-bin column table holds these values (same window,
        // same arithmetic): read them while the (e, w) loads are in flight
        constexpr int REC = col_rec(NBX);
        const float4 *r4 = reinterpret_cast<const float4 *>(
            P.colt + (((size_t)f * P.N_x * P.N_y + (size_t)yy * P.N_x + xx) * P.colt_nv +
                      (act ? q : 0)) * REC);
        float rr[REC];
#pragma unroll
        for (int k = 0; k < REC / 4; k++) {
          const float4 x4 = r4[k];
          rr[4 * k] = x4.x;
          rr[4 * k + 1] = x4.y;
          rr[4 * k + 2] = x4.z;
          rr[4 * k + 3] = x4.w;
        }
#pragma unroll
        for (int j = 0; j < NBX; j++) a[j] = w0 >= 0 ? rr[j] : 0.0f;
      } else {
#pragma unroll
        for (int j = 0; j < NBX; j++) a[j] = w0 >= 0 ? col_value(P, H, u, ws + j, loc, hic) : 0.0f;
      }
      // theta sums (XT_VoxUpdate.c:61-92) with the generalized-Huber weights
      // (XT_VoxUpdate.c:66-80)
      double t1 = 0.0;
      float t2 = 0.0f;
#pragma unroll
      for (int j = 0; j < NBX; j++) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const float ee = e[j][r], ww = w[j][r];
          const float cw = huber_sel(P, ee, ww)
                               ? ww
                               : P.Td * __builtin_amdgcn_sqrtf(ww) * __builtin_amdgcn_rcpf(fabsf(ee));
          const float ac = a[j] * cw;
          t1 = t1 + (double)(-ee * ac);
          t2 = t2 + a[j] * ac;
        }
        nent += a[j] > 0.0f ? R : 0;
      }
      wave_total_df(t1, t2);
      if (lane == 0) {
        S.t1[par][wave] = t1;
        S.t2[par][wave] = t2;
      }
      __syncthreads();
      float V;
      bool skip;
      const float xn = solve_step<MAXTH>(P, S, par, nth, &V, &skip);
      const float dx = xn - V;
      if (dx != 0.0f) {
#pragma unroll
        for (int j = 0; j < NBX; j++) {
          if (a[j] != 0.0f) {
            float o[R];
#pragma unroll
            for (int r = 0; r < R; r++) o[r] = fmaf(-a[j], dx, e[j][r]);
            if constexpr (R == 1) ebw[(size_t)(ws + j) * P.mz] = o[0];
            else *reinterpret_cast<float2 *>(ebw + (size_t)(ws + j) * P.mz) = make_float2(o[0], o[1]);
          }
        }
      }
    }
    const double nsum = wave_sum_d((double)nent);
    if (lane == 0) S.nent[wave] = (unsigned long long)nsum;
  } else {
    // ---------------- prior wave ----------------
    prior_wave<MAXTH>(P, S, list, n0, cnt, f, z, iter, seq, count_mag, nth, lane);
  }
  __syncthreads();
  if (threadIdx.x == 0 && seq != 2) {
    unsigned long long ne = 0;
    for (int w = 0; w < nth; w++) ne += S.nent[w];
    P.cell_ent[cell] += ne;
  }
}

// one problem: the stage's cell list
template <int NBX, int R, int MAXTH>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_wide(KP P, const int2 *__restrict__ cells,
    int iter, int chunk, int seq, int count_mag, int ngr) {
  k_icd_wide_body<NBX, R, MAXTH>(P, cells[blockIdx.x], iter, chunk, seq, count_mag, ngr);
}
// many problems: cells (frame, slice, problem, last argument), parameters
// from the problems' table
template <int NBX, int R, int MAXTH>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_wide_m(const KP *__restrict__ tab,
    const int4 *__restrict__ cells, int iter, int chunk, int seq, int count_mag) {
  const int4 c = cells[blockIdx.x];
  k_icd_wide_body<NBX, R, MAXTH>(tab[c.z], make_int2(c.x, c.y), iter, chunk, seq, count_mag, c.w);
}


// up to 8 theta waves: with more, the 17 x R (e, w) registers spill
template __global__ void k_icd_wide<17, 1, 8>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wide_m<17, 1, 8>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wide<17, 2, 8>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wide_m<17, 2, 8>(const KP *, const int4 *, int, int, int, int);

// ---------------------------------------------------------------------------
// k_icd_split: the wide-column voxel update with each cell's views split
// over S workgroups, for stages whose colours hold few cells (single-frame
// scans' coarse stages: 2 cells per colour at (8,8) leave 254 of 256 CUs
// idle).  Grid: cells x S workgroups, launched cooperatively so that all are
// resident.  Workgroup (cell, s) owns views [s nv / S, (s + 1) nv / S) of
// the cell's frame, with k_icd_wide's lane layout over them; per voxel
//   its theta waves reduce their partial theta sums in LDS (barrier 1); its
//   last wave, which has evaluated the voxel's qGGMRF surrogate from the
//   neighbours, adds them in a fixed tree and publishes the workgroup's
//   partial as a record tagged with the voxel's sequence number; it then
//   reads all S records (lane s polls workgroup s's until its tags match),
//   adds them in a fixed tree over the lanes, solves, stores the voxel and
//   hands the step to its theta waves in LDS (barrier 2), which update
//   their entries.
// Every workgroup of the cell solves every voxel from the same S partials in
// the same order and the same neighbour values, so all find the same step
// bit for bit and all store the same voxel value; the neighbours a
// workgroup reads are the values it stored itself earlier in the launch, or
// values from before it (no other workgroup can have stored a voxel the
// reader has not: each stores voxel n only after every workgroup has
// published its partial of voxel n, which follows its reading of voxel n's
// neighbours).  Zero-skipped voxels are decided from those values alike and
// exchange nothing.
// Every wait is bounded (spin_limit polls) and also ends when the host's
// abort word is nonzero (tb_stage_abort; the reset before each launch
// clears it): then the workgroup records the timeout in err, takes
// dx = 0 for the rest of the sequence and stops waiting, so every wave still
// reaches the end; the error word collects the timeouts of every launch
// until the host reads and clears it.  Records are double-buffered by voxel parity (a workgroup
// publishes voxel n + 2's record only after every workgroup has published
// voxel n + 1's, which each does after it has read voxel n's) and zeroed
// before each launch.
// ---------------------------------------------------------------------------
// The records are written and read only with device-scope atomic 64-bit
// accesses, which go to the device's coherence point past the XCDs' L2s;
// a word carries its own tag, so no ordering between words and no cache-wide
// fence (an L2 invalidate or write-back per poll) is needed.
__global__ void k_split_reset(SplitBuf B, int slots) {
  const size_t n = (size_t)slots * 2 * B.smax * 4;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    B.rec[i] = 0ull;
  // the error word is left alone: it collects every launch's timeouts until
  // the host reads it
  if (blockIdx.x == 0 && threadIdx.x == 0) *const_cast<unsigned *>(B.abort) = 0u;
}

template <int NBX, int R, int MAXTH>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_split(KP P, const int2 *__restrict__ cells,
                                                             int iter, int chunk, int seq,
                                                             int count_mag, int ngr, SplitBuf B) {
  __shared__ StepIn S;
  __shared__ float s_dx[2];
  extern __shared__ float s_H[];  // [views of the segment][kHrStride]
  const int slot = blockIdx.x / B.S, seg = blockIdx.x - slot * B.S;
  const int2 cl = cells[slot];
  const int f = cl.x, z = cl.y;
  const int cell = f * P.N_z + z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nth = (int)(blockDim.x >> 6) - 1;
  const int vf = P.view_off[f], nvf = P.view_off[f + 1] - vf;
  const int qs = (int)((long long)nvf * seg / B.S), qe = (int)((long long)nvf * (seg + 1) / B.S);
  const int vb = vf + qs, nv = qe - qs;

  const int *list;
  int n0 = 0, cnt = 0;
  if (seq == 0) {
    list = P.ord + (size_t)cell * P.ord_cap;
    n0 = P.ord_off[(size_t)cell * (P.K + 1) + chunk];
    cnt = P.ord_off[(size_t)cell * (P.K + 1) + chunk + 1] - n0;
  } else if (seq == 1) {
    list = P.nh_ord + (size_t)cell * P.nh_cap;
    cnt = P.nh_ord_count[cell];
  } else {
    list = P.single;
    cnt = 1;
  }
  if (cnt <= 0) return;  // the same for every workgroup of the cell
  for (int i = threadIdx.x; i < nv * kHrStride; i += blockDim.x)
    s_H[i] = P.hr[(size_t)P.view_idx[vb + i / kHrStride] * kHrStride + i % kHrStride];
  if (threadIdx.x < 32) {
    (&S.t1[0][0])[threadIdx.x] = 0.0;
    (&S.t2[0][0])[threadIdx.x] = 0.0f;
  }
  __syncthreads();
  const int *lst = list + n0;

  if (wave < nth) {
    // ---------------- theta wave (k_icd_wide's, over the segment) ----------------
    // Each lane keeps one view and one row group for the whole sequence, so
    // the entries voxel n updates that voxel n + 1 also reads are this lane's
    // own: with PF, voxel n + 1's rows are loaded while voxel n's exchange
    // runs and those entries are then set to the values this lane stores,
    // which is what a load after the store would return.
    constexpr bool PF = NBX <= 9;
    const int t = wave * 64 + lane;
    const int q = t / ngr, g = t - q * ngr;
    const bool act = q < nv;
    const int v = P.view_idx[nv > 0 ? vb + (act ? q : 0) : vf];  // a segment may hold no view
    double Au, Bu, Cu;
    col_coef(P, v, Au, Bu, Cu);
    const float *H = s_H + (act ? q : 0) * kHrStride;
    const float *eb = P.e + sidx(P, z, v, 0, g * R);
    const float *wb = P.w + sidx(P, z, v, 0, g * R);
    float *ebw = P.e + sidx(P, z, v, 0, g * R);
    unsigned long long nent = 0;
    // the window, (e, w) rows and A values of the voxel at pix
    auto load_rows = [&](int pix, int &ws, float (&e)[NBX][R], float (&w)[NBX][R],
                         float (&a)[NBX]) {
      const int yy = pix >> 16, xx = pix & 0xffff;
      const double u = col_u(xx, yy, Au, Bu, Cu);
      int loc, hic;
      const int w0 = act ? col_window(P, u, NBX, loc, hic) : -1;
      ws = max(w0, 0);
#pragma unroll
      for (int j = 0; j < NBX; j++) {
        ld_r<R>(eb + (size_t)(ws + j) * P.mz, e[j]);
        ld_r<R>(wb + (size_t)(ws + j) * P.mz, w[j]);
      }
      if (P.colt && P.colt_rec == col_rec(NBX)) {
        constexpr int REC = col_rec(NBX);
        const float4 *r4 = reinterpret_cast<const float4 *>(
            P.colt + (((size_t)f * P.N_x * P.N_y + (size_t)yy * P.N_x + xx) * P.colt_nv +
                      (act ? qs + q : 0)) * REC);
        float rr[REC];
#pragma unroll
        for (int k = 0; k < REC / 4; k++) {
          const float4 x4 = r4[k];
          rr[4 * k] = x4.x;
          rr[4 * k + 1] = x4.y;
          rr[4 * k + 2] = x4.z;
          rr[4 * k + 3] = x4.w;
        }
#pragma unroll
        for (int j = 0; j < NBX; j++) a[j] = w0 >= 0 ? rr[j] : 0.0f;
      } else {
#pragma unroll
        for (int j = 0; j < NBX; j++) a[j] = w0 >= 0 ? col_value(P, H, u, ws + j, loc, hic) : 0.0f;
      }
    };
    float ec[NBX][R], wc[NBX][R], ac[NBX];
    int wsc;
    load_rows(lst[0], wsc, ec, wc, ac);
    int pix_next = lst[min(1, cnt - 1)];
    for (int n = 0; n < cnt; n++) {
      const int par = n & 1;
      const bool more = n + 1 < cnt;
      double t1 = 0.0;
      float t2 = 0.0f;
#pragma unroll
      for (int j = 0; j < NBX; j++) {
#pragma unroll
        for (int r = 0; r < R; r++) {
          const float ee = ec[j][r], ww = wc[j][r];
          const float cw = huber_sel(P, ee, ww)
                               ? ww
                               : P.Td * __builtin_amdgcn_sqrtf(ww) * __builtin_amdgcn_rcpf(fabsf(ee));
          const float ac_ = ac[j] * cw;
          t1 = t1 + (double)(-ee * ac_);
          t2 = t2 + ac[j] * ac_;
        }
        nent += ac[j] > 0.0f ? R : 0;
      }
      wave_total_df(t1, t2);
      if (lane == 0) {
        S.t1[par][wave] = t1;
        S.t2[par][wave] = t2;
      }
      __syncthreads();  // 1: the workgroup's partial sums are in LDS
      float en[NBX][R], wn[NBX][R], an[NBX];
      int wsn = 0;
      const int pix1 = pix_next;
      if constexpr (PF) {
        if (more) {
          load_rows(pix1, wsn, en, wn, an);
          pix_next = lst[min(n + 2, cnt - 1)];
        }
      }
      __syncthreads();  // 2: the voxel's step is in LDS
      const float dx = s_dx[par];
      if (dx != 0.0f) {
        float o[NBX][R];
#pragma unroll
        for (int j = 0; j < NBX; j++)
#pragma unroll
          for (int r = 0; r < R; r++) o[j][r] = fmaf(-ac[j], dx, ec[j][r]);
#pragma unroll
        for (int j = 0; j < NBX; j++) {
          if (ac[j] != 0.0f) {
            if constexpr (R == 1) ebw[(size_t)(wsc + j) * P.mz] = o[j][0];
            else *reinterpret_cast<float2 *>(ebw + (size_t)(wsc + j) * P.mz) = make_float2(o[j][0], o[j][1]);
          }
        }
        if constexpr (PF) {
          // entry ws' + j' of the next window is entry wsc + j of this one
          // when j - j' = ws' - wsc
          const int d = wsn - wsc;
          if (more && __ballot(d > -NBX && d < NBX) != 0ull) {
#pragma unroll
            for (int jp = 0; jp < NBX; jp++)
#pragma unroll
              for (int j = 0; j < NBX; j++) {
                const bool m = (d == j - jp) && ac[j] != 0.0f;
#pragma unroll
                for (int r = 0; r < R; r++) en[jp][r] = m ? o[j][r] : en[jp][r];
              }
          }
        }
      }
      if (!more) break;
      if constexpr (!PF) {
        load_rows(pix1, wsn, en, wn, an);
        pix_next = lst[min(n + 2, cnt - 1)];
      }
      wsc = wsn;
#pragma unroll
      for (int j = 0; j < NBX; j++) {
        ac[j] = an[j];
#pragma unroll
        for (int r = 0; r < R; r++) {
          ec[j][r] = en[j][r];
          wc[j][r] = wn[j][r];
        }
      }
    }
    const double nsum = wave_sum_d((double)nent);
    if (lane == 0) S.nent[wave] = (unsigned long long)nsum;
  } else {
    // ---------------- exchange wave: the prior, the exchange, the solve ----------------
    const bool lead = seg == 0;  // the workgroup that keeps the cell's update counts
    const int ln = lane;
    PriorLane C;
    C.ln = ln;
    int ndz = 0, ndy = 0, ndx = 0, ndt = 0;
    C.ng = 0.0f;
    C.nsq = 1.0f;
    C.nisqp = 1.0f;
    C.nc = 1.0f;
    if (ln < 29) {
      if (ln < 27) {
        ndz = ln / 9 - 1;
        ndy = (ln / 3) % 3 - 1;
        ndx = ln % 3 - 1;
        C.ng = P.sfilt[ln];
        C.nsq = P.sSq, C.nisqp = 1.0f / P.sSqp, C.nc = P.c_s;
      } else {
        ndt = ln == 27 ? -1 : 1;
        C.ng = P.tfilt;
        C.nsq = P.sTq, C.nisqp = 1.0f / P.sTqp, C.nc = P.c_t;
      }
    }
    C.two_c = 2.0f * C.nc;
    const int zz = z + ndz, ff = f + ndt;
    const bool zok = ln < 29 && zz >= 0 && zz < P.N_z && ff >= 0 && ff < P.N_time;
    const int zc = min(max(zz, 0), P.N_z - 1), fc = min(max(ff, 0), P.N_time - 1);
    const int nplane = (fc * P.N_z + zc) * P.N_x * P.N_y;
    const int cbase = cell * P.N_x * P.N_y;
    auto nb_load = [&](int pix) -> float {
      const int yy = pix >> 16, xx = pix & 0xffff;
      const int jj = yy + ndy, kk = xx + ndx;
      const bool nok = zok & (jj >= 0) & (jj < P.N_y) & (kk >= 0) & (kk < P.N_x);
      return P.obj[(unsigned)(nok ? nplane + jj * P.N_x + kk : P.nobj + 64)];
    };
    unsigned long long *rec = B.rec + (size_t)slot * 2 * B.smax * 4;
    bool dead = false;
    double vmacc = 0.0;
    unsigned long long nupd = 0, nzero = 0;
    float nbv = nb_load(lst[0]);
    int pix = lst[0], pix1 = lst[min(1, cnt - 1)];
    for (int n = 0; n < cnt; n++) {
      const int par = n & 1;
      const bool more = n + 1 < cnt;
      // the voxel's surrogate (XT_ICD_update.c:134-183) and zero-skip
      // (XT_VoxUpdate.c:236-259), from the neighbour values as they are
      // after the previous voxel's store
      const float V = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(nbv), 13));
      const bool anypos = (__ballot((ln != 13) & (ln < 29) & (nbv > 0.0f)) != 0ull);
      float qb, qbx;
      prior_pair(C, nbv, V, qb, qbx);
      const float pb = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qb), 0));
      const float pbx = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(qbx), 0));
      const bool skip = (V == 0.0f) && iter > 1 && !anypos;
      __syncthreads();  // 1
      // the next voxel's neighbours, loaded while this voxel is exchanged
      // (this voxel, if it is one of them, is set after its solve)
      float nbn = 0.0f;
      int pix2 = pix1;
      if (more) {
        nbn = nb_load(pix1);
        pix2 = lst[min(n + 2, cnt - 1)];
      }
      float xn = V, dx = 0.0f;
      if (!skip) {
        // the workgroup's partial: a fixed tree over the MAXTH wave slots
        double a1[MAXTH];
        float a2[MAXTH];
#pragma unroll
        for (int w = 0; w < MAXTH; w++) {
          a1[w] = S.t1[par][w];
          a2[w] = S.t2[par][w];
        }
#pragma unroll
        for (int m = MAXTH; m > 1; m = (m + 1) / 2) {
          const int h = (m + 1) / 2;
#pragma unroll
          for (int w = 0; w + h < m; w++) {
            a1[w] = a1[w] + a1[w + h];
            a2[w] = a2[w] + a2[w + h];
          }
        }
        // publish: theta1 as a high and a low float (48 significant bits),
        // theta2, each word tagged n + 1
        const unsigned long long tag = (unsigned long long)(unsigned)(n + 1) << 32;
        unsigned long long *rp = rec + (size_t)par * B.smax * 4;
        if (lane < 3) {
          const float hi = (float)a1[0];
          const float lo = (float)(a1[0] - (double)hi);
          const float pay = lane == 0 ? hi : lane == 1 ? lo : a2[0];
          __hip_atomic_store(rp + seg * 4 + lane, tag | __float_as_uint(pay), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
        // every workgroup's record: lane s polls workgroup s's
        unsigned long long r0 = 0, r1 = 0, r2 = 0;
        bool have = lane >= B.S;
        for (unsigned k = 0; !dead; k++) {
          if (!have) {
            const unsigned long long *q = rp + lane * 4;
            r0 = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            r1 = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            r2 = __hip_atomic_load(q + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            have = ((r0 & r1 & r2) >> 32) == (tag >> 32) && ((r0 | r1 | r2) >> 32) == (tag >> 32);
          }
          if (__ballot(!have) == 0ull) break;
          const unsigned ab = __builtin_amdgcn_readfirstlane(
              __hip_atomic_load(B.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
          if (k >= B.spin_limit || ab != 0u) dead = true;
          else __builtin_amdgcn_s_sleep(1);
        }
        double p1 = 0.0;
        float p2 = 0.0f;
        if (!dead && lane < B.S) {
          p1 = (double)__uint_as_float((unsigned)r0) + (double)__uint_as_float((unsigned)r1);
          p2 = __uint_as_float((unsigned)r2);
        }
        // the S partials in a fixed tree over the lanes: the same sums, in
        // the same order, in every workgroup
        wave_total_df(p1, p2);
        if (!dead) {
#pragma clang fp contract(off)
          const float th1 = (float)p1;
          const float num = P.alpha * (pbx - th1);
          const float step = num * __builtin_amdgcn_rcpf(pb + p2);
          xn = V + step;
          dx = xn - V;
        }
      }
      // the object and update map (the same values from every workgroup),
      // then the step for the theta waves
      if (lane == 0) {
        const int lin = cbase + (pix >> 16) * P.N_x + (pix & 0xffff);
        P.obj[lin] = xn;
        P.mag[lin] = skip ? 0.0f : fabsf(dx);
        s_dx[par] = dx;
      }
      if (lead) {
        vmacc += (lane == 0 && count_mag && !skip) ? fabs((double)xn) : 0.0;
        nupd += (lane == 0 && !skip) ? 1ull : 0ull;
        nzero += (lane == 0 && skip) ? 1ull : 0ull;
      }
      if (more && ln >= 9 && ln < 18 && (pix1 >> 16) + ndy == (pix >> 16) &&
          (pix1 & 0xffff) + ndx == (pix & 0xffff))
        nbn = xn;  // an in-plane neighbour: the value just stored
      __syncthreads();  // 2
      nbv = nbn;
      pix = pix1;
      pix1 = pix2;
    }
    if (dead && lane == 0) atomicOr(B.err, 1u);
    const double vm = wave_sum_d(vmacc);
    const double nu = wave_sum_d((double)nupd), nz0 = wave_sum_d((double)nzero);
    if (lead && lane == 0 && seq != 2) {
      P.cell_voxmag[cell] += vm;
      P.cell_upd[cell] += (unsigned long long)nu;
      P.cell_zero[cell] += (unsigned long long)nz0;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && seq != 2) {
    unsigned long long ne = 0;
    for (int w = 0; w < nth; w++) ne += S.nent[w];
    atomicAdd(P.cell_ent + cell, ne);
  }
}


// the split-cell kernel: column widths 17, 9, 5 (two rows per lane) and 3
// (single-row slices), up to 8 theta waves per segment
template __global__ void k_icd_split<17, 2, 8>(KP, const int2 *, int, int, int, int, int, SplitBuf);
template __global__ void k_icd_split<17, 1, 8>(KP, const int2 *, int, int, int, int, int, SplitBuf);
template __global__ void k_icd_split<9, 2, 8>(KP, const int2 *, int, int, int, int, int, SplitBuf);
template __global__ void k_icd_split<9, 1, 8>(KP, const int2 *, int, int, int, int, int, SplitBuf);
template __global__ void k_icd_split<5, 2, 8>(KP, const int2 *, int, int, int, int, int, SplitBuf);
template __global__ void k_icd_split<5, 1, 8>(KP, const int2 *, int, int, int, int, int, SplitBuf);
template __global__ void k_icd_split<3, 1, 8>(KP, const int2 *, int, int, int, int, int, SplitBuf);

// ---------------------------------------------------------------------------
// Multi-wavefront voxel update for wide columns and many views: the coarse
// multires stages whose cells neither the column-table kernels nor
// k_icd_wide can hold (hundreds of views per cell).  Theta wave w owns a
// contiguous block of the cell's views and is the only wave that reads or
// writes their e rows, so the sequential sweep follows from program order,
// as in k_icd_mwg.  Per voxel the wave walks its views' entries in passes
// of 64 lanes: a lane takes one (view, bin of an NBX-bin window, R detector
// rows) and consecutive lanes cover consecutive bins and rows of a view, so
// every (e, w) access of a pass is contiguous per view.  The lane evaluates
// its A value with the column-table arithmetic (col_u / col_window /
// col_value) from per-view coefficients in LDS.  The prior wave, the
// exchange and the step are those of k_icd_mw.
// ---------------------------------------------------------------------------
struct WideView {
  double Au, Bu, Cu;
  int v;
  int pad;
};

template <int NBX, int R, int PMAX, int MAXTH, bool HL>
__device__ __forceinline__ void k_icd_wpass_body(const KP &P, const int2 cl, int iter,
                                                int chunk, int seq, int count_mag, int vpw) {
  __shared__ StepIn S;
  // [views of the frame] coefficients, then (HL) their detector responses
  extern __shared__ WideView s_view[];
  const int f = cl.x, z = cl.y;
  const int cell = f * P.N_z + z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nth = (int)(blockDim.x >> 6) - 1;
  const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;
  const int ngr = P.mz / R;      // row groups of a view
  const int EPV = NBX * ngr;     // lane entries of a view

  const int *list;
  int n0 = 0, cnt = 0;
  if (seq == 0) {
    list = P.ord + (size_t)cell * P.ord_cap;
    n0 = P.ord_off[(size_t)cell * (P.K + 1) + chunk];
    cnt = P.ord_off[(size_t)cell * (P.K + 1) + chunk + 1] - n0;
  } else if (seq == 1) {
    list = P.nh_ord + (size_t)cell * P.nh_cap;
    cnt = P.nh_ord_count[cell];
  } else {
    list = P.single;
    cnt = 1;
  }
  if (cnt <= 0 || nv <= 0) return;
  float *s_H = reinterpret_cast<float *>(s_view + nv);
  // per theta wave: its views' A values at the NBX window bins, then the
  // window start w0 (int bits), NBX + 1 floats per view
  float *s_A = s_H + (HL ? nv * kHrStride : 0);
  for (int q = threadIdx.x; q < nv; q += blockDim.x) {
    WideView w;
    w.v = P.view_idx[vb + q];
    col_coef(P, w.v, w.Au, w.Bu, w.Cu);
    w.pad = q;
    s_view[q] = w;
  }
  if constexpr (HL)
    for (int i = threadIdx.x; i < nv * kHrStride; i += blockDim.x)
      s_H[i] = P.hr[(size_t)P.view_idx[vb + i / kHrStride] * kHrStride + i % kHrStride];
  if (threadIdx.x < 32) {
    (&S.t1[0][0])[threadIdx.x] = 0.0;
    (&S.t2[0][0])[threadIdx.x] = 0.0f;
  }
  __syncthreads();

  if (wave < nth) {
    // ---------------- theta wave ----------------
    const int q0 = wave * vpw, nq = max(0, min(vpw, nv - q0));
    const int npair = nq * EPV;
    const int *lst = list + n0;
    unsigned long long nent = 0;
    int pix_next = lst[0];
    for (int n = 0; n < cnt; n++) {
      const int par = n & 1;
      const int pix = pix_next;
      pix_next = lst[min(n + 1, cnt - 1)];
      const int yy = pix >> 16, xx = pix & 0xffff;
      float e[PMAX][R], w[PMAX][R], a[PMAX];
      int off[PMAX];
      // the wave's views: window and A values once per (view, bin), in LDS
      float *sa = s_A + wave * vpw * (NBX + 1);
      if (P.colt && P.colt_rec == col_rec(NBX)) {
        // the stage's NBX-bin column table holds these values, [A at w0 ..
        // w0 + NBX - 1, w0] per view slot, with the same arithmetic: copy
        constexpr int REC = col_rec(NBX);
        const float *rec = P.colt + (((size_t)f * P.N_x * P.N_y + (size_t)yy * P.N_x + xx) *
                                         P.colt_nv + q0) * REC;
        if constexpr (NBX + 1 == REC) {
          // whole records: one 16-byte load per view
          const float4 *r4 = reinterpret_cast<const float4 *>(rec);
          for (int lq = lane; lq < nq; lq += 64) {
            const float4 r = r4[lq];
            sa[lq * 4 + 0] = r.x;
            sa[lq * 4 + 1] = r.y;
            sa[lq * 4 + 2] = r.z;
            sa[lq * 4 + 3] = r.w;
          }
        } else {
          // 8 loads per lane in flight before their LDS stores
          const int n = nq * (NBX + 1);
          for (int i0 = 0; i0 < n; i0 += 64 * 8) {
            float t[8];
#pragma unroll
            for (int b = 0; b < 8; b++) {
              const int i = min(i0 + b * 64 + lane, n - 1);
              const int lq = i / (NBX + 1);
              t[b] = rec[lq * REC + (i - lq * (NBX + 1))];
            }
#pragma unroll
            for (int b = 0; b < 8; b++) {
              const int i = i0 + b * 64 + lane;
              if (i < n) sa[i] = t[b];
            }
          }
        }
      } else
      for (int i = lane; i < nq * NBX; i += 64) {
        const int lq = i / NBX, j = i - lq * NBX;
        const WideView &wv = s_view[q0 + lq];
        const double u = col_u(xx, yy, wv.Au, wv.Bu, wv.Cu);
        int loc, hic;
        const int w0 = col_window(P, u, NBX, loc, hic);
        const float *H = (HL ? s_H + (size_t)(q0 + lq) * kHrStride
                             : P.hr + (size_t)wv.v * kHrStride);
        sa[lq * (NBX + 1) + j] = w0 >= 0 ? col_value(P, H, u, w0 + j, loc, hic) : 0.0f;
        if (j == 0) sa[lq * (NBX + 1) + NBX] = __int_as_float(w0);
      }
      // passes: this wave's LDS writes above complete before its reads
#pragma unroll
      for (int p = 0; p < PMAX; p++) {
        // this lane's entry of pass p: view q0 + lq, window bin j, row group g
        const int i = p * 64 + lane;
        const bool on = i < npair;
        const int lq = on ? i / EPV : 0, rem = i - lq * EPV;
        const int j = on ? rem / ngr : 0, g = on ? rem - (rem / ngr) * ngr : 0;
        const int v = s_view[on ? q0 + lq : 0].v;
        const int w0 = on ? __float_as_int(sa[lq * (NBX + 1) + NBX]) : -1;
        a[p] = w0 >= 0 ? sa[lq * (NBX + 1) + j] : 0.0f;
        off[p] = (int)(sidx(P, z, v, max(w0, 0) + j, g * R) - sidx(P, z, 0, 0, 0));
        if (p * 64 < npair) {
          ld_r<R>(P.e + sidx(P, z, 0, 0, 0) + off[p], e[p]);
          ld_r<R>(P.w + sidx(P, z, 0, 0, 0) + off[p], w[p]);
        } else {
#pragma unroll
          for (int r = 0; r < R; r++) e[p][r] = w[p][r] = 0.0f;
        }
      }
      // theta sums (XT_VoxUpdate.c:61-92) with the generalized-Huber weights
      // (XT_VoxUpdate.c:66-80)
      double t1 = 0.0;
      float t2 = 0.0f;
      // full unrolling keeps the pass arrays in registers: no early exit
#pragma unroll
      for (int p = 0; p < PMAX; p++) {
        if (p * 64 < npair) {
#pragma unroll
          for (int r = 0; r < R; r++) {
            const float ee = e[p][r], ww = w[p][r];
            const float cw = huber_sel(P, ee, ww)
                                 ? ww
                                 : P.Td * __builtin_amdgcn_sqrtf(ww) *
                                       __builtin_amdgcn_rcpf(fabsf(ee));
            const float ac = a[p] * cw;
            t1 = t1 + (double)(-ee * ac);
            t2 = t2 + a[p] * ac;
          }
          nent += a[p] > 0.0f ? R : 0;
        }
      }
      wave_total_df(t1, t2);
      if (lane == 0) {
        S.t1[par][wave] = t1;
        S.t2[par][wave] = t2;
      }
      __syncthreads();
      float V;
      bool skip;
      const float xn = solve_step<MAXTH>(P, S, par, nth, &V, &skip);
      const float dx = xn - V;
      if (dx != 0.0f) {
#pragma unroll
        for (int p = 0; p < PMAX; p++) {
          if (p * 64 < npair && a[p] != 0.0f) {
            float *q = P.e + sidx(P, z, 0, 0, 0) + off[p];
            if constexpr (R == 1) {
              q[0] = fmaf(-a[p], dx, e[p][0]);
            } else {
              *reinterpret_cast<float2 *>(q) =
                  make_float2(fmaf(-a[p], dx, e[p][0]), fmaf(-a[p], dx, e[p][1]));
            }
          }
        }
      }
    }
    const double nsum = wave_sum_d((double)nent);
    if (lane == 0) S.nent[wave] = (unsigned long long)nsum;
  } else {
    // ---------------- prior wave ----------------
    prior_wave<MAXTH>(P, S, list, n0, cnt, f, z, iter, seq, count_mag, nth, lane);
  }
  __syncthreads();
  if (threadIdx.x == 0 && seq != 2) {
    unsigned long long ne = 0;
    for (int w = 0; w < nth; w++) ne += S.nent[w];
    P.cell_ent[cell] += ne;
  }
}

// one problem: the stage's cell list
template <int NBX, int R, int PMAX, int MAXTH, bool HL>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_wpass(KP P, const int2 *__restrict__ cells,
    int iter, int chunk, int seq, int count_mag, int vpw) {
  k_icd_wpass_body<NBX, R, PMAX, MAXTH, HL>(P, cells[blockIdx.x], iter, chunk, seq, count_mag, vpw);
}
// many problems: cells (frame, slice, problem, last argument), parameters
// from the problems' table
template <int NBX, int R, int PMAX, int MAXTH, bool HL>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_wpass_m(const KP *__restrict__ tab,
    const int4 *__restrict__ cells, int iter, int chunk, int seq, int count_mag) {
  const int4 c = cells[blockIdx.x];
  k_icd_wpass_body<NBX, R, PMAX, MAXTH, HL>(tab[c.z], make_int2(c.x, c.y), iter, chunk, seq, count_mag, c.w);
}


// ---------------------------------------------------------------------------
// k_icd_wloop: k_icd_wpass for any number of passes.  The wave walks its
// entries in groups of GP passes (one group's e, w and A values in
// registers at a time) and, after the step, reads e again to write
// e - a dx.  No other wave writes the wave's entries between the two reads:
// the columns of one voxel in one view are disjoint across views and each
// (view, bin, row group) entry has one lane.  Serves the coarse stages of
// scans with many views per frame (more passes than k_icd_wpass keeps in
// registers); the detector responses stay in global memory.
// ---------------------------------------------------------------------------
template <int NBX, int R, int GP, int MAXTH, bool HL>
__device__ __forceinline__ void k_icd_wloop_body(const KP &P, const int2 cl, int iter,
                                                int chunk, int seq, int count_mag, int vpw) {
  __shared__ StepIn S;
  // [views of the frame] coefficients, then (HL) their detector responses
  extern __shared__ WideView s_view[];
  const int f = cl.x, z = cl.y;
  const int cell = f * P.N_z + z;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nth = (int)(blockDim.x >> 6) - 1;
  const int vb = P.view_off[f], nv = P.view_off[f + 1] - vb;
  const int ngr = P.mz / R;      // row groups of a view
  const int EPV = NBX * ngr;     // lane entries of a view

  const int *list;
  int n0 = 0, cnt = 0;
  if (seq == 0) {
    list = P.ord + (size_t)cell * P.ord_cap;
    n0 = P.ord_off[(size_t)cell * (P.K + 1) + chunk];
    cnt = P.ord_off[(size_t)cell * (P.K + 1) + chunk + 1] - n0;
  } else if (seq == 1) {
    list = P.nh_ord + (size_t)cell * P.nh_cap;
    cnt = P.nh_ord_count[cell];
  } else {
    list = P.single;
    cnt = 1;
  }
  if (cnt <= 0 || nv <= 0) return;
  float *s_H = reinterpret_cast<float *>(s_view + nv);
  // per theta wave: its views' A values at the NBX window bins, then the
  // window start w0 (int bits), NBX + 1 floats per view
  float *s_A = s_H + (HL ? nv * kHrStride : 0);
  for (int q = threadIdx.x; q < nv; q += blockDim.x) {
    WideView w;
    w.v = P.view_idx[vb + q];
    col_coef(P, w.v, w.Au, w.Bu, w.Cu);
    w.pad = q;
    s_view[q] = w;
  }
  if constexpr (HL)
    for (int i = threadIdx.x; i < nv * kHrStride; i += blockDim.x)
      s_H[i] = P.hr[(size_t)P.view_idx[vb + i / kHrStride] * kHrStride + i % kHrStride];
  if (threadIdx.x < 32) {
    (&S.t1[0][0])[threadIdx.x] = 0.0;
    (&S.t2[0][0])[threadIdx.x] = 0.0f;
  }
  __syncthreads();

  if (wave < nth) {
    // ---------------- theta wave ----------------
    const int q0 = wave * vpw, nq = max(0, min(vpw, nv - q0));
    const int npair = nq * EPV;
    const int *lst = list + n0;
    unsigned long long nent = 0;
    int pix_next = lst[0];
    for (int n = 0; n < cnt; n++) {
      const int par = n & 1;
      const int pix = pix_next;
      pix_next = lst[min(n + 1, cnt - 1)];
      const int yy = pix >> 16, xx = pix & 0xffff;
      // the wave's views: window and A values once per (view, bin), in LDS
      float *sa = s_A + wave * vpw * (NBX + 1);
      if (P.colt && P.colt_rec == col_rec(NBX)) {
        // the stage's NBX-bin column table holds these values, [A at w0 ..
        // w0 + NBX - 1, w0] per view slot, with the same arithmetic: copy
        constexpr int REC = col_rec(NBX);
        const float *rec = P.colt + (((size_t)f * P.N_x * P.N_y + (size_t)yy * P.N_x + xx) *
                                         P.colt_nv + q0) * REC;
        if constexpr (NBX + 1 == REC) {
          // whole records: one 16-byte load per view
          const float4 *r4 = reinterpret_cast<const float4 *>(rec);
          for (int lq = lane; lq < nq; lq += 64) {
            const float4 r = r4[lq];
            sa[lq * 4 + 0] = r.x;
            sa[lq * 4 + 1] = r.y;
            sa[lq * 4 + 2] = r.z;
            sa[lq * 4 + 3] = r.w;
          }
        } else {
          // 8 loads per lane in flight before their LDS stores
          const int n = nq * (NBX + 1);
          for (int i0 = 0; i0 < n; i0 += 64 * 8) {
            float t[8];
#pragma unroll
            for (int b = 0; b < 8; b++) {
              const int i = min(i0 + b * 64 + lane, n - 1);
              const int lq = i / (NBX + 1);
              t[b] = rec[lq * REC + (i - lq * (NBX + 1))];
            }
#pragma unroll
            for (int b = 0; b < 8; b++) {
              const int i = i0 + b * 64 + lane;
              if (i < n) sa[i] = t[b];
            }
          }
        }
      } else
      for (int i = lane; i < nq * NBX; i += 64) {
        const int lq = i / NBX, j = i - lq * NBX;
        const WideView &wv = s_view[q0 + lq];
        const double u = col_u(xx, yy, wv.Au, wv.Bu, wv.Cu);
        int loc, hic;
        const int w0 = col_window(P, u, NBX, loc, hic);
        const float *H = (HL ? s_H + (size_t)(q0 + lq) * kHrStride
                             : P.hr + (size_t)wv.v * kHrStride);
        sa[lq * (NBX + 1) + j] = w0 >= 0 ? col_value(P, H, u, w0 + j, loc, hic) : 0.0f;
        if (j == 0) sa[lq * (NBX + 1) + NBX] = __int_as_float(w0);
      }
      // entry i of the wave: view q0 + lq, window bin j, row group g
      auto entry = [&](int i, float &a, int &off) {
        const bool on = i < npair;
        const int lq = on ? i / EPV : 0, rem = i - lq * EPV;
        const int j = on ? rem / ngr : 0, g = on ? rem - (rem / ngr) * ngr : 0;
        const int v = s_view[on ? q0 + lq : 0].v;
        const int w0 = on ? __float_as_int(sa[lq * (NBX + 1) + NBX]) : -1;
        a = w0 >= 0 ? sa[lq * (NBX + 1) + j] : 0.0f;
        off = (int)(sidx(P, z, v, max(w0, 0) + j, g * R) - sidx(P, z, 0, 0, 0));
      };
      const float *eb = P.e + sidx(P, z, 0, 0, 0);
      const float *wb = P.w + sidx(P, z, 0, 0, 0);
      // theta sums (XT_VoxUpdate.c:61-92) with the generalized-Huber weights
      // (XT_VoxUpdate.c:66-80), GP passes of 64 entries at a time
      double t1 = 0.0;
      float t2 = 0.0f;
      for (int base = 0; base < npair; base += 64 * GP) {
        float e[GP][R], w[GP][R], a[GP];
        int off[GP];
#pragma unroll
        for (int p = 0; p < GP; p++) {
          entry(base + p * 64 + lane, a[p], off[p]);
          ld_r<R>(eb + off[p], e[p]);
          ld_r<R>(wb + off[p], w[p]);
        }
#pragma unroll
        for (int p = 0; p < GP; p++) {
#pragma unroll
          for (int r = 0; r < R; r++) {
            const float ee = e[p][r], ww = w[p][r];
            const float cw = huber_sel(P, ee, ww)
                                 ? ww
                                 : P.Td * __builtin_amdgcn_sqrtf(ww) *
                                       __builtin_amdgcn_rcpf(fabsf(ee));
            const float ac = a[p] * cw;
            t1 = t1 + (double)(-ee * ac);
            t2 = t2 + a[p] * ac;
          }
          nent += a[p] > 0.0f ? R : 0;
        }
      }
      wave_total_df(t1, t2);
      if (lane == 0) {
        S.t1[par][wave] = t1;
        S.t2[par][wave] = t2;
      }
      __syncthreads();
      float V;
      bool skip;
      const float xn = solve_step<MAXTH>(P, S, par, nth, &V, &skip);
      const float dx = xn - V;
      if (dx != 0.0f) {
        // this wave read those entries above and no other wave writes them:
        // read them again and write e - a dx
        for (int base = 0; base < npair; base += 64 * GP) {
          float e[GP][R], a[GP];
          int off[GP];
#pragma unroll
          for (int p = 0; p < GP; p++) {
            entry(base + p * 64 + lane, a[p], off[p]);
            ld_r<R>(eb + off[p], e[p]);
          }
#pragma unroll
          for (int p = 0; p < GP; p++) {
            if (a[p] != 0.0f) {
              float *q = P.e + sidx(P, z, 0, 0, 0) + off[p];
              if constexpr (R == 1) {
                q[0] = fmaf(-a[p], dx, e[p][0]);
              } else {
                *reinterpret_cast<float2 *>(q) =
                    make_float2(fmaf(-a[p], dx, e[p][0]), fmaf(-a[p], dx, e[p][1]));
              }
            }
          }
        }
      }
    }
    const double nsum = wave_sum_d((double)nent);
    if (lane == 0) S.nent[wave] = (unsigned long long)nsum;
  } else {
    // ---------------- prior wave ----------------
    prior_wave<MAXTH>(P, S, list, n0, cnt, f, z, iter, seq, count_mag, nth, lane);
  }
  __syncthreads();
  if (threadIdx.x == 0 && seq != 2) {
    unsigned long long ne = 0;
    for (int w = 0; w < nth; w++) ne += S.nent[w];
    P.cell_ent[cell] += ne;
  }
}

// one problem: the stage's cell list
template <int NBX, int R, int GP, int MAXTH, bool HL>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_wloop(KP P, const int2 *__restrict__ cells,
    int iter, int chunk, int seq, int count_mag, int vpw) {
  k_icd_wloop_body<NBX, R, GP, MAXTH, HL>(P, cells[blockIdx.x], iter, chunk, seq, count_mag, vpw);
}
// many problems: cells (frame, slice, problem, last argument), parameters
// from the problems' table
template <int NBX, int R, int GP, int MAXTH, bool HL>
__global__ __launch_bounds__(64 * (MAXTH + 1)) void k_icd_wloop_m(const KP *__restrict__ tab,
    const int4 *__restrict__ cells, int iter, int chunk, int seq, int count_mag) {
  const int4 c = cells[blockIdx.x];
  k_icd_wloop_body<NBX, R, GP, MAXTH, HL>(tab[c.z], make_int2(c.x, c.y), iter, chunk, seq, count_mag, c.w);
}


// 15 theta waves: a group of 8 passes fits the registers of 4 waves per SIMD
template __global__ void k_icd_wloop<17, 2, 8, 15, false>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wloop_m<17, 2, 8, 15, false>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wloop<9, 2, 8, 15, false>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wloop_m<9, 2, 8, 15, false>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wloop<5, 2, 8, 15, false>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wloop_m<5, 2, 8, 15, false>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wloop<3, 1, 8, 15, false>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wloop_m<3, 1, 8, 15, false>(const KP *, const int4 *, int, int, int, int);

// 11 theta waves keep the passes' registers unspilled
template __global__ void k_icd_wpass<17, 2, 16, 11, false>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wpass_m<17, 2, 16, 11, false>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wpass<17, 2, 24, 11, false>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wpass_m<17, 2, 24, 11, false>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wpass<9, 2, 16, 11, false>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wpass_m<9, 2, 16, 11, false>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wpass<5, 2, 16, 11, false>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wpass_m<5, 2, 16, 11, false>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wpass<17, 2, 16, 11, true>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wpass_m<17, 2, 16, 11, true>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wpass<17, 2, 24, 11, true>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wpass_m<17, 2, 24, 11, true>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wpass<9, 2, 16, 11, true>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wpass_m<9, 2, 16, 11, true>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wpass<5, 2, 16, 11, true>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wpass_m<5, 2, 16, 11, true>(const KP *, const int4 *, int, int, int, int);
// single-row slices (the finest stage) with more views than k_icd_mwg holds
template __global__ void k_icd_wpass<3, 1, 16, 11, false>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wpass_m<3, 1, 16, 11, false>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_wpass<3, 1, 16, 11, true>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_wpass_m<3, 1, 16, 11, true>(const KP *, const int4 *, int, int, int, int);

// the deep-pipeline global-e kernel: 3 voxels' rows in flight, column
// entries 4 voxels further (the finest and second-finest stages; with two
// views per lane the rings would spill); 1 and 7 rows for comparison
template __global__ void k_icd_mwd<3, 1, 1, 8, 3, 4>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwd<3, 1, 1, 15, 3, 4>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwd<5, 2, 1, 8, 3, 4>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwd<5, 2, 1, 15, 3, 4>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwd<3, 1, 1, 8, 1, 2>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwd<3, 1, 1, 8, 7, 1>(KP, const int2 *, int, int, int, int, int);
// the batched finest-stage kernel: up to 8 or 15 theta waves
template __global__ void k_icd_bat<8>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_bat<15>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<3, 1, 1, 8, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<3, 1, 1, 8, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<3, 1, 1, 15, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<3, 1, 1, 15, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 1, 8, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<5, 2, 1, 8, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 1, 15, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<5, 2, 1, 15, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<9, 4, 1, 15, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<9, 4, 1, 15, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<9, 4, 2, 15, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<9, 4, 2, 15, 0>(const KP *, const int4 *, int, int, int, int);
// more views per lane for cells with many views (single-frame scans); with
// 4 views per lane the registers fit only 11 theta waves without spilling
template __global__ void k_icd_mwg<3, 1, 2, 15, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<3, 1, 2, 15, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<3, 1, 4, 11, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<3, 1, 4, 11, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 2, 15, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<5, 2, 2, 15, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 4, 11, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<5, 2, 4, 11, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<9, 2, 1, 15, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<9, 2, 1, 15, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<9, 2, 2, 15, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<9, 2, 2, 15, 0>(const KP *, const int4 *, int, int, int, int);
template __global__ void k_icd_mwg<9, 2, 4, 11, 0>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg_m<9, 2, 4, 11, 0>(const KP *, const int4 *, int, int, int, int);

// FP32 theta sums, compensated (TF 1) and plain (TF 2), for the 3- and 5-bin
// columns (the finest and second-finest stages)
template __global__ void k_icd_mwg<3, 1, 1, 8, 1>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<3, 1, 1, 15, 1>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<3, 1, 2, 15, 1>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<3, 1, 4, 11, 1>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 1, 8, 1>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 1, 15, 1>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 2, 15, 1>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 4, 11, 1>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<3, 1, 1, 8, 2>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<3, 1, 1, 15, 2>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<3, 1, 2, 15, 2>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<3, 1, 4, 11, 2>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 1, 8, 2>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 1, 15, 2>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 2, 15, 2>(KP, const int2 *, int, int, int, int, int);
template __global__ void k_icd_mwg<5, 2, 4, 11, 2>(KP, const int2 *, int, int, int, int, int);

// finest stage (m_xy = m_z = 1: at most 3 bins per view, one detector row)
template __global__ void k_icd_lds<1, 3, 1, 4>(KP, const int2 *, int, int, int, int);
template __global__ void k_icd_lds<2, 3, 1, 4>(KP, const int2 *, int, int, int, int);
template __global__ void k_icd_lds<4, 3, 1, 2>(KP, const int2 *, int, int, int, int);
// second finest (m_xy = m_z = 2: at most 5 bins, two rows)
template __global__ void k_icd_lds<1, 5, 2, 4>(KP, const int2 *, int, int, int, int);
template __global__ void k_icd_lds<2, 5, 2, 2>(KP, const int2 *, int, int, int, int);

}  // namespace tb
