// Kernels of the learned initialiser: a 3D U-Net whose convolutions run on
// gfx950's BF16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32 accumulate).
//
// The reference has no counterpart: its stage 0 starts from zero
// (XT_ICD_update.c:549-553) and its initICD = 1 path reads an object from a
// file without interpolation (XT_ICD_update.c:558-571).  The network's output
// enters the ICD stage the same way, as the stage's starting object.
//
// Layouts.  Activations and their gradients are bf16, channels last:
// [N][D][H][W][C] with C a multiple of 32.  Convolution weights are bf16
// [27][Cout][Cin] (tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1)), so one
// lane's eight consecutive input channels are one 16-byte load.  Biases,
// master weights, optimiser state and weight gradients are fp32.  Every
// 3x3x3 convolution pads with zeros ("same" size).
//
// Matrix-core lane layout of v_mfma_f32_16x16x32_bf16 (64 lanes, l = lane):
//   A (16 x 32, M x K): lane l holds A[l % 16][8 (l / 16) + i], i = 0..7
//   B (32 x 16, K x N): lane l holds B[8 (l / 16) + i][l % 16]
//   D (16 x 16):        lane l holds D[4 (l / 16) + j][l % 16], j = 0..3
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "timbir_dev.h"

namespace tbnn {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kTaps = 27;
constexpr int kWavesPerBlock = 4;

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 zero8() {
  bf16x8 z;
#pragma unroll
  for (int i = 0; i < 8; i++) z[i] = (bf16)0.0f;
  return z;
}

// ---------------------------------------------------------------------------
// 3x3x3 convolution, forward: out[v][co] = bias[co] + sum_tap sum_ci
// in[v + off(tap)][ci] w[tap][co][ci], then ReLU when relu != 0.  A wave owns
// 16 MB voxels of one (n, d, h) row and 16 NB output channels; its K loop runs
// over the 27 taps and the input channels in chunks of 32.  Taps whose
// (d, h) neighbour row lies outside the volume are skipped (uniform over the
// wave); voxels whose x neighbour lies outside read zeros.
// ---------------------------------------------------------------------------
template <int MB, int NB>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_conv3d_fwd(
    const bf16 *__restrict__ in, const bf16 *__restrict__ w, const float *__restrict__ bias,
    void *__restrict__ out, int N, int D, int H, int W, int Cin, int Cout, int relu,
    int out_f32) {
  const int lane = threadIdx.x & 63;
  const long long g = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const int WT = (W + 16 * MB - 1) / (16 * MB);
  const int CT = Cout / (16 * NB);
  const long long total = (long long)N * D * H * WT * CT;
  if (g >= total) return;
  long long r = g;
  const int ct = (int)(r % CT); r /= CT;
  const int wt = (int)(r % WT); r /= WT;
  const int h = (int)(r % H); r /= H;
  const int d = (int)(r % D);
  const int n = (int)(r / D);
  const int row = lane & 15, kq = lane >> 4;
  const int w0 = wt * 16 * MB, co0 = ct * 16 * NB;
  f32x4 acc[MB][NB];
#pragma unroll
  for (int a = 0; a < MB; a++)
#pragma unroll
    for (int b = 0; b < NB; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int nchunk = Cin / 32;
  for (int tap = 0; tap < kTaps; tap++) {
    const int zz = d + tap / 9 - 1, yy = h + (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    if (zz < 0 || zz >= D || yy < 0 || yy >= H) continue;
    const bf16 *inrow = in + (((size_t)n * D + zz) * H + yy) * (size_t)W * Cin;
    const bf16 *wt_tap = w + (size_t)tap * Cout * Cin;
    for (int c = 0; c < nchunk; c++) {
      bf16x8 av[MB], bv[NB];
#pragma unroll
      for (int a = 0; a < MB; a++) {
        const int xx = w0 + 16 * a + row + dx;
        av[a] = (xx >= 0 && xx < W)
                    ? *reinterpret_cast<const bf16x8 *>(inrow + (size_t)xx * Cin + c * 32 + kq * 8)
                    : zero8();
      }
#pragma unroll
      for (int b = 0; b < NB; b++)
        bv[b] = *reinterpret_cast<const bf16x8 *>(wt_tap + (size_t)(co0 + 16 * b + row) * Cin +
                                                  c * 32 + kq * 8);
#pragma unroll
      for (int a = 0; a < MB; a++)
#pragma unroll
        for (int b = 0; b < NB; b++) acc[a][b] = mfma(av[a], bv[b], acc[a][b]);
    }
  }
  const size_t orow = (((size_t)n * D + d) * H + h) * (size_t)W;
#pragma unroll
  for (int b = 0; b < NB; b++) {
    const int co = co0 + 16 * b + row;
    const float bb = bias ? bias[co] : 0.f;
#pragma unroll
    for (int a = 0; a < MB; a++)
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int x = w0 + 16 * a + 4 * kq + j;
        if (x >= W) continue;
        float v = acc[a][b][j] + bb;
        if (relu) v = fmaxf(v, 0.f);
        const size_t o = (orow + x) * Cout + co;
        if (out_f32) static_cast<float *>(out)[o] = v;
        else static_cast<bf16 *>(out)[o] = (bf16)v;
      }
  }
}

// wt[t][ci][co] = w[26 - t][co][ci]: the weights of the gradient with respect
// to the input, which is the forward convolution of the output gradient with
// the spatially flipped, transposed kernel
__global__ void k_conv_flip(const bf16 *__restrict__ w, bf16 *__restrict__ wt, int Cin, int Cout) {
  const size_t n = (size_t)kTaps * Cin * Cout;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int co = (int)(i % Cout);
    const int ci = (int)((i / Cout) % Cin);
    const int t = (int)(i / ((size_t)Cout * Cin));
    wt[i] = w[((size_t)(kTaps - 1 - t) * Cout + co) * Cin + ci];
  }
}

// ---------------------------------------------------------------------------
// 3x3x3 convolution, weight gradient: gw[tap][co][ci] = sum_v gout[v][co]
// in[v + off(tap)][ci].  M = co, N = ci, K = voxels in chunks of 32.  A wave
// owns one tap, 16 MBO output and 16 NBI input channels and one of S slices
// of the voxels; it writes its partial sums to part[s][tap][co][ci] (the
// reduction over s is a separate pass, so the result does not depend on the
// order waves finish).
// ---------------------------------------------------------------------------
template <int MBO, int NBI>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_conv3d_wgrad(
    const bf16 *__restrict__ in, const bf16 *__restrict__ gout, float *__restrict__ part,
    int N, int D, int H, int W, int Cin, int Cout, int S) {
  const int lane = threadIdx.x & 63;
  const long long g = (long long)blockIdx.x * kWavesPerBlock + (threadIdx.x >> 6);
  const int OT = Cout / (16 * MBO), IT = Cin / (16 * NBI);
  const long long total = (long long)S * kTaps * OT * IT;
  if (g >= total) return;
  long long r = g;
  const int it = (int)(r % IT); r /= IT;
  const int ot = (int)(r % OT); r /= OT;
  const int tap = (int)(r % kTaps);
  const int s = (int)(r / kTaps);
  const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
  const int row = lane & 15, kq = lane >> 4;
  const int co0 = ot * 16 * MBO, ci0 = it * 16 * NBI;
  const long long nvox = (long long)N * D * H * W;
  const long long nch = (nvox + 31) / 32;
  const long long c_begin = nch * s / S, c_end = nch * (s + 1) / S;
  f32x4 acc[MBO][NBI];
#pragma unroll
  for (int a = 0; a < MBO; a++)
#pragma unroll
    for (int b = 0; b < NBI; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (long long c = c_begin; c < c_end; c++) {
    bf16x8 av[MBO], bv[NBI];
#pragma unroll
    for (int a = 0; a < MBO; a++) av[a] = zero8();
#pragma unroll
    for (int b = 0; b < NBI; b++) bv[b] = zero8();
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const long long v = c * 32 + kq * 8 + i;
      if (v >= nvox) break;
      long long q = v;
      const int x = (int)(q % W); q /= W;
      const int y = (int)(q % H); q /= H;
      const int z = (int)(q % D);
      const int nn = (int)(q / D);
#pragma unroll
      for (int a = 0; a < MBO; a++) av[a][i] = gout[(size_t)v * Cout + co0 + 16 * a + row];
      const int zz = z + dz, yy = y + dy, xx = x + dx;
      if (zz >= 0 && zz < D && yy >= 0 && yy < H && xx >= 0 && xx < W) {
        const size_t u = (((size_t)nn * D + zz) * H + yy) * (size_t)W + xx;
#pragma unroll
        for (int b = 0; b < NBI; b++) bv[b][i] = in[u * Cin + ci0 + 16 * b + row];
      }
    }
#pragma unroll
    for (int a = 0; a < MBO; a++)
#pragma unroll
      for (int b = 0; b < NBI; b++) acc[a][b] = mfma(av[a], bv[b], acc[a][b]);
  }
  float *dst = part + ((size_t)s * kTaps + tap) * Cout * Cin;
#pragma unroll
  for (int a = 0; a < MBO; a++)
#pragma unroll
    for (int b = 0; b < NBI; b++)
#pragma unroll
      for (int j = 0; j < 4; j++)
        dst[(size_t)(co0 + 16 * a + 4 * kq + j) * Cin + ci0 + 16 * b + row] = acc[a][b][j];
}

// out[i] = sum_s part[s][i], s in order
__global__ void k_sum_parts(const float *__restrict__ part, float *__restrict__ out, size_t n,
                            int S) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    float acc = 0.f;
    for (int s = 0; s < S; s++) acc += part[(size_t)s * n + i];
    out[i] = acc;
  }
}

// part[s][c] = sum over the voxels of slice s of g[v][c] (bias gradient)
__global__ void k_channel_sum(const bf16 *__restrict__ g, float *__restrict__ part,
                              long long nvox, int C, int S) {
  const int s = blockIdx.x;
  const long long v0 = nvox * s / S, v1 = nvox * (s + 1) / S;
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    float acc = 0.f;
    for (long long v = v0; v < v1; v++) acc += (float)g[(size_t)v * C + c];
    part[(size_t)s * C + c] = acc;
  }
}

// g[i] = y[i] > 0 ? g[i] : 0 (ReLU backward through the stored output)
__global__ void k_relu_bwd(const bf16 *__restrict__ y, bf16 *__restrict__ g, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    if (!((float)y[i] > 0.f)) g[i] = (bf16)0.f;
}

// 2x2x2 average pooling, forward and backward
__global__ void k_avgpool2(const bf16 *__restrict__ in, bf16 *__restrict__ out, int N, int D,
                           int H, int W, int C) {
  const int Do = D / 2, Ho = H / 2, Wo = W / 2;
  const size_t n = (size_t)N * Do * Ho * Wo * C;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    size_t q = i;
    const int c = (int)(q % C); q /= C;
    const int x = (int)(q % Wo); q /= Wo;
    const int y = (int)(q % Ho); q /= Ho;
    const int z = (int)(q % Do);
    const int nn = (int)(q / Do);
    float s = 0.f;
    for (int k = 0; k < 8; k++) {
      const size_t u = (((size_t)nn * D + 2 * z + (k >> 2)) * H + 2 * y + ((k >> 1) & 1)) * W +
                       2 * x + (k & 1);
      s += (float)in[u * C + c];
    }
    out[i] = (bf16)(s * 0.125f);
  }
}

__global__ void k_avgpool2_bwd(const bf16 *__restrict__ gout, bf16 *__restrict__ gin, int N,
                               int D, int H, int W, int C) {
  const size_t n = (size_t)N * D * H * W * C;
  const int Do = D / 2, Ho = H / 2, Wo = W / 2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    size_t q = i;
    const int c = (int)(q % C); q /= C;
    const int x = (int)(q % W); q /= W;
    const int y = (int)(q % H); q /= H;
    const int z = (int)(q % D);
    const int nn = (int)(q / D);
    const size_t p = (((size_t)nn * Do + z / 2) * Ho + y / 2) * Wo + x / 2;
    gin[i] = (bf16)((float)gout[p * C + c] * 0.125f);
  }
}

// out[v][0:Ca] = a[parent(v)][:] (nearest x2 upsampling of the coarse
// features), out[v][Ca:Ca+Cb] = b[v][:] (the skip connection)
__global__ void k_up2_concat(const bf16 *__restrict__ a, const bf16 *__restrict__ b,
                             bf16 *__restrict__ out, int N, int D, int H, int W, int Ca, int Cb) {
  const int C = Ca + Cb;
  const size_t n = (size_t)N * D * H * W * C;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    size_t q = i;
    const int c = (int)(q % C); q /= C;
    const size_t v = q;
    const int x = (int)(q % W); q /= W;
    const int y = (int)(q % H); q /= H;
    const int z = (int)(q % D);
    const int nn = (int)(q / D);
    if (c < Ca) {
      const size_t p = (((size_t)nn * (D / 2) + z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2;
      out[i] = a[p * Ca + c];
    } else {
      out[i] = b[v * Cb + (c - Ca)];
    }
  }
}

// backward of k_up2_concat: ga[p][c] = sum over the 8 children of g[v][c],
// gb[v][c] = g[v][Ca + c]
__global__ void k_up2_concat_bwd(const bf16 *__restrict__ g, bf16 *__restrict__ ga,
                                 bf16 *__restrict__ gb, int N, int D, int H, int W, int Ca,
                                 int Cb) {
  const int C = Ca + Cb;
  const int Dc = D / 2, Hc = H / 2, Wc = W / 2;
  const size_t na = (size_t)N * Dc * Hc * Wc * Ca;
  const size_t nb = (size_t)N * D * H * W * Cb;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < na + nb;
       i += (size_t)gridDim.x * blockDim.x) {
    if (i < na) {
      size_t q = i;
      const int c = (int)(q % Ca); q /= Ca;
      const int x = (int)(q % Wc); q /= Wc;
      const int y = (int)(q % Hc); q /= Hc;
      const int z = (int)(q % Dc);
      const int nn = (int)(q / Dc);
      float s = 0.f;
      for (int k = 0; k < 8; k++) {
        const size_t u = (((size_t)nn * D + 2 * z + (k >> 2)) * H + 2 * y + ((k >> 1) & 1)) * W +
                         2 * x + (k & 1);
        s += (float)g[u * C + c];
      }
      ga[i] = (bf16)s;
    } else {
      const size_t j = i - na;
      const int c = (int)(j % Cb);
      const size_t v = j / Cb;
      gb[j] = g[v * C + Ca + c];
    }
  }
}

// network input: x[v][0] = s * fbp[v], x[v][1] = s * mean over frames of
// fbp[v], channels 2..31 zero; the volumes are fp32 [N][D][H][W] (one
// sample per frame) and mean is [D][H][W] (shared by all samples)
__global__ void k_make_input(const float *__restrict__ fbp, const float *__restrict__ mean,
                             bf16 *__restrict__ x, long long nvox, long long per, float scale) {
  const size_t n = (size_t)nvox * 32;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % 32);
    const size_t v = i / 32;
    float val = 0.f;
    if (c == 0) val = scale * fbp[v];
    else if (c == 1 && mean) val = scale * mean[v % per];
    x[i] = (bf16)val;
  }
}

// mean[q] = sum over frames f of x[f * per + q] / nframes (fp32, frames in
// order): the temporal mean of a 4D object, the network's second input
__global__ void k_frame_mean(const float *__restrict__ x, float *__restrict__ mean,
                             long long nframes, long long per) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < (size_t)per;
       q += (size_t)gridDim.x * blockDim.x) {
    float s = 0.f;
    for (long long f = 0; f < nframes; f++) s += x[(size_t)f * per + q];
    mean[q] = s / (float)nframes;
  }
}

// output head: out[v] = fbp[v] + (b + sum_c w[c] h[v][c]) / scale (a 1x1x1
// convolution to one channel, a residual on the filtered back-projection)
__global__ void k_head_fwd(const bf16 *__restrict__ h, const float *__restrict__ w,
                           const float *__restrict__ b, const float *__restrict__ fbp,
                           float *__restrict__ out, long long nvox, int C, float inv_scale) {
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < (size_t)nvox;
       v += (size_t)gridDim.x * blockDim.x) {
    float s = b[0];
    for (int c = 0; c < C; c++) s += w[c] * (float)h[v * C + c];
    out[v] = fbp[v] + s * inv_scale;
  }
}

// mean squared error over nvox voxels: part[block] = partial sums of
// (out - target)^2; gout[v] = 2 (out - target) / nvox (fp32)
__global__ void k_mse(const float *__restrict__ out, const float *__restrict__ target,
                      float *__restrict__ gout, double *__restrict__ part, long long nvox) {
  __shared__ double sh[256];
  double s = 0.0;
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < (size_t)nvox;
       v += (size_t)gridDim.x * blockDim.x) {
    const float e = out[v] - target[v];
    s += (double)e * e;
    gout[v] = 2.f * e / (float)nvox;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// head backward: gh[v][c] = gout[v] w[c] / scale (bf16); the head's weight
// and bias gradients per slice of voxels: part[s][c] = sum gout[v] h[v][c] /
// scale (c < C), part[s][C] = sum gout[v] / scale
__global__ void k_head_bwd(const bf16 *__restrict__ h, const float *__restrict__ w,
                           const float *__restrict__ gout, bf16 *__restrict__ gh,
                           float *__restrict__ part, long long nvox, int C, int S,
                           float inv_scale) {
  const size_t n = (size_t)nvox * C;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    gh[i] = (bf16)(gout[i / C] * inv_scale * w[c]);
  }
  // weight gradient partials: block b < S takes slice b
  if ((int)blockIdx.x < S) {
    const int s = blockIdx.x;
    const long long v0 = nvox * s / S, v1 = nvox * (s + 1) / S;
    for (int c = threadIdx.x; c <= C; c += blockDim.x) {
      float acc = 0.f;
      for (long long v = v0; v < v1; v++)
        acc += gout[v] * inv_scale * (c < C ? (float)h[(size_t)v * C + c] : 1.f);
      part[(size_t)s * (C + 1) + c] = acc;
    }
  }
}

// Adam with bias correction on fp32 master weights; bf16 copies for the
// kernels (wb may be null)
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ gr, float *__restrict__ m,
                       float *__restrict__ v, bf16 *__restrict__ wb, size_t n, float lr,
                       float b1, float b2, float eps, float c1, float c2) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    p[i] = pi;
    if (wb) wb[i] = (bf16)pi;
  }
}

// a[i] = a[i] + b[i] (bf16, the sum rounded once): gradients reaching one
// activation from two consumers
__global__ void k_add_bf16(bf16 *__restrict__ a, const bf16 *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    a[i] = (bf16)((float)a[i] + (float)b[i]);
}

__global__ void k_to_bf16(const float *__restrict__ a, bf16 *__restrict__ b, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    b[i] = (bf16)a[i];
}

}  // namespace tbnn

using namespace tbnn;

// ---------------------------------------------------------------------------
// C ABI (declared in include/timbir_dev.h)
// ---------------------------------------------------------------------------
static thread_local std::string g_nn_err;

static int nn_fail(const char *msg, hipError_t e = hipSuccess) {
  g_nn_err = msg;
  if (e != hipSuccess) {
    g_nn_err += ": ";
    g_nn_err += hipGetErrorString(e);
  }
  return -1;
}

#define NNCHK(what)                                     \
  do {                                                  \
    hipError_t e_ = hipGetLastError();                  \
    if (e_ != hipSuccess) return nn_fail(what, e_);     \
  } while (0)

static int grid1d(size_t n) {
  size_t g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 65535 * 8) g = 65535 * 8;
  return (int)g;
}

extern "C" const char *tb_nn_last_error(void) { return g_nn_err.c_str(); }

// Device memory for the callers of these kernels, from this library's HIP
// runtime (so that the kernels and the copies are ordered on its streams).
extern "C" int tb_dev_malloc(uint64_t bytes, void **ptr) {
  *ptr = nullptr;
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess) return nn_fail("tb_dev_malloc", e);
  return 0;
}

extern "C" int tb_dev_free(void *ptr) {
  hipError_t e = hipFree(ptr);
  if (e != hipSuccess) return nn_fail("tb_dev_free", e);
  return 0;
}

extern "C" int tb_dev_copy(void *dst, const void *src, uint64_t bytes, int32_t kind) {
  // kind: 0 host to device, 1 device to host, 2 device to device; blocking
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice
                          : kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  hipError_t e = hipMemcpy(dst, src, bytes, k);
  if (e != hipSuccess) return nn_fail("tb_dev_copy", e);
  return 0;
}

extern "C" int tb_dev_zero(void *ptr, uint64_t bytes, void *stream) {
  hipError_t e = hipMemsetAsync(ptr, 0, bytes, static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return nn_fail("tb_dev_zero", e);
  return 0;
}

extern "C" int tb_dev_synchronize(void) {
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return nn_fail("tb_dev_synchronize", e);
  return 0;
}

extern "C" int tb_nn_conv3d(const void *in, const void *w, const float *bias, void *out,
                            int32_t n, int32_t d, int32_t h, int32_t wd, int32_t cin,
                            int32_t cout, int32_t relu, int32_t out_f32, void *stream) {
  if (n <= 0 || d <= 0 || h <= 0 || wd <= 0 || cin % 32 || cout % 32 || cin <= 0 || cout <= 0)
    return nn_fail("tb_nn_conv3d: sizes must be positive, channels multiples of 32");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int MB = 4, NB = 2;
  const long long waves = (long long)n * d * h * ((wd + 16 * MB - 1) / (16 * MB)) * (cout / (16 * NB));
  const long long blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
  if (blocks > 0x7fffffffLL) return nn_fail("tb_nn_conv3d: volume too large");
  hipLaunchKernelGGL((k_conv3d_fwd<4, 2>), dim3((unsigned)blocks), dim3(64 * kWavesPerBlock), 0,
                     s, static_cast<const bf16 *>(in), static_cast<const bf16 *>(w), bias, out, n,
                     d, h, wd, cin, cout, relu, out_f32);
  NNCHK("tb_nn_conv3d");
  return 0;
}

extern "C" int tb_nn_conv3d_flip(const void *w, void *wt, int32_t cin, int32_t cout,
                                 void *stream) {
  hipLaunchKernelGGL(k_conv_flip, dim3(grid1d((size_t)kTaps * cin * cout)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const bf16 *>(w),
                     static_cast<bf16 *>(wt), cin, cout);
  NNCHK("tb_nn_conv3d_flip");
  return 0;
}

extern "C" int tb_nn_conv3d_wgrad(const void *in, const void *gout, float *gw, float *part,
                                  int32_t slices, int32_t n, int32_t d, int32_t h, int32_t wd,
                                  int32_t cin, int32_t cout, void *stream) {
  if (cin % 32 || cout % 32 || slices <= 0)
    return nn_fail("tb_nn_conv3d_wgrad: channels must be multiples of 32");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long waves = (long long)slices * kTaps * (cout / 32) * (cin / 32);
  const long long blocks = (waves + kWavesPerBlock - 1) / kWavesPerBlock;
  hipLaunchKernelGGL((k_conv3d_wgrad<2, 2>), dim3((unsigned)blocks), dim3(64 * kWavesPerBlock),
                     0, s, static_cast<const bf16 *>(in), static_cast<const bf16 *>(gout), part,
                     n, d, h, wd, cin, cout, slices);
  const size_t nw = (size_t)kTaps * cin * cout;
  hipLaunchKernelGGL(k_sum_parts, dim3(grid1d(nw)), dim3(256), 0, s, part, gw, nw, slices);
  NNCHK("tb_nn_conv3d_wgrad");
  return 0;
}

extern "C" int tb_nn_channel_sum(const void *g, float *out, float *part, int32_t slices,
                                 int64_t nvox, int32_t c, void *stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_channel_sum, dim3(slices), dim3(64), 0, s, static_cast<const bf16 *>(g),
                     part, (long long)nvox, c, slices);
  hipLaunchKernelGGL(k_sum_parts, dim3(grid1d(c)), dim3(256), 0, s, part, out, (size_t)c, slices);
  NNCHK("tb_nn_channel_sum");
  return 0;
}

extern "C" int tb_nn_relu_bwd(const void *y, void *g, int64_t count, void *stream) {
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid1d(count)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const bf16 *>(y),
                     static_cast<bf16 *>(g), (size_t)count);
  NNCHK("tb_nn_relu_bwd");
  return 0;
}

extern "C" int tb_nn_avgpool2(const void *in, void *out, int32_t n, int32_t d, int32_t h,
                              int32_t wd, int32_t c, int32_t backward, void *stream) {
  if (d % 2 || h % 2 || wd % 2) return nn_fail("tb_nn_avgpool2: odd size");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (backward)
    hipLaunchKernelGGL(k_avgpool2_bwd, dim3(grid1d((size_t)n * d * h * wd * c)), dim3(256), 0, s,
                       static_cast<const bf16 *>(in), static_cast<bf16 *>(out), n, d, h, wd, c);
  else
    hipLaunchKernelGGL(k_avgpool2, dim3(grid1d((size_t)n * d * h * wd * c / 8)), dim3(256), 0, s,
                       static_cast<const bf16 *>(in), static_cast<bf16 *>(out), n, d, h, wd, c);
  NNCHK("tb_nn_avgpool2");
  return 0;
}

extern "C" int tb_nn_up2_concat(const void *a, const void *b, void *out, int32_t n, int32_t d,
                                int32_t h, int32_t wd, int32_t ca, int32_t cb, void *stream) {
  hipLaunchKernelGGL(k_up2_concat, dim3(grid1d((size_t)n * d * h * wd * (ca + cb))), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const bf16 *>(a),
                     static_cast<const bf16 *>(b), static_cast<bf16 *>(out), n, d, h, wd, ca, cb);
  NNCHK("tb_nn_up2_concat");
  return 0;
}

extern "C" int tb_nn_up2_concat_bwd(const void *g, void *ga, void *gb, int32_t n, int32_t d,
                                    int32_t h, int32_t wd, int32_t ca, int32_t cb, void *stream) {
  const size_t tot = (size_t)n * (d / 2) * (h / 2) * (wd / 2) * ca + (size_t)n * d * h * wd * cb;
  hipLaunchKernelGGL(k_up2_concat_bwd, dim3(grid1d(tot)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const bf16 *>(g),
                     static_cast<bf16 *>(ga), static_cast<bf16 *>(gb), n, d, h, wd, ca, cb);
  NNCHK("tb_nn_up2_concat_bwd");
  return 0;
}

extern "C" int tb_nn_add(void *a, const void *b, int64_t count, void *stream) {
  hipLaunchKernelGGL(k_add_bf16, dim3(grid1d(count)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<bf16 *>(a),
                     static_cast<const bf16 *>(b), (size_t)count);
  NNCHK("tb_nn_add");
  return 0;
}

extern "C" int tb_nn_make_input(const float *fbp, const float *mean, void *x, int64_t nvox,
                                int64_t per_sample, float scale, void *stream) {
  hipLaunchKernelGGL(k_make_input, dim3(grid1d((size_t)nvox * 32)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), fbp, mean, static_cast<bf16 *>(x),
                     (long long)nvox, (long long)per_sample, scale);
  NNCHK("tb_nn_make_input");
  return 0;
}

extern "C" int tb_nn_frame_mean(const float *x, float *mean, int64_t nframes, int64_t per,
                                void *stream) {
  hipLaunchKernelGGL(k_frame_mean, dim3(grid1d(per)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), x, mean, (long long)nframes,
                     (long long)per);
  NNCHK("tb_nn_frame_mean");
  return 0;
}

extern "C" int tb_nn_head(const void *h, const float *w, const float *b, const float *fbp,
                          float *out, int64_t nvox, int32_t c, float scale, void *stream) {
  hipLaunchKernelGGL(k_head_fwd, dim3(grid1d(nvox)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), static_cast<const bf16 *>(h), w, b, fbp,
                     out, (long long)nvox, c, 1.f / scale);
  NNCHK("tb_nn_head");
  return 0;
}

extern "C" int tb_nn_head_bwd(const void *h, const float *w, const float *gout, void *gh,
                              float *part, int32_t slices, int64_t nvox, int32_t c, float scale,
                              void *stream) {
  const int blocks = std::max(grid1d((size_t)nvox * c), (int)slices);
  hipLaunchKernelGGL(k_head_bwd, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     static_cast<const bf16 *>(h), w, gout, static_cast<bf16 *>(gh), part,
                     (long long)nvox, c, slices, 1.f / scale);
  NNCHK("tb_nn_head_bwd");
  return 0;
}

extern "C" int tb_nn_mse(const float *out, const float *target, float *gout, double *part,
                         int32_t blocks, int64_t nvox, void *stream) {
  hipLaunchKernelGGL(k_mse, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), out,
                     target, gout, part, (long long)nvox);
  NNCHK("tb_nn_mse");
  return 0;
}

extern "C" int tb_nn_sum_parts(const float *part, float *out, int64_t count, int32_t slices,
                               void *stream) {
  hipLaunchKernelGGL(k_sum_parts, dim3(grid1d(count)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), part, out, (size_t)count, slices);
  NNCHK("tb_nn_sum_parts");
  return 0;
}

extern "C" int tb_nn_adam(float *p, const float *g, float *m, float *v, void *wb, int64_t count,
                          float lr, float b1, float b2, float eps, int32_t step, void *stream) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid1d(count)), dim3(256), 0, static_cast<hipStream_t>(stream),
                     p, g, m, v, static_cast<bf16 *>(wb), (size_t)count, lr, b1, b2, eps, c1, c2);
  NNCHK("tb_nn_adam");
  return 0;
}

extern "C" int tb_nn_to_bf16(const float *a, void *b, int64_t count, void *stream) {
  hipLaunchKernelGGL(k_to_bf16, dim3(grid1d(count)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), a, static_cast<bf16 *>(b), (size_t)count);
  NNCHK("tb_nn_to_bf16");
  return 0;
}
