// psv_conv.hip -- bfloat16 MFMA kernels that run and train the learned
// gradient preconditioner (a U-Net over the velocity kernels).
//
// Activations are NHWC bfloat16 bit patterns, images stacked along the rows
// (rows = batch * H).  A convolution is an implicit GEMM
//   out[m][n] = sum_k A[m][k] * B[k][n],   B[k][n] = w[n * K + k],
// with m = output pixel and A gathered from the input by one of three modes:
//   kConv       ks x ks "same" convolution, k = tap * cin + ci;
//   kScatter    2x2 stride-2 transposed convolution: K = cin, n = q * cout + co,
//               result written to pixel (2y + q / 2, 2x + q % 2);
//   kGather2x2  space-to-depth gather (the input gradient of the transposed
//               convolution): the input has 2H x 2W pixels,
//               k = q * cin + ci reads pixel (2y + q / 2, 2x + q % 2).
// One wavefront computes a 32 x 32 tile of out with v_mfma_f32_32x32x8_bf16
// (fp32 accumulation); bias, the erf GELU (optionally storing the
// pre-activation) and the GELU derivative of a stored pre-activation (the
// backward pass) are fused into the epilogue.
//
// The weight gradient dW[i][j] = sum_p L[p][i] R[p][j] is a second MFMA
// kernel whose reduction runs over pixels, split across wavefronts into
// fp32 partial sums that a fixed-order pass adds up (deterministic).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "psv_kernels.h"

namespace psv {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ inline uint16_t to_bf16(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float gelu(float v) { return 0.5f * v * (1.0f + erff(v * 0.70710678118654752f)); }

// d/dv [v Phi(v)] = Phi(v) + v phi(v)
__device__ inline float gelu_grad(float v) {
  return 0.5f * (1.0f + erff(v * 0.70710678118654752f)) +
         v * 0.39894228040143268f * __expf(-0.5f * v * v);
}

// Accumulator r of lane (half, l32) of v_mfma_f32_32x32x8_bf16 holds row
// 8 (r / 4) + 4 half + r % 4, column l32 of the 32 x 32 tile.
__device__ inline int acc_row(int r, int half) { return 8 * (r >> 2) + 4 * half + (r & 3); }

// The input channel count (cin) and in_stride are multiples of 8, so each
// 8-wide step of K lies in one tap and a lane's 4 values are 8-byte aligned.
__global__ void __launch_bounds__(256) conv_mfma_kernel(ConvGemm a) {
  const int lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
  const int M = a.rows * a.W, N = a.mode == kScatter ? 4 * a.cout : a.cout;
  const int ntaps = a.mode == kGather2x2 ? 4 : a.ks * a.ks, K = ntaps * a.cin;
  const int ntm = (M + 31) / 32, ntn = (N + 31) / 32;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= ntm * ntn) return;  // whole wavefronts exit together
  const int tm = tile % ntm, tn = tile / ntm;
  const int m = tm * 32 + l32, n = tn * 32 + l32;
  const int yg = m / a.W, x = m - yg * a.W, yi = yg % a.H, pad = a.ks / 2;
  const bool mok = m < M, nok = n < N;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  for (int tap = 0; tap < ntaps; ++tap) {
    long long src;
    bool ok = mok;
    if (a.mode == kGather2x2) {
      src = (long long)(2 * yg + (tap >> 1)) * (2 * a.W) + 2 * x + (tap & 1);
    } else {
      const int dy = tap / a.ks - pad, dx = tap % a.ks - pad;
      ok = ok && yi + dy >= 0 && yi + dy < a.H && x + dx >= 0 && x + dx < a.W;
      src = (long long)(yg + dy) * a.W + x + dx;
    }
    const uint16_t *ap = a.in + (ok ? src * a.in_stride : 0) + 4 * half;
    const uint16_t *bp = a.w + (nok ? (long long)n * K : 0) + tap * a.cin + 4 * half;
    for (int c0 = 0; c0 < a.cin; c0 += 8) {
      bf16x4 av = {0, 0, 0, 0}, bv = {0, 0, 0, 0};
      if (ok) av = *(const bf16x4 *)(ap + c0);
      if (nok) bv = *(const bf16x4 *)(bp + c0);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
    }
  }
  if (!nok) return;
  const int co = a.mode == kScatter ? n % a.cout : n;
  const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int mm = tm * 32 + acc_row(r, half);
    if (mm >= M) continue;
    long long pix = mm;
    if (a.mode == kScatter) {
      const int q = n / a.cout, py = mm / a.W, px = mm - py * a.W;
      pix = (long long)(2 * py + (q >> 1)) * (2 * a.W) + (2 * px + (q & 1));
    }
    float v = acc[r] + b;
    if (a.pre16) {  // training: keep the rounded pre-activation for the backward pass
      const uint16_t h = to_bf16(v);
      a.pre16[pix * a.cout + co] = h;
      v = from_bf16(h);
    }
    if (a.gelu) v = gelu(v);
    if (a.dz) v *= gelu_grad(from_bf16(a.dz[pix * a.dz_stride + co]));
    const long long o = pix * a.out_stride + a.out_off + co;
    if (a.out32)
      a.out32[o] = v;
    else
      a.out16[o] = to_bf16(v);
  }
}

// Weight gradient: out[s][i][j] = sum over pixels p of split s of
// L[p][i] R[p][j].  L: the output gradient at the output pixels
// (l_s2d = 0: L[p][i] = l[p * l_stride + i]; l_s2d = 1: i = q * lc + c reads
// fine pixel (2y + q / 2, 2x + q % 2)); R: the layer input gathered as the
// forward convolution gathers it (j = tap * rc + c, zero outside the image).
__global__ void __launch_bounds__(256) wgrad_mfma_kernel(Wgrad a) {
  const int lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
  const int nti = (a.I + 31) / 32, ntj = (a.J + 31) / 32;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= nti * ntj * a.nsplit) return;
  const int ti = tile % nti, tj = (tile / nti) % ntj, s = tile / (nti * ntj);
  const int i = ti * 32 + l32, j = tj * 32 + l32;
  const bool iok = i < a.I, jok = j < a.J;
  const long long P = (long long)a.rows * a.W;
  const long long p_begin = (long long)s * a.chunk;
  const long long p_end = p_begin + a.chunk < P ? p_begin + a.chunk : P;
  // per-lane constants of the two gathers
  int lq = 0, lci = i;
  if (a.l_s2d) {
    lq = i / a.lc;
    lci = i - lq * a.lc;
  }
  const int tap = j / a.rc, rci = j - tap * a.rc, pad = a.ks / 2;
  const int dy = tap / a.ks - pad, dx = tap % a.ks - pad;
  // the 4 pixels this lane feeds per step: p0 + 4 half + jj
  int px[4], pyg[4], pyi[4];
#pragma unroll
  for (int jj = 0; jj < 4; ++jj) {
    const long long p = p_begin + 4 * half + jj;
    pyg[jj] = (int)(p / a.W);
    px[jj] = (int)(p - (long long)pyg[jj] * a.W);
    pyi[jj] = pyg[jj] % a.H;
  }
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  for (long long p0 = p_begin; p0 < p_end; p0 += 8) {
    bf16x4 av, bv;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const long long p = p0 + 4 * half + jj;
      const bool pok = p < p_end;
      uint16_t lv = 0, rv = 0;
      if (pok && iok) {
        long long li;
        if (a.l_s2d)
          li = ((long long)(2 * pyg[jj] + (lq >> 1)) * (2 * a.W) + 2 * px[jj] + (lq & 1)) *
                   a.l_stride + lci;
        else
          li = p * a.l_stride + i;
        lv = a.l[li];
      }
      const int yy = pyi[jj] + dy, xx = px[jj] + dx;
      if (pok && jok && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
        rv = a.r[((long long)(pyg[jj] + dy) * a.W + xx) * a.r_stride + rci];
      av[jj] = (short)lv;
      bv[jj] = (short)rv;
      // advance this pixel by 8
      px[jj] += 8;
      while (px[jj] >= a.W) {
        px[jj] -= a.W;
        ++pyg[jj];
        if (++pyi[jj] == a.H) pyi[jj] = 0;
      }
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
  }
  if (!jok) return;
  float *out = a.out + (long long)s * a.I * a.J;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int ii = ti * 32 + acc_row(r, half);
    if (ii < a.I) out[(long long)ii * a.J + j] = acc[r];
  }
}

// out[e] = sum_s part[s * n + e], s in order (deterministic).
__global__ void sum_splits_kernel(const float *__restrict__ part, float *out, long long n,
                                  int nsplit) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float v = 0.f;
  for (int s = 0; s < nsplit; ++s) v += part[(long long)s * n + e];
  out[e] = v;
}

// part[s][c] = sum of src[p * stride + off + c] over the pixels of split s;
// block = 64 channels x 4 pixel lanes.
__global__ void __launch_bounds__(256) channel_sum_kernel(const uint16_t *__restrict__ src,
                                                          int stride, int off, int C,
                                                          long long P, long long chunk,
                                                          float *part) {
  __shared__ float red[4][64];
  const int c = blockIdx.x * 64 + (threadIdx.x & 63), r = threadIdx.x >> 6;
  const long long p_begin = (long long)blockIdx.y * chunk;
  const long long p_end = p_begin + chunk < P ? p_begin + chunk : P;
  float v = 0.f;
  if (c < C)
    for (long long p = p_begin + r; p < p_end; p += 4) v += from_bf16(src[p * stride + off + c]);
  red[r][threadIdx.x & 63] = v;
  __syncthreads();
  if (r == 0 && c < C)
    part[(long long)blockIdx.y * C + c] =
        ((red[0][threadIdx.x] + red[1][threadIdx.x]) + red[2][threadIdx.x]) + red[3][threadIdx.x];
}

__global__ void maxpool2_kernel(const uint16_t *__restrict__ in, uint16_t *out, int H,
                                int W, int C) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2;
  if (i >= (long long)Ho * Wo * C) return;
  const int c = (int)(i % C);
  const long long p = i / C;
  const int y = (int)(p / Wo), x = (int)(p % Wo);
  float v = -INFINITY;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
      v = fmaxf(v, from_bf16(in[((long long)(2 * y + a) * W + (2 * x + b)) * C + c]));
  out[i] = to_bf16(v);
}

// Gradient at a block output that feeds both a 2x2 max pooling and a skip
// connection, times the GELU derivative of the block's pre-activation:
//   out = (route(dpool) + dskip) * gelu'(z),
// route(dpool) sending each pooled gradient to the first maximum of its
// window in row-major order (the pooling's own choice).
__global__ void pool_backward_kernel(const uint16_t *__restrict__ act,
                                     const uint16_t *__restrict__ z,
                                     const float *__restrict__ dpool,
                                     const uint16_t *__restrict__ dskip, int skip_stride,
                                     int skip_off, uint16_t *out, int rows, int W, int C) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)rows * W * C) return;
  const int c = (int)(i % C);
  const long long pix = i / C;
  const int yg = (int)(pix / W), x = (int)(pix - (long long)yg * W);
  float g = 0.f;
  if (dpool) {
    const int cy = yg >> 1, cx = x >> 1;
    int best = 0;
    float bv = -INFINITY;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float v =
          from_bf16(act[((long long)(2 * cy + (q >> 1)) * W + 2 * cx + (q & 1)) * C + c]);
      if (q == 0 || v > bv) {
        bv = v;
        best = q;
      }
    }
    if (best == ((yg & 1) << 1) + (x & 1)) g = dpool[((long long)cy * (W >> 1) + cx) * C + c];
  }
  if (dskip) g += from_bf16(dskip[pix * skip_stride + skip_off + c]);
  out[i] = to_bf16(g * gelu_grad(from_bf16(z[i])));
}

unsigned blocks_of(long long n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

hipError_t launch_conv_gemm(const ConvGemm &a, hipStream_t s) {
  if (a.cin <= 0 || a.cin % 8 || a.in_stride % 8 || a.in_stride < a.cin || a.cout <= 0 ||
      a.rows <= 0 || a.H <= 0 || a.rows % a.H || a.W <= 0 ||
      (a.mode == kConv && a.ks != 1 && a.ks != 3) || (a.mode != kConv && a.ks != 1) ||
      a.mode < kConv || a.mode > kGather2x2 || (a.pre16 && a.mode != kConv))
    return hipErrorInvalidValue;
  const long long M = (long long)a.rows * a.W, N = a.mode == kScatter ? 4LL * a.cout : a.cout;
  const long long tiles = ((M + 31) / 32) * ((N + 31) / 32);
  hipLaunchKernelGGL(conv_mfma_kernel, dim3(blocks_of(tiles, 4)), dim3(256), 0, s, a);
  return hipGetLastError();
}

long long wgrad_plan(Wgrad &a, long long workspace_floats) {
  const long long P = (long long)a.rows * a.W;
  const long long tiles = (long long)((a.I + 31) / 32) * ((a.J + 31) / 32);
  long long ns = (2048 + tiles - 1) / tiles;
  const long long max_ns = (P + 255) / 256;
  if (ns > max_ns) ns = max_ns;
  const long long fit = workspace_floats / ((long long)a.I * a.J);
  if (ns > fit) ns = fit;
  if (ns < 1) ns = 1;
  long long chunk = (P + ns - 1) / ns;
  chunk = (chunk + 7) / 8 * 8;
  a.nsplit = (int)((P + chunk - 1) / chunk);
  a.chunk = chunk;
  return a.nsplit;
}

hipError_t launch_wgrad(Wgrad a, float *dw, float *workspace, long long workspace_floats,
                        hipStream_t s) {
  if (a.I <= 0 || a.J <= 0 || a.rows <= 0 || a.H <= 0 || a.rows % a.H || a.W <= 0 ||
      (a.ks != 1 && a.ks != 3) || a.rc <= 0 || a.J != a.ks * a.ks * a.rc ||
      (a.l_s2d && (a.lc <= 0 || a.I != 4 * a.lc)))
    return hipErrorInvalidValue;
  wgrad_plan(a, workspace ? workspace_floats : 0);
  a.out = a.nsplit > 1 ? workspace : dw;
  const long long tiles = (long long)((a.I + 31) / 32) * ((a.J + 31) / 32) * a.nsplit;
  hipLaunchKernelGGL(wgrad_mfma_kernel, dim3(blocks_of(tiles, 4)), dim3(256), 0, s, a);
  if (a.nsplit > 1) {
    const long long n = (long long)a.I * a.J;
    hipLaunchKernelGGL(sum_splits_kernel, dim3(blocks_of(n, 256)), dim3(256), 0, s,
                       workspace, dw, n, a.nsplit);
  }
  return hipGetLastError();
}

hipError_t launch_channel_sum(const uint16_t *src, int stride, int off, int c, long long P,
                              float *out, float *workspace, long long workspace_floats,
                              hipStream_t s) {
  if (c <= 0 || P <= 0 || off < 0 || off + c > stride) return hipErrorInvalidValue;
  long long ns = (P + 4095) / 4096;
  if (ns > 256) ns = 256;
  if (!workspace || ns > workspace_floats / c) ns = 1;
  const long long chunk = (P + ns - 1) / ns;
  ns = (P + chunk - 1) / chunk;
  hipLaunchKernelGGL(channel_sum_kernel, dim3(blocks_of(c, 64), (unsigned)ns), dim3(256), 0,
                     s, src, stride, off, c, P, chunk, ns > 1 ? workspace : out);
  if (ns > 1)
    hipLaunchKernelGGL(sum_splits_kernel, dim3(blocks_of(c, 256)), dim3(256), 0, s,
                       workspace, out, (long long)c, (int)ns);
  return hipGetLastError();
}

hipError_t launch_conv_bf16(const uint16_t *in, const uint16_t *w, const float *bias,
                            void *out, bool out_f32, int h, int wd, int cin, int cout,
                            int ks, bool transposed, bool gelu, int out_stride,
                            int out_off, hipStream_t s) {
  ConvGemm a{};
  a.in = in;
  a.in_stride = cin;
  a.w = w;
  a.bias = bias;
  a.out16 = out_f32 ? nullptr : (uint16_t *)out;
  a.out32 = out_f32 ? (float *)out : nullptr;
  a.out_stride = out_stride;
  a.out_off = out_off;
  a.rows = h;
  a.H = h;
  a.W = wd;
  a.cin = cin;
  a.cout = cout;
  a.ks = ks;
  a.mode = transposed ? kScatter : kConv;
  a.gelu = gelu ? 1 : 0;
  return launch_conv_gemm(a, s);
}

hipError_t launch_maxpool2_bf16(const uint16_t *in, uint16_t *out, int h, int w, int c,
                                hipStream_t s) {
  if (h % 2 || w % 2 || c <= 0) return hipErrorInvalidValue;
  const long long n = (long long)(h / 2) * (w / 2) * c;
  hipLaunchKernelGGL(maxpool2_kernel, dim3(blocks_of(n, 256)), dim3(256), 0, s, in, out, h,
                     w, c);
  return hipGetLastError();
}

hipError_t launch_pool_backward(const uint16_t *act, const uint16_t *z, const float *dpool,
                                const uint16_t *dskip, int skip_stride, int skip_off,
                                uint16_t *out, int rows, int w, int c, hipStream_t s) {
  if (rows % 2 || w % 2 || c <= 0 || (dpool && !act) ||
      (dskip && (skip_off < 0 || skip_off + c > skip_stride)))
    return hipErrorInvalidValue;
  const long long n = (long long)rows * w * c;
  hipLaunchKernelGGL(pool_backward_kernel, dim3(blocks_of(n, 256)), dim3(256), 0, s, act, z,
                     dpool, dskip, skip_stride, skip_off, out, rows, w, c);
  return hipGetLastError();
}

}  // namespace psv
