// psv_conv.hip -- bfloat16 MFMA kernels for the inference of the learned
// gradient preconditioner (a U-Net over the velocity kernels).
//
// Activations are NHWC (one image), bfloat16 bit patterns; weights are laid
// out [N][K] with K = (ky * KS + kx) * Cin + ci, so a convolution is an
// implicit GEMM  out[m][n] = sum_k A[m][k] * B[k][n]  with m = pixel,
// n = output channel (or, for the 2x2 stride-2 transposed convolution,
// n = (2a + b) * Cout + co and the result scattered to pixel (2y+a, 2x+b)).
// One wavefront computes a 32 x 32 tile of out with
// v_mfma_f32_32x32x8_bf16 (fp32 accumulation), reading its A and B
// fragments straight from memory (8 bytes per lane per step); bias and the
// erf GELU are fused into the epilogue.
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

// Cin is a multiple of 8, so each 8-wide step of K lies in one filter tap
// and each lane's 4 values are 8-byte aligned.
__global__ void __launch_bounds__(256) conv_mfma_kernel(
    const uint16_t *__restrict__ in, const uint16_t *__restrict__ w,
    const float *__restrict__ bias, uint16_t *out16, float *out32, int H, int W,
    int Cin, int Cout, int KS, int transposed, int gelu, int out_stride,
    int out_off) {
  const int lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
  const int M = H * W, N = transposed ? 4 * Cout : Cout, K = KS * KS * Cin;
  const int ntm = (M + 31) / 32, ntn = (N + 31) / 32;
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= ntm * ntn) return;  // whole wavefronts exit together
  const int tm = tile % ntm, tn = tile / ntm;
  const int m = tm * 32 + l32, n = tn * 32 + l32;
  const int y = m / W, x = m - (m / W) * W, pad = KS / 2;
  const bool mok = m < M, nok = n < N;
  f32x16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  for (int k0 = 0; k0 < K; k0 += 8) {
    const int tap = k0 / Cin, ci = k0 - tap * Cin + 4 * half;
    const int yy = y + tap / KS - pad, xx = x + tap % KS - pad;
    bf16x4 av = {0, 0, 0, 0}, bv = {0, 0, 0, 0};
    if (mok && yy >= 0 && yy < H && xx >= 0 && xx < W)
      av = *(const bf16x4 *)(in + ((long long)yy * W + xx) * Cin + ci);
    if (nok) bv = *(const bf16x4 *)(w + (long long)n * K + k0 + 4 * half);
    acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc, 0, 0, 0);
  }
  // accumulator r of lane (half, l32) holds row 8 (r / 4) + 4 half + r % 4,
  // column l32 of the tile
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int mm = tm * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
    if (mm >= M || !nok) continue;
    const int co = transposed ? n % Cout : n;
    float v = acc[r] + bias[co];
    if (gelu) v = 0.5f * v * (1.0f + erff(v * 0.70710678118654752f));
    long long pix;
    if (transposed) {
      const int q = n / Cout, py = mm / W, px = mm - (mm / W) * W;
      pix = (long long)(2 * py + (q >> 1)) * (2 * W) + (2 * px + (q & 1));
    } else {
      pix = mm;
    }
    const long long o = pix * out_stride + out_off + co;
    if (out32)
      out32[o] = v;
    else
      out16[o] = to_bf16(v);
  }
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

}  // namespace

hipError_t launch_conv_bf16(const uint16_t *in, const uint16_t *w, const float *bias,
                            void *out, bool out_f32, int h, int wd, int cin, int cout,
                            int ks, bool transposed, bool gelu, int out_stride,
                            int out_off, hipStream_t s) {
  if (cin % 8 != 0 || (ks != 1 && ks != 3) || (transposed && ks != 1) || h <= 0 ||
      wd <= 0 || cout <= 0)
    return hipErrorInvalidValue;
  const long long M = (long long)h * wd, N = transposed ? 4LL * cout : cout;
  const long long tiles = ((M + 31) / 32) * ((N + 31) / 32);
  hipLaunchKernelGGL(conv_mfma_kernel, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, s,
                     in, w, bias, out_f32 ? nullptr : (uint16_t *)out,
                     out_f32 ? (float *)out : nullptr, h, wd, cin, cout, ks,
                     transposed ? 1 : 0, gelu ? 1 : 0, out_stride, out_off);
  return hipGetLastError();
}

hipError_t launch_maxpool2_bf16(const uint16_t *in, uint16_t *out, int h, int w, int c,
                                hipStream_t s) {
  if (h % 2 || w % 2 || c <= 0) return hipErrorInvalidValue;
  const long long n = (long long)(h / 2) * (w / 2) * c;
  hipLaunchKernelGGL(maxpool2_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     in, out, h, w, c);
  return hipGetLastError();
}

}  // namespace psv
