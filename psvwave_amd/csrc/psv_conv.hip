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
// One wavefront computes a 64 x 64 (or 128 x 32) tile of out with
// v_mfma_f32_32x32x16_bf16 (fp32 accumulation); bias, the erf GELU (optionally storing the
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

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
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

// A lane's 8 values of K (k = k0 + 8 half ... + 7) lie in one tap because
// cin and in_stride are multiples of 8, so they are one aligned 16-byte load.
// Source element offset of pixel (yg, x) (yi = row within its image) for a
// tap, or -1 outside the image.
__device__ inline long long conv_src(const ConvGemm &a, int tap, int yg, int x, int yi) {
  if (a.mode == kGather2x2)
    return ((long long)(2 * yg + (tap >> 1)) * (2 * a.W) + 2 * x + (tap & 1)) * a.in_stride;
  const int pad = a.ks / 2, dy = tap / a.ks - pad, dx = tap % a.ks - pad;
  if (yi + dy < 0 || yi + dy >= a.H || x + dx < 0 || x + dx >= a.W) return -1;
  return ((long long)(yg + dy) * a.W + x + dx) * a.in_stride;
}

// Epilogue of one output value (bias already added): keep the rounded
// pre-activation (training), GELU, times the GELU derivative of a stored
// pre-activation (backward), store.
__device__ inline void conv_store(const ConvGemm &a, long long pix, int co, float v) {
  if (a.pre16) {
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

// One wavefront computes a (32 WM) x (32 WN) tile of out with
// v_mfma_f32_32x32x16_bf16: per step of 16 along K each lane loads 16 bytes
// per 32-row slice of A and of B, and issues WM x WN matrix instructions.
template <int WM, int WN>
__global__ void __launch_bounds__(256) conv_mfma_kernel(ConvGemm a) {
  const int lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
  const int M = a.rows * a.W, N = a.mode == kScatter ? 4 * a.cout : a.cout;
  const int ntaps = a.mode == kGather2x2 ? 4 : a.ks * a.ks, K = ntaps * a.cin;
  const int ntm = (M + 32 * WM - 1) / (32 * WM), ntn = (N + 32 * WN - 1) / (32 * WN);
  const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (tile >= ntm * ntn) return;  // whole wavefronts exit together
  const int tm = tile % ntm, tn = tile / ntm;
  int yg[WM], xx[WM], yi[WM];
  bool mok[WM];
#pragma unroll
  for (int i = 0; i < WM; ++i) {
    const int m = (tm * WM + i) * 32 + l32;
    mok[i] = m < M;
    yg[i] = m / a.W;
    xx[i] = m - yg[i] * a.W;
    yi[i] = yg[i] % a.H;
  }
  const uint16_t *bp[WN];
  bool nok[WN];
#pragma unroll
  for (int j = 0; j < WN; ++j) {
    const int n = (tn * WN + j) * 32 + l32;
    nok[j] = n < N;
    bp[j] = a.w + (nok[j] ? (long long)n * K : 0) + 8 * half;
  }
  // this lane's position along K: tap and channel of k0 + 8 half
  int tap = (8 * half) / a.cin, ci = 8 * half - tap * a.cin;
  long long src[WM];
#pragma unroll
  for (int i = 0; i < WM; ++i) src[i] = mok[i] ? conv_src(a, tap, yg[i], xx[i], yi[i]) : -1;
  f32x16 acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  const bf16x8 zero = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    const bool kok = tap < ntaps;
    bf16x8 av[WM], bv[WN];
#pragma unroll
    for (int i = 0; i < WM; ++i)
      av[i] = kok && src[i] >= 0 ? *(const bf16x8 *)(a.in + src[i] + ci) : zero;
#pragma unroll
    for (int j = 0; j < WN; ++j)
      bv[j] = kok && nok[j] ? *(const bf16x8 *)(bp[j] + k0) : zero;
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av[i], bv[j], acc[i][j], 0, 0, 0);
    ci += 16;
    if (ci >= a.cin) {
      do {
        ci -= a.cin;
        ++tap;
      } while (ci >= a.cin);
      if (tap < ntaps) {
#pragma unroll
        for (int i = 0; i < WM; ++i)
          src[i] = mok[i] ? conv_src(a, tap, yg[i], xx[i], yi[i]) : -1;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < WN; ++j) {
    if (!nok[j]) continue;
    const int n = (tn * WN + j) * 32 + l32;
    const int co = a.mode == kScatter ? n % a.cout : n;
    const float b = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for (int i = 0; i < WM; ++i) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int mm = (tm * WM + i) * 32 + acc_row(r, half);
        if (mm >= M) continue;
        long long pix = mm;
        if (a.mode == kScatter) {
          const int q = n / a.cout, py = mm / a.W, px = mm - py * a.W;
          pix = (long long)(2 * py + (q >> 1)) * (2 * a.W) + (2 * px + (q & 1));
        }
        conv_store(a, pix, co, acc[i][j][r] + b);
      }
    }
  }
}

// 3 x 3 convolution (mode 0, cin a multiple of 16, H a multiple of 4) with
// the operands staged in LDS.  A block computes 4 output rows x 32 columns
// x (32 WN) channels; each wavefront one row.  Per 16 input channels the
// block loads the input tile with its halo (6 x 34 pixels, zero outside the
// image) and the weights of all 9 taps for its channels into LDS; the 9
// taps then read their A fragments from the same tile shifted by (dy, dx),
// so every input value is fetched from L2 once per block instead of 9 times.
constexpr int kC3Rows = 4, kC3Cols = 32, kC3TileW = kC3Cols + 2, kC3TileP = (kC3Rows + 2) * kC3TileW;
template <int WN>
__global__ void __launch_bounds__(256) conv3_lds_kernel(ConvGemm a) {
  __shared__ __attribute__((aligned(16))) uint16_t s_in[kC3TileP * 16];
  __shared__ __attribute__((aligned(16))) uint16_t s_w[9 * 32 * WN * 16];
  const int t = threadIdx.x, wave = t >> 6, lane = t & 63, half = lane >> 5, l32 = lane & 31;
  const int N = a.cout, NB = 32 * WN, K = 9 * a.cin;
  const int ntn = (N + NB - 1) / NB, ncol = (a.W + kC3Cols - 1) / kC3Cols;
  const int nt = blockIdx.x % ntn, rest = blockIdx.x / ntn;
  const int ct = rest % ncol, band = rest / ncol;
  const int y0 = band * kC3Rows, x0 = ct * kC3Cols, n0 = nt * NB;
  const int yi0 = y0 % a.H;  // image row of the band's first row (bands never straddle images)
  f32x16 acc[WN];
#pragma unroll
  for (int j = 0; j < WN; ++j)
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[j][r] = 0.f;
  const u32x4 zero = {0u, 0u, 0u, 0u};
  for (int c0 = 0; c0 < a.cin; c0 += 16) {
    __syncthreads();  // the previous chunk's fragments have been read
    for (int item = t; item < kC3TileP * 2; item += 256) {
      const int p = item >> 1, h = item & 1, r = p / kC3TileW, c = p - r * kC3TileW;
      const int yl = yi0 - 1 + r, x = x0 - 1 + c;
      u32x4 v = zero;
      if (yl >= 0 && yl < a.H && x >= 0 && x < a.W)
        v = *(const u32x4 *)(a.in + ((long long)(y0 - 1 + r) * a.W + x) * a.in_stride + c0 +
                             8 * h);
      *(u32x4 *)(s_in + p * 16 + 8 * h) = v;
    }
    for (int item = t; item < 9 * NB * 2; item += 256) {
      const int h = item & 1, tn = item >> 1, tap = tn / NB, n = tn - tap * NB;
      u32x4 v = zero;
      if (n0 + n < N)
        v = *(const u32x4 *)(a.w + (long long)(n0 + n) * K + tap * a.cin + c0 + 8 * h);
      *(u32x4 *)(s_w + tn * 16 + 8 * h) = v;
    }
    __syncthreads();
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
      const int dy = tap / 3, dx = tap % 3;
      const bf16x8 av =
          *(const bf16x8 *)(s_in + ((wave + dy) * kC3TileW + l32 + dx) * 16 + 8 * half);
#pragma unroll
      for (int j = 0; j < WN; ++j) {
        const bf16x8 bv = *(const bf16x8 *)(s_w + (tap * NB + 32 * j + l32) * 16 + 8 * half);
        acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc[j], 0, 0, 0);
      }
    }
  }
  const int y = y0 + wave;
#pragma unroll
  for (int j = 0; j < WN; ++j) {
    const int n = n0 + 32 * j + l32;
    if (n >= N) continue;
    const float b = a.bias ? a.bias[n] : 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int x = x0 + acc_row(r, half);
      if (x < a.W) conv_store(a, (long long)y * a.W + x, n, acc[j][r] + b);
    }
  }
}

// Weight gradient: out[s][i][j] = sum over pixels p of split s of
// L[p][i] R[p][j].  L: the output gradient at the output pixels
// (l_s2d = 0: L[p][i] = l[p * l_stride + i]; l_s2d = 1: i = q * lc + c reads
// fine pixel (2y + q / 2, 2x + q % 2)); R: the layer input gathered as the
// forward convolution gathers it (j = tap * rc + c, zero outside the image).
// One wavefront computes a 64 x 64 tile of out.  Per step of 16 pixels each
// lane loads 16 bytes (8 consecutive channels of one pixel) from two
// 8-column chunks of L and of R, and writes them transposed into this
// wavefront's part of LDS ([column][pixel]); the MFMA fragments
// (8 consecutive pixels of one column) are then single 16-byte LDS reads.
constexpr int kWgRow = 24;  // LDS row pitch in bf16 (16 pixels + 8 pad): 16-byte aligned rows
__global__ void __launch_bounds__(256) wgrad_mfma_kernel(Wgrad a) {
  __shared__ __attribute__((aligned(16))) uint16_t lds[4][2][64 * kWgRow];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, half = lane >> 5, l32 = lane & 31;
  const int nti = (a.I + 63) / 64, ntj = (a.J + 63) / 64;
  const int tile = blockIdx.x * 4 + wave;
  if (tile >= nti * ntj * a.nsplit) return;  // whole wavefronts exit together
  const int ti = tile % nti, tj = (tile / nti) % ntj, s = tile / (nti * ntj);
  const int i0 = ti * 64, j0 = tj * 64;
  const long long P = (long long)a.rows * a.W;
  const long long p_begin = (long long)s * a.chunk;
  const long long p_end = p_begin + a.chunk < P ? p_begin + a.chunk : P;
  uint16_t *la = lds[wave][0], *lb = lds[wave][1];
  // loading role: pixel slot ps of each step, column chunks 2 cg and 2 cg + 1
  const int ps = lane >> 2, cg = lane & 3;
  int lq[2], lcol[2], dy[2], dx[2], rcol[2];
  bool lok[2], rok[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const int i = i0 + 8 * (2 * cg + c), j = j0 + 8 * (2 * cg + c);
    lok[c] = i < a.I;
    lq[c] = a.l_s2d ? i / a.lc : 0;
    lcol[c] = a.l_s2d ? i - lq[c] * a.lc : i;
    rok[c] = j < a.J;
    const int tap = j / a.rc, pad = a.ks / 2;
    rcol[c] = j - tap * a.rc;
    dy[c] = tap / a.ks - pad;
    dx[c] = tap % a.ks - pad;
  }
  long long p = p_begin + ps;
  int yg = (int)(p / a.W), x = (int)(p - (long long)yg * a.W), yi = yg % a.H;
  f32x16 acc[2][2];
#pragma unroll
  for (int u = 0; u < 2; ++u)
#pragma unroll
    for (int v = 0; v < 2; ++v)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[u][v][r] = 0.f;
  const bool bias_tile = a.bias_out && tj == 0;  // also sums L over the pixels
  float lsum = 0.f;
  const u32x4 zero = {0u, 0u, 0u, 0u};
  for (long long p0 = p_begin; p0 < p_end; p0 += 16, p += 16) {
    const bool pok = p < p_end;
    u32x4 lv[2], rv[2];  // 8 bf16 each, element t in the (t & 1) half of word t / 2
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      lv[c] = zero;
      if (pok && lok[c]) {
        long long li;
        if (a.l_s2d)
          li = ((long long)(2 * yg + (lq[c] >> 1)) * (2 * a.W) + 2 * x + (lq[c] & 1)) *
                   a.l_stride + lcol[c];
        else
          li = p * a.l_stride + lcol[c];
        lv[c] = *(const u32x4 *)(a.l + li);
      }
      rv[c] = zero;
      const int yy = yi + dy[c], xx = x + dx[c];
      if (pok && rok[c] && yy >= 0 && yy < a.H && xx >= 0 && xx < a.W)
        rv[c] = *(const u32x4 *)(a.r + ((long long)(yg + dy[c]) * a.W + xx) * a.r_stride +
                                 rcol[c]);
    }
    // transpose into LDS: column 8 (2 cg + c) + t, pixel ps
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int row = 8 * (2 * cg + c) + t;
        la[row * kWgRow + ps] = (uint16_t)(lv[c][t >> 1] >> (16 * (t & 1)));
        lb[row * kWgRow + ps] = (uint16_t)(rv[c][t >> 1] >> (16 * (t & 1)));
      }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    bf16x8 af[2], bf[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      af[u] = *(const bf16x8 *)(la + (32 * u + l32) * kWgRow + 8 * half);
      bf[u] = *(const bf16x8 *)(lb + (32 * u + l32) * kWgRow + 8 * half);
    }
    if (bias_tile) {
      const bf16x8 r0 = *(const bf16x8 *)(la + lane * kWgRow);
      const bf16x8 r1 = *(const bf16x8 *)(la + lane * kWgRow + 8);
      float t0 = 0.f;
#pragma unroll
      for (int t = 0; t < 8; ++t) t0 += (float)r0[t];
#pragma unroll
      for (int t = 0; t < 8; ++t) t0 += (float)r1[t];
      lsum += t0;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();  // the next step overwrites this wavefront's LDS rows
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v)
        acc[u][v] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[u], bf[v], acc[u][v], 0, 0, 0);
    x += 16;
    while (x >= a.W) {
      x -= a.W;
      ++yg;
      if (++yi == a.H) yi = 0;
    }
  }
  if (bias_tile && i0 + lane < a.I) a.bias_out[(long long)s * a.I + i0 + lane] = lsum;
  float *out = a.out + (long long)s * a.I * a.J;
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    const int j = j0 + 32 * v + l32;
    if (j >= a.J) continue;
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int ii = i0 + 32 * u + acc_row(r, half);
        if (ii < a.I) out[(long long)ii * a.J + j] = acc[u][v][r];
      }
  }
}

// out[e] = sum over s < nsplit of part[s * n + e]: `so` split lanes per
// output (a power of two dividing 256), each summing every so-th split, then
// a tree over the lanes -- a fixed order for a given (n, nsplit).
__global__ void __launch_bounds__(256) reduce_splits_kernel(const float *__restrict__ part,
                                                            float *out, long long n,
                                                            int nsplit, int so) {
  __shared__ float red[256];
  const int eo = 256 / so, el = threadIdx.x % eo, sl = threadIdx.x / eo;
  const long long e = (long long)blockIdx.x * eo + el;
  float v0 = 0.f, v1 = 0.f, v2 = 0.f, v3 = 0.f;
  if (e < n) {
    int sp = sl;
    for (; sp + 3 * so < nsplit; sp += 4 * so) {
      v0 += part[(long long)sp * n + e];
      v1 += part[(long long)(sp + so) * n + e];
      v2 += part[(long long)(sp + 2 * so) * n + e];
      v3 += part[(long long)(sp + 3 * so) * n + e];
    }
    for (; sp < nsplit; sp += so) v0 += part[(long long)sp * n + e];
  }
  red[threadIdx.x] = (v0 + v1) + (v2 + v3);
  __syncthreads();
  for (int w = so / 2; w >= 1; w /= 2) {
    if (sl < w) red[sl * eo + el] += red[(sl + w) * eo + el];
    __syncthreads();
  }
  if (sl == 0 && e < n) out[e] = red[el];
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
  float v0 = 0.f, v1 = 0.f;
  if (c < C) {
    long long p = p_begin + r;
    for (; p + 4 < p_end; p += 8) {
      v0 += from_bf16(src[p * stride + off + c]);
      v1 += from_bf16(src[(p + 4) * stride + off + c]);
    }
    if (p < p_end) v0 += from_bf16(src[p * stride + off + c]);
  }
  red[r][threadIdx.x & 63] = v0 + v1;
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
  if (((uintptr_t)a.in | (uintptr_t)a.w) & 15) return hipErrorInvalidValue;  // 16-byte loads
  if (a.mode == kConv && a.ks == 3 && a.cin % 16 == 0 && a.H % kC3Rows == 0) {
    const long long blocks = (long long)(a.rows / kC3Rows) * ((a.W + kC3Cols - 1) / kC3Cols);
    if (a.cout > 32) {
      hipLaunchKernelGGL(HIP_KERNEL_NAME(conv3_lds_kernel<2>),
                         dim3((unsigned)(blocks * ((a.cout + 63) / 64))), dim3(256), 0, s, a);
    } else {
      hipLaunchKernelGGL(HIP_KERNEL_NAME(conv3_lds_kernel<1>),
                         dim3((unsigned)(blocks * ((a.cout + 31) / 32))), dim3(256), 0, s, a);
    }
    return hipGetLastError();
  }
  const long long M = (long long)a.rows * a.W, N = a.mode == kScatter ? 4LL * a.cout : a.cout;
  if (N > 32) {  // 64 x 64 per wavefront
    const long long tiles = ((M + 63) / 64) * ((N + 63) / 64);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(conv_mfma_kernel<2, 2>), dim3(blocks_of(tiles, 4)),
                       dim3(256), 0, s, a);
  } else {  // 128 x 32
    const long long tiles = ((M + 127) / 128) * ((N + 31) / 32);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(conv_mfma_kernel<4, 1>), dim3(blocks_of(tiles, 4)),
                       dim3(256), 0, s, a);
  }
  return hipGetLastError();
}

static void launch_reduce_splits(const float *part, float *out, long long n, int nsplit,
                          hipStream_t s) {
  int so = 1;
  while (so < 256 && (long long)so * 8 < nsplit) so *= 2;
  hipLaunchKernelGGL(reduce_splits_kernel, dim3(blocks_of(n, 256 / so)), dim3(256), 0, s,
                     part, out, n, nsplit, so);
}

static long long wgrad_plan(Wgrad &a, long long workspace_floats) {
  const long long P = (long long)a.rows * a.W;
  const long long tiles = (long long)((a.I + 63) / 64) * ((a.J + 63) / 64);
  long long ns = (2048 + tiles - 1) / tiles;
  const long long max_ns = (P + 511) / 512;
  if (ns > max_ns) ns = max_ns;
  const long long fit = workspace_floats / ((long long)a.I * a.J + a.I);
  if (ns > fit) ns = fit;
  if (ns < 1) ns = 1;
  long long chunk = (P + ns - 1) / ns;
  chunk = (chunk + 15) / 16 * 16;
  a.nsplit = (int)((P + chunk - 1) / chunk);
  a.chunk = chunk;
  return a.nsplit;
}

hipError_t launch_wgrad(Wgrad a, float *dw, float *db, float *workspace,
                        long long workspace_floats, hipStream_t s) {
  if (a.I <= 0 || a.J <= 0 || a.rows <= 0 || a.H <= 0 || a.rows % a.H || a.W <= 0 ||
      (a.ks != 1 && a.ks != 3) || a.rc <= 0 || a.rc % 8 || a.J != a.ks * a.ks * a.rc ||
      a.l_stride % 8 || a.r_stride % 8 || (((uintptr_t)a.l | (uintptr_t)a.r) & 15) ||
      (a.l_s2d && (a.lc <= 0 || a.lc % 8 || a.I != 4 * a.lc)))
    return hipErrorInvalidValue;
  wgrad_plan(a, workspace ? workspace_floats : 0);
  const long long n = (long long)a.I * a.J;
  const bool split = a.nsplit > 1;
  a.out = split ? workspace : dw;
  a.bias_out = db ? (split ? workspace + (long long)a.nsplit * n : db) : nullptr;
  const long long tiles = (long long)((a.I + 63) / 64) * ((a.J + 63) / 64) * a.nsplit;
  hipLaunchKernelGGL(wgrad_mfma_kernel, dim3(blocks_of(tiles, 4)), dim3(256), 0, s, a);
  if (split) {
    launch_reduce_splits(workspace, dw, n, a.nsplit, s);
    if (db) launch_reduce_splits(a.bias_out, db, a.I, a.nsplit, s);
  }
  return hipGetLastError();
}

hipError_t launch_channel_sum(const uint16_t *src, int stride, int off, int c, long long P,
                              float *out, float *workspace, long long workspace_floats,
                              hipStream_t s) {
  if (c <= 0 || P <= 0 || off < 0 || off + c > stride) return hipErrorInvalidValue;
  long long ns = (P + 511) / 512;
  if (ns > 4096) ns = 4096;
  if (!workspace || ns > workspace_floats / c) ns = workspace ? workspace_floats / c : 1;
  if (ns < 1) ns = 1;
  const long long chunk = (P + ns - 1) / ns;
  ns = (P + chunk - 1) / chunk;
  hipLaunchKernelGGL(channel_sum_kernel, dim3(blocks_of(c, 64), (unsigned)ns), dim3(256), 0,
                     s, src, stride, off, c, P, chunk, ns > 1 ? workspace : out);
  if (ns > 1) launch_reduce_splits(workspace, out, (long long)c, (int)ns, s);
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
