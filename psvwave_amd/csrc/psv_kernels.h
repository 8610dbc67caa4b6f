// psv_kernels.h -- device data layout, per-point stencil arithmetic and the
// launch interface of the MI355X fdModel kernels.
//
// Grids on the device are x-fastest ("[iz][ix]"): wavefront lanes run along
// x, so every row segment a wave touches is one contiguous, coalesced
// transaction.  The stencil arrays (state and materials) carry a zero margin
// of kHaloX columns / kHaloZ rows around the tiled extent so the LDS halo
// loads need no bounds checks; all other grids (parameters, kernels,
// snapshots) are compact, index iz*nx + ix.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace psv {

constexpr int kTileX = 64;   // tile width = one wavefront of lanes
constexpr int kHaloX = 16;   // margin columns left of x = 0 (rows stay
                             // 128-byte aligned at tile starts)
constexpr int kHaloZ = 8;    // margin rows above z = 0
constexpr int kBlock = 256;  // threads per block for the step kernels

// 4th-order staggered coefficients (src/fdModel.h:201-202).
constexpr double kC1 = 9.0 / 8.0;
constexpr double kC2 = 1.0 / 24.0;

// One injection term applied by the owner of a grid point after its
// velocity update.  Forward: a moment-tensor term (src/fdModel.cpp:681-751),
// value ((coef * stf[src][it]) * dt) * b / D4 with coef = +-M or +-0.25 M.
// Adjoint: an adjoint source (src/fdModel.cpp:951-958), value
// ((dt * b) * a[src][it]) / (dx * dz) with src = receiver index.
struct InjEntry {
  int local;  // (z - z0) * kTileX + (x - x0) inside the owning tile
  int comp;   // 0 = vx, 1 = vz
  int src;    // source (forward) or receiver (adjoint) index
  int pad;
  double coef;
};

// Arguments of one time step of one shot.
struct StepArgs {
  int nx, nz, pitch, ntx, ntz, tz;  // grid, row pitch, tiles, tile height
  long long origin;                  // padded index of (0, 0)
  const double *vx_i, *vz_i, *txx_i, *tzz_i, *txz_i;  // state at step start
  double *vx_o, *vz_o, *txx_o, *tzz_o, *txz_o;        // state after the step
  const double *lm, *mu, *b;                          // padded materials
  const double *taper_lut;                            // np_boundary + 1 values
  int np;
  double dts;  // +dt forward, -dt adjoint (stencil sign)
  double dt, dx, dz, dxdz, d4x, d4z;
  double rdx, rdz;  // RN(1/dx), RN(1/dz) for RecipDiv
  // FastDiv split reciprocals: zh = RD(1/d), zl = RN(1/d - zh) >= 0
  double zhx, zlx, zhz, zlz;
  // taper factors by column and by row: taper(x, z) = min(taper_x[x],
  // taper_z[z]) (the table is nondecreasing in the nesting count); valid
  // (margins hold 1) for x in [-kHaloX, pitch - kHaloX) and z in
  // [-kHaloZ, rows - kHaloZ)
  const double *taper_x, *taper_z;
  int it, nt, adjoint;
  // snapshot slot of this step (compact [iz][ix]); written forward, read
  // adjoint when do_snap is set
  int do_snap;
  double *snap[5];  // vx, vz, txx, tzz, txz
  // kernel correlation (adjoint, do_snap steps): Lame kernels, compact
  double *lam_k, *mu_k, *rho_k;
  // kfresh: the three arrays start at zero for this step (a deferred
  // adjoint's per-slot terms), so the correlation writes its terms without
  // reading them (0 - e == -e, 0 + e == e exactly)
  int kfresh;
  int kx0, kx1, kz0, kz1;
  double wcorr;  // snapshot_interval * dt
  // injection lists for this shot, CSR by tile
  const int *tile_off;
  const InjEntry *entries;
  const double *stf;             // [n_sources][nt] (forward)
  const double *adj_x, *adj_z;   // this shot's adjoint sources [nr][nt]
  // receiver record (forward): this shot's traces [nr][nt]
  int nr;
  const long long *rec_idx;      // padded index of each receiver
  double *rtf_x, *rtf_z;
  int rec_increments;            // store dt v / (dx dz) only; a scan follows
  int xcd_remap;                 // map consecutive tiles to one XCD's L2
  int ntiles;
  // two-step launches: the second step of the pair
  int it1, do_snap1;
  double *snap1[5];
  // injection terms of every point of a tile's 4-point halo region (CSR by
  // tile, local index over the (64 + 8) x (tz + 8) region)
  const int *halo_off;
  const InjEntry *halo_entries;
  // receivers by owning tile (forward record of two-step launches)
  const int *rec_off;
  const InjEntry *rec_entries;
};

// Pointwise arithmetic, written once for the device kernels.  Expression
// trees follow the reference operation by operation; with contraction off
// the results are bitwise those of the reference's double build.
//
// Division by a grid spacing.  IeeeDiv is the plain IEEE quotient.
// RecipDiv multiplies by y = RN(1/d) and corrects once with an FMA
// (Markstein): for |x| in [2^-960, 2^1000] the result is the correctly
// rounded quotient, i.e. bit-identical to x / d; zeros keep their sign via
// x * y, and every other x (tiny, huge, inf, NaN) takes the IEEE division.
struct IeeeDiv {
  double d;
  __host__ __device__ double operator()(double x) const { return x / d; }
};
struct RecipDiv {
  double d, y;
  __host__ __device__ double operator()(double x) const {
    double q = x * y;
    double r = __builtin_fma(-q, d, x);
    double res = __builtin_fma(r, y, q);
    double ax = __builtin_fabs(x);
    if (x == 0.0) res = q;
    if (!(ax >= 0x1p-960 && ax <= 0x1p1000) && x != 0.0) res = x / d;
    return res;
  }
};

// Stress at one point from velocities v*, accessed at offsets sx (one column)
// and sz (one row).  src/fdModel.cpp:615-642 (forward), :882-909 (adjoint).
template <class DX, class DZ>
__host__ __device__ inline void stress_point(const double *vx,
                                             const double *vz, long long i,
                                             long long sx, long long sz,
                                             double lm, double mu,
                                             double taper, double dts,
                                             DX divx, DZ divz, double &txx,
                                             double &tzz, double &txz) {
  double dvx_x = kC1 * (vx[i + sx] - vx[i]) + kC2 * (vx[i - sx] - vx[i + 2 * sx]);
  double dvz_z = kC1 * (vz[i] - vz[i - sz]) + kC2 * (vz[i - 2 * sz] - vz[i + sz]);
  double dvx_z = kC1 * (vx[i + sz] - vx[i]) + kC2 * (vx[i - sz] - vx[i + 2 * sz]);
  double dvz_x = kC1 * (vz[i] - vz[i - sx]) + kC2 * (vz[i - 2 * sx] - vz[i + sx]);
  double la = lm - 2 * mu;  // la array of src/fdModel.cpp:1046, recomputed
  txx = taper * (txx + dts * (divx(lm * dvx_x) + divz(la * dvz_z)));
  tzz = taper * (tzz + dts * (divx(la * dvx_x) + divz(lm * dvz_z)));
  txz = taper * (txz + dts * mu * (divz(dvx_z) + divx(dvz_x)));
}

// Velocity at one point from stresses.  src/fdModel.cpp:661-676 (forward),
// :931-946 (adjoint); b_vx == b_vz == 1/rho (src/fdModel.cpp:1047-1048).
template <class DX, class DZ>
__host__ __device__ inline void velocity_point(
    const double *txx, const double *tzz, const double *txz, long long i,
    long long sx, long long sz, double b, double taper, double dts, DX divx,
    DZ divz, double &vx, double &vz) {
  double ax = divx(kC1 * (txx[i] - txx[i - sx]) + kC2 * (txx[i - 2 * sx] - txx[i + sx])) +
              divz(kC1 * (txz[i] - txz[i - sz]) + kC2 * (txz[i - 2 * sz] - txz[i + sz]));
  double az = divx(kC1 * (txz[i + sx] - txz[i]) + kC2 * (txz[i - sx] - txz[i + 2 * sx])) +
              divz(kC1 * (tzz[i + sz] - tzz[i]) + kC2 * (tzz[i - sz] - tzz[i + 2 * sz]));
  vx = taper * (vx + b * dts * ax);
  vz = taper * (vz + b * dts * az);
}

// Correlation of the adjoint state with a forward snapshot at one point of
// the kernel region.  src/fdModel.cpp:832-858.
__host__ __device__ inline void correlate_point(
    double w, double lm, double mu, double vx, double vz, double txx,
    double tzz, double txz, double avx, double avz, double atxx, double atzz,
    double atxz, double &rho_k, double &lam_k, double &mu_k) {
  double la = lm - 2 * mu;
  rho_k -= w * (avx * vx + avz * vz);
  double qa1 = atxx - (atzz * la) / lm;
  double qa2 = atzz - (atxx * la) / lm;
  double q1 = txx - (tzz * la) / lm;
  double q2 = tzz - (txx * la) / lm;
  double d = lm - ((la * la) / (lm));
  lam_k += w * ((qa1 + qa2) * (q1 + q2)) / (d * d);
  mu_k += w * 2 * ((q1 * qa1 + q2 * qa2) / (d * d) + 2 * (txz * atxz / (4 * mu * mu)));
}

// Taper of a point from its nesting count (src/fdModel.cpp:263-298):
// count = min(np, 1 + distance to the nearest grid edge).
__host__ __device__ inline int taper_count(int x, int z, int nx, int nz, int np) {
  int d = x < nx - 1 - x ? x : nx - 1 - x;
  int dz = z < nz - 1 - z ? z : nz - 1 - z;
  d = d < dz ? d : dz;
  return d + 1 < np ? d + 1 : np;
}

// Step-kernel variants: 0-2 divide by the spacings with IEEE divisions,
// 3-5 are the same kernels with RecipDiv (bit-identical results), 6-7
// advance two time steps per launch, 8-13 are the lean one-step kernels
// (64 x 16, 8, 12, 20, 24, 32 tiles, FastDiv with IEEE fallback), 14 the
// lean two-step kernel (64 x 8 tiles), 15-17 the pipelined lean kernels,
// 18 the lean kernel on 64 x 4 tiles (7 waves per SIMD), 19 and 20 the
// 64 x 8 and 64 x 12 lean kernels held to 6 and 5 waves per SIMD, 21 the
// 64 x 8 lean kernel with the velocity halo columns in the row transfers,
// 22 / 23 the 64 x 8 / 64 x 4 lean kernels with separate velocity and
// stress LDS stages, 24-26 lean kernels with ring operands by DMA, 27 / 28
// variants 24 / 26 with their loads capped to the grid.
enum StepVariant {
  kFusedTz8 = 0,
  kFusedTz16 = 1,
  kSplitTz8 = 2,
  kFusedTz8Recip = 3,
  kFusedTz16Recip = 4,
  kSplitTz8Recip = 5,
  kTwoStepTz8 = 6,
  kTwoStepTz16 = 7,
  kLeanTz16 = 8,
  kLeanTz8 = 9,
  kLeanTz12 = 10,
  kLeanTz20 = 11,
  kLeanTz24 = 12,
  kLeanTz32 = 13,
  kLeanTwoTz8 = 14,
  kLeanPipe2Tz8 = 15,
  kLeanPipe3Tz8 = 16,
  kLeanPipe4Tz8 = 17,
  kLeanTz4 = 18,
  kLeanTz8W6 = 19,
  kLeanTz12W5 = 20,
  kLeanTz8Halo = 21,
  kLeanTz8Sep = 22,
  kLeanTz4Sep = 23,
  kLeanTz8Ring = 24,  // lean 64 x 8, halo columns and ring operands by DMA
  kLeanTz8RingOnly = 25,  // lean 64 x 8, ring operands by DMA
  kLeanTz16Ring = 26,     // lean 64 x 16, ring operands by DMA (4 waves per SIMD)
  kLeanTz8RingClamp = 27,   // 24 with loads capped to the grid
  kLeanTz16RingClamp = 28,  // 26 with loads capped to the grid
  kNumVariants = 29
};
constexpr int kMaxTileHeight = 32;

// Lean one-step kernels (variants 8-13): lane = column, FastDiv, reduced
// stress ring.  They need FastDiv's divisor range and a taper table that is
// nondecreasing in the nesting count; the host checks both.
bool step_is_lean(int variant);       // lean one-step kernels (8-13, 18-23)
bool step_uses_fastdiv(int variant);  // lean one- and two-step kernels
// Pipelined lean kernels (15-17): 64 x 8 tiles, a block advances 2, 3 or 4
// tiles of one column strip in turn, fetching the next while one computes.
bool step_is_lean_pipe(int variant);
// Step variants batched gradients can use (the correlating batch kernels).
bool batch_gradient_variant_ok(int variant);
bool fast_division_ok(double d);
void fast_division_constants(double d, double *zh, double *zl);

// Host launchers (psv_kernels.hip).  All enqueue on `stream`.
hipError_t launch_step(int variant, const StepArgs &a, hipStream_t stream,
                       hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
int step_tile_height(int variant);
bool step_is_inplace(int variant);
bool step_is_two_step(int variant);
// The single-step variant a two-step variant uses for an odd last step.
int step_single_variant(int variant);
// Two time steps (a.it, a.it1) in one launch; variant 6 or 7.
hipError_t launch_two_steps(int variant, const StepArgs &a, hipStream_t stream,
                            hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// One time step (or two, for the two-step variants) of every tile of a batch
// of models: descs holds per (model, shot, parity) launch arguments, tiles
// the (model, tile) pairs of all models.
// With `snapshots`, step it of a member stores (forward) or correlates with
// (adjoint; launched with corr) snapshot slot it / si when si divides it:
// the descriptor's snap pointers are the bases of a one-shot snapshot store.
// One batched launch: descs[(member * n_shots + shot) * 2 + parity] are the
// members' step arguments, tiles the (member, tile) pairs.  shot >= 0
// advances that shot (ntiles = pairs); shot < 0 advances every shot at once
// (lean steps): ntiles counts (pair, shot) items, item g being shot
// g % n_shots of pair g / n_shots, so the shots of one tile are neighbours
// in the grid (one XCD, close in time) and share their material reads in
// its L2.
struct BatchLaunch {
  const StepArgs *descs;
  const int2 *tiles;
  int ntiles, n_shots, shot, parity, it, it1, xcd_remap;
  int snapshots, si;
};
hipError_t launch_batch_step(int variant, bool two, const BatchLaunch &b,
                             hipStream_t stream, hipEvent_t e0 = nullptr,
                             hipEvent_t e1 = nullptr, bool corr = false);
// Steps [k0, k1) of shot b.shot of every member in one cooperative launch
// (lean 64 x 8 tiles; a grid barrier between steps).
hipError_t launch_batch_persistent(const BatchLaunch &b, int k0, int k1,
                                   hipStream_t stream);
// bfloat16 MFMA kernels of the preconditioner network (psv_conv.hip).
// NHWC bfloat16 activations with images stacked along the rows.
enum ConvMode { kConv = 0, kScatter = 1, kGather2x2 = 2 };
struct ConvGemm {
  const uint16_t *in;   // gathered input, in_stride channels per pixel
  const uint16_t *w;    // [N][K], K = taps * cin
  const float *bias;    // [cout] or null
  uint16_t *out16;      // bfloat16 output, or
  float *out32;         // fp32 output
  uint16_t *pre16;      // kConv only: pre-activation [pixels][cout], or null
  const uint16_t *dz;   // multiply by gelu'(dz[pixel * dz_stride + co]), or null
  int in_stride, out_stride, out_off, dz_stride;
  int rows, H, W;       // rows = batch * H of the output (kScatter: of the input)
  int cin, cout, ks, mode, gelu;
};
hipError_t launch_conv_gemm(const ConvGemm &a, hipStream_t s);
// Weight gradient out[i][j] = sum_p L[p][i] R[p][j] over the rows x W
// pixels p (see psv_conv.hip); fp32 result in dw and, when db is given,
// db[i] = sum_p L[p][i]; split partial sums in workspace (fewer splits when
// it is small or null).
struct Wgrad {
  const uint16_t *l;
  const uint16_t *r;
  float *out;
  float *bias_out;  // per-split sums of L over the pixels [nsplit][I], or null
  long long chunk;
  int l_stride, l_s2d, lc, r_stride, rc, ks, I, J, rows, H, W, nsplit;
};
hipError_t launch_wgrad(Wgrad a, float *dw, float *db, float *workspace,
                        long long workspace_floats, hipStream_t s);
hipError_t launch_channel_sum(const uint16_t *src, int stride, int off, int c, long long P,
                              float *out, float *workspace, long long workspace_floats,
                              hipStream_t s);
// One image: cin a multiple of 8, weights [N][K], K = (ky * ks + kx) * cin +
// ci; `transposed`: 2x2 stride-2 transposed convolution (ks = 1, N = 4
// cout).  Output bf16 or fp32 at channel offset out_off of rows of
// out_stride channels; optional fused erf GELU.
hipError_t launch_conv_bf16(const uint16_t *in, const uint16_t *w, const float *bias,
                            void *out, bool out_f32, int h, int wd, int cin, int cout,
                            int ks, bool transposed, bool gelu, int out_stride,
                            int out_off, hipStream_t s);
hipError_t launch_maxpool2_bf16(const uint16_t *in, uint16_t *out, int h, int w, int c,
                                hipStream_t s);
hipError_t launch_pool_backward(const uint16_t *act, const uint16_t *z, const float *dpool,
                                const uint16_t *dskip, int skip_stride, int skip_off,
                                uint16_t *out, int rows, int w, int c, hipStream_t s);
// Running sums of recorded receiver increments, in place: [nr][nt] x 2.
hipError_t launch_receiver_scan(double *rx, double *rz, int nr, int nt,
                                hipStream_t s);

hipError_t launch_fill(double *p, long long n, double v, hipStream_t s);
// Zeroes arrays ptrs[k] of lens[k] doubles, k < n (device tables), in one
// launch of blocks_x x n blocks.
hipError_t launch_zero_many(double *const *ptrs, const long long *lens, int n, int blocks_x,
                            hipStream_t s);
hipError_t launch_update_from_velocity(const double *vp, const double *vs,
                                       const double *rho, double *lm,
                                       double *mu, double *b, int nx, int nz,
                                       int pitch, long long origin,
                                       hipStream_t s);
hipError_t launch_adjoint_sources(const double *sx, const double *sz,
                                  const double *ox, const double *oz,
                                  double *ax, double *az, long long n,
                                  hipStream_t s);
// Adds deferred correlation terms [slot][rho, lam, mu][nx * nz] to the
// kernels over the kernel region, slots from last to first.
hipError_t launch_replay_correlation(double *rho_k, double *lam_k, double *mu_k,
                                     const double *d, int nslots, long long ngrid, int nx,
                                     int kx0, int kx1, int kz0, int kz1, hipStream_t s);
hipError_t launch_map_kernels(const double *vp, const double *vs,
                              const double *rho, const double *lam_k,
                              const double *mu_k, const double *rho_l,
                              double *vp_k, double *vs_k, double *rho_v,
                              long long n, hipStream_t s);
// misfit: partials must hold 2 * kMisfitBlocks doubles, out 1 double.
constexpr int kMisfitBlocks = 512;
hipError_t launch_misfit(const double *sx, const double *sz, const double *ox,
                         const double *oz, long long n, double half_dt,
                         double *partials, double *out, hipStream_t s);
// The same reduction with its partial sums in a fixed device array of the
// library (no allocation; 2 * kMisfitBlocks doubles per device, whatever n).
hipError_t launch_misfit_noalloc(const double *sx, const double *sz, const double *ox,
                                 const double *oz, long long n, double half_dt,
                                 double *out, hipStream_t s);
hipError_t launch_model_vector(int mode, double *vp, double *vs, double *rho,
                               double *vec, int nx, int nxf, int nzf, int npx,
                               int npz, int gx0, int gz0, hipStream_t s);
hipError_t launch_gradient_vector(const double *vp_k, const double *vs_k,
                                  const double *rho_k, double *vec, int nx,
                                  int nxf, int nzf, int npx, int npz, int gx0,
                                  int gz0, hipStream_t s);
// Batched layout change between compact [slice][iz][ix] (device) and the
// reference's [slice][ix][iz].  to_ref = 1: device -> reference.
hipError_t launch_transpose(const double *in, double *out, int nx, int nz,
                            int slices, int to_ref, hipStream_t s);
// Streaming copy of n doubles (n even), for bandwidth and counter
// calibration.
hipError_t launch_copy(const double *in, double *out, long long n,
                       hipStream_t s);
// Compact [iz][ix] -> padded stencil array (grid points only).
hipError_t launch_pad(const double *in, double *out, int nx, int nz,
                      int pitch, long long origin, hipStream_t s);
// Padded stencil array -> compact [iz][ix].
hipError_t launch_unpad(const double *in, double *out, int nx, int nz,
                        int pitch, long long origin, hipStream_t s);

}  // namespace psv
