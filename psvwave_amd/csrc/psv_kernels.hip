// psv_kernels.hip -- hand-written gfx950 kernels for the fdModel hot path.
//
// The time step (stress then velocity, src/fdModel.cpp:599-678 forward,
// :864-948 adjoint) is one launch per step in the fused variant:
//
//   phase A  stage vx, vz of the tile plus a 4-point halo in LDS;
//   phase B  update the three stresses on the tile plus a 2-point ring
//            (the ring is recomputed redundantly by neighbouring tiles), keep
//            them in LDS and store the tile's own stresses;
//   phase C  update vx, vz on the tile from the LDS stresses, apply this
//            tile's injection terms in reference order, store.
//
// A step reads one buffer set and writes the other (ping-pong), so no tile
// ever reads what another tile of the same launch writes.  Snapshot stores
// (forward), kernel correlation (adjoint) and the receiver record (forward)
// ride in the same launch on the owning threads.
//
// The split variant does the same work as two in-place launches per step
// (stress, velocity) without redundant compute; it is kept as an
// independent cross-check of the fused kernel and as a tuning alternative.
#include "psv_kernels.h"

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_ext.h>

#include <cmath>
#include <mutex>

namespace psv {

namespace {

__device__ inline long long pidx(const StepArgs &a, int x, int z) {
  return a.origin + (long long)z * a.pitch + x;
}

__device__ inline double taper_at(const StepArgs &a, int x, int z) {
  int c = taper_count(x, z, a.nx, a.nz, a.np);
  return c >= a.np ? 1.0 : a.taper_lut[c];
}

__device__ inline bool interior(const StepArgs &a, int x, int z) {
  return x >= 2 && x < a.nx - 2 && z >= 2 && z < a.nz - 2;
}

// Tile handled by this block.  With xcd_remap, blocks that the dispatcher
// sends to the same XCD (block id mod 8) get a contiguous run of tiles, so
// halo re-reads between neighbouring tiles hit that XCD's L2.
__device__ inline int block_tile(const StepArgs &a) {
  int b = blockIdx.x;
  if (!a.xcd_remap) return b;
  int chunk = (a.ntiles + 7) / 8;
  return (b % 8) * chunk + b / 8;
}

// Receiver record at the start of step it (src/fdModel.cpp:576-596):
// running sum of dt * v / (dx * dz).
__device__ inline void record_receivers(const StepArgs &a, int it) {
  for (int r = threadIdx.x; r < a.nr; r += blockDim.x) {
    long long li = a.rec_idx[r];
    double incx = a.dt * a.vx_i[li] / a.dxdz;
    double incz = a.dt * a.vz_i[li] / a.dxdz;
    long long k = (long long)r * a.nt + it;
    if (it == 0 || a.rec_increments) {
      a.rtf_x[k] = incx;
      a.rtf_z[k] = incz;
    } else {
      a.rtf_x[k] = a.rtf_x[k - 1] + incx;
      a.rtf_z[k] = a.rtf_z[k - 1] + incz;
    }
  }
}

// Application of a tile's injection terms to one point, in list order:
// entries[off[tile] .. off[tile+1]) whose local index is `local`, at time
// step it.
__device__ inline void apply_terms(const StepArgs &a, const int *off,
                                   const InjEntry *entries, int tile,
                                   int local, int it, double b, double &vx,
                                   double &vz) {
  int e0 = off[tile], e1 = off[tile + 1];
  for (int e = e0; e < e1; ++e) {
    const InjEntry &E = entries[e];
    if (E.local != local) continue;
    double term;
    if (!a.adjoint) {
      double S = a.stf[(long long)E.src * a.nt + it];
      term = E.coef * S * a.dt * b / (E.comp == 0 ? a.d4x : a.d4z);
    } else {
      const double *src = E.comp == 0 ? a.adj_x : a.adj_z;
      term = a.dt * b * src[(long long)E.src * a.nt + it] / a.dxdz;
    }
    if (E.comp == 0)
      vx += term;
    else
      vz += term;
  }
}

// Owner-side application of this tile's injection terms at step a.it.
__device__ inline void apply_injection(const StepArgs &a, int tile, int local,
                                       double b, double &vx, double &vz) {
  apply_terms(a, a.tile_off, a.entries, tile, local, a.it, b, vx, vz);
}

// Forward receiver record by the owner of a point: dt v / (dx dz) at step it
// for every receiver entry at `local` (two-step launches; a scan forms the
// running sums afterwards).
__device__ inline void record_owned(const StepArgs &a, int tile, int local,
                                    int it, double vx, double vz) {
  int e0 = a.rec_off[tile], e1 = a.rec_off[tile + 1];
  for (int e = e0; e < e1; ++e) {
    const InjEntry &E = a.rec_entries[e];
    if (E.local != local) continue;
    long long k = (long long)E.src * a.nt + it;
    if (E.comp == 0)
      a.rtf_x[k] = a.dt * vx / a.dxdz;
    else
      a.rtf_z[k] = a.dt * vz / a.dxdz;
  }
}

// Everything the owner of (x, z) does with the state at the start of the
// step: snapshot store (forward, src/fdModel.cpp:555-572) or correlation
// (adjoint, src/fdModel.cpp:817-861).
__device__ inline void owner_snapshot(const StepArgs &a, double *const *snap,
                                      int x, int z, double vx, double vz,
                                      double txx, double tzz, double txz) {
  long long ci = (long long)z * a.nx + x;
  if (!a.adjoint) {
    snap[0][ci] = vx;
    snap[1][ci] = vz;
    snap[2][ci] = txx;
    snap[3][ci] = tzz;
    snap[4][ci] = txz;
  } else if (x >= a.kx0 && x < a.kx1 && z >= a.kz0 && z < a.kz1) {
    long long pi = pidx(a, x, z);
    double rk = 0.0, lk = 0.0, mk = 0.0;
    if (!a.kfresh) rk = a.rho_k[ci], lk = a.lam_k[ci], mk = a.mu_k[ci];
    correlate_point(a.wcorr, a.lm[pi], a.mu[pi], vx, vz, txx, tzz, txz,
                    snap[0][ci], snap[1][ci], snap[2][ci], snap[3][ci],
                    snap[4][ci], rk, lk, mk);
    a.rho_k[ci] = rk;
    a.lam_k[ci] = lk;
    a.mu_k[ci] = mk;
  }
}

template <bool RECIP>
struct Divs;
template <>
struct Divs<false> {
  IeeeDiv x, z;
  __device__ explicit Divs(const StepArgs &a) : x{a.dx}, z{a.dz} {}
};
template <>
struct Divs<true> {
  RecipDiv x, z;
  __device__ explicit Divs(const StepArgs &a) : x{a.dx, a.rdx}, z{a.dz, a.rdz} {}
};

// One time step of one 64 x TZ tile (the fused one-step kernel's body).
template <int TZ, bool RECIP, int NT>
__device__ __forceinline__ void fused_tile(const StepArgs &a, const int tile) {
  constexpr int WV = kTileX + 8, HV = TZ + 8;  // velocity halo region
  constexpr int WS = kTileX + 4, HS = TZ + 4;  // stress ring region
  constexpr int NB = (HS * WS + NT - 1) / NT;  // stress items per thread
  constexpr int NC = (TZ * kTileX + NT - 1) / NT;  // velocity items per thread
  __shared__ double s_vx[HV * WV], s_vz[HV * WV];
  __shared__ double s_txx[HS * WS], s_tzz[HS * WS], s_txz[HS * WS];
  const Divs<RECIP> dv(a);
  const int x0 = (tile % a.ntx) * kTileX, z0 = (tile / a.ntx) * TZ;
  const int tid = threadIdx.x;

  // phase A: issue every global load of the block before the first wait.
  // Velocities with a 4-point halo (inside the zero margin) and stresses on
  // the ring are loaded into registers by fully unrolled loops, then written
  // to LDS; the materials of this thread's items stay in registers.  Points
  // outside the grid stage zeros.
  constexpr int NV = (HV * WV + NT - 1) / NT;
  double vx_r[NV], vz_r[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    int i = tid + k * NT;
    vx_r[k] = vz_r[k] = 0.0;
    if (i < HV * WV) {
      int r = i / WV, c = i - r * WV;
      long long g = pidx(a, x0 - 4 + c, z0 - 4 + r);
      vx_r[k] = a.vx_i[g];
      vz_r[k] = a.vz_i[g];
    }
  }
  double txx_r[NB], tzz_r[NB], txz_r[NB], lm_r[NB], mu_r[NB], b_r[NC];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    int i = tid + k * NT;
    txx_r[k] = tzz_r[k] = txz_r[k] = lm_r[k] = mu_r[k] = 0.0;
    if (i < HS * WS) {
      int r = i / WS, c = i - r * WS;
      int x = x0 - 2 + c, z = z0 - 2 + r;
      if (x >= 0 && x < a.nx && z >= 0 && z < a.nz) {
        long long g = pidx(a, x, z);
        txx_r[k] = a.txx_i[g];
        tzz_r[k] = a.tzz_i[g];
        txz_r[k] = a.txz_i[g];
        if (interior(a, x, z)) {
          lm_r[k] = a.lm[g];
          mu_r[k] = a.mu[g];
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    int i = tid + k * NT;
    b_r[k] = 0.0;
    if (i < TZ * kTileX) {
      int r = i / kTileX, c = i - r * kTileX;
      int x = x0 + c, z = z0 + r;
      if (x < a.nx && z < a.nz) b_r[k] = a.b[pidx(a, x, z)];
    }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    int i = tid + k * NT;
    if (i < HV * WV) {
      s_vx[i] = vx_r[k];
      s_vz[i] = vz_r[k];
    }
  }
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    int i = tid + k * NT;
    if (i < HS * WS) {
      s_txx[i] = txx_r[k];
      s_tzz[i] = tzz_r[k];
      s_txz[i] = txz_r[k];
    }
  }
  if (!a.adjoint && tile == 0) record_receivers(a, a.it);
  __syncthreads();

  // phase B: stresses on the tile plus a 2-point ring, in place in LDS (a
  // stress update reads only its own old stress).
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    int i = tid + k * NT;
    if (i >= HS * WS) break;
    int r = i / WS, c = i - r * WS;
    int x = x0 - 2 + c, z = z0 - 2 + r;
    if (x < 0 || x >= a.nx || z < 0 || z >= a.nz) continue;
    double txx = txx_r[k], tzz = tzz_r[k], txz = txz_r[k];
    bool own = r >= 2 && r < TZ + 2 && c >= 2 && c < kTileX + 2;
    if (own && a.do_snap && !a.adjoint) {
      long long ci = (long long)z * a.nx + x;
      a.snap[2][ci] = txx;
      a.snap[3][ci] = tzz;
      a.snap[4][ci] = txz;
    }
    if (interior(a, x, z)) {
      stress_point(s_vx, s_vz, (long long)(r + 2) * WV + (c + 2), 1, WV,
                   lm_r[k], mu_r[k], taper_at(a, x, z), a.dts, dv.x, dv.z,
                   txx, tzz, txz);
      s_txx[i] = txx;
      s_tzz[i] = tzz;
      s_txz[i] = txz;
    }
    if (own) {
      long long g = pidx(a, x, z);
      a.txx_o[g] = txx;
      a.tzz_o[g] = tzz;
      a.txz_o[g] = txz;
    }
  }
  __syncthreads();

  // phase C: velocities on the tile.
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    int i = tid + k * NT;
    if (i >= TZ * kTileX) break;
    int r = i / kTileX, c = i - r * kTileX;
    int x = x0 + c, z = z0 + r;
    if (x >= a.nx || z >= a.nz) continue;
    int lv = (r + 4) * WV + (c + 4);
    double vx = s_vx[lv], vz = s_vz[lv];
    long long g = pidx(a, x, z);
    if (a.do_snap) {
      if (!a.adjoint) {
        long long ci = (long long)z * a.nx + x;
        a.snap[0][ci] = vx;
        a.snap[1][ci] = vz;
      } else {
        owner_snapshot(a, a.snap, x, z, vx, vz, a.txx_i[g], a.tzz_i[g], a.txz_i[g]);
      }
    }
    double b = b_r[k];
    if (interior(a, x, z)) {
      velocity_point(s_txx, s_tzz, s_txz, (long long)(r + 2) * WS + (c + 2), 1,
                     WS, b, taper_at(a, x, z), a.dts, dv.x, dv.z, vx, vz);
    }
    apply_injection(a, tile, r * kTileX + c, b, vx, vz);
    a.vx_o[g] = vx;
    a.vz_o[g] = vz;
  }
}


template <int TZ, bool RECIP, int NT = kBlock>
__global__ void __launch_bounds__(NT) step_fused(StepArgs a) {
  const int tile = block_tile(a);
  if (tile >= a.ntiles) return;
  fused_tile<TZ, RECIP, NT>(a, tile);
}

// FastDiv: correctly rounded x / d for a constant d > 0 in four FMA-class
// operations and no reciprocal instruction.  With zh = RD(1/d),
// zl = RN(1/d - zh) and y = RN(1/d):
//   q0 = RN(x zl + RN(x zh))   is within half an ulp plus ~2^-104 |x/d| of
//                              x / d, hence a faithful quotient;
//   r  = RN(q0 d - x)          is exact (residual of a faithful quotient);
//   q1 = RN(q0 - r y)          is RN(x / d) (Markstein's correction theorem:
//                              y within half an ulp of 1/d, q0 faithful).
// zl >= 0 makes q0 keep the sign of a zero x, and the residual's sign
// convention keeps it through q1.  The theorem needs no underflow or
// overflow: that holds for d in [2^-40, 2^40] and x = 0 or
// 2^-960 <= |x| <= 2^980.  Every call folds x into two flags (smallest
// binary exponent, sum of magnitudes -- inf and NaN poison the sum); a
// caller whose flags fail recomputes with IeeeDiv.
struct DivFlags {
  int emin;
  double asum;
  __device__ DivFlags() : emin(0), asum(0.0) {}
  __device__ bool ok() const { return emin >= -959 && asum <= 0x1p980; }
};
struct FastDiv {
  double d, y, zh, zl;
  DivFlags *f;
  __device__ double operator()(double x) const {
    f->emin = min(f->emin, __builtin_amdgcn_frexp_exp(x));
    f->asum += __builtin_fabs(x);
    double q0 = __builtin_fma(x, zl, x * zh);
    double r = __builtin_fma(q0, d, -x);
    return __builtin_fma(-r, y, q0);
  }
};

// ---------------------------------------------------------------------------
// Lean one-step kernel.  A block of four waves owns a 64 x TZ tile; lane l of
// wave w owns column x0 + l on rows w, w+4, .., w + TZ - 4, so every
// row a wave touches is one coalesced 512-byte segment, row offsets are
// wave-uniform (scalar) and every LDS access is an immediate offset from one
// per-lane address.  Global loads address a per-field scalar base plus one
// 32-bit per-lane row offset shared by all fields.
//
// The velocity update at (x, z) reads txx at x-2..x+1 and txz at x-1..x+2 on
// its row, txz at z-2..z+1 and tzz at z-1..z+2 on its column -- no corner
// values.  Besides its own points the block therefore updates only
//   rows    z0-2: txz   z0-1: txz, tzz   z0+TZ: tzz, txz   z0+TZ+1: tzz
//   columns x0-2: txx   x0-1: txx, txz   x0+64: txx, txz   x0+65: txz
// (one ring row per wave; the 4*TZ ring-column items in ceil(TZ/16) passes).
// Velocities are staged in LDS on rows z0-3..z0+TZ+2, columns x0-3..x0+66.
//
// Divisions use FastDiv; a lane whose flags fail redoes its pass with IEEE
// divisions, so results are bitwise those of the IEEE kernels.
enum : int { kXX = 1, kZZ = 2, kXZ = 4 };
constexpr int kInjStage = 128;  // injection entries per tile staged in LDS

template <int TZ>
struct Lean {
  static constexpr int R = TZ / 4;          // own rows per wave
  static constexpr int VR = TZ + 6, VW = 70;  // v: rows -3.., cols -3..
  static constexpr int XXW = 67;            // txx: rows 0..TZ-1, cols -2..64
  static constexpr int ZZR = TZ + 3;        // tzz: rows -1..TZ+1, cols 0..63
  static constexpr int XZR = TZ + 3, XZW = 67;  // txz: rows -2..TZ, cols -1..65
  static constexpr int NCI = 4 * TZ;        // ring-column items
  static constexpr int NCP = (NCI + 63) / 64;  // ring-column passes
  static constexpr int KV = (TZ + 2) / 4 + 2;  // v rows per wave: w + 4m, m >= -1
  static constexpr int NSIDE = 6 * VR;      // v side items (3 + 3 columns)
  __device__ static int V(int r, int c) { return (r + 3) * VW + (c + 3); }
  __device__ static int XX(int r, int c) { return r * XXW + (c + 2); }
  __device__ static int ZZ(int r, int c) { return (r + 1) * 64 + c; }
  __device__ static int XZ(int r, int c) { return (r + 2) * XZW + (c + 1); }
};

// Stress components MASK at the point whose v-stage index is v, expression
// trees of stress_point.
template <int MASK, int VW, class DX, class DZ>
__device__ __forceinline__ void lean_stress(const double *sx, const double *sz,
                                            int v, double lm, double mu,
                                            double taper, double dts, DX divx,
                                            DZ divz, double &txx, double &tzz,
                                            double &txz) {
  if (MASK & (kXX | kZZ)) {
    double dvx_x = kC1 * (sx[v + 1] - sx[v]) + kC2 * (sx[v - 1] - sx[v + 2]);
    double dvz_z = kC1 * (sz[v] - sz[v - VW]) + kC2 * (sz[v - 2 * VW] - sz[v + VW]);
    double la = __builtin_fma(-2.0, mu, lm);  // == lm - 2 * mu (2 mu is exact)
    if (MASK & kXX) txx = taper * (txx + dts * (divx(lm * dvx_x) + divz(la * dvz_z)));
    if (MASK & kZZ) tzz = taper * (tzz + dts * (divx(la * dvx_x) + divz(lm * dvz_z)));
  }
  if (MASK & kXZ) {
    double dvx_z = kC1 * (sx[v + VW] - sx[v]) + kC2 * (sx[v - VW] - sx[v + 2 * VW]);
    double dvz_x = kC1 * (sz[v] - sz[v - 1]) + kC2 * (sz[v - 2] - sz[v + 1]);
    txz = taper * (txz + dts * mu * (divz(dvx_z) + divx(dvz_x)));
  }
}

// Stresses of wave w's ring row: txz (w = 0), tzz and txz (w = 1, 2), tzz
// (w = 3).
template <int VW, class DX, class DZ>
__device__ __forceinline__ void ring_row_stress(int w, const double *sx,
                                                const double *sz, int v,
                                                double lm, double mu,
                                                double taper, double dts,
                                                DX dx, DZ dz, double &tzz,
                                                double &txz) {
  double hxx = 0.0, hzz = tzz, hxz = txz;
  if (w == 0)
    lean_stress<kXZ, VW>(sx, sz, v, lm, mu, taper, dts, dx, dz, hxx, hzz, hxz);
  else if (w == 3)
    lean_stress<kZZ, VW>(sx, sz, v, lm, mu, taper, dts, dx, dz, hxx, hzz, hxz);
  else
    lean_stress<kZZ | kXZ, VW>(sx, sz, v, lm, mu, taper, dts, dx, dz, hxx, hzz, hxz);
  tzz = hzz;
  txz = hxz;
}

// Velocity at own point (r, c) from the staged stresses (velocity_point).
template <int TZ, class DX, class DZ>
__device__ __forceinline__ void lean_velocity(const double *sxx,
                                              const double *szz,
                                              const double *sxz, int r, int c,
                                              double b, double taper,
                                              double dts, DX divx, DZ divz,
                                              double &vx, double &vz) {
  using L = Lean<TZ>;
  const int ix = L::XX(r, c), iz = L::ZZ(r, c), ixz = L::XZ(r, c);
  constexpr int W = L::XZW;
  double ax = divx(kC1 * (sxx[ix] - sxx[ix - 1]) + kC2 * (sxx[ix - 2] - sxx[ix + 1])) +
              divz(kC1 * (sxz[ixz] - sxz[ixz - W]) + kC2 * (sxz[ixz - 2 * W] - sxz[ixz + W]));
  double az = divx(kC1 * (sxz[ixz + 1] - sxz[ixz]) + kC2 * (sxz[ixz - 1] - sxz[ixz + 2])) +
              divz(kC1 * (szz[iz + 64] - szz[iz]) + kC2 * (szz[iz - 64] - szz[iz + 128]));
  vx = taper * (vx + b * dts * ax);
  vz = taper * (vz + b * dts * az);
}

// Keeps the scheduler from interleaving the stencils of several rows (whose
// LDS operands would then all be live at once).
__device__ __forceinline__ void row_fence() { __builtin_amdgcn_sched_barrier(0); }

// A value the compiler cannot see through: reloads through it are not
// merged with earlier loads of the same address (whose results would then
// stay live in registers).
__device__ __forceinline__ unsigned opaque(unsigned v) {
  asm volatile("" : "+v"(v));
  return v;
}

typedef __attribute__((address_space(3))) void lds_void;
// LDS pointers for the DMA transfers below.  They are formed from a
// __shared__ array's own address and offset in the LDS address space, never
// converted from generic pointers (a generic-to-LDS conversion carries a
// null test for every transfer).
typedef __attribute__((address_space(3))) double lds_dbl;
// The LDS address of a pointer into a __shared__ array: the low 32 bits of
// its generic address.
__device__ __forceinline__ lds_dbl *lds_ptr(double *p) {
  return (lds_dbl *)(unsigned)(unsigned long long)p;
}

// One 64-double row from memory straight into LDS: two transfers of one
// dword per lane (bytes 0-255, then 256-511).  lane_off is this lane's byte
// offset of column `lane`; lds_row (column 0 of the row) is wave-uniform.
__device__ __forceinline__ void dma_row(const double *base, unsigned lane_off,
                                        lds_dbl *lds_row) {
  const unsigned lane = threadIdx.x & 63;
  const char *g = (const char *)base + (lane_off - lane * 4u);
  __builtin_amdgcn_global_load_lds((const void *)g, (lds_void *)lds_row, 4, 0, 0);
  __builtin_amdgcn_global_load_lds((const void *)(g + 256), (lds_void *)(lds_row + 32),
                                   4, 0, 0);
}

// One staged velocity row, columns -3 .. 66 (70 doubles = 140 dwords),
// from memory straight into LDS: dwords 0-63, 64-127 and 128-139 (12
// lanes).  row_base is the global address of column -3, lds_row the LDS
// address of column -3 (both wave-uniform).
__device__ __forceinline__ void dma_row70(const double *row_base, lds_dbl *lds_row) {
  const unsigned lane = threadIdx.x & 63;
  const char *g = (const char *)row_base + lane * 4u;
  __builtin_amdgcn_global_load_lds((const void *)g, (lds_void *)lds_row, 4, 0, 0);
  __builtin_amdgcn_global_load_lds((const void *)(g + 256), (lds_void *)(lds_row + 32), 4,
                                   0, 0);
  if (lane < 12)
    __builtin_amdgcn_global_load_lds((const void *)(g + 512), (lds_void *)(lds_row + 64), 4,
                                     0, 0);
}

// dma_row and dma_row70 with every lane's dword index capped at dlim: the
// lanes past the grid's last column re-read that column's dword (a line the
// row fetches anyway) instead of the zero padding beyond it.  row_base is
// the global address of the row's first staged column.
__device__ __forceinline__ void dma_row_capped(const double *row_base, lds_dbl *lds_row,
                                               unsigned dlim) {
  const unsigned lane = threadIdx.x & 63;
  const char *g = (const char *)row_base;
  __builtin_amdgcn_global_load_lds((const void *)(g + min(lane, dlim) * 4u),
                                   (lds_void *)lds_row, 4, 0, 0);
  __builtin_amdgcn_global_load_lds((const void *)(g + min(lane + 64u, dlim) * 4u),
                                   (lds_void *)(lds_row + 32), 4, 0, 0);
}
__device__ __forceinline__ void dma_row70_capped(const double *row_base, lds_dbl *lds_row,
                                                 unsigned dlim) {
  const unsigned lane = threadIdx.x & 63;
  const char *g = (const char *)row_base;
  __builtin_amdgcn_global_load_lds((const void *)(g + min(lane, dlim) * 4u),
                                   (lds_void *)lds_row, 4, 0, 0);
  __builtin_amdgcn_global_load_lds((const void *)(g + min(lane + 64u, dlim) * 4u),
                                   (lds_void *)(lds_row + 32), 4, 0, 0);
  if (lane < 12)
    __builtin_amdgcn_global_load_lds((const void *)(g + min(lane + 128u, dlim) * 4u),
                                     (lds_void *)(lds_row + 64), 4, 0, 0);
}

__device__ __forceinline__ double ldb(const double *base, unsigned off) {
  return *(const double *)((const char *)base + off);
}
__device__ __forceinline__ void stb(double *base, unsigned off, double v) {
  *(double *)((char *)base + off) = v;
}

// Source sample of an injection entry at step a.it: the source time
// function (forward) or this shot's adjoint source (adjoint).
__device__ __forceinline__ double inj_sample(const StepArgs &a, const InjEntry &E,
                                            int it) {
  const long long k = (long long)E.src * a.nt + it;
  const double *ax = a.adj_x, *az = a.adj_z;  // select values, not members
  return !a.adjoint ? a.stf[k] : (E.comp == 0 ? ax : az)[k];
}

// Forward snapshot store (src/fdModel.cpp:555-572) of the R points
// (x, z0w + 4j) of a lean-tile thread from its start-of-step registers.
template <int R>
__device__ __forceinline__ void lean_store_snapshot(const StepArgs &a, int x, int z0w,
                                                    bool xin, const double *vx,
                                                    const double *vz, const double *txx,
                                                    const double *tzz,
                                                    const double *txz) {
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int z = z0w + 4 * j;
    if (xin && z < a.nz) {
      const long long ci = (long long)z * a.nx + x;
      __builtin_nontemporal_store(vx[j], a.snap[0] + ci);
      __builtin_nontemporal_store(vz[j], a.snap[1] + ci);
      __builtin_nontemporal_store(txx[j], a.snap[2] + ci);
      __builtin_nontemporal_store(tzz[j], a.snap[3] + ci);
      __builtin_nontemporal_store(txz[j], a.snap[4] + ci);
    }
  }
}

// Byte offset of tile-relative (row r, column c) from the lean tile's field
// bases (row -3, column -8).
__device__ __forceinline__ unsigned lean_off(int r, int c, unsigned rowb) {
  return (unsigned)(r + 3) * rowb + (unsigned)(c + 8) * 8u;
}

// HALO_DMA: the halo columns of the velocity rows come in the rows' own
// transfers (70-column rows) instead of through registers.
// SEP: the stresses get an LDS stage of their own beside the velocities, so
// they are written as soon as computed (no barrier between the two uses of
// one stage) and the velocity pass reads the start-of-step velocities back
// from LDS instead of keeping them in registers.
// RDMA: the ring row's and ring columns' operands (old stresses, lm, mu)
// also go straight into an LDS stage by DMA instead of waiting in
// registers from the first load to the stress pass; with HALO_DMA this
// frees enough registers for six waves per SIMD.
// CLAMP (with RDMA): loads of rows below the grid's last row or above its
// first, and of columns right of its last, read that last (first) row or
// column instead of the zero padding.  No grid point's stencil reads a value
// outside the grid (interior points read at most two rows or columns away,
// the two-point rim is copied), so those values are never used; the tiles
// on the grid's right and bottom edges then fetch only the lines that hold
// grid points.  Stores are unchanged.
template <int TZ, bool CORR, bool HALO_DMA = false, bool SEP = false, bool RDMA = false,
          bool CLAMP = false>
__device__ __forceinline__ void lean_tile(const StepArgs &a, const int tile, const int it) {
  static_assert(!CLAMP || RDMA, "clamped loads are written for the ring-DMA layout");
  using L = Lean<TZ>;
  constexpr int R = L::R;
  // One LDS buffer: velocities while the stresses are computed, then the
  // stresses for the velocity update (a barrier separates the two uses);
  // two stages side by side with SEP.
  constexpr int NV = 2 * L::VR * L::VW;
  constexpr int NS = TZ * L::XXW + L::ZZR * 64 + L::XZR * L::XZW;
  constexpr int NI = SEP ? 16 : RDMA ? 32 : kInjStage;  // staged injection entries
  // the tile's taper factors: columns x0-2 .. x0+65, then rows z0-2 ..
  // z0+TZ+1, staged by DMA like the velocity rows, after the stages
  constexpr int TXN = 68, TZN = TZ + 4;
  constexpr int NST = SEP ? NV + NS : (NV > NS ? NV : NS);
  // RDMA ring stages: [wave][txz, tzz, lm, mu][64] for the ring rows, then
  // [txx, txz, lm, mu][NCQ] for the ring-column items (two lanes per value
  // in the transfers, so NCQ rounds NCI up to 32)
  constexpr int NCQ = (L::NCI + 31) / 32 * 32;
  constexpr int NRG = RDMA ? 4 * 4 * 64 + 4 * NCQ : 0;
  __shared__ double lds[NST + TXN + TZN + NRG];
  double *const s_vx = lds, *const s_vz = lds + L::VR * L::VW;
  double *const s_xx = SEP ? lds + NV : lds, *const s_zz = s_xx + TZ * L::XXW;
  double *const s_xz = s_zz + L::ZZR * 64;
  double *const s_tap = lds + NST;
  lds_dbl *const d_vx = lds_ptr(lds), *const d_vz = d_vx + L::VR * L::VW;
  lds_dbl *const d_tap = d_vx + NST;
  double *const s_gr = lds + NST + TXN + TZN, *const s_gc = s_gr + 4 * 4 * 64;
  lds_dbl *const d_gr = d_tap + TXN + TZN, *const d_gc = d_gr + 4 * 4 * 64;

  // This tile's injection list bounds, before any memory access of the
  // kernel: a uniform load that nothing can have clobbered yet is a scalar
  // load, which the vector-memory counters do not see.
  const int tile_u = __builtin_amdgcn_readfirstlane(tile);
  const int e0 = a.tile_off[tile_u], ne = a.tile_off[tile_u + 1] - e0;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x0 = (tile % a.ntx) * kTileX, z0 = (tile / a.ntx) * TZ;
  // Field bases at tile-relative (row -3, column -8): every offset below is
  // a nonnegative 32-bit byte count.
  const long long tb = a.origin + (long long)(z0 - 3) * a.pitch + (x0 - 8);
  const unsigned rowb = (unsigned)a.pitch * 8u;
  auto off = [&](int r, int c) -> unsigned { return lean_off(r, c, rowb); };
  const double *vxi = a.vx_i + tb, *vzi = a.vz_i + tb;
  const double *txxi = a.txx_i + tb, *tzzi = a.tzz_i + tb, *txzi = a.txz_i + tb;
  const double *lmp = a.lm + tb, *mup = a.mu + tb, *bp = a.b + tb;
  // CLAMP: row and column of a load, capped to the grid (tile-relative
  // bounds: rows -z0 .. nz-1-z0, columns up to nx-1-x0); lc = the lane's
  // column so capped
  const int rlo = -z0, rhi = a.nz - 1 - z0, cmax = a.nx - 1 - x0;
  auto rc = [&](int r) { return CLAMP ? min(max(r, rlo), rhi) : r; };
  auto cc = [&](int c) { return CLAMP ? min(c, cmax) : c; };
  const int lc = cc(lane);
  // the last dword a row transfer may read: column cmax of a 64-column row
  // (columns 0..63) or of a 70-column row (columns -3..66)
  const unsigned dl64 = 2u * (unsigned)cmax + 1u, dl70 = 2u * (unsigned)cmax + 7u;

  // ---- phase A: every global load before the first wait ----
  {
    const unsigned lane4 = (threadIdx.x & 63) * 4u;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (wv == 0) {  // 136 dwords of taper_x
      const char *g = (const char *)(a.taper_x + x0 - 2) + lane4;
      __builtin_amdgcn_global_load_lds((const void *)g, (lds_void *)d_tap, 4, 0, 0);
      __builtin_amdgcn_global_load_lds((const void *)(g + 256), (lds_void *)(d_tap + 32), 4,
                                       0, 0);
      if (lane4 < 32)
        __builtin_amdgcn_global_load_lds((const void *)(g + 512), (lds_void *)(d_tap + 64), 4,
                                         0, 0);
    } else if (wv == 1) {  // 2 TZN dwords of taper_z
      const char *g = (const char *)(a.taper_z + z0 - 2) + lane4;
      if (lane4 < 8u * TZN)
        __builtin_amdgcn_global_load_lds((const void *)g, (lds_void *)(d_tap + TXN), 4, 0, 0);
      if (2 * TZN > 64 && lane4 + 256 < 8u * TZN)
        __builtin_amdgcn_global_load_lds((const void *)(g + 256), (lds_void *)(d_tap + TXN + 32),
                                         4, 0, 0);
    }
  }
  // Velocity rows go straight from memory into LDS (no registers): wave w
  // stages rows r = w + 4m, m = -1 ..; its own rows are r = w + 4j,
  // j = 0 .. R-1.
#pragma unroll
  for (int k = 0; k < L::KV; ++k) {
    const int r = w + 4 * (k - 1);
    if (r >= -3 && r <= TZ + 2) {
      if (CLAMP) {
        const long long ro = (long long)(rc(r) + 3) * a.pitch;  // row rc(r), column -8
        if (HALO_DMA) {
          dma_row70_capped(vxi + ro + 5, d_vx + L::V(r, -3), dl70);
          dma_row70_capped(vzi + ro + 5, d_vz + L::V(r, -3), dl70);
        } else {
          dma_row_capped(vxi + ro + 8, d_vx + L::V(r, 0), dl64);
          dma_row_capped(vzi + ro + 8, d_vz + L::V(r, 0), dl64);
        }
      } else if (HALO_DMA) {
        const long long ro = tb + (long long)(r + 3) * a.pitch + 5;  // column -3
        dma_row70(a.vx_i + ro, d_vx + L::V(r, -3));
        dma_row70(a.vz_i + ro, d_vz + L::V(r, -3));
      } else {
        const unsigned o = off(r, lane);
        dma_row(vxi, o, d_vx + L::V(r, 0));
        dma_row(vzi, o, d_vz + L::V(r, 0));
      }
    }
  }
  double vxs[2], vzs[2];
  int sidx[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int t = threadIdx.x + 256 * k;
    sidx[k] = -1;
    vxs[k] = vzs[k] = 0.0;
    if (!HALO_DMA && k * 256 < L::NSIDE && t < L::NSIDE) {
      int r = t / 6 - 3, cs = t - (t / 6) * 6;
      int c = cs < 3 ? cs - 3 : 61 + cs;
      unsigned o = off(rc(r), cc(c));
      vxs[k] = ldb(vxi, o);
      vzs[k] = ldb(vzi, o);
      sidx[k] = L::V(r, c);
    }
  }
  double oxx[R], ozz[R], oxz[R], olm[R], omu[R], ob[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    unsigned o = off(rc(w + 4 * j), lc);
    oxx[j] = ldb(txxi, o);
    ozz[j] = ldb(tzzi, o);
    oxz[j] = ldb(txzi, o);
    olm[j] = ldb(lmp, o);
    omu[j] = ldb(mup, o);
    ob[j] = ldb(bp, o);
  }
  // adjoint snapshot steps: the forward snapshot and the kernels of the own
  // points (src/fdModel.cpp:817-861), in flight with everything else
  bool cin[R];
  double csv[CORR ? R : 1][5], crk[CORR ? R : 1], clk[CORR ? R : 1], cmk[CORR ? R : 1];
  if (CORR) {
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int z = z0 + w + 4 * j, xx = x0 + lane;
      cin[j] = xx < a.nx && z < a.nz && xx >= a.kx0 && xx < a.kx1 && z >= a.kz0 &&
               z < a.kz1;
      if (cin[j]) {
        const long long ci = (long long)z * a.nx + xx;
#pragma unroll
        for (int f = 0; f < 5; ++f) csv[j][f] = __builtin_nontemporal_load(a.snap[f] + ci);
        crk[j] = clk[j] = cmk[j] = 0.0;
        if (!a.kfresh) {
          crk[j] = __builtin_nontemporal_load(a.rho_k + ci);
          clk[j] = __builtin_nontemporal_load(a.lam_k + ci);
          cmk[j] = __builtin_nontemporal_load(a.mu_k + ci);
        }
      }
    }
  }
  // this wave's ring row: only the fields its components read
  const int rr = w == 0 ? -2 : w == 1 ? -1 : w == 2 ? TZ : TZ + 1;
  double gzz = 0.0, gxz = 0.0, glm = 0.0, gmu = 0.0;
  // ring-column items of this wave: item i = p * 64 + lane
  double cxx[L::NCP], cxz[L::NCP], clm[L::NCP], cmu[L::NCP];
#pragma unroll
  for (int p = 0; p < L::NCP; ++p) cxx[p] = cxz[p] = clm[p] = cmu[p] = 0.0;
  if (RDMA) {
    // ring row fields into s_gr[w][field], ring-column items (wave 3, two
    // lanes per value: lane 2k + h moves dword h of item 32 m + k) into
    // s_gc[field]
    lds_dbl *const g = d_gr + w * 4 * 64;
    if (CLAMP) {
      const long long ro = (long long)(rc(rr) + 3) * a.pitch + 8;  // row rc(rr), column 0
      if (w != 3) dma_row_capped(txzi + ro, g, dl64);
      if (w != 0) {
        dma_row_capped(tzzi + ro, g + 64, dl64);
        dma_row_capped(lmp + ro, g + 128, dl64);
      }
      dma_row_capped(mup + ro, g + 192, dl64);
    } else {
      const unsigned o = off(rr, lane);
      if (w != 3) dma_row(txzi, o, g);
      if (w != 0) {
        dma_row(tzzi, o, g + 64);
        dma_row(lmp, o, g + 128);
      }
      dma_row(mup, o, g + 192);
    }
    if (w == 3) {
#pragma unroll
      for (int m = 0; m < NCQ / 32; ++m) {
        const int i = 32 * m + (lane >> 1);
        if (i < L::NCI) {
          const int q = i / TZ, r = i - q * TZ;
          const int c = q < 2 ? q - 2 : 62 + q;
          const unsigned oc = off(rc(r), cc(c)) + (lane & 1) * 4u;
          const double *const src[4] = {txxi, txzi, lmp, mup};
#pragma unroll
          for (int f = 0; f < 4; ++f)
            __builtin_amdgcn_global_load_lds((const void *)((const char *)src[f] + oc),
                                             (lds_void *)(d_gc + f * NCQ + 32 * m), 4, 0, 0);
        }
      }
    }
  } else {
    {
      unsigned o = off(rr, lane);
      if (w != 3) gxz = ldb(txzi, o);
      if (w != 0) {
        gzz = ldb(tzzi, o);
        glm = ldb(lmp, o);
      }
      gmu = ldb(mup, o);
    }
#pragma unroll
    for (int p = 0; p < L::NCP; ++p) {
      const int i = p * 64 + lane;
      if ((p & 3) == w && i < L::NCI) {
        const int q = i / TZ, r = i - q * TZ;
        const int c = q < 2 ? q - 2 : 62 + q;
        unsigned o = off(r, c);
        cxx[p] = ldb(txxi, o);
        cxz[p] = ldb(txzi, o);
        clm[p] = ldb(lmp, o);
        cmu[p] = ldb(mup, o);
      }
    }
  }
  const int x = x0 + lane;
  const bool xin = x < a.nx;
  const bool xint = x >= 2 && x < a.nx - 2;
  // taper factors from the LDS segments: column c, row r of the tile
  auto tap_x = [&](int c) { return s_tap[c + 2]; };
  auto tap_z = [&](int r) { return s_tap[TXN + r + 2]; };
  if (!a.adjoint && tile == 0) record_receivers(a, it);
  // This tile's injection entries and their source samples at this step,
  // fetched by all threads at once and staged in LDS; phase C applies them
  // in list order.
  __shared__ double s_ival[NI], s_icoef[NI];
  __shared__ int s_iloc[NI];
  for (int t = threadIdx.x; t < ne && t < NI; t += 256) {
    const InjEntry E = a.entries[e0 + t];
    s_iloc[t] = E.local | (E.comp << 16);
    s_icoef[t] = E.coef;
    s_ival[t] = inj_sample(a, E, it);
  }

#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (sidx[k] >= 0) {
      s_vx[sidx[k]] = vxs[k];
      s_vz[sidx[k]] = vzs[k];
    }
  __builtin_amdgcn_s_waitcnt(0);  // the LDS DMA of this wave has landed
  __syncthreads();
  const double tx = tap_x(lane);
  double trow[R];
#pragma unroll
  for (int j = 0; j < R; ++j) trow[j] = tap_z(w + 4 * j);
  // start-of-step velocities of the own points
  double vx0[R], vz0[R];
#pragma unroll
  for (int j = 0; j < R; ++j) {
    vx0[j] = s_vx[L::V(w + 4 * j, lane)];
    vz0[j] = s_vz[L::V(w + 4 * j, lane)];
  }

  // ---- phase B: stresses into registers ----
  // The IEEE redo of a lane whose FastDiv flags fail reloads its operands
  // from memory (rare), so only the results stay live across the phase.
  const double dts = a.dts;
  const IeeeDiv ix{a.dx}, iz{a.dz};
  auto zint = [&](int z) { return z >= 2 && z < a.nz - 2; };
  if (CORR) {
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (cin[j]) {
        const long long ci = (long long)(z0 + w + 4 * j) * a.nx + x;
        correlate_point(a.wcorr, olm[j], omu[j], vx0[j], vz0[j], oxx[j], ozz[j], oxz[j],
                        csv[j][0], csv[j][1], csv[j][2], csv[j][3], csv[j][4], crk[j],
                        clk[j], cmk[j]);
        __builtin_nontemporal_store(crk[j], a.rho_k + ci);
        __builtin_nontemporal_store(clk[j], a.lam_k + ci);
        __builtin_nontemporal_store(cmk[j], a.mu_k + ci);
      }
  } else if (a.do_snap && !a.adjoint) {
    lean_store_snapshot<R>(a, x, z0 + w, xin, vx0, vz0, oxx, ozz, oxz);
  }
  {
    DivFlags fl;
    const FastDiv fx{a.dx, a.rdx, a.zhx, a.zlx, &fl}, fz{a.dz, a.rdz, a.zhz, a.zlz, &fl};
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int r = w + 4 * j;
      if (xint && zint(z0 + r))
        lean_stress<kXX | kZZ | kXZ, L::VW>(s_vx, s_vz, L::V(r, lane), olm[j], omu[j],
                                            fmin(tx, trow[j]), dts, fx, fz, oxx[j],
                                            ozz[j], oxz[j]);
      row_fence();
    }
    if (!fl.ok()) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = w + 4 * j;
        const unsigned o = opaque(off(r, lane));
        oxx[j] = ldb(txxi, o), ozz[j] = ldb(tzzi, o), oxz[j] = ldb(txzi, o);
        if (xint && zint(z0 + r))
          lean_stress<kXX | kZZ | kXZ, L::VW>(s_vx, s_vz, L::V(r, lane), ldb(lmp, o),
                                              ldb(mup, o), fmin(tx, trow[j]), dts, ix,
                                              iz, oxx[j], ozz[j], oxz[j]);
      }
      // the reloads have landed before the paths join, so that later code
      // never waits on them (counters would otherwise make it wait for every
      // store issued since)
      __builtin_amdgcn_s_waitcnt(0);
    }
  }
  // ring row (its new values replace the loaded ones)
  {
    const double *const g = s_gr + w * 4 * 64 + lane;  // RDMA stage
    if (RDMA) {
      if (w != 3) gxz = g[0];
      if (w != 0) gzz = g[64], glm = g[128];
      gmu = g[192];
    }
    const int z = z0 + rr;
    if (xint && zint(z)) {
      const double t = fmin(tx, tap_z(rr));
      const int v = L::V(rr, lane);
      DivFlags gf;
      const FastDiv gx{a.dx, a.rdx, a.zhx, a.zlx, &gf}, gz{a.dz, a.rdz, a.zhz, a.zlz, &gf};
      ring_row_stress<L::VW>(w, s_vx, s_vz, v, glm, gmu, t, dts, gx, gz, gzz, gxz);
      if (!gf.ok()) {
        if (RDMA) {
          gzz = w != 0 ? g[64] : 0.0, gxz = w != 3 ? g[0] : 0.0;
          ring_row_stress<L::VW>(w, s_vx, s_vz, v, g[128], g[192], t, dts, ix, iz, gzz, gxz);
        } else {
          const unsigned o = opaque(off(rr, lane));
          gzz = ldb(tzzi, o), gxz = ldb(txzi, o);
          ring_row_stress<L::VW>(w, s_vx, s_vz, v, ldb(lmp, o), ldb(mup, o), t, dts, ix, iz,
                                 gzz, gxz);
          __builtin_amdgcn_s_waitcnt(0);
        }
      }
    }
  }
  // ring columns
#pragma unroll
  for (int p = 0; p < L::NCP; ++p) {
    const int i = p * 64 + lane;
    if ((p & 3) == w && i < L::NCI) {
      if (RDMA) {
        cxx[p] = s_gc[i], cxz[p] = s_gc[NCQ + i];
        clm[p] = s_gc[2 * NCQ + i], cmu[p] = s_gc[3 * NCQ + i];
      }
      const int q = i / TZ, r = i - q * TZ;
      const int c = q < 2 ? q - 2 : 62 + q;
      const int xc = x0 + c, z = z0 + r;
      if (xc >= 2 && xc < a.nx - 2 && zint(z)) {
        const double t = fmin(tap_x(c), tap_z(r));
        DivFlags cf;
        const FastDiv cx{a.dx, a.rdx, a.zhx, a.zlx, &cf}, cz{a.dz, a.rdz, a.zhz, a.zlz, &cf};
        double hxx = cxx[p], hxz = cxz[p], hzz = 0.0;
        lean_stress<kXX | kXZ, L::VW>(s_vx, s_vz, L::V(r, c), clm[p], cmu[p], t, dts, cx,
                                      cz, hxx, hzz, hxz);
        if (!cf.ok()) {
          if (RDMA) {
            hxx = s_gc[i], hxz = s_gc[NCQ + i];
            lean_stress<kXX | kXZ, L::VW>(s_vx, s_vz, L::V(r, c), s_gc[2 * NCQ + i],
                                          s_gc[3 * NCQ + i], t, dts, ix, iz, hxx, hzz, hxz);
          } else {
            const unsigned o = opaque(off(r, c));
            hxx = ldb(txxi, o), hxz = ldb(txzi, o);
            lean_stress<kXX | kXZ, L::VW>(s_vx, s_vz, L::V(r, c), ldb(lmp, o), ldb(mup, o),
                                          t, dts, ix, iz, hxx, hzz, hxz);
            __builtin_amdgcn_s_waitcnt(0);
          }
        }
        cxx[p] = hxx, cxz[p] = hxz;
      }
    }
  }
  // own stresses to memory; start-of-step owner work (snapshot store or
  // kernel correlation) from the start-of-step values
  double *txxo = a.txx_o + tb, *tzzo = a.tzz_o + tb, *txzo = a.txz_o + tb;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int r = w + 4 * j, z = z0 + r;
    if (xin && z < a.nz) {
      const unsigned o = off(r, lane);
      stb(txxo, o, oxx[j]);
      stb(tzzo, o, ozz[j]);
      stb(txzo, o, oxz[j]);
    }
  }
  if (!SEP) __syncthreads();  // every wave is done reading the velocity stage
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int r = w + 4 * j;
    s_xx[L::XX(r, lane)] = oxx[j];
    s_zz[L::ZZ(r, lane)] = ozz[j];
    s_xz[L::XZ(r, lane)] = oxz[j];
  }
  if (w != 3) s_xz[L::XZ(rr, lane)] = gxz;
  if (w != 0) s_zz[L::ZZ(rr, lane)] = gzz;
#pragma unroll
  for (int p = 0; p < L::NCP; ++p) {
    const int i = p * 64 + lane;
    if ((p & 3) == w && i < L::NCI) {
      const int q = i / TZ, r = i - q * TZ;
      const int c = q < 2 ? q - 2 : 62 + q;
      if (q != 3) s_xx[L::XX(r, c)] = cxx[p];
      if (q != 0) s_xz[L::XZ(r, c)] = cxz[p];
    }
  }
  __syncthreads();

  // ---- phase C: velocities ----
  if (SEP) {  // the velocity stage still holds the start-of-step values
#pragma unroll
    for (int j = 0; j < R; ++j) {
      vx0[j] = s_vx[L::V(w + 4 * j, lane)];
      vz0[j] = s_vz[L::V(w + 4 * j, lane)];
    }
  }
  double vx[R], vz[R];
  {
    DivFlags vf;
    const FastDiv ux{a.dx, a.rdx, a.zhx, a.zlx, &vf}, uz{a.dz, a.rdz, a.zhz, a.zlz, &vf};
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int r = w + 4 * j;
      vx[j] = vx0[j];
      vz[j] = vz0[j];
      if (xint && zint(z0 + r))
        lean_velocity<TZ>(s_xx, s_zz, s_xz, r, lane, ob[j], fmin(tx, trow[j]), dts, ux,
                          uz, vx[j], vz[j]);
      row_fence();
    }
    if (!vf.ok()) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = w + 4 * j;
        vx[j] = vx0[j];
        vz[j] = vz0[j];
        if (xint && zint(z0 + r))
          lean_velocity<TZ>(s_xx, s_zz, s_xz, r, lane, ob[j], fmin(tx, trow[j]), dts, ix,
                            iz, vx[j], vz[j]);
      }
    }
  }
  // injection terms of this tile, in list order, applied by their owners
  for (int e = 0; e < ne; ++e) {
    int lc;
    double coef, val;
    if (e < NI) {
      lc = s_iloc[e], coef = s_icoef[e], val = s_ival[e];
    } else {
      const InjEntry E = a.entries[e0 + e];
      lc = E.local | (E.comp << 16), coef = E.coef, val = inj_sample(a, E, it);
    }
    const int local = lc & 0xffff, comp = lc >> 16;
    const int r = local / kTileX, c = local - r * kTileX;
    if ((r & 3) != w || c != lane) continue;
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (j == (r >> 2)) {
        const double term = !a.adjoint
                                ? coef * val * a.dt * ob[j] / (comp == 0 ? a.d4x : a.d4z)
                                : a.dt * ob[j] * val / a.dxdz;
        if (comp == 0)
          vx[j] += term;
        else
          vz[j] += term;
      }
  }
  double *vxo = a.vx_o + tb, *vzo = a.vz_o + tb;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int r = w + 4 * j;
    if (xin && z0 + r < a.nz) {
      const unsigned o = off(r, lane);
      stb(vxo, o, vx[j]);
      stb(vzo, o, vz[j]);
    }
  }
}

// CORR: instantiation for adjoint snapshot steps (kernel correlation; one
// wave per SIMD fewer, for the correlation operands in flight).
template <int TZ, int WPE, bool CORR, bool HALO_DMA = false, bool SEP = false,
          bool RDMA = false, bool CLAMP = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE)))
step_lean(StepArgs a) {
  const int tile = block_tile(a);
  if (tile >= a.ntiles) return;
  lean_tile<TZ, CORR, HALO_DMA, SEP, RDMA, CLAMP>(a, tile, a.it);
}

// ---------------------------------------------------------------------------
// Pipelined lean kernel: a block advances P tiles of 64 x TZ that are
// consecutive in z within one 64-column strip, one after the other, with
// the loads of tile k + 1 in flight while tile k computes.  Vector memory
// counters retire in issue order, so tile k's stress and velocity passes
// contain no global loads (but the rare IEEE redo): the velocity rows of a
// tile, halo columns included, go by DMA into one of two LDS buffers --
// tile k + 1's as soon as tile k's have landed -- and every other operand
// (own stresses and materials, ring row and ring-column items, taper
// factors, correlation operands) is fetched into registers in one batch,
// tile k + 1's once tile k's stresses sit in LDS, so the batch overlaps
// tile k's velocity pass.  A tile's injection entries are staged at its
// first barrier, before the next tile's transfers are queued; its velocity
// stores are issued after the next tile's first barrier.  Per point the
// arithmetic, and so every result, is that of lean_tile.
template <int TZ, bool CORR>
struct LeanRegs {
  static constexpr int R = TZ / 4, NCP = Lean<TZ>::NCP;
  double oxx[R], ozz[R], oxz[R], olm[R], omu[R], ob[R];
  double trow[R], tx;
  double gzz, gxz, glm, gmu, gt;
  double cxx[NCP], cxz[NCP], clm[NCP], cmu[NCP], ct[NCP];
  bool cin[R];
  double csv[CORR ? R : 1][5], crk[CORR ? R : 1], clk[CORR ? R : 1], cmk[CORR ? R : 1];
};

// Velocity rows z0-3 .. z0+TZ+2, columns x0-3 .. x0+66 of `tile` into the
// LDS buffer `buf` (vx, then vz) -- the halo columns come in the same
// transfers as the row.
template <int TZ>
__device__ __forceinline__ void lean_dma_tile(const StepArgs &a, int tile, lds_dbl *buf) {
  using L = Lean<TZ>;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x0 = (tile % a.ntx) * kTileX, z0 = (tile / a.ntx) * TZ;
  const long long tb = a.origin + (long long)(z0 - 3) * a.pitch + (x0 - 3);
  lds_dbl *const s_vx = buf, *const s_vz = buf + L::VR * L::VW;
#pragma unroll
  for (int k = 0; k < L::KV; ++k) {
    const int r = w + 4 * (k - 1);
    if (r >= -3 && r <= TZ + 2) {
      const long long ro = tb + (long long)(r + 3) * a.pitch;
      dma_row70(a.vx_i + ro, s_vx + L::V(r, -3));
      dma_row70(a.vz_i + ro, s_vz + L::V(r, -3));
    }
  }
}

// Every register operand of `tile` at step `it` (see LeanRegs).
template <int TZ, bool CORR>
__device__ __forceinline__ void lean_fetch_tile(const StepArgs &a, int tile, int it,
                                                LeanRegs<TZ, CORR> &g) {
  using L = Lean<TZ>;
  constexpr int R = L::R;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x0 = (tile % a.ntx) * kTileX, z0 = (tile / a.ntx) * TZ;
  const long long tb = a.origin + (long long)(z0 - 3) * a.pitch + (x0 - 8);
  const unsigned rowb = (unsigned)a.pitch * 8u;
  auto off = [&](int r, int c) -> unsigned { return lean_off(r, c, rowb); };
  const double *txxi = a.txx_i + tb, *tzzi = a.tzz_i + tb, *txzi = a.txz_i + tb;
  const double *lmp = a.lm + tb, *mup = a.mu + tb, *bp = a.b + tb;
  auto taper_row = [&](int z) {
    return a.taper_z[z < 0 ? 0 : (z >= a.nz ? a.nz - 1 : z)];
  };
#pragma unroll
  for (int j = 0; j < R; ++j) {
    unsigned o = off(w + 4 * j, lane);
    g.oxx[j] = ldb(txxi, o);
    g.ozz[j] = ldb(tzzi, o);
    g.oxz[j] = ldb(txzi, o);
    g.olm[j] = ldb(lmp, o);
    g.omu[j] = ldb(mup, o);
    g.ob[j] = ldb(bp, o);
    g.trow[j] = taper_row(z0 + w + 4 * j);
  }
  const int x = x0 + lane;
  g.tx = a.taper_x[x < a.nx ? x : 0];
  if (CORR) {
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int z = z0 + w + 4 * j;
      g.cin[j] = x < a.nx && z < a.nz && x >= a.kx0 && x < a.kx1 && z >= a.kz0 && z < a.kz1;
      if (g.cin[j]) {
        const long long ci = (long long)z * a.nx + x;
#pragma unroll
        for (int f = 0; f < 5; ++f) g.csv[j][f] = a.snap[f][ci];
        g.crk[j] = g.clk[j] = g.cmk[j] = 0.0;
        if (!a.kfresh) {
          g.crk[j] = a.rho_k[ci];
          g.clk[j] = a.lam_k[ci];
          g.cmk[j] = a.mu_k[ci];
        }
      }
    }
  }
  const int rr = w == 0 ? -2 : w == 1 ? -1 : w == 2 ? TZ : TZ + 1;
  {
    unsigned o = off(rr, lane);
    g.gzz = g.gxz = g.glm = 0.0;
    if (w != 3) g.gxz = ldb(txzi, o);
    if (w != 0) {
      g.gzz = ldb(tzzi, o);
      g.glm = ldb(lmp, o);
    }
    g.gmu = ldb(mup, o);
    g.gt = taper_row(z0 + rr);
  }
#pragma unroll
  for (int p = 0; p < L::NCP; ++p) {
    g.cxx[p] = g.cxz[p] = g.clm[p] = g.cmu[p] = g.ct[p] = 0.0;
    const int i = p * 64 + lane;
    if ((p & 3) == w && i < L::NCI) {
      const int q = i / TZ, r = i - q * TZ;
      const int c = q < 2 ? q - 2 : 62 + q;
      unsigned o = off(r, c);
      g.cxx[p] = ldb(txxi, o);
      g.cxz[p] = ldb(txzi, o);
      g.clm[p] = ldb(lmp, o);
      g.cmu[p] = ldb(mup, o);
      const int xc = x0 + c;
      g.ct[p] = fmin(a.taper_x[xc < 0 ? 0 : (xc >= a.nx ? a.nx - 1 : xc)], taper_row(z0 + r));
    }
  }
}

template <int TZ, int P, bool CORR>
__device__ __forceinline__ void lean_pipe_strip(const StepArgs &a, const int strip,
                                                const int seg, const int it) {
  using L = Lean<TZ>;
  constexpr int R = L::R;
  constexpr int NV = 2 * L::VR * L::VW;
  constexpr int NS = TZ * L::XXW + L::ZZR * 64 + L::XZR * L::XZW;
  constexpr int NB = NV > NS ? NV : NS;
  __shared__ double lds[2 * NB];
  __shared__ double s_ival[kInjStage], s_icoef[kInjStage];
  __shared__ int s_iloc[kInjStage];

  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tz0 = seg * P;
  const int nk = min(P, a.ntz - tz0);
  const unsigned rowb = (unsigned)a.pitch * 8u;
  auto off = [&](int r, int c) -> unsigned { return lean_off(r, c, rowb); };
  const int x0 = strip * kTileX, x = x0 + lane;
  const bool xin = x < a.nx;
  const bool xint = x >= 2 && x < a.nx - 2;
  const double dts = a.dts;
  const IeeeDiv ix{a.dx}, iz{a.dz};
  auto zint = [&](int z) { return z >= 2 && z < a.nz - 2; };

  LeanRegs<TZ, CORR> g;
  const int t_first = tz0 * a.ntx + strip;
  lean_dma_tile<TZ>(a, t_first, lds_ptr(lds));
  lean_fetch_tile<TZ, CORR>(a, t_first, it, g);
  if (!a.adjoint && t_first == 0) record_receivers(a, it);
  // velocities of the previous tile, stored after the next tile's barrier
  double pvx[R], pvz[R];
  long long ptb = 0;
  int pz0 = 0;
  bool pend = false;

  for (int k = 0; k < nk; ++k) {
    const int tile = (tz0 + k) * a.ntx + strip;
    const int z0 = (tz0 + k) * TZ;
    double *const buf = lds + (k & 1) * NB;
    double *const s_vx = buf, *const s_vz = buf + L::VR * L::VW;
    double *const s_xx = buf, *const s_zz = buf + TZ * L::XXW;
    double *const s_xz = s_zz + L::ZZR * 64;
    const long long tb = a.origin + (long long)(z0 - 3) * a.pitch + (x0 - 8);
    __builtin_amdgcn_s_waitcnt(0);  // this tile's LDS DMA and operands have landed
    __syncthreads();
    // this tile's injection entries and their samples into LDS, before the
    // next tile's transfers are queued behind them (most tiles have none)
    const int e0 = a.tile_off[tile], ne = a.tile_off[tile + 1] - e0;
    if ((int)threadIdx.x < ne && (int)threadIdx.x < kInjStage) {
      const InjEntry E = a.entries[e0 + threadIdx.x];
      s_iloc[threadIdx.x] = E.local | (E.comp << 16);
      s_icoef[threadIdx.x] = E.coef;
      s_ival[threadIdx.x] = inj_sample(a, E, it);
    }
    if (k + 1 < nk) lean_dma_tile<TZ>(a, tile + a.ntx, lds_ptr(lds) + ((k + 1) & 1) * NB);
    if (pend) {
      double *vxo = a.vx_o + ptb, *vzo = a.vz_o + ptb;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = w + 4 * j;
        if (xin && pz0 + r < a.nz) {
          const unsigned o = off(r, lane);
          stb(vxo, o, pvx[j]);
          stb(vzo, o, pvz[j]);
        }
      }
    }
    double vx0[R], vz0[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
      vx0[j] = s_vx[L::V(w + 4 * j, lane)];
      vz0[j] = s_vz[L::V(w + 4 * j, lane)];
    }
    const double *txxi = a.txx_i + tb, *tzzi = a.tzz_i + tb, *txzi = a.txz_i + tb;
    const double *lmp = a.lm + tb, *mup = a.mu + tb;
    if (CORR) {
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (g.cin[j]) {
          const long long ci = (long long)(z0 + w + 4 * j) * a.nx + x;
          correlate_point(a.wcorr, g.olm[j], g.omu[j], vx0[j], vz0[j], g.oxx[j], g.ozz[j],
                          g.oxz[j], g.csv[j][0], g.csv[j][1], g.csv[j][2], g.csv[j][3],
                          g.csv[j][4], g.crk[j], g.clk[j], g.cmk[j]);
          a.rho_k[ci] = g.crk[j];
          a.lam_k[ci] = g.clk[j];
          a.mu_k[ci] = g.cmk[j];
        }
    } else if (a.do_snap && !a.adjoint) {
      lean_store_snapshot<R>(a, x, z0 + w, xin, vx0, vz0, g.oxx, g.ozz, g.oxz);
    }
    // ---- stresses (no global loads but the rare IEEE redo) ----
    {
      DivFlags fl;
      const FastDiv fx{a.dx, a.rdx, a.zhx, a.zlx, &fl}, fz{a.dz, a.rdz, a.zhz, a.zlz, &fl};
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = w + 4 * j;
        if (xint && zint(z0 + r))
          lean_stress<kXX | kZZ | kXZ, L::VW>(s_vx, s_vz, L::V(r, lane), g.olm[j], g.omu[j],
                                              fmin(g.tx, g.trow[j]), dts, fx, fz, g.oxx[j],
                                              g.ozz[j], g.oxz[j]);
        row_fence();
      }
      if (!fl.ok()) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const int r = w + 4 * j;
          const unsigned o = opaque(off(r, lane));
          g.oxx[j] = ldb(txxi, o), g.ozz[j] = ldb(tzzi, o), g.oxz[j] = ldb(txzi, o);
          if (xint && zint(z0 + r))
            lean_stress<kXX | kZZ | kXZ, L::VW>(s_vx, s_vz, L::V(r, lane), ldb(lmp, o),
                                                ldb(mup, o), fmin(g.tx, g.trow[j]), dts, ix,
                                                iz, g.oxx[j], g.ozz[j], g.oxz[j]);
        }
      }
    }
    const int rr = w == 0 ? -2 : w == 1 ? -1 : w == 2 ? TZ : TZ + 1;
    {
      const int z = z0 + rr;
      if (xint && zint(z)) {
        const double t = fmin(g.tx, g.gt);
        const int v = L::V(rr, lane);
        DivFlags gf;
        const FastDiv gx{a.dx, a.rdx, a.zhx, a.zlx, &gf}, gz{a.dz, a.rdz, a.zhz, a.zlz, &gf};
        double hzz = g.gzz, hxz = g.gxz;
        ring_row_stress<L::VW>(w, s_vx, s_vz, v, g.glm, g.gmu, t, dts, gx, gz, hzz, hxz);
        if (!gf.ok()) {
          const unsigned o = opaque(off(rr, lane));
          hzz = ldb(tzzi, o), hxz = ldb(txzi, o);
          ring_row_stress<L::VW>(w, s_vx, s_vz, v, ldb(lmp, o), ldb(mup, o), t, dts, ix, iz,
                                 hzz, hxz);
        }
        g.gzz = hzz, g.gxz = hxz;
      }
    }
#pragma unroll
    for (int p = 0; p < L::NCP; ++p) {
      const int i = p * 64 + lane;
      if ((p & 3) == w && i < L::NCI) {
        const int q = i / TZ, r = i - q * TZ;
        const int c = q < 2 ? q - 2 : 62 + q;
        const int xc = x0 + c, z = z0 + r;
        if (xc >= 2 && xc < a.nx - 2 && zint(z)) {
          DivFlags cf;
          const FastDiv cx{a.dx, a.rdx, a.zhx, a.zlx, &cf}, cz{a.dz, a.rdz, a.zhz, a.zlz, &cf};
          double hxx = g.cxx[p], hxz = g.cxz[p], hzz = 0.0;
          lean_stress<kXX | kXZ, L::VW>(s_vx, s_vz, L::V(r, c), g.clm[p], g.cmu[p], g.ct[p],
                                        dts, cx, cz, hxx, hzz, hxz);
          if (!cf.ok()) {
            const unsigned o = opaque(off(r, c));
            hxx = ldb(txxi, o), hxz = ldb(txzi, o);
            lean_stress<kXX | kXZ, L::VW>(s_vx, s_vz, L::V(r, c), ldb(lmp, o), ldb(mup, o),
                                          g.ct[p], dts, ix, iz, hxx, hzz, hxz);
          }
          g.cxx[p] = hxx, g.cxz[p] = hxz;
        }
      }
    }
    double *txxo = a.txx_o + tb, *tzzo = a.tzz_o + tb, *txzo = a.txz_o + tb;
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int r = w + 4 * j, z = z0 + r;
      if (xin && z < a.nz) {
        const unsigned o = off(r, lane);
        stb(txxo, o, g.oxx[j]);
        stb(tzzo, o, g.ozz[j]);
        stb(txzo, o, g.oxz[j]);
      }
    }
    __syncthreads();  // every wave is done reading this tile's velocity stage
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int r = w + 4 * j;
      s_xx[L::XX(r, lane)] = g.oxx[j];
      s_zz[L::ZZ(r, lane)] = g.ozz[j];
      s_xz[L::XZ(r, lane)] = g.oxz[j];
    }
    if (w != 3) s_xz[L::XZ(rr, lane)] = g.gxz;
    if (w != 0) s_zz[L::ZZ(rr, lane)] = g.gzz;
#pragma unroll
    for (int p = 0; p < L::NCP; ++p) {
      const int i = p * 64 + lane;
      if ((p & 3) == w && i < L::NCI) {
        const int q = i / TZ, r = i - q * TZ;
        const int c = q < 2 ? q - 2 : 62 + q;
        if (q != 3) s_xx[L::XX(r, c)] = g.cxx[p];
        if (q != 0) s_xz[L::XZ(r, c)] = g.cxz[p];
      }
    }
    __syncthreads();
    // the velocity pass needs only b, the tapers and the injection entries
    // of this tile; the next tile's operands are fetched now
    double ob[R], trow[R];
#pragma unroll
    for (int j = 0; j < R; ++j) ob[j] = g.ob[j], trow[j] = g.trow[j];
    const double tx = g.tx;
    if (k + 1 < nk) lean_fetch_tile<TZ, CORR>(a, tile + a.ntx, it, g);
    {
      DivFlags vf;
      const FastDiv ux{a.dx, a.rdx, a.zhx, a.zlx, &vf}, uz{a.dz, a.rdz, a.zhz, a.zlz, &vf};
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = w + 4 * j;
        pvx[j] = vx0[j];
        pvz[j] = vz0[j];
        if (xint && zint(z0 + r))
          lean_velocity<TZ>(s_xx, s_zz, s_xz, r, lane, ob[j], fmin(tx, trow[j]), dts, ux, uz,
                            pvx[j], pvz[j]);
        row_fence();
      }
      if (!vf.ok()) {
#pragma unroll
        for (int j = 0; j < R; ++j) {
          const int r = w + 4 * j;
          pvx[j] = vx0[j];
          pvz[j] = vz0[j];
          if (xint && zint(z0 + r))
            lean_velocity<TZ>(s_xx, s_zz, s_xz, r, lane, ob[j], fmin(tx, trow[j]), dts, ix,
                              iz, pvx[j], pvz[j]);
        }
      }
    }
    for (int e = 0; e < ne; ++e) {
      int lc;
      double coef, val;
      if (e < kInjStage) {
        lc = s_iloc[e], coef = s_icoef[e], val = s_ival[e];
      } else {
        const InjEntry E = a.entries[e0 + e];
        lc = E.local | (E.comp << 16), coef = E.coef, val = inj_sample(a, E, it);
      }
      const int local = lc & 0xffff, comp = lc >> 16;
      const int r = local / kTileX, c = local - r * kTileX;
      if ((r & 3) != w || c != lane) continue;
#pragma unroll
      for (int j = 0; j < R; ++j)
        if (j == (r >> 2)) {
          const double term = !a.adjoint
                                  ? coef * val * a.dt * ob[j] / (comp == 0 ? a.d4x : a.d4z)
                                  : a.dt * ob[j] * val / a.dxdz;
          if (comp == 0)
            pvx[j] += term;
          else
            pvz[j] += term;
        }
    }
    ptb = tb;
    pz0 = z0;
    pend = true;
    if (k + 1 == nk) {
      double *vxo = a.vx_o + tb, *vzo = a.vz_o + tb;
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = w + 4 * j;
        if (xin && z0 + r < a.nz) {
          const unsigned o = off(r, lane);
          stb(vxo, o, pvx[j]);
          stb(vzo, o, pvz[j]);
        }
      }
    }
  }
}

// Block -> (strip, segment of P tiles); with xcd_remap, blocks that the
// dispatcher sends to one XCD get a contiguous run of (strip, segment)
// items, x fastest.
template <int TZ, int P, int WPE, bool CORR>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE)))
step_lean_pipe(StepArgs a) {
  const int nseg = (a.ntz + P - 1) / P;
  const int nitems = a.ntx * nseg;
  int b = blockIdx.x;
  if (a.xcd_remap) b = (b % 8) * ((nitems + 7) / 8) + b / 8;
  if (b >= nitems) return;
  lean_pipe_strip<TZ, P, CORR>(a, b % a.ntx, b / a.ntx, a.it);
}

// ---------------------------------------------------------------------------
// Lean two-step kernel: steps s0 = a.it and s1 = a.it1 of one 64 x TZ tile
// per launch (temporal blocking: the state crosses HBM once per two steps).
// Four LDS stages updated in place, with the tile-relative boxes
//   v0  rows -8 .. TZ+7, cols -8 .. 71   (A, loaded)
//   σ1  rows -6 .. TZ+5, cols -6 .. 69   (B: σ0 loaded, updated in place)
//   v1  rows -4 .. TZ+3, cols -4 .. 67   (A, in place; step s0's injection
//                                          terms on the whole box)
//   σ2  rows -2 .. TZ+1, cols -2 .. 65   (B, in place)
//   v2  rows  0 .. TZ-1, cols  0 .. 63   (to memory; step s1's injection)
// Every stage maps lane l of wave w to column l on rows w + 4k of its box,
// plus one pass of "side" items (the box columns outside 0 .. 63, lane ->
// (column, row)).  An update reads only neighbours from the other buffer and
// its own old value, so the in-place writes need no extra barrier.  Owners
// of the tile's points record receivers at both steps (increments; a scan
// forms the running sums after the loop), store forward snapshots and
// correlate adjoint ones from the start-of-step values in LDS.
template <int TZ>
struct Lean2 {
  static constexpr int AH = TZ + 16, AW = 80;  // v box
  static constexpr int BH = TZ + 12, BW = 76;  // σ box
  static constexpr int KA = AH / 4;            // A rows per wave
  static constexpr int K1 = BH / 4;            // σ1 rows per wave
  static constexpr int KV = (TZ + 8) / 4;      // v1 rows per wave
  static constexpr int NSA = 16 * AH;          // A side items (16 columns)
  static constexpr int NS1 = 12 * BH;          // σ1 side items (12 columns)
  static constexpr int NV1 = 8 * (TZ + 8);     // v1 side items (8 columns)
  static constexpr int NS2 = 4 * (TZ + 4);     // σ2 side items (4 columns)
  __device__ static int A(int r, int c) { return (r + 8) * AW + (c + 8); }
  __device__ static int B(int r, int c) { return (r + 6) * BW + (c + 6); }
  // side columns of a box of half-width h (h columns left of 0, h right of 63)
  __device__ static int side_col(int s, int h) { return s < h ? s - h : 64 + (s - h); }
};

// Velocity from stresses in a box of width W (velocity_point's trees).
template <int W, class DX, class DZ>
__device__ __forceinline__ void l2_velocity(const double *sxx, const double *szz,
                                            const double *sxz, int i, double b,
                                            double taper, double dts, DX divx,
                                            DZ divz, double &vx, double &vz) {
  double ax = divx(kC1 * (sxx[i] - sxx[i - 1]) + kC2 * (sxx[i - 2] - sxx[i + 1])) +
              divz(kC1 * (sxz[i] - sxz[i - W]) + kC2 * (sxz[i - 2 * W] - sxz[i + W]));
  double az = divx(kC1 * (sxz[i + 1] - sxz[i]) + kC2 * (sxz[i - 1] - sxz[i + 2])) +
              divz(kC1 * (szz[i + W] - szz[i]) + kC2 * (szz[i - W] - szz[i + 2 * W]));
  vx = taper * (vx + b * dts * ax);
  vz = taper * (vz + b * dts * az);
}

template <int TZ>
__device__ __forceinline__ void lean2_tile(const StepArgs &a, const int tile) {
  using L = Lean2<TZ>;
  constexpr int R = TZ / 4;
  __shared__ double s_vx[L::AH * L::AW], s_vz[L::AH * L::AW];
  __shared__ double s_xx[L::BH * L::BW], s_zz[L::BH * L::BW], s_xz[L::BH * L::BW];
  __shared__ double s_ival[2][kInjStage], s_icoef[2][kInjStage];
  __shared__ int s_iloc[2][kInjStage];

  const int lane = threadIdx.x & 63, tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int x0 = (tile % a.ntx) * kTileX, z0 = (tile / a.ntx) * TZ;
  // field bases at tile-relative (row -8, column -8)
  const long long tb = a.origin + (long long)(z0 - 8) * a.pitch + (x0 - 8);
  const unsigned rowb = (unsigned)a.pitch * 8u;
  auto off = [&](int r, int c) -> unsigned {
    return (unsigned)(r + 8) * rowb + (unsigned)(c + 8) * 8u;
  };
  const double *vxi = a.vx_i + tb, *vzi = a.vz_i + tb;
  const double *txxi = a.txx_i + tb, *tzzi = a.tzz_i + tb, *txzi = a.txz_i + tb;
  const double *lmp = a.lm + tb, *mup = a.mu + tb, *bp = a.b + tb;
  auto interior = [&](int r, int c) {
    const int x = x0 + c, z = z0 + r;
    return x >= 2 && x < a.nx - 2 && z >= 2 && z < a.nz - 2;
  };
  auto taper = [&](int r, int c) {
    int x = x0 + c, z = z0 + r;
    x = x < 0 ? 0 : (x >= a.nx ? a.nx - 1 : x);
    z = z < 0 ? 0 : (z >= a.nz ? a.nz - 1 : z);
    return fmin(a.taper_x[x], a.taper_z[z]);
  };

  // ---- phase A: every load before the first wait ----
  // main columns 0..63 of A and B straight into LDS
#pragma unroll
  for (int k = 0; k < L::KA; ++k) {
    const int r = -8 + w + 4 * k;
    const unsigned o = off(r, lane);
    dma_row(vxi, o, lds_ptr(s_vx) + L::A(r, 0));
    dma_row(vzi, o, lds_ptr(s_vz) + L::A(r, 0));
  }
#pragma unroll
  for (int k = 0; k < L::K1; ++k) {
    const int r = -6 + w + 4 * k;
    const unsigned o = off(r, lane);
    dma_row(txxi, o, lds_ptr(s_xx) + L::B(r, 0));
    dma_row(tzzi, o, lds_ptr(s_zz) + L::B(r, 0));
    dma_row(txzi, o, lds_ptr(s_xz) + L::B(r, 0));
  }
  // side columns through registers
  double avx[2], avz[2];
  int aidx[2];
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    const int t = tid + 256 * k;
    aidx[k] = -1;
    avx[k] = avz[k] = 0.0;
    if (k * 256 < L::NSA && t < L::NSA) {
      const int s = t / L::AH, r = t - s * L::AH - 8, c = L::side_col(s, 8);
      const unsigned o = off(r, c);
      avx[k] = ldb(vxi, o);
      avz[k] = ldb(vzi, o);
      aidx[k] = L::A(r, c);
    }
  }
  double bxx = 0.0, bzz = 0.0, bxz = 0.0, slm = 0.0, smu = 0.0;
  const bool has1 = tid < L::NS1;
  const int s1r = has1 ? tid % L::BH - 6 : 0, s1c = has1 ? L::side_col(tid / L::BH, 6) : 0;
  if (has1) {
    const unsigned o = off(s1r, s1c);
    bxx = ldb(txxi, o), bzz = ldb(tzzi, o), bxz = ldb(txzi, o);
    slm = ldb(lmp, o), smu = ldb(mup, o);
  }
  // materials of this thread's main items
  double lm1[L::K1], mu1[L::K1], bv[L::KV];
#pragma unroll
  for (int k = 0; k < L::K1; ++k) {
    const unsigned o = off(-6 + w + 4 * k, lane);
    lm1[k] = ldb(lmp, o);
    mu1[k] = ldb(mup, o);
  }
#pragma unroll
  for (int k = 0; k < L::KV; ++k) bv[k] = ldb(bp, off(-4 + w + 4 * k, lane));
  const bool hasv = tid < L::NV1;
  const int v1r = hasv ? tid % (TZ + 8) - 4 : 0, v1c = hasv ? L::side_col(tid / (TZ + 8), 4) : 0;
  const double vb = hasv ? ldb(bp, off(v1r, v1c)) : 0.0;
  const bool has2 = tid < L::NS2;
  const int s2r = has2 ? tid % (TZ + 4) - 2 : 0, s2c = has2 ? L::side_col(tid / (TZ + 4), 2) : 0;
  double s2lm = 0.0, s2mu = 0.0;
  if (has2) {
    const unsigned o = off(s2r, s2c);
    s2lm = ldb(lmp, o), s2mu = ldb(mup, o);
  }
  // injection entries: step s0 on the v1 box (halo table, local index over
  // (64 + 8) x (TZ + 8) from (x0 - 4, z0 - 4)), step s1 on the tile
  const int h0 = a.halo_off[tile], nh = a.halo_off[tile + 1] - h0;
  const int e0 = a.tile_off[tile], ne = a.tile_off[tile + 1] - e0;
  for (int t = tid; t < nh && t < kInjStage; t += 256) {
    const InjEntry E = a.halo_entries[h0 + t];
    s_iloc[0][t] = E.local | (E.comp << 16);
    s_icoef[0][t] = E.coef;
    s_ival[0][t] = inj_sample(a, E, a.it);
  }
  for (int t = tid; t < ne && t < kInjStage; t += 256) {
    const InjEntry E = a.entries[e0 + t];
    s_iloc[1][t] = E.local | (E.comp << 16);
    s_icoef[1][t] = E.coef;
    s_ival[1][t] = inj_sample(a, E, a.it1);
  }
#pragma unroll
  for (int k = 0; k < 2; ++k)
    if (aidx[k] >= 0) {
      s_vx[aidx[k]] = avx[k];
      s_vz[aidx[k]] = avz[k];
    }
  if (has1) {
    const int i = L::B(s1r, s1c);
    s_xx[i] = bxx, s_zz[i] = bzz, s_xz[i] = bxz;
  }
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();

  const double dts = a.dts;
  const IeeeDiv ix{a.dx}, iz{a.dz};
  const int xo = x0 + lane;  // this lane's column in main items
  const bool xin = xo < a.nx;
  // owner work at the start of a step for own point (r, lane): forward
  // snapshot store or adjoint correlation (src/fdModel.cpp:555-572, 817-861)
  auto owner_start = [&](double *const *snap, bool do_snap, int r) {
    const int z = z0 + r;
    if (!do_snap || !xin || z >= a.nz) return;
    const int ia = L::A(r, lane), ib = L::B(r, lane);
    owner_snapshot(a, snap, xo, z, s_vx[ia], s_vz[ia], s_xx[ib], s_zz[ib], s_xz[ib]);
  };

  // ---- stage 1: σ1 on the σ box ----
  {
    double n1[L::K1][3];
    DivFlags fl;
    const FastDiv fx{a.dx, a.rdx, a.zhx, a.zlx, &fl}, fz{a.dz, a.rdz, a.zhz, a.zlz, &fl};
#pragma unroll
    for (int k = 0; k < L::K1; ++k) {
      const int r = -6 + w + 4 * k, ib = L::B(r, lane);
      n1[k][0] = s_xx[ib], n1[k][1] = s_zz[ib], n1[k][2] = s_xz[ib];
      if (r >= 0 && r < TZ) owner_start(a.snap, a.do_snap, r);
      if (interior(r, lane))
        lean_stress<kXX | kZZ | kXZ, L::AW>(s_vx, s_vz, L::A(r, lane), lm1[k], mu1[k],
                                            taper(r, lane), dts, fx, fz, n1[k][0],
                                            n1[k][1], n1[k][2]);
      row_fence();
    }
    if (!fl.ok()) {
#pragma unroll
      for (int k = 0; k < L::K1; ++k) {
        const int r = -6 + w + 4 * k, ib = L::B(r, lane);
        n1[k][0] = s_xx[ib], n1[k][1] = s_zz[ib], n1[k][2] = s_xz[ib];
        if (interior(r, lane))
          lean_stress<kXX | kZZ | kXZ, L::AW>(s_vx, s_vz, L::A(r, lane), lm1[k], mu1[k],
                                              taper(r, lane), dts, ix, iz, n1[k][0],
                                              n1[k][1], n1[k][2]);
      }
    }
    double h[3] = {bxx, bzz, bxz};
    if (has1 && interior(s1r, s1c)) {
      DivFlags sf;
      const FastDiv gx{a.dx, a.rdx, a.zhx, a.zlx, &sf}, gz{a.dz, a.rdz, a.zhz, a.zlz, &sf};
      const double t = taper(s1r, s1c);
      lean_stress<kXX | kZZ | kXZ, L::AW>(s_vx, s_vz, L::A(s1r, s1c), slm, smu, t, dts, gx,
                                          gz, h[0], h[1], h[2]);
      if (!sf.ok()) {
        h[0] = bxx, h[1] = bzz, h[2] = bxz;
        lean_stress<kXX | kZZ | kXZ, L::AW>(s_vx, s_vz, L::A(s1r, s1c), slm, smu, t, dts,
                                            ix, iz, h[0], h[1], h[2]);
      }
    }
#pragma unroll
    for (int k = 0; k < L::K1; ++k) {
      const int ib = L::B(-6 + w + 4 * k, lane);
      s_xx[ib] = n1[k][0], s_zz[ib] = n1[k][1], s_xz[ib] = n1[k][2];
    }
    if (has1) {
      const int ib = L::B(s1r, s1c);
      s_xx[ib] = h[0], s_zz[ib] = h[1], s_xz[ib] = h[2];
    }
  }
  __syncthreads();

  // ---- stage 2: v1 on the v box, step s0's injection, receivers ----
  {
    double vx[L::KV], vz[L::KV];
    DivFlags fl;
    const FastDiv fx{a.dx, a.rdx, a.zhx, a.zlx, &fl}, fz{a.dz, a.rdz, a.zhz, a.zlz, &fl};
#pragma unroll
    for (int k = 0; k < L::KV; ++k) {
      const int r = -4 + w + 4 * k, ia = L::A(r, lane);
      vx[k] = s_vx[ia], vz[k] = s_vz[ia];
      if (interior(r, lane))
        l2_velocity<L::BW>(s_xx, s_zz, s_xz, L::B(r, lane), bv[k], taper(r, lane), dts,
                           fx, fz, vx[k], vz[k]);
      row_fence();
    }
    if (!fl.ok()) {
#pragma unroll
      for (int k = 0; k < L::KV; ++k) {
        const int r = -4 + w + 4 * k, ia = L::A(r, lane);
        vx[k] = s_vx[ia], vz[k] = s_vz[ia];
        if (interior(r, lane))
          l2_velocity<L::BW>(s_xx, s_zz, s_xz, L::B(r, lane), bv[k], taper(r, lane), dts,
                             ix, iz, vx[k], vz[k]);
      }
    }
    double svx = 0.0, svz = 0.0;
    if (hasv) {
      const int ia = L::A(v1r, v1c);
      svx = s_vx[ia], svz = s_vz[ia];
      if (interior(v1r, v1c)) {
        DivFlags sf;
        const FastDiv gx{a.dx, a.rdx, a.zhx, a.zlx, &sf}, gz{a.dz, a.rdz, a.zhz, a.zlz, &sf};
        const double t = taper(v1r, v1c);
        l2_velocity<L::BW>(s_xx, s_zz, s_xz, L::B(v1r, v1c), vb, t, dts, gx, gz, svx, svz);
        if (!sf.ok()) {
          svx = s_vx[ia], svz = s_vz[ia];
          l2_velocity<L::BW>(s_xx, s_zz, s_xz, L::B(v1r, v1c), vb, t, dts, ix, iz, svx,
                             svz);
        }
      }
    }
    // step s0's injection terms on the v box, in list order, by owners
    for (int e = 0; e < nh; ++e) {
      int lc;
      double coef, val;
      if (e < kInjStage) {
        lc = s_iloc[0][e], coef = s_icoef[0][e], val = s_ival[0][e];
      } else {
        const InjEntry E = a.halo_entries[h0 + e];
        lc = E.local | (E.comp << 16), coef = E.coef, val = inj_sample(a, E, a.it);
      }
      const int local = lc & 0xffff, comp = lc >> 16;
      const int r = local / (kTileX + 8) - 4, c = local % (kTileX + 8) - 4;
      if (c >= 0 && c < 64) {
        if (c != lane || ((r + 4) & 3) != w) continue;
#pragma unroll
        for (int k = 0; k < L::KV; ++k)
          if (k == ((r + 4) >> 2)) {
            const double term = !a.adjoint
                                    ? coef * val * a.dt * bv[k] / (comp == 0 ? a.d4x : a.d4z)
                                    : a.dt * bv[k] * val / a.dxdz;
            if (comp == 0)
              vx[k] += term;
            else
              vz[k] += term;
          }
      } else {
        if (!hasv || r != v1r || c != v1c) continue;
        const double term = !a.adjoint ? coef * val * a.dt * vb / (comp == 0 ? a.d4x : a.d4z)
                                       : a.dt * vb * val / a.dxdz;
        if (comp == 0)
          svx += term;
        else
          svz += term;
      }
    }
    // owners: receivers at the start of s0 and s1, forward snapshot of v1
#pragma unroll
    for (int k = 0; k < L::KV; ++k) {
      const int r = -4 + w + 4 * k;
      if (r < 0 || r >= TZ || !xin || z0 + r >= a.nz) continue;
      if (!a.adjoint) {
        const int ia = L::A(r, lane), local = r * kTileX + lane;
        record_owned(a, tile, local, a.it, s_vx[ia], s_vz[ia]);
        record_owned(a, tile, local, a.it1, vx[k], vz[k]);
        if (a.do_snap1) {
          const long long ci = (long long)(z0 + r) * a.nx + xo;
          a.snap1[0][ci] = vx[k];
          a.snap1[1][ci] = vz[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < L::KV; ++k) {
      const int ia = L::A(-4 + w + 4 * k, lane);
      s_vx[ia] = vx[k], s_vz[ia] = vz[k];
    }
    if (hasv) {
      const int ia = L::A(v1r, v1c);
      s_vx[ia] = svx, s_vz[ia] = svz;
    }
  }
  __syncthreads();

  // ---- stage 3: σ2 on the σ2 box (main rows are stage-1 rows k = 1..3) ----
  double *txxo = a.txx_o + tb, *tzzo = a.tzz_o + tb, *txzo = a.txz_o + tb;
  {
    constexpr int K2 = (TZ + 4) / 4;
    double n2[K2][3];
    DivFlags fl;
    const FastDiv fx{a.dx, a.rdx, a.zhx, a.zlx, &fl}, fz{a.dz, a.rdz, a.zhz, a.zlz, &fl};
#pragma unroll
    for (int k = 0; k < K2; ++k) {
      const int r = -2 + w + 4 * k, ib = L::B(r, lane);
      n2[k][0] = s_xx[ib], n2[k][1] = s_zz[ib], n2[k][2] = s_xz[ib];
      if (r >= 0 && r < TZ) {
        if (a.adjoint) {
          owner_start(a.snap1, a.do_snap1, r);
        } else if (a.do_snap1 && xin && z0 + r < a.nz) {
          const long long ci = (long long)(z0 + r) * a.nx + xo;
          a.snap1[2][ci] = n2[k][0];
          a.snap1[3][ci] = n2[k][1];
          a.snap1[4][ci] = n2[k][2];
        }
      }
      if (interior(r, lane))
        lean_stress<kXX | kZZ | kXZ, L::AW>(s_vx, s_vz, L::A(r, lane), lm1[k + 1],
                                            mu1[k + 1], taper(r, lane), dts, fx, fz,
                                            n2[k][0], n2[k][1], n2[k][2]);
      row_fence();
    }
    if (!fl.ok()) {
#pragma unroll
      for (int k = 0; k < K2; ++k) {
        const int r = -2 + w + 4 * k, ib = L::B(r, lane);
        n2[k][0] = s_xx[ib], n2[k][1] = s_zz[ib], n2[k][2] = s_xz[ib];
        if (interior(r, lane))
          lean_stress<kXX | kZZ | kXZ, L::AW>(s_vx, s_vz, L::A(r, lane), lm1[k + 1],
                                              mu1[k + 1], taper(r, lane), dts, ix, iz,
                                              n2[k][0], n2[k][1], n2[k][2]);
      }
    }
    double h[3] = {0.0, 0.0, 0.0};
    if (has2) {
      const int ib = L::B(s2r, s2c);
      h[0] = s_xx[ib], h[1] = s_zz[ib], h[2] = s_xz[ib];
      if (interior(s2r, s2c)) {
        DivFlags sf;
        const FastDiv gx{a.dx, a.rdx, a.zhx, a.zlx, &sf}, gz{a.dz, a.rdz, a.zhz, a.zlz, &sf};
        const double t = taper(s2r, s2c);
        lean_stress<kXX | kZZ | kXZ, L::AW>(s_vx, s_vz, L::A(s2r, s2c), s2lm, s2mu, t, dts,
                                            gx, gz, h[0], h[1], h[2]);
        if (!sf.ok()) {
          h[0] = s_xx[ib], h[1] = s_zz[ib], h[2] = s_xz[ib];
          lean_stress<kXX | kZZ | kXZ, L::AW>(s_vx, s_vz, L::A(s2r, s2c), s2lm, s2mu, t,
                                              dts, ix, iz, h[0], h[1], h[2]);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < K2; ++k) {
      const int r = -2 + w + 4 * k, ib = L::B(r, lane);
      s_xx[ib] = n2[k][0], s_zz[ib] = n2[k][1], s_xz[ib] = n2[k][2];
      if (r >= 0 && r < TZ && xin && z0 + r < a.nz) {
        const unsigned o = off(r, lane);
        stb(txxo, o, n2[k][0]);
        stb(tzzo, o, n2[k][1]);
        stb(txzo, o, n2[k][2]);
      }
    }
    if (has2) {
      const int ib = L::B(s2r, s2c);
      s_xx[ib] = h[0], s_zz[ib] = h[1], s_xz[ib] = h[2];
    }
  }
  __syncthreads();

  // ---- stage 4: v2 on the tile, step s1's injection, store ----
  double vx[R], vz[R];
  {
    DivFlags fl;
    const FastDiv fx{a.dx, a.rdx, a.zhx, a.zlx, &fl}, fz{a.dz, a.rdz, a.zhz, a.zlz, &fl};
#pragma unroll
    for (int j = 0; j < R; ++j) {
      const int r = w + 4 * j, ia = L::A(r, lane);
      vx[j] = s_vx[ia], vz[j] = s_vz[ia];
      if (interior(r, lane))
        l2_velocity<L::BW>(s_xx, s_zz, s_xz, L::B(r, lane), bv[j + 1], taper(r, lane), dts,
                           fx, fz, vx[j], vz[j]);
      row_fence();
    }
    if (!fl.ok()) {
#pragma unroll
      for (int j = 0; j < R; ++j) {
        const int r = w + 4 * j, ia = L::A(r, lane);
        vx[j] = s_vx[ia], vz[j] = s_vz[ia];
        if (interior(r, lane))
          l2_velocity<L::BW>(s_xx, s_zz, s_xz, L::B(r, lane), bv[j + 1], taper(r, lane),
                             dts, ix, iz, vx[j], vz[j]);
      }
    }
  }
  for (int e = 0; e < ne; ++e) {
    int lc;
    double coef, val;
    if (e < kInjStage) {
      lc = s_iloc[1][e], coef = s_icoef[1][e], val = s_ival[1][e];
    } else {
      const InjEntry E = a.entries[e0 + e];
      lc = E.local | (E.comp << 16), coef = E.coef, val = inj_sample(a, E, a.it1);
    }
    const int local = lc & 0xffff, comp = lc >> 16;
    const int r = local / kTileX, c = local - r * kTileX;
    if ((r & 3) != w || c != lane) continue;
#pragma unroll
    for (int j = 0; j < R; ++j)
      if (j == (r >> 2)) {
        const double term = !a.adjoint
                                ? coef * val * a.dt * bv[j + 1] / (comp == 0 ? a.d4x : a.d4z)
                                : a.dt * bv[j + 1] * val / a.dxdz;
        if (comp == 0)
          vx[j] += term;
        else
          vz[j] += term;
      }
  }
  double *vxo = a.vx_o + tb, *vzo = a.vz_o + tb;
#pragma unroll
  for (int j = 0; j < R; ++j) {
    const int r = w + 4 * j;
    if (xin && z0 + r < a.nz) {
      const unsigned o = off(r, lane);
      stb(vxo, o, vx[j]);
      stb(vzo, o, vz[j]);
    }
  }
}

template <int TZ>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
step_lean2(StepArgs a) {
  const int tile = block_tile(a);
  if (tile >= a.ntiles) return;
  lean2_tile<TZ>(a, tile);
}

// Two time steps per launch (s0 = a.it, s1 = a.it1).  The tile stages
// velocities with an 8-point halo and stresses with a 6-point ring, then
// updates in place in LDS: stresses of step s0 on the 6-ring, velocities on
// the 4-ring (applying step s0's injection terms on the whole 4-ring, so the
// halo copies are exact), stresses of step s1 on the 2-ring, velocities of
// step s1 on the tile.  Global traffic per point-update is about half that
// of a one-step launch; the stencil work on the rings is recomputed.
template <int TZ, int NT>
__device__ __forceinline__ void two_step_tile(const StepArgs &a, const int tile) {
  constexpr int TX = kTileX;
  constexpr int WV = TX + 16, HV = TZ + 16;  // v on the 8-ring
  constexpr int WS = TX + 12, HS = TZ + 12;  // stresses on the 6-ring
  constexpr int W4 = TX + 8, H4 = TZ + 8;    // velocity items of step s0
  constexpr int W2 = TX + 4, H2 = TZ + 4;    // stress items of step s1
  constexpr int NB = (HS * WS + NT - 1) / NT;
  constexpr int NC = (H4 * W4 + NT - 1) / NT;
  constexpr int ND = (H2 * W2 + NT - 1) / NT;
  constexpr int NE = (TZ * TX + NT - 1) / NT;
  __shared__ double s_vx[HV * WV], s_vz[HV * WV];
  __shared__ double s_txx[HS * WS], s_tzz[HS * WS], s_txz[HS * WS];
  const IeeeDiv dvx{a.dx}, dvz{a.dz};
  const int x0 = (tile % a.ntx) * TX, z0 = (tile / a.ntx) * TZ;
  const int tid = threadIdx.x;
  auto in_grid = [&](int x, int z) {
    return x >= 0 && x < a.nx && z >= 0 && z < a.nz;
  };
  auto is_own = [&](int x, int z) {
    return x >= x0 && x < x0 + TX && z >= z0 && z < z0 + TZ;
  };

  // phase A: every global load before the first wait (fully unrolled loops
  // into registers, then LDS).  v on the 8-ring (inside the zero margin),
  // stresses on the 6-ring (zero outside the grid); the materials of this
  // thread's items of phases B-E stay in registers.
  constexpr int NV = (HV * WV + NT - 1) / NT;
  double vx_r[NV], vz_r[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    int i = tid + k * NT;
    vx_r[k] = vz_r[k] = 0.0;
    if (i < HV * WV) {
      int r = i / WV, c = i - r * WV;
      long long g = pidx(a, x0 - 8 + c, z0 - 8 + r);
      vx_r[k] = a.vx_i[g];
      vz_r[k] = a.vz_i[g];
    }
  }
  double txx_r[NB], tzz_r[NB], txz_r[NB];
  double lmB[NB], muB[NB], bC[NC], lmD[ND], muD[ND], bE[NE];
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    int i = tid + k * NT;
    txx_r[k] = tzz_r[k] = txz_r[k] = lmB[k] = muB[k] = 0.0;
    if (i < HS * WS) {
      int r = i / WS, c = i - r * WS;
      int x = x0 - 6 + c, z = z0 - 6 + r;
      if (in_grid(x, z)) {
        long long g = pidx(a, x, z);
        txx_r[k] = a.txx_i[g];
        tzz_r[k] = a.tzz_i[g];
        txz_r[k] = a.txz_i[g];
        if (interior(a, x, z)) {
          lmB[k] = a.lm[g];
          muB[k] = a.mu[g];
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    int i = tid + k * NT;
    bC[k] = 0.0;
    if (i < H4 * W4) {
      int r = i / W4, c = i - r * W4;
      int x = x0 - 4 + c, z = z0 - 4 + r;
      if (in_grid(x, z)) bC[k] = a.b[pidx(a, x, z)];
    }
  }
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    int i = tid + k * NT;
    lmD[k] = muD[k] = 0.0;
    if (i < H2 * W2) {
      int r = i / W2, c = i - r * W2;
      int x = x0 - 2 + c, z = z0 - 2 + r;
      if (in_grid(x, z) && interior(a, x, z)) {
        long long g = pidx(a, x, z);
        lmD[k] = a.lm[g];
        muD[k] = a.mu[g];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    int i = tid + k * NT;
    bE[k] = 0.0;
    if (i < TZ * TX) {
      int r = i / TX, c = i - r * TX;
      int x = x0 + c, z = z0 + r;
      if (in_grid(x, z)) bE[k] = a.b[pidx(a, x, z)];
    }
  }
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    int i = tid + k * NT;
    if (i < HV * WV) {
      s_vx[i] = vx_r[k];
      s_vz[i] = vz_r[k];
    }
  }
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    int i = tid + k * NT;
    if (i < HS * WS) {
      s_txx[i] = txx_r[k];
      s_tzz[i] = tzz_r[k];
      s_txz[i] = txz_r[k];
    }
  }
  __syncthreads();

  // phase B: stresses of step s0 on the 6-ring (in place).  Owners store
  // the forward snapshot of step s0's stresses or correlate (adjoint).
#pragma unroll
  for (int k = 0; k < NB; ++k) {
    int i = tid + k * NT;
    if (i >= HS * WS) break;
    int r = i / WS, c = i - r * WS;
    int x = x0 - 6 + c, z = z0 - 6 + r;
    if (!in_grid(x, z)) continue;
    double txx = txx_r[k], tzz = tzz_r[k], txz = txz_r[k];
    int lv = (r + 2) * WV + (c + 2);
    if (a.do_snap && is_own(x, z)) {
      if (!a.adjoint) {
        long long ci = (long long)z * a.nx + x;
        a.snap[2][ci] = txx;
        a.snap[3][ci] = tzz;
        a.snap[4][ci] = txz;
      } else {
        owner_snapshot(a, a.snap, x, z, s_vx[lv], s_vz[lv], txx, tzz, txz);
      }
    }
    if (interior(a, x, z)) {
      stress_point(s_vx, s_vz, lv, 1, WV, lmB[k], muB[k], taper_at(a, x, z),
                   a.dts, dvx, dvz, txx, tzz, txz);
      s_txx[i] = txx;
      s_tzz[i] = tzz;
      s_txz[i] = txz;
    }
  }
  __syncthreads();

  // phase C: velocities of step s0 on the 4-ring (in place), with step s0's
  // injection terms for every point of the ring; owners record receivers
  // (forward) before and after, and store the forward snapshot of v.
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    int i = tid + k * NT;
    if (i >= H4 * W4) break;
    int r = i / W4, c = i - r * W4;
    int x = x0 - 4 + c, z = z0 - 4 + r;
    if (!in_grid(x, z)) continue;
    int lv = (r + 4) * WV + (c + 4);
    double vx = s_vx[lv], vz = s_vz[lv];
    bool own = is_own(x, z);
    int local = (z - z0) * TX + (x - x0);
    if (own && !a.adjoint) {
      if (a.do_snap) {
        long long ci = (long long)z * a.nx + x;
        a.snap[0][ci] = vx;
        a.snap[1][ci] = vz;
      }
      record_owned(a, tile, local, a.it, vx, vz);
    }
    double b = bC[k];
    if (interior(a, x, z)) {
      velocity_point(s_txx, s_tzz, s_txz, (r + 2) * WS + (c + 2), 1, WS, b,
                     taper_at(a, x, z), a.dts, dvx, dvz, vx, vz);
    }
    apply_terms(a, a.halo_off, a.halo_entries, tile, i, a.it, b, vx, vz);
    s_vx[lv] = vx;
    s_vz[lv] = vz;
    if (own && !a.adjoint) record_owned(a, tile, local, a.it1, vx, vz);
  }
  __syncthreads();

  // phase D: stresses of step s1 on the 2-ring (in place); owners handle the
  // step-s1 snapshot of stresses / correlation and store the new stresses.
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    int i = tid + k * NT;
    if (i >= H2 * W2) break;
    int r = i / W2, c = i - r * W2;
    int x = x0 - 2 + c, z = z0 - 2 + r;
    if (!in_grid(x, z)) continue;
    int ls = (r + 4) * WS + (c + 4);
    int lv = (r + 6) * WV + (c + 6);
    double txx = s_txx[ls], tzz = s_tzz[ls], txz = s_txz[ls];
    bool own = is_own(x, z);
    if (own && a.do_snap1) {
      if (!a.adjoint) {
        long long ci = (long long)z * a.nx + x;
        a.snap1[2][ci] = txx;
        a.snap1[3][ci] = tzz;
        a.snap1[4][ci] = txz;
      } else {
        owner_snapshot(a, a.snap1, x, z, s_vx[lv], s_vz[lv], txx, tzz, txz);
      }
    }
    if (interior(a, x, z)) {
      stress_point(s_vx, s_vz, lv, 1, WV, lmD[k], muD[k], taper_at(a, x, z),
                   a.dts, dvx, dvz, txx, tzz, txz);
      s_txx[ls] = txx;
      s_tzz[ls] = tzz;
      s_txz[ls] = txz;
    }
    if (own) {
      long long g = pidx(a, x, z);
      a.txx_o[g] = txx;
      a.tzz_o[g] = tzz;
      a.txz_o[g] = txz;
    }
  }
  __syncthreads();

  // phase E: velocities of step s1 on the tile, step s1's injection terms.
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    int i = tid + k * NT;
    if (i >= TZ * TX) break;
    int r = i / TX, c = i - r * TX;
    int x = x0 + c, z = z0 + r;
    if (!in_grid(x, z)) continue;
    int lv = (r + 8) * WV + (c + 8);
    double vx = s_vx[lv], vz = s_vz[lv];
    if (a.do_snap1 && !a.adjoint) {
      long long ci = (long long)z * a.nx + x;
      a.snap1[0][ci] = vx;
      a.snap1[1][ci] = vz;
    }
    double b = bE[k];
    if (interior(a, x, z)) {
      velocity_point(s_txx, s_tzz, s_txz, (r + 6) * WS + (c + 6), 1, WS, b,
                     taper_at(a, x, z), a.dts, dvx, dvz, vx, vz);
    }
    apply_terms(a, a.tile_off, a.entries, tile, r * TX + c, a.it1, b, vx, vz);
    long long g = pidx(a, x, z);
    a.vx_o[g] = vx;
    a.vz_o[g] = vz;
  }
}

template <int TZ, int NT>
__global__ void __launch_bounds__(NT) step_two(StepArgs a) {
  const int tile = block_tile(a);
  if (tile >= a.ntiles) return;
  two_step_tile<TZ, NT>(a, tile);
}

__device__ inline int batch_tile(const BatchLaunch &b) {
  int blk = blockIdx.x;
  if (!b.xcd_remap) return blk;
  int chunk = (b.ntiles + 7) / 8;
  return (blk % 8) * chunk + blk / 8;
}

// The snapshot slot of step b.it in a member's one-shot store.
__device__ inline void batch_snapshot_slot(StepArgs &a, const BatchLaunch &b) {
  a.do_snap = b.it % b.si == 0;
  const long long off = (long long)(b.it / b.si) * a.nx * a.nz;
#pragma unroll
  for (int f = 0; f < 5; ++f) a.snap[f] += off;
}

template <int TZ, int NT>
__global__ void __launch_bounds__(NT) batch_fused(BatchLaunch b) {
  const int g = batch_tile(b);
  if (g >= b.ntiles) return;
  const int2 mt = b.tiles[g];
  StepArgs a = b.descs[((long long)mt.x * b.n_shots + b.shot) * 2 + b.parity];
  a.it = b.it;
  if (b.snapshots) batch_snapshot_slot(a, b);
  fused_tile<TZ, false, NT>(a, mt.y);
}

template <int TZ, int NT>
__global__ void __launch_bounds__(NT) batch_two(BatchLaunch b) {
  const int g = batch_tile(b);
  if (g >= b.ntiles) return;
  const int2 mt = b.tiles[g];
  StepArgs a = b.descs[((long long)mt.x * b.n_shots + b.shot) * 2 + b.parity];
  a.it = b.it;
  a.it1 = b.it1;
  two_step_tile<TZ, NT>(a, mt.y);
}

template <int TZ, int WPE, bool CORR = false, bool HALO_DMA = false, bool SEP = false,
          bool RDMA = false, bool CLAMP = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE)))
batch_lean(BatchLaunch b) {
  const int g = batch_tile(b);
  if (g >= b.ntiles) return;
  const int all = b.shot < 0;
  const int2 mt = b.tiles[all ? g / b.n_shots : g];
  const int shot = all ? g % b.n_shots : b.shot;
  StepArgs a = b.descs[((long long)mt.x * b.n_shots + shot) * 2 + b.parity];
  if (b.snapshots) batch_snapshot_slot(a, b);
  lean_tile<TZ, CORR, HALO_DMA, SEP, RDMA, CLAMP>(a, mt.y, b.it);
}

// Persistent batched forward: one cooperative launch advances steps
// [k0, k1) of one shot for every member.  Each block walks its share of the
// (member, tile) pairs -- the blocks of one XCD a contiguous range, as
// block_tile maps them -- and the grid meets at a barrier between steps,
// since a step reads what every block wrote in the previous one.  The grid
// holds no more blocks than fit on the device at once (the cooperative
// launch checks it), so every block reaches every barrier.
// One tile of one member at step k.
template <int TZ>
__device__ __forceinline__ void persistent_tile(const BatchLaunch &b, int g, int k) {
  const int2 mt = b.tiles[g];
  const StepArgs a = b.descs[((long long)mt.x * b.n_shots + b.shot) * 2 + (k & 1)];
  lean_tile<TZ, false>(a, mt.y, k);
}

template <int TZ, int WPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE)))
batch_lean_persistent(BatchLaunch b, int k0, int k1) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int per_xcd = gridDim.x / 8;  // the grid is a multiple of 8 blocks
  const int chunk = (b.ntiles + 7) / 8;
  const int t0 = (blockIdx.x % 8) * chunk;
  const int t1 = min(b.ntiles, t0 + chunk);
  for (int k = k0; k < k1; ++k) {
    for (int g = t0 + (int)blockIdx.x / 8; g < t1; g += per_xcd) {
      persistent_tile<TZ>(b, g, k);
      __syncthreads();  // the next tile reuses the LDS stage
    }
    grid.sync();
  }
}

// Running sums of receiver increments (src/fdModel.cpp:582-595 order): one
// thread per trace.
__global__ void receiver_scan_kernel(double *rx, double *rz, int nr, int nt) {
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * nr) return;
  double *p = (t < nr ? rx : rz) + (long long)(t % nr) * nt;
  double acc = p[0];
  for (int it = 1; it < nt; ++it) {
    acc = acc + p[it];
    p[it] = acc;
  }
}

// Split variant, launch 1 of 2: start-of-step owner work plus the in-place
// stress update.  Velocities are read-only here.
template <int TZ, bool RECIP>
__global__ void __launch_bounds__(kBlock) step_split_stress(StepArgs a) {
  const Divs<RECIP> dv(a);
  const int tile = block_tile(a);
  if (tile >= a.ntiles) return;
  const int x0 = (tile % a.ntx) * kTileX, z0 = (tile / a.ntx) * TZ;
  if (!a.adjoint && tile == 0) record_receivers(a, a.it);
  for (int i = threadIdx.x; i < TZ * kTileX; i += kBlock) {
    int r = i / kTileX, c = i - r * kTileX;
    int x = x0 + c, z = z0 + r;
    if (x >= a.nx || z >= a.nz) continue;
    long long g = pidx(a, x, z);
    double txx = a.txx_i[g], tzz = a.tzz_i[g], txz = a.txz_i[g];
    if (a.do_snap)
      owner_snapshot(a, a.snap, x, z, a.vx_i[g], a.vz_i[g], txx, tzz, txz);
    if (interior(a, x, z)) {
      stress_point(a.vx_i, a.vz_i, g, 1, a.pitch, a.lm[g], a.mu[g],
                   taper_at(a, x, z), a.dts, dv.x, dv.z, txx, tzz, txz);
      a.txx_o[g] = txx;
      a.tzz_o[g] = tzz;
      a.txz_o[g] = txz;
    }
  }
}

// Split variant, launch 2 of 2: in-place velocity update and injection.
template <int TZ, bool RECIP>
__global__ void __launch_bounds__(kBlock) step_split_velocity(StepArgs a) {
  const Divs<RECIP> dv(a);
  const int tile = block_tile(a);
  if (tile >= a.ntiles) return;
  const int x0 = (tile % a.ntx) * kTileX, z0 = (tile / a.ntx) * TZ;
  for (int i = threadIdx.x; i < TZ * kTileX; i += kBlock) {
    int r = i / kTileX, c = i - r * kTileX;
    int x = x0 + c, z = z0 + r;
    if (x >= a.nx || z >= a.nz) continue;
    long long g = pidx(a, x, z);
    double vx = a.vx_i[g], vz = a.vz_i[g], b = a.b[g];
    if (interior(a, x, z)) {
      velocity_point(a.txx_o, a.tzz_o, a.txz_o, g, 1, a.pitch, b,
                     taper_at(a, x, z), a.dts, dv.x, dv.z, vx, vz);
    }
    apply_injection(a, tile, r * kTileX + c, b, vx, vz);
    a.vx_o[g] = vx;
    a.vz_o[g] = vz;
  }
}

// Zeroes n arrays at once: blockIdx.y picks the array, the blocks along x
// stride over it.
__global__ void zero_many_kernel(double *const *ptrs, const long long *lens) {
  double *const p = ptrs[blockIdx.y];
  const long long n = lens[blockIdx.y];
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = 0.0;
}

__global__ void fill_kernel(double *p, long long n, double v) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

// src/fdModel.cpp:1034-1051, compact parameters -> padded materials.
__global__ void update_from_velocity_kernel(const double *vp, const double *vs,
                                            const double *rho, double *lm,
                                            double *mu, double *b, int nx,
                                            int nz, int pitch,
                                            long long origin) {
  long long n = (long long)nx * nz;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int z = (int)(i / nx), x = (int)(i - (long long)z * nx);
  long long g = origin + (long long)z * pitch + x;
  double r = rho[i];
  mu[g] = (vs[i] * vs[i]) * r;
  lm[g] = (vp[i] * vp[i]) * r;
  b[g] = 1.0 / r;
}

// src/fdModel.cpp:1147-1162
__global__ void adjoint_sources_kernel(const double *sx, const double *sz,
                                       const double *ox, const double *oz,
                                       double *ax, double *az, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  ax[i] = sx[i] - ox[i];
  az[i] = sz[i] - oz[i];
}

// src/fdModel.cpp:1164-1183 (b_vx = 1/rho recomputed bitwise)
__global__ void map_kernels_kernel(const double *vp, const double *vs,
                                   const double *rho, const double *lam_k,
                                   const double *mu_k, const double *rho_l,
                                   double *vp_k, double *vs_k, double *rho_v,
                                   long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double p = vp[i], s = vs[i], lk = lam_k[i], mk = mu_k[i];
  double b = 1.0 / rho[i];
  vp_k[i] = 2 * p * lk / b;
  vs_k[i] = (2 * s * mk - 4 * s * lk) / b;
  rho_v[i] = rho_l[i] + (p * p - 2 * s * s) * lk + s * s * mk;
}

// Compensated (TwoSum) accumulation: the misfit is the sum of millions of
// positive terms, and a fixed-order compensated tree keeps the result within
// a few ulp of the exact sum, independent of launch shape.
struct Acc {
  double s, c;
};

__device__ inline void acc_add(Acc &a, double x) {
  double t = a.s + x;
  double bb = t - a.s;
  a.c += (a.s - (t - bb)) + (x - bb);
  a.s = t;
}

__device__ inline Acc acc_merge(Acc a, Acc b) {
  Acc r{a.s, a.c + b.c};
  acc_add(r, b.s);
  return r;
}

__device__ inline Acc block_reduce(Acc v) {
  __shared__ double ss[kBlock], sc[kBlock];
  ss[threadIdx.x] = v.s;
  sc[threadIdx.x] = v.c;
  __syncthreads();
  for (int h = kBlock / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      Acc o = acc_merge(Acc{ss[threadIdx.x], sc[threadIdx.x]},
                        Acc{ss[threadIdx.x + h], sc[threadIdx.x + h]});
      ss[threadIdx.x] = o.s;
      sc[threadIdx.x] = o.c;
    }
    __syncthreads();
  }
  return Acc{ss[0], sc[0]};
}

// Terms of src/fdModel.cpp:1140-1141: (0.5 * dt) * (obs - syn)^2.
__global__ void __launch_bounds__(kBlock)
    misfit_partial_kernel(const double *sx, const double *sz, const double *ox,
                          const double *oz, long long n, double half_dt,
                          double *partials) {
  Acc a{0.0, 0.0};
  long long stride = (long long)gridDim.x * kBlock;
  for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += stride) {
    double d1 = ox[i] - sx[i];
    double d2 = oz[i] - sz[i];
    acc_add(a, half_dt * (d1 * d1));
    acc_add(a, half_dt * (d2 * d2));
  }
  Acc r = block_reduce(a);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = r.s;
    partials[2 * blockIdx.x + 1] = r.c;
  }
}

__global__ void __launch_bounds__(kBlock)
    misfit_final_kernel(const double *partials, int nparts, double *out) {
  Acc a{0.0, 0.0};
  for (int i = threadIdx.x; i < nparts; i += kBlock)
    a = acc_merge(a, Acc{partials[2 * i], partials[2 * i + 1]});
  Acc r = block_reduce(a);
  if (threadIdx.x == 0) out[0] = r.s + r.c;
}

// Model vector <-> parameter fields (src/fdModel.cpp:1358-1457).  One thread
// per parameter cell; the parameter index is x-fastest.
__global__ void model_vector_kernel(int mode, double *vp, double *vs,
                                    double *rho, double *vec, int nx, int nxf,
                                    int nzf, int npx, int npz, int gx0,
                                    int gz0) {
  int ncx = nxf / npx, ncz = nzf / npz;
  int ip = blockIdx.x * blockDim.x + threadIdx.x;
  int npp = ncx * ncz;
  if (ip >= npp) return;
  int iz = ip / ncx, ix = ip - iz * ncx;
  int gix = ix * npx + gx0, giz = iz * npz + gz0;
  if (mode == 0) {  // get: cell average
    double a = 0, b = 0, c = 0;
    for (int sx = 0; sx < npx; ++sx)
      for (int sz = 0; sz < npz; ++sz) {
        long long i = (long long)(giz + sz) * nx + (gix + sx);
        a += vp[i] / (npx * npz);
        b += vs[i] / (npx * npz);
        c += rho[i] / (npx * npz);
      }
    vec[ip] = a;
    vec[ip + npp] = b;
    vec[ip + 2 * npp] = c;
  } else {  // set: fill the cell
    for (int sx = 0; sx < npx; ++sx)
      for (int sz = 0; sz < npz; ++sz) {
        long long i = (long long)(giz + sz) * nx + (gix + sx);
        vp[i] = vec[ip];
        vs[i] = vec[ip + npp];
        rho[i] = vec[ip + 2 * npp];
      }
  }
}

// src/fdModel.cpp:1459-1515: cell sum of the velocity kernels.
__global__ void gradient_vector_kernel(const double *vp_k, const double *vs_k,
                                       const double *rho_k, double *vec,
                                       int nx, int nxf, int nzf, int npx,
                                       int npz, int gx0, int gz0) {
  int ncx = nxf / npx, ncz = nzf / npz;
  int ip = blockIdx.x * blockDim.x + threadIdx.x;
  int npp = ncx * ncz;
  if (ip >= npp) return;
  int iz = ip / ncx, ix = ip - iz * ncx;
  int gix = ix * npx + gx0, giz = iz * npz + gz0;
  double a = 0, b = 0, c = 0;
  for (int sx = 0; sx < npx; ++sx)
    for (int sz = 0; sz < npz; ++sz) {
      long long i = (long long)(giz + sz) * nx + (gix + sx);
      a += vp_k[i];
      b += vs_k[i];
      c += rho_k[i];
    }
  vec[ip] = a;
  vec[ip + npp] = b;
  vec[ip + 2 * npp] = c;
}

// 32x32 LDS tile transpose between [slice][iz][ix] and [slice][ix][iz].
__global__ void transpose_kernel(const double *in, double *out, int nx, int nz,
                                 int to_ref) {
  __shared__ double t[32][33];
  long long slice = (long long)blockIdx.z * nx * nz;
  in += slice;
  out += slice;
  // rows/cols of the input matrix
  int rows = to_ref ? nz : nx, cols = to_ref ? nx : nz;
  int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int r = r0 + k, c = c0 + threadIdx.x;
    if (r < rows && c < cols) t[k][threadIdx.x] = in[(long long)r * cols + c];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    int c = c0 + k, r = r0 + threadIdx.x;
    if (r < rows && c < cols) out[(long long)c * rows + r] = t[threadIdx.x][k];
  }
}

__global__ void unpad_kernel(const double *in, double *out, int nx, int nz,
                             int pitch, long long origin) {
  long long n = (long long)nx * nz;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int z = (int)(i / nx), x = (int)(i - (long long)z * nx);
  out[i] = in[origin + (long long)z * pitch + x];
}

// Plain streaming copy, the reference point for byte counters and for the
// sustainable bandwidth of a read+write stream: each thread keeps four
// independent 16-byte loads in flight.
__global__ void __launch_bounds__(256) copy_kernel(const double2 *__restrict__ in,
                                                   double2 *__restrict__ out,
                                                   long long n2) {
  long long base = (long long)blockIdx.x * 1024 + threadIdx.x;
  double2 v[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k * 256 < n2) v[k] = in[base + k * 256];
#pragma unroll
  for (int k = 0; k < 4; ++k)
    if (base + k * 256 < n2) out[base + k * 256] = v[k];
}

__global__ void pad_kernel(const double *in, double *out, int nx, int nz,
                           int pitch, long long origin) {
  long long n = (long long)nx * nz;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int z = (int)(i / nx), x = (int)(i - (long long)z * nx);
  out[origin + (long long)z * pitch + x] = in[i];
}

inline unsigned blocks_for(long long n, int bs) {
  return (unsigned)((n + bs - 1) / bs);
}

}  // namespace

int step_tile_height(int variant) {
  switch (variant) {
    case kFusedTz16:
    case kFusedTz16Recip:
    case kTwoStepTz16:
    case kLeanTz16:
    case kLeanTz16Ring:
    case kLeanTz16RingClamp:
      return 16;
    case kLeanTz12:
      return 12;
    case kLeanTz20:
      return 20;
    case kLeanTz24:
      return 24;
    case kLeanTz32:
      return 32;
    case kLeanTz4:
    case kLeanTz4Sep:
      return 4;
    case kLeanTz12W5:
      return 12;
    default:
      return 8;
  }
}
bool step_is_lean(int variant) {
  return (variant >= kLeanTz16 && variant <= kLeanTz32) ||
         (variant >= kLeanTz4 && variant <= kLeanTz16RingClamp);
}
bool fast_division_ok(double d) {
  return d >= 0x1p-40 && d <= 0x1p40;  // false for NaN
}
void fast_division_constants(double d, double *zh, double *zl) {
  double y = 1.0 / d;
  // 1 - y d is exact (y is within an ulp of 1/d); negative means y > 1/d
  if (std::fma(-y, d, 1.0) < 0.0) y = std::nextafter(y, 0.0);
  *zh = y;
  *zl = std::fma(-y, d, 1.0) / d;  // RN of the exact (1 - zh d) / d, >= 0
}
bool step_is_two_step(int variant) {
  return variant == kTwoStepTz8 || variant == kTwoStepTz16 || variant == kLeanTwoTz8;
}
int step_single_variant(int variant) {
  if (variant == kTwoStepTz8) return kFusedTz8;
  if (variant == kTwoStepTz16) return kFusedTz16;
  if (variant == kLeanTwoTz8) return kLeanTz8;
  return variant;
}
bool step_uses_fastdiv(int variant) {
  return step_is_lean(variant) || variant == kLeanTwoTz8 || step_is_lean_pipe(variant);
}
bool batch_gradient_variant_ok(int variant) {
  return variant == kFusedTz8 || variant == kFusedTz16 || step_is_lean(variant);
}
bool step_is_lean_pipe(int variant) {
  return variant >= kLeanPipe2Tz8 && variant <= kLeanPipe4Tz8;
}

hipError_t launch_two_steps(int variant, const StepArgs &a, hipStream_t s,
                            hipEvent_t e0, hipEvent_t e1) {
  int nblocks = a.xcd_remap ? 8 * ((a.ntiles + 7) / 8) : a.ntiles;
  dim3 g(nblocks);
  if (variant == kTwoStepTz8)
    hipExtLaunchKernelGGL((step_two<8, 256>), g, dim3(256), 0, s, e0, e1, 0, a);
  else if (variant == kTwoStepTz16)
    hipExtLaunchKernelGGL((step_two<16, 512>), g, dim3(512), 0, s, e0, e1, 0, a);
  else if (variant == kLeanTwoTz8)
    hipExtLaunchKernelGGL((step_lean2<8>), g, dim3(256), 0, s, e0, e1, 0, a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_batch_step(int variant, bool two, const BatchLaunch &b,
                             hipStream_t s, hipEvent_t e0, hipEvent_t e1, bool corr) {
  int nblocks = b.xcd_remap ? 8 * ((b.ntiles + 7) / 8) : b.ntiles;
  dim3 g(nblocks), t(256);
  const bool tall = step_tile_height(variant) == 16;
  // one-step lean kernels; adjoint snapshot steps use the correlating
  // instantiation (one wave per SIMD fewer, as for single models)
#define PSV_BATCH_LEAN(TZ, W, WC)                                                  \
  if (corr)                                                                        \
    hipExtLaunchKernelGGL((batch_lean<TZ, WC, true>), g, t, 0, s, e0, e1, 0, b);   \
  else                                                                             \
    hipExtLaunchKernelGGL((batch_lean<TZ, W, false>), g, t, 0, s, e0, e1, 0, b);
  if (step_is_lean(variant)) {
    switch (variant) {
      case kLeanTz16: PSV_BATCH_LEAN(16, 3, 2) break;
      case kLeanTz8: PSV_BATCH_LEAN(8, 5, 4) break;
      case kLeanTz12: PSV_BATCH_LEAN(12, 4, 3) break;
      case kLeanTz20: PSV_BATCH_LEAN(20, 3, 2) break;
      case kLeanTz24: PSV_BATCH_LEAN(24, 2, 1) break;
      case kLeanTz32: PSV_BATCH_LEAN(32, 2, 1) break;
      case kLeanTz4: PSV_BATCH_LEAN(4, 7, 5) break;
      case kLeanTz8W6: PSV_BATCH_LEAN(8, 6, 4) break;
      case kLeanTz12W5: PSV_BATCH_LEAN(12, 5, 3) break;
      case kLeanTz8Halo:
        if (corr)
          hipExtLaunchKernelGGL((batch_lean<8, 4, true, true>), g, t, 0, s, e0, e1, 0, b);
        else
          hipExtLaunchKernelGGL((batch_lean<8, 6, false, true>), g, t, 0, s, e0, e1, 0, b);
        break;
      case kLeanTz8Sep:
        if (corr)
          hipExtLaunchKernelGGL((batch_lean<8, 4, true, false, true>), g, t, 0, s, e0, e1, 0,
                                b);
        else
          hipExtLaunchKernelGGL((batch_lean<8, 5, false, false, true>), g, t, 0, s, e0, e1, 0,
                                b);
        break;
      case kLeanTz4Sep:
        if (corr)
          hipExtLaunchKernelGGL((batch_lean<4, 5, true, false, true>), g, t, 0, s, e0, e1, 0,
                                b);
        else
          hipExtLaunchKernelGGL((batch_lean<4, 7, false, false, true>), g, t, 0, s, e0, e1, 0,
                                b);
        break;
      case kLeanTz8Ring:
        if (corr)
          hipExtLaunchKernelGGL((batch_lean<8, 5, true, true, false, true>), g, t, 0, s, e0,
                                e1, 0, b);
        else
          hipExtLaunchKernelGGL((batch_lean<8, 6, false, true, false, true>), g, t, 0, s, e0,
                                e1, 0, b);
        break;
      case kLeanTz8RingOnly:
        if (corr)
          hipExtLaunchKernelGGL((batch_lean<8, 5, true, false, false, true>), g, t, 0, s, e0,
                                e1, 0, b);
        else
          hipExtLaunchKernelGGL((batch_lean<8, 6, false, false, false, true>), g, t, 0, s, e0,
                                e1, 0, b);
        break;
      case kLeanTz16Ring:
        if (corr)
          hipExtLaunchKernelGGL((batch_lean<16, 3, true, false, false, true>), g, t, 0, s, e0,
                                e1, 0, b);
        else
          hipExtLaunchKernelGGL((batch_lean<16, 4, false, false, false, true>), g, t, 0, s, e0,
                                e1, 0, b);
        break;
      case kLeanTz8RingClamp:
        if (corr)
          hipExtLaunchKernelGGL((batch_lean<8, 5, true, true, false, true, true>), g, t, 0, s,
                                e0, e1, 0, b);
        else
          hipExtLaunchKernelGGL((batch_lean<8, 6, false, true, false, true, true>), g, t, 0, s,
                                e0, e1, 0, b);
        break;
      case kLeanTz16RingClamp:
        if (corr)
          hipExtLaunchKernelGGL((batch_lean<16, 3, true, false, false, true, true>), g, t, 0, s,
                                e0, e1, 0, b);
        else
          hipExtLaunchKernelGGL((batch_lean<16, 4, false, false, false, true, true>), g, t, 0,
                                s, e0, e1, 0, b);
        break;
      default:
        return hipErrorInvalidValue;
    }
  } else if (two) {
    if (tall)
      hipExtLaunchKernelGGL((batch_two<16, 512>), g, dim3(512), 0, s, e0, e1, 0, b);
    else
      hipExtLaunchKernelGGL((batch_two<8, 256>), g, t, 0, s, e0, e1, 0, b);
  } else {
    if (tall)
      hipExtLaunchKernelGGL((batch_fused<16, 512>), g, dim3(512), 0, s, e0, e1, 0, b);
    else
      hipExtLaunchKernelGGL((batch_fused<8, 256>), g, t, 0, s, e0, e1, 0, b);
  }
#undef PSV_BATCH_LEAN
  return hipGetLastError();
}

hipError_t launch_batch_persistent(const BatchLaunch &b, int k0, int k1, hipStream_t s) {
  auto kernel = batch_lean_persistent<8, 3>;
  int per_cu = 0, cus = 0, dev = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, 0);
  if (e == hipSuccess) e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  int grid = per_cu * cus;
  grid -= grid % 8;
  grid = std::min(grid, 8 * ((b.ntiles + 7) / 8));
  if (grid < 8) return hipErrorInvalidConfiguration;
  BatchLaunch bl = b;
  void *args[] = {(void *)&bl, (void *)&k0, (void *)&k1};
  return hipLaunchCooperativeKernel((const void *)kernel, dim3(grid), dim3(256), args, 0, s);
}

hipError_t launch_receiver_scan(double *rx, double *rz, int nr, int nt,
                                hipStream_t s) {
  if (nr <= 0 || nt <= 1) return hipSuccess;
  hipLaunchKernelGGL(receiver_scan_kernel, dim3((2 * nr + 63) / 64), dim3(64),
                     0, s, rx, rz, nr, nt);
  return hipGetLastError();
}
bool step_is_inplace(int variant) {
  return variant == kSplitTz8 || variant == kSplitTz8Recip;
}

// With e0/e1 non-null the launch records the kernel's own start and end
// timestamps into them (hipExtLaunchKernel), i.e. the kernel duration without
// the queueing around it.
hipError_t launch_step(int variant, const StepArgs &a, hipStream_t s,
                       hipEvent_t e0, hipEvent_t e1) {
  int nblocks = a.xcd_remap ? 8 * ((a.ntiles + 7) / 8) : a.ntiles;
  const bool corr = a.do_snap && a.adjoint;
  dim3 g(nblocks), b(kBlock), b512(512);
  switch (variant) {
    case kFusedTz8:
      hipExtLaunchKernelGGL((step_fused<8, false>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kFusedTz16:
      hipExtLaunchKernelGGL((step_fused<16, false, 512>), g, b512, 0, s, e0, e1, 0, a);
      break;
    case kSplitTz8:
      hipExtLaunchKernelGGL((step_split_stress<8, false>), g, b, 0, s, e0, nullptr, 0, a);
      hipExtLaunchKernelGGL((step_split_velocity<8, false>), g, b, 0, s, nullptr, e1, 0, a);
      break;
    case kFusedTz8Recip:
      hipExtLaunchKernelGGL((step_fused<8, true>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kFusedTz16Recip:
      hipExtLaunchKernelGGL((step_fused<16, true, 512>), g, b512, 0, s, e0, e1, 0, a);
      break;
    case kSplitTz8Recip:
      hipExtLaunchKernelGGL((step_split_stress<8, true>), g, b, 0, s, e0, nullptr, 0, a);
      hipExtLaunchKernelGGL((step_split_velocity<8, true>), g, b, 0, s, nullptr, e1, 0, a);
      break;
    case kLeanTz16:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<16, 2, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<16, 3, false>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz8:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<8, 4, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<8, 5, false>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz12:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<12, 3, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<12, 4, false>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz20:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<20, 2, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<20, 3, false>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz24:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<24, 1, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<24, 2, false>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz32:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<32, 1, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<32, 2, false>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz4:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<4, 5, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<4, 7, false>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz8W6:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<8, 4, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<8, 6, false>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz12W5:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<12, 3, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<12, 5, false>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz8Halo:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<8, 4, true, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<8, 6, false, true>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz8Sep:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<8, 4, true, false, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<8, 5, false, false, true>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz4Sep:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<4, 5, true, false, true>), g, b, 0, s, e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<4, 7, false, false, true>), g, b, 0, s, e0, e1, 0, a);
      break;
    case kLeanTz8Ring:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<8, 5, true, true, false, true>), g, b, 0, s, e0, e1,
                              0, a);
      else
        hipExtLaunchKernelGGL((step_lean<8, 6, false, true, false, true>), g, b, 0, s, e0, e1,
                              0, a);
      break;
    case kLeanTz8RingOnly:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<8, 5, true, false, false, true>), g, b, 0, s, e0, e1,
                              0, a);
      else
        hipExtLaunchKernelGGL((step_lean<8, 6, false, false, false, true>), g, b, 0, s, e0, e1,
                              0, a);
      break;
    case kLeanTz16Ring:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<16, 3, true, false, false, true>), g, b, 0, s, e0, e1,
                              0, a);
      else
        hipExtLaunchKernelGGL((step_lean<16, 4, false, false, false, true>), g, b, 0, s, e0, e1,
                              0, a);
      break;
    case kLeanTz8RingClamp:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<8, 5, true, true, false, true, true>), g, b, 0, s, e0,
                              e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<8, 6, false, true, false, true, true>), g, b, 0, s, e0,
                              e1, 0, a);
      break;
    case kLeanTz16RingClamp:
      if (corr)
        hipExtLaunchKernelGGL((step_lean<16, 3, true, false, false, true, true>), g, b, 0, s,
                              e0, e1, 0, a);
      else
        hipExtLaunchKernelGGL((step_lean<16, 4, false, false, false, true, true>), g, b, 0, s,
                              e0, e1, 0, a);
      break;
    case kLeanPipe2Tz8:
    case kLeanPipe3Tz8:
    case kLeanPipe4Tz8: {
      const int P = variant == kLeanPipe2Tz8 ? 2 : variant == kLeanPipe3Tz8 ? 3 : 4;
      const int items = a.ntx * ((a.ntz + P - 1) / P);
      dim3 gp(a.xcd_remap ? 8 * ((items + 7) / 8) : items);
      if (P == 2) {
        if (corr)
          hipExtLaunchKernelGGL((step_lean_pipe<8, 2, 3, true>), gp, b, 0, s, e0, e1, 0, a);
        else
          hipExtLaunchKernelGGL((step_lean_pipe<8, 2, 4, false>), gp, b, 0, s, e0, e1, 0, a);
      } else if (P == 3) {
        if (corr)
          hipExtLaunchKernelGGL((step_lean_pipe<8, 3, 3, true>), gp, b, 0, s, e0, e1, 0, a);
        else
          hipExtLaunchKernelGGL((step_lean_pipe<8, 3, 4, false>), gp, b, 0, s, e0, e1, 0, a);
      } else {
        if (corr)
          hipExtLaunchKernelGGL((step_lean_pipe<8, 4, 3, true>), gp, b, 0, s, e0, e1, 0, a);
        else
          hipExtLaunchKernelGGL((step_lean_pipe<8, 4, 4, false>), gp, b, 0, s, e0, e1, 0, a);
      }
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Deferred kernel correlation: an adjoint whose kernel pointers pointed at
// a zeroed buffer per snapshot slot ([slot][rho, lam, mu][nx * nz]) left there
// exactly the terms it would have applied (0 - e = -e for the density kernel,
// 0 + e = e for the others).  Adding them in the adjoint's own order -- slots
// from last to first -- repeats its roundings: K - e == K + (-e).  Only the
// kernel region [kx0, kx1) x [kz0, kz1) is ever correlated.
__global__ void replay_correlation_kernel(double *rho_k, double *lam_k, double *mu_k,
                                          const double *__restrict__ d, int nslots,
                                          long long ngrid, int nx, int kx0, int kx1,
                                          int kz0, int kz1) {
  const long long w = kx1 - kx0;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w * (kz1 - kz0)) return;
  const long long ci = (long long)(kz0 + i / w) * nx + kx0 + i % w;
  double r = rho_k[ci], l = lam_k[ci], m = mu_k[ci];
  for (int s = nslots - 1; s >= 0; --s) {
    const double *ds = d + (long long)s * 3 * ngrid + ci;
    r = r + __builtin_nontemporal_load(ds);
    l = l + __builtin_nontemporal_load(ds + ngrid);
    m = m + __builtin_nontemporal_load(ds + 2 * ngrid);
  }
  rho_k[ci] = r;
  lam_k[ci] = l;
  mu_k[ci] = m;
}

hipError_t launch_replay_correlation(double *rho_k, double *lam_k, double *mu_k,
                                     const double *d, int nslots, long long ngrid, int nx,
                                     int kx0, int kx1, int kz0, int kz1, hipStream_t s) {
  const long long n = (long long)(kx1 - kx0) * (kz1 - kz0);
  if (n <= 0 || nslots <= 0) return hipSuccess;
  hipLaunchKernelGGL(replay_correlation_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                     rho_k, lam_k, mu_k, d, nslots, ngrid, nx, kx0, kx1, kz0, kz1);
  return hipGetLastError();
}

hipError_t launch_zero_many(double *const *ptrs, const long long *lens, int n, int blocks_x,
                            hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_many_kernel, dim3(blocks_x, n), dim3(256), 0, s, ptrs, lens);
  return hipGetLastError();
}

hipError_t launch_fill(double *p, long long n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  long long nb = (n + 255) / 256;
  unsigned grid = (unsigned)(nb < 65536 ? nb : 65536);
  hipLaunchKernelGGL(fill_kernel, dim3(grid), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_update_from_velocity(const double *vp, const double *vs,
                                       const double *rho, double *lm,
                                       double *mu, double *b, int nx, int nz,
                                       int pitch, long long origin,
                                       hipStream_t s) {
  long long n = (long long)nx * nz;
  hipLaunchKernelGGL(update_from_velocity_kernel, dim3(blocks_for(n, 256)),
                     dim3(256), 0, s, vp, vs, rho, lm, mu, b, nx, nz, pitch,
                     origin);
  return hipGetLastError();
}

hipError_t launch_adjoint_sources(const double *sx, const double *sz,
                                  const double *ox, const double *oz,
                                  double *ax, double *az, long long n,
                                  hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_sources_kernel, dim3(blocks_for(n, 256)),
                     dim3(256), 0, s, sx, sz, ox, oz, ax, az, n);
  return hipGetLastError();
}

hipError_t launch_map_kernels(const double *vp, const double *vs,
                              const double *rho, const double *lam_k,
                              const double *mu_k, const double *rho_l,
                              double *vp_k, double *vs_k, double *rho_v,
                              long long n, hipStream_t s) {
  hipLaunchKernelGGL(map_kernels_kernel, dim3(blocks_for(n, 256)), dim3(256),
                     0, s, vp, vs, rho, lam_k, mu_k, rho_l, vp_k, vs_k, rho_v,
                     n);
  return hipGetLastError();
}

hipError_t launch_misfit(const double *sx, const double *sz, const double *ox,
                         const double *oz, long long n, double half_dt,
                         double *partials, double *out, hipStream_t s) {
  hipLaunchKernelGGL(misfit_partial_kernel, dim3(kMisfitBlocks), dim3(kBlock),
                     0, s, sx, sz, ox, oz, n, half_dt, partials);
  hipLaunchKernelGGL(misfit_final_kernel, dim3(1), dim3(kBlock), 0, s,
                     partials, kMisfitBlocks, out);
  return hipGetLastError();
}

// Partial sums of the allocation-free misfit (launch_misfit_noalloc): the
// same kMisfitBlocks compensated pairs that launch_misfit keeps in a model's
// own buffer, here in a fixed array that every GPU running the library holds
// once until the process ends.  Each call overwrites every pair before the
// final kernel reads them, so nothing has to be cleared; calls that overlap
// on one device race on it.
static __device__ __align__(256) double scratch[2 * kMisfitBlocks];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (the host-side name is only a stand-in for the symbol).
hipError_t scratch_on_current_device(double **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = static_cast<double *>(p);
  }
  *out = g_scratch[dev];
  return hipSuccess;
}
}  // namespace

hipError_t launch_misfit_noalloc(const double *sx, const double *sz, const double *ox,
                                 const double *oz, long long n, double half_dt,
                                 double *out, hipStream_t s) {
  double *partials = nullptr;
  hipError_t e = scratch_on_current_device(&partials);
  if (e != hipSuccess) return e;
  return launch_misfit(sx, sz, ox, oz, n, half_dt, partials, out, s);
}

hipError_t launch_model_vector(int mode, double *vp, double *vs, double *rho,
                               double *vec, int nx, int nxf, int nzf, int npx,
                               int npz, int gx0, int gz0, hipStream_t s) {
  int npp = (nxf / npx) * (nzf / npz);
  hipLaunchKernelGGL(model_vector_kernel, dim3(blocks_for(npp, 256)),
                     dim3(256), 0, s, mode, vp, vs, rho, vec, nx, nxf, nzf,
                     npx, npz, gx0, gz0);
  return hipGetLastError();
}

hipError_t launch_gradient_vector(const double *vp_k, const double *vs_k,
                                  const double *rho_k, double *vec, int nx,
                                  int nxf, int nzf, int npx, int npz, int gx0,
                                  int gz0, hipStream_t s) {
  int npp = (nxf / npx) * (nzf / npz);
  hipLaunchKernelGGL(gradient_vector_kernel, dim3(blocks_for(npp, 256)),
                     dim3(256), 0, s, vp_k, vs_k, rho_k, vec, nx, nxf, nzf,
                     npx, npz, gx0, gz0);
  return hipGetLastError();
}

hipError_t launch_transpose(const double *in, double *out, int nx, int nz,
                            int slices, int to_ref, hipStream_t s) {
  int rows = to_ref ? nz : nx, cols = to_ref ? nx : nz;
  for (int s0 = 0; s0 < slices; s0 += 65535) {
    int cnt = slices - s0 < 65535 ? slices - s0 : 65535;
    long long off = (long long)s0 * nx * nz;
    dim3 grid((cols + 31) / 32, (rows + 31) / 32, cnt);
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(32, 8), 0, s, in + off,
                       out + off, nx, nz, to_ref);
  }
  return hipGetLastError();
}

hipError_t launch_copy(const double *in, double *out, long long n,
                       hipStream_t s) {
  long long n2 = n / 2;
  unsigned grid = (unsigned)((n2 + 1023) / 1024);
  hipLaunchKernelGGL(copy_kernel, dim3(grid), dim3(256), 0, s,
                     (const double2 *)in, (double2 *)out, n2);
  return hipGetLastError();
}

hipError_t launch_pad(const double *in, double *out, int nx, int nz,
                      int pitch, long long origin, hipStream_t s) {
  long long n = (long long)nx * nz;
  hipLaunchKernelGGL(pad_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                     in, out, nx, nz, pitch, origin);
  return hipGetLastError();
}

hipError_t launch_unpad(const double *in, double *out, int nx, int nz,
                        int pitch, long long origin, hipStream_t s) {
  long long n = (long long)nx * nz;
  hipLaunchKernelGGL(unpad_kernel, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                     in, out, nx, nz, pitch, origin);
  return hipGetLastError();
}

}  // namespace psv
