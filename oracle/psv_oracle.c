/*
 * psv_oracle.c -- CPU restatement of psvWave's fdModel hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Every function names the reference lines it restates.  Floating-point
 * expressions keep the reference's evaluation order (C left-to-right
 * association, no contraction: build with -ffp-contract=off), so results are
 * bitwise those of an IEEE-754 double build of the reference.  pow(x, 2) in
 * the reference is written x*x here: GCC folds pow(x, 2.0) to x*x, which is
 * the correctly rounded square either way.
 *
 * Loops over grid points are elementwise (each point reads only the previous
 * phase's arrays), so OpenMP threading does not change any result bit.
 */
#include "psv_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_PI 3.14159265 /* src/fdModel.cpp:14 (truncated on purpose) */

struct oracle_model {
  /* configuration (src/fdModel.h:263-302) */
  int nt, nx_inner, nz_inner, nxib, nzib;
  double dx, dz, dt;
  int np;
  double np_factor, scalar_rho, scalar_vp, scalar_vs;
  int npx, npz;
  double peak_frequency, t0, delay;
  int n_sources, n_shots, nr, si;
  int *ix_src, *iz_src, *ix_rec, *iz_rec;
  double *angles;
  int *which_off, *which_src;
  /* derived (src/fdModel.cpp:398-461) */
  int nx, nz, nxf, nzf, free_parameters, snapshots;
  double alpha;
  /* arrays (src/fdModel.h:209-252) */
  double *vx, *vz, *txx, *tzz, *txz;
  double *lm, *la, *mu, *b_vx, *b_vz, *rho, *vp, *vs;
  double *lambda_k, *mu_k, *rho_l_k, *vp_k, *vs_k, *rho_v_k;
  double *taper, *t, *stf, *moment;
  double *rtf_ux, *rtf_uz, *rtf_ux_true, *rtf_uz_true, *a_ux, *a_uz;
  double *accu[5]; /* vx, vz, txx, tzz, txz; allocated on first store */
};

static char g_err[512];
const char *oracle_last_error(void) { return g_err; }

void oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* linear_IDX, src/contiguous_arrays.cpp:43-74 (row-major, last index fastest) */
#define I2(ix, iz) ((long)(ix) * m->nz + (iz))

static double *zalloc(size_t n) {
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  return p;
}

static int *icopy(const int *src, int n) {
  int *p = (int *)malloc(sizeof(int) * (n ? n : 1));
  if (n) memcpy(p, src, sizeof(int) * n);
  return p;
}

/* src/fdModel.cpp:1034-1051 */
static void update_from_velocity(oracle_model *m) {
  long n = (long)m->nx * m->nz;
#pragma omp parallel for
  for (long i = 0; i < n; ++i) {
    m->mu[i] = (m->vs[i] * m->vs[i]) * m->rho[i];
    m->lm[i] = (m->vp[i] * m->vp[i]) * m->rho[i];
    m->la[i] = m->lm[i] - 2 * m->mu[i];
    m->b_vx[i] = 1.0 / m->rho[i];
    m->b_vz[i] = m->b_vx[i];
  }
}

/* src/fdModel.cpp:193-299 */
static void initialize_arrays(oracle_model *m) {
  for (int ir = 0; ir < m->nr; ++ir) { /* :197-204 */
    m->ix_rec[ir] += m->np;
    m->iz_rec[ir] += m->np;
  }
  for (int is = 0; is < m->n_sources; ++is) { /* :205-212 */
    m->ix_src[is] += m->np;
    m->iz_src[is] += m->np;
  }
  /* Ricker wavelet per source, delayed by its position in the shot list
   * (:214-231). */
  for (int i_shot = 0; i_shot < m->n_shots; ++i_shot) {
    int n_in_shot = m->which_off[i_shot + 1] - m->which_off[i_shot];
    for (int i_source = 0; i_source < n_in_shot; ++i_source) {
      int src = m->which_src[m->which_off[i_shot] + i_source];
      for (unsigned int it = 0; it < (unsigned int)m->nt; ++it) {
        m->t[it] = it * m->dt;
        double f = 1.0 / m->alpha;
        double shifted = m->t[it] - 1.4 / f - m->delay * i_source / f;
        double a = M_PI * f * shifted;
        m->stf[(long)src * m->nt + it] = (1 - 2 * (a * a)) * exp(-(a * a));
      }
    }
  }
  /* moment tensors from the dip angle (:233-246) */
  for (int s = 0; s < m->n_sources; ++s) {
    double arg = m->angles[s] * ORACLE_PI / 180.0;
    m->moment[s * 4 + 0] = cos(arg) * 1e15;
    m->moment[s * 4 + 1] = -sin(arg) * 1e15;
    m->moment[s * 4 + 2] = -sin(arg) * 1e15;
    m->moment[s * 4 + 3] = -cos(arg) * 1e15;
  }
  /* homogeneous background (:249-261) */
  long n = (long)m->nx * m->nz;
  for (long i = 0; i < n; ++i) {
    m->vp[i] = m->scalar_vp;
    m->vs[i] = m->scalar_vs;
    m->rho[i] = m->scalar_rho;
  }
  update_from_velocity(m);
  /* taper: count the nested rectangles a point lies in, then Gaussian
   * (:263-298) */
  for (long i = 0; i < n; ++i) m->taper[i] = 0.0;
  for (int id = 0; id < m->np; ++id)
    for (int ix = id; ix < m->nx - id; ++ix)
      for (int iz = id; iz < m->nz - id; ++iz) m->taper[I2(ix, iz)]++;
  for (long i = 0; i < n; ++i) {
    double x = m->np_factor * (m->np - m->taper[i]);
    m->taper[i] = exp(-(x * x));
  }
}

static void alloc_arrays(oracle_model *m) {
  long n = (long)m->nx * m->nz;
  double **grid[] = {&m->vx, &m->vz, &m->txx, &m->tzz, &m->txz, &m->lm,
                     &m->la, &m->mu, &m->b_vx, &m->b_vz, &m->rho, &m->vp,
                     &m->vs, &m->lambda_k, &m->mu_k, &m->rho_l_k, &m->vp_k,
                     &m->vs_k, &m->rho_v_k, &m->taper};
  for (size_t k = 0; k < sizeof(grid) / sizeof(grid[0]); ++k)
    *grid[k] = zalloc(n);
  m->t = zalloc(m->nt);
  m->stf = zalloc((long)m->n_sources * m->nt);
  m->moment = zalloc((long)m->n_sources * 4);
  long nrec = (long)m->n_shots * m->nr * m->nt;
  m->rtf_ux = zalloc(nrec);
  m->rtf_uz = zalloc(nrec);
  m->rtf_ux_true = zalloc(nrec);
  m->rtf_uz_true = zalloc(nrec);
  m->a_ux = zalloc(nrec);
  m->a_uz = zalloc(nrec);
  for (int k = 0; k < 5; ++k) m->accu[k] = NULL;
}

/* src/fdModel.cpp:390-468 (validation) and :29-62 (variable constructor) */
oracle_model *oracle_create(const oracle_params *p) {
  g_err[0] = 0;
  if (p->n_shots < 0 || p->n_sources < 0 || p->nr < 0 ||
      p->snapshot_interval <= 0 || p->npx <= 0 || p->npz <= 0) {
    snprintf(g_err, sizeof g_err, "invalid sizes");
    return NULL;
  }
  int total = p->which_offsets[p->n_shots];
  if (total != p->n_sources) {
    snprintf(g_err, sizeof g_err,
             "Mismatch between n_sources and "
             "sources.which_source_to_fire_in_which_shot.");
    return NULL;
  }
  oracle_model *m = (oracle_model *)calloc(1, sizeof(oracle_model));
  m->nt = p->nt;
  m->nx_inner = p->nx_inner;
  m->nz_inner = p->nz_inner;
  m->nxib = p->nx_inner_boundary;
  m->nzib = p->nz_inner_boundary;
  m->dx = p->dx;
  m->dz = p->dz;
  m->dt = p->dt;
  m->np = p->np_boundary;
  m->np_factor = p->np_factor;
  m->scalar_rho = p->scalar_rho;
  m->scalar_vp = p->scalar_vp;
  m->scalar_vs = p->scalar_vs;
  m->npx = p->npx;
  m->npz = p->npz;
  m->peak_frequency = p->peak_frequency;
  m->t0 = p->source_timeshift;
  m->delay = p->delay_cycles_per_shot;
  m->n_sources = p->n_sources;
  m->n_shots = p->n_shots;
  m->nr = p->nr;
  m->si = p->snapshot_interval;
  m->ix_src = icopy(p->ix_sources, p->n_sources);
  m->iz_src = icopy(p->iz_sources, p->n_sources);
  m->angles = (double *)malloc(sizeof(double) * (p->n_sources ? p->n_sources : 1));
  if (p->n_sources) memcpy(m->angles, p->moment_angles, sizeof(double) * p->n_sources);
  m->which_off = icopy(p->which_offsets, p->n_shots + 1);
  m->which_src = icopy(p->which_sources, total);
  m->ix_rec = icopy(p->ix_receivers, p->nr);
  m->iz_rec = icopy(p->iz_receivers, p->nr);

  m->nx = m->nx_inner + m->np * 2;
  m->nz = m->nz_inner + m->np * 2;
  m->nxf = m->nx_inner - m->nxib * 2;
  m->nzf = m->nz_inner - m->nzib * 2;
  m->free_parameters = 3 * m->nxf * m->nzf / (m->npx * m->npz);
  m->alpha = 1.0 / m->peak_frequency;
  m->snapshots = m->nt / m->si;

  alloc_arrays(m);
  initialize_arrays(m);
  return m;
}

/* src/fdModel.cpp:64-97 and :301-388 (deep copy) */
oracle_model *oracle_copy(const oracle_model *s) {
  oracle_model *m = (oracle_model *)malloc(sizeof(oracle_model));
  *m = *s;
  long n = (long)s->nx * s->nz;
  long nrec = (long)s->n_shots * s->nr * s->nt;
  m->ix_src = icopy(s->ix_src, s->n_sources);
  m->iz_src = icopy(s->iz_src, s->n_sources);
  m->ix_rec = icopy(s->ix_rec, s->nr);
  m->iz_rec = icopy(s->iz_rec, s->nr);
  m->angles = (double *)malloc(sizeof(double) * (s->n_sources ? s->n_sources : 1));
  if (s->n_sources) memcpy(m->angles, s->angles, sizeof(double) * s->n_sources);
  m->which_off = icopy(s->which_off, s->n_shots + 1);
  m->which_src = icopy(s->which_src, s->which_off[s->n_shots]);
#define DUP(f, cnt)                                   \
  do {                                                \
    m->f = zalloc(cnt);                               \
    memcpy(m->f, s->f, sizeof(double) * (size_t)(cnt)); \
  } while (0)
  DUP(vx, n); DUP(vz, n); DUP(txx, n); DUP(tzz, n); DUP(txz, n);
  DUP(lm, n); DUP(la, n); DUP(mu, n); DUP(b_vx, n); DUP(b_vz, n);
  DUP(rho, n); DUP(vp, n); DUP(vs, n); DUP(lambda_k, n); DUP(mu_k, n);
  DUP(rho_l_k, n); DUP(vp_k, n); DUP(vs_k, n); DUP(rho_v_k, n);
  DUP(taper, n); DUP(t, s->nt); DUP(stf, (long)s->n_sources * s->nt);
  DUP(moment, (long)s->n_sources * 4);
  DUP(rtf_ux, nrec); DUP(rtf_uz, nrec); DUP(rtf_ux_true, nrec);
  DUP(rtf_uz_true, nrec); DUP(a_ux, nrec); DUP(a_uz, nrec);
#undef DUP
  long na = (long)s->n_shots * s->snapshots * n;
  for (int k = 0; k < 5; ++k) {
    if (s->accu[k]) {
      m->accu[k] = zalloc(na);
      memcpy(m->accu[k], s->accu[k], sizeof(double) * (size_t)na);
    } else {
      m->accu[k] = NULL;
    }
  }
  return m;
}

void oracle_destroy(oracle_model *m) {
  if (!m) return;
  double *ptrs[] = {m->vx, m->vz, m->txx, m->tzz, m->txz, m->lm, m->la,
                    m->mu, m->b_vx, m->b_vz, m->rho, m->vp, m->vs,
                    m->lambda_k, m->mu_k, m->rho_l_k, m->vp_k, m->vs_k,
                    m->rho_v_k, m->taper, m->t, m->stf, m->moment,
                    m->rtf_ux, m->rtf_uz, m->rtf_ux_true, m->rtf_uz_true,
                    m->a_ux, m->a_uz, m->accu[0], m->accu[1], m->accu[2],
                    m->accu[3], m->accu[4], m->angles};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
  free(m->ix_src); free(m->iz_src); free(m->ix_rec); free(m->iz_rec);
  free(m->which_off); free(m->which_src);
  free(m);
}

void oracle_dims(const oracle_model *m, int *o) {
  o[0] = m->nx; o[1] = m->nz; o[2] = m->nt; o[3] = m->snapshots;
  o[4] = m->free_parameters; o[5] = m->n_shots; o[6] = m->nr;
  o[7] = m->n_sources;
}

void oracle_set_parameter_fields(oracle_model *m, const double *vp,
                                 const double *vs, const double *rho) {
  size_t n = (size_t)m->nx * m->nz * sizeof(double);
  memcpy(m->vp, vp, n);
  memcpy(m->vs, vs, n);
  memcpy(m->rho, rho, n);
  update_from_velocity(m);
}

void oracle_get_parameter_fields(const oracle_model *m, double *vp, double *vs,
                                 double *rho) {
  size_t n = (size_t)m->nx * m->nz * sizeof(double);
  memcpy(vp, m->vp, n);
  memcpy(vs, m->vs, n);
  memcpy(rho, m->rho, n);
}

/* Stress phase of one step, src/fdModel.cpp:599-644 (forward, +dt) and
 * :864-911 (adjoint, -dt). */
void oracle_stress_phase(oracle_model *m, int adjoint) {
  const int nx = m->nx, nz = m->nz;
  const double c1 = 9.0 / 8.0, c2 = 1.0 / 24.0;
  const double dt = m->dt, dx = m->dx, dz = m->dz;
  double *vx = m->vx, *vz = m->vz, *txx = m->txx, *tzz = m->tzz, *txz = m->txz;
#pragma omp parallel for
  for (int ix = 2; ix < nx - 2; ++ix) {
    for (int iz = 2; iz < nz - 2; ++iz) {
      long i = I2(ix, iz);
      long xp1 = I2(ix + 1, iz), xp2 = I2(ix + 2, iz);
      long xm1 = I2(ix - 1, iz), xm2 = I2(ix - 2, iz);
      long zp1 = I2(ix, iz + 1), zp2 = I2(ix, iz + 2);
      long zm1 = I2(ix, iz - 1), zm2 = I2(ix, iz - 2);
      double dvx_x = c1 * (vx[xp1] - vx[i]) + c2 * (vx[xm1] - vx[xp2]);
      double dvz_z = c1 * (vz[i] - vz[zm1]) + c2 * (vz[zm2] - vz[zp1]);
      double dvx_z = c1 * (vx[zp1] - vx[i]) + c2 * (vx[zm1] - vx[zp2]);
      double dvz_x = c1 * (vz[i] - vz[xm1]) + c2 * (vz[xm2] - vz[xp1]);
      double sxx = m->lm[i] * dvx_x / dx + m->la[i] * dvz_z / dz;
      double szz = m->la[i] * dvx_x / dx + (m->lm[i]) * dvz_z / dz;
      double sxz = dvx_z / dz + dvz_x / dx;
      if (!adjoint) {
        txx[i] = m->taper[i] * (txx[i] + dt * sxx);
        tzz[i] = m->taper[i] * (tzz[i] + dt * szz);
        txz[i] = m->taper[i] * (txz[i] + dt * m->mu[i] * sxz);
      } else {
        txx[i] = m->taper[i] * (txx[i] - dt * sxx);
        tzz[i] = m->taper[i] * (tzz[i] - dt * szz);
        txz[i] = m->taper[i] * (txz[i] - dt * m->mu[i] * sxz);
      }
    }
  }
}

/* Velocity phase of one step, src/fdModel.cpp:646-678 (forward) and
 * :913-948 (adjoint). */
void oracle_velocity_phase(oracle_model *m, int adjoint) {
  const int nx = m->nx, nz = m->nz;
  const double c1 = 9.0 / 8.0, c2 = 1.0 / 24.0;
  const double dt = m->dt, dx = m->dx, dz = m->dz;
  double *vx = m->vx, *vz = m->vz, *txx = m->txx, *tzz = m->tzz, *txz = m->txz;
#pragma omp parallel for
  for (int ix = 2; ix < nx - 2; ++ix) {
    for (int iz = 2; iz < nz - 2; ++iz) {
      long i = I2(ix, iz);
      long xp1 = I2(ix + 1, iz), xp2 = I2(ix + 2, iz);
      long xm1 = I2(ix - 1, iz), xm2 = I2(ix - 2, iz);
      long zp1 = I2(ix, iz + 1), zp2 = I2(ix, iz + 2);
      long zm1 = I2(ix, iz - 1), zm2 = I2(ix, iz - 2);
      double ax = (c1 * (txx[i] - txx[xm1]) + c2 * (txx[xm2] - txx[xp1])) / dx +
                  (c1 * (txz[i] - txz[zm1]) + c2 * (txz[zm2] - txz[zp1])) / dz;
      double az = (c1 * (txz[xp1] - txz[i]) + c2 * (txz[xm1] - txz[xp2])) / dx +
                  (c1 * (tzz[zp1] - tzz[i]) + c2 * (tzz[zm1] - tzz[zp2])) / dz;
      if (!adjoint) {
        vx[i] = m->taper[i] * (vx[i] + m->b_vx[i] * dt * ax);
        vz[i] = m->taper[i] * (vz[i] + m->b_vz[i] * dt * az);
      } else {
        vx[i] = m->taper[i] * (vx[i] - m->b_vx[i] * dt * ax);
        vz[i] = m->taper[i] * (vz[i] - m->b_vz[i] * dt * az);
      }
    }
  }
}

/* Serial moment-tensor injection, src/fdModel.cpp:681-751.  Every term uses
 * b_vz and dx^4 for x-couples, dz^4 for z-couples. */
static void inject_sources(oracle_model *m, int i_shot, int it) {
  const int nz = m->nz;
  const double dt = m->dt, dx = m->dx, dz = m->dz;
  double *vx = m->vx, *vz = m->vz, *b = m->b_vz;
  (void)nz;
  for (int k = m->which_off[i_shot]; k < m->which_off[i_shot + 1]; ++k) {
    int s = m->which_src[k];
    int x = m->ix_src[s], z = m->iz_src[s];
    double S = m->stf[(long)s * m->nt + it];
    const double *M = m->moment + s * 4;
    long i = I2(x, z), xm1 = I2(x - 1, z), zm1 = I2(x, z - 1);
    long xp1 = I2(x + 1, z), zp1 = I2(x, z + 1);
    long xp1zm1 = I2(x + 1, z - 1), xm1zp1 = I2(x - 1, z + 1);
    long xm1zm1 = I2(x - 1, z - 1);
    double Dx = dx * dx * dx * dx, Dz = dz * dz * dz * dz;
    vx[xm1] -= M[0] * S * dt * b[xm1] / Dx;
    vx[i] += M[0] * S * dt * b[i] / Dx;
    vz[zm1] -= M[3] * S * dt * b[zm1] / Dz;
    vz[i] += M[3] * S * dt * b[i] / Dz;
    vx[xm1zp1] += 0.25 * M[1] * S * dt * b[xm1zp1] / Dx;
    vx[zp1] += 0.25 * M[1] * S * dt * b[zp1] / Dx;
    vx[xm1zm1] -= 0.25 * M[1] * S * dt * b[xm1zm1] / Dx;
    vx[zm1] -= 0.25 * M[1] * S * dt * b[zm1] / Dx;
    vz[xp1zm1] += 0.25 * M[2] * S * dt * b[xp1zm1] / Dz;
    vz[xp1] += 0.25 * M[2] * S * dt * b[xp1] / Dz;
    vz[xm1zm1] -= 0.25 * M[2] * S * dt * b[xm1zm1] / Dz;
    vz[xm1] -= 0.25 * M[2] * S * dt * b[xm1] / Dz;
  }
}

static int ensure_accu(oracle_model *m) {
  long na = (long)m->n_shots * m->snapshots * m->nx * m->nz;
  for (int k = 0; k < 5; ++k) {
    if (!m->accu[k]) {
      m->accu[k] = zalloc(na);
      if (!m->accu[k]) {
        snprintf(g_err, sizeof g_err, "out of memory for snapshots");
        return -1;
      }
    }
  }
  return 0;
}

static int check_shot(oracle_model *m, int i_shot) {
  if (i_shot < 0 || i_shot >= m->n_shots) {
    snprintf(g_err, sizeof g_err, "i_shot out of range");
    return -1;
  }
  return 0;
}

/* src/fdModel.cpp:526-787 */
int oracle_forward(oracle_model *m, int i_shot, int store_fields) {
  if (check_shot(m, i_shot)) return -1;
  if (store_fields && m->nt % m->si != 0) {
    snprintf(g_err, sizeof g_err, "nt must be a multiple of snapshot_interval");
    return -1;
  }
  if (store_fields && ensure_accu(m)) return -1;
  const long n = (long)m->nx * m->nz;
  memset(m->vx, 0, n * sizeof(double)); /* :530-542 */
  memset(m->vz, 0, n * sizeof(double));
  memset(m->txx, 0, n * sizeof(double));
  memset(m->tzz, 0, n * sizeof(double));
  memset(m->txz, 0, n * sizeof(double));
  const double dt = m->dt, dx = m->dx, dz = m->dz;
  for (int it = 0; it < m->nt; ++it) {
    if (it % m->si == 0 && store_fields) { /* :555-572 */
      long off = ((long)i_shot * m->snapshots + it / m->si) * n;
      memcpy(m->accu[0] + off, m->vx, n * sizeof(double));
      memcpy(m->accu[1] + off, m->vz, n * sizeof(double));
      memcpy(m->accu[2] + off, m->txx, n * sizeof(double));
      memcpy(m->accu[3] + off, m->tzz, n * sizeof(double));
      memcpy(m->accu[4] + off, m->txz, n * sizeof(double));
    }
    for (int r = 0; r < m->nr; ++r) { /* :576-596 */
      long k = ((long)i_shot * m->nr + r) * m->nt + it;
      long loc = I2(m->ix_rec[r], m->iz_rec[r]);
      if (it == 0) {
        m->rtf_ux[k] = dt * m->vx[loc] / (dx * dz);
        m->rtf_uz[k] = dt * m->vz[loc] / (dx * dz);
      } else {
        m->rtf_ux[k] = m->rtf_ux[k - 1] + dt * m->vx[loc] / (dx * dz);
        m->rtf_uz[k] = m->rtf_uz[k - 1] + dt * m->vz[loc] / (dx * dz);
      }
    }
    oracle_stress_phase(m, 0);
    oracle_velocity_phase(m, 0);
    inject_sources(m, i_shot, it);
  }
  return 0;
}

/* Kernel correlation at a snapshot step, src/fdModel.cpp:817-861. */
static void correlate(oracle_model *m, int i_shot, int it) {
  const long n = (long)m->nx * m->nz;
  const long off = ((long)i_shot * m->snapshots + it / m->si) * n;
  const double *Avx = m->accu[0] + off, *Avz = m->accu[1] + off;
  const double *Atxx = m->accu[2] + off, *Atzz = m->accu[3] + off;
  const double *Atxz = m->accu[4] + off;
  const int si = m->si;
  const double dt = m->dt;
  const int x0 = m->np + m->nxib, x1 = m->np + m->nx_inner - m->nxib;
  const int z0 = m->np + m->nzib, z1 = m->np + m->nz_inner - m->nzib;
#pragma omp parallel for
  for (int ix = x0; ix < x1; ++ix) {
    for (int iz = z0; iz < z1; ++iz) {
      long i = I2(ix, iz);
      double la = m->la[i], lm = m->lm[i], mu = m->mu[i];
      double txx = m->txx[i], tzz = m->tzz[i], txz = m->txz[i];
      m->rho_l_k[i] -= si * dt * (Avx[i] * m->vx[i] + Avz[i] * m->vz[i]);
      double D = lm - ((la * la) / (lm));
      m->lambda_k[i] +=
          si * dt *
          (((Atxx[i] - (Atzz[i] * la) / lm) + (Atzz[i] - (Atxx[i] * la) / lm)) *
           ((txx - (tzz * la) / lm) + (tzz - (txx * la) / lm))) /
          (D * D);
      m->mu_k[i] +=
          si * dt * 2 *
          ((((txx - (tzz * la) / lm) * (Atxx[i] - (Atzz[i] * la) / lm)) +
            ((tzz - (txx * la) / lm) * (Atzz[i] - (Atxx[i] * la) / lm))) /
               (D * D) +
           2 * (txz * Atxz[i] / (4 * mu * mu)));
    }
  }
}

/* Test hooks: write one stored snapshot (allocating the store if needed)
 * and apply the correlation of time step `it` with the current state. */
int oracle_set_snapshot(oracle_model *m, int i_shot, int i_snap, const double *vx,
                        const double *vz, const double *txx, const double *tzz,
                        const double *txz) {
  if (i_shot < 0 || i_shot >= m->n_shots || i_snap < 0 || i_snap >= m->snapshots)
    return -1;
  const long n = (long)m->nx * m->nz;
  if (!m->accu[0])
    for (int k = 0; k < 5; ++k)
      m->accu[k] = calloc((size_t)m->n_shots * m->snapshots * n, sizeof(double));
  const long off = ((long)i_shot * m->snapshots + i_snap) * n;
  const double *ins[5] = {vx, vz, txx, tzz, txz};
  for (int k = 0; k < 5; ++k) memcpy(m->accu[k] + off, ins[k], n * sizeof(double));
  return 0;
}
int oracle_correlate_step(oracle_model *m, int i_shot, int it) {
  if (!m->accu[0] || i_shot < 0 || i_shot >= m->n_shots || it < 0 || it >= m->nt ||
      it % m->si != 0)
    return -1;
  correlate(m, i_shot, it);
  return 0;
}

/* src/fdModel.cpp:789-969 */
int oracle_adjoint(oracle_model *m, int i_shot) {
  if (check_shot(m, i_shot)) return -1;
  if (m->nt % m->si != 0) {
    snprintf(g_err, sizeof g_err, "nt must be a multiple of snapshot_interval");
    return -1;
  }
  if (!m->accu[0]) {
    snprintf(g_err, sizeof g_err, "no stored forward wavefield");
    return -1;
  }
  const long n = (long)m->nx * m->nz;
  memset(m->vx, 0, n * sizeof(double)); /* :792-805 */
  memset(m->vz, 0, n * sizeof(double));
  memset(m->txx, 0, n * sizeof(double));
  memset(m->tzz, 0, n * sizeof(double));
  memset(m->txz, 0, n * sizeof(double));
  const double dt = m->dt, dx = m->dx, dz = m->dz;
  for (int it = m->nt - 1; it >= 0; --it) {
    if (it % m->si == 0) correlate(m, i_shot, it);
    oracle_stress_phase(m, 1);
    oracle_velocity_phase(m, 1);
    for (int r = 0; r < m->nr; ++r) { /* :951-958 */
      long loc = I2(m->ix_rec[r], m->iz_rec[r]);
      long k = ((long)i_shot * m->nr + r) * m->nt + it;
      m->vx[loc] += dt * m->b_vx[loc] * m->a_ux[k] / (dx * dz);
      m->vz[loc] += dt * m->b_vz[loc] * m->a_uz[k] / (dx * dz);
    }
  }
  return 0;
}

/* src/fdModel.cpp:1128-1145: serial sum, x term then z term */
double oracle_l2_misfit(oracle_model *m) {
  double misfit = 0;
  long nrec = (long)m->n_shots * m->nr * m->nt;
  for (long k = 0; k < nrec; ++k) {
    double dxv = m->rtf_ux_true[k] - m->rtf_ux[k];
    double dzv = m->rtf_uz_true[k] - m->rtf_uz[k];
    misfit += 0.5 * m->dt * (dxv * dxv);
    misfit += 0.5 * m->dt * (dzv * dzv);
  }
  return misfit;
}

/* src/fdModel.cpp:1147-1162 */
void oracle_l2_adjoint_sources(oracle_model *m) {
  long nrec = (long)m->n_shots * m->nr * m->nt;
  for (long k = 0; k < nrec; ++k) {
    m->a_ux[k] = m->rtf_ux[k] - m->rtf_ux_true[k];
    m->a_uz[k] = m->rtf_uz[k] - m->rtf_uz_true[k];
  }
}

/* src/fdModel.cpp:1310-1324 */
void oracle_reset_kernels(oracle_model *m) {
  size_t n = (size_t)m->nx * m->nz * sizeof(double);
  memset(m->lambda_k, 0, n);
  memset(m->mu_k, 0, n);
  memset(m->rho_l_k, 0, n);
}

/* src/fdModel.cpp:1164-1183 */
void oracle_map_kernels(oracle_model *m) {
  long n = (long)m->nx * m->nz;
  for (long i = 0; i < n; ++i) {
    double vp = m->vp[i], vs = m->vs[i], lk = m->lambda_k[i], mk = m->mu_k[i];
    m->vp_k[i] = 2 * vp * lk / m->b_vx[i];
    m->vs_k[i] = (2 * vs * mk - 4 * vs * lk) / m->b_vx[i];
    m->rho_v_k[i] = m->rho_l_k[i] + (vp * vp - 2 * vs * vs) * lk + vs * vs * mk;
  }
}

void oracle_get_kernels(const oracle_model *m, double *a, double *b, double *c) {
  size_t n = (size_t)m->nx * m->nz * sizeof(double);
  memcpy(a, m->vp_k, n);
  memcpy(b, m->vs_k, n);
  memcpy(c, m->rho_v_k, n);
}

void oracle_get_lame_kernels(const oracle_model *m, double *a, double *b,
                             double *c) {
  size_t n = (size_t)m->nx * m->nz * sizeof(double);
  memcpy(a, m->lambda_k, n);
  memcpy(b, m->mu_k, n);
  memcpy(c, m->rho_l_k, n);
}

void oracle_set_lame_kernels(oracle_model *m, const double *a, const double *b,
                             const double *c) {
  size_t n = (size_t)m->nx * m->nz * sizeof(double);
  memcpy(m->lambda_k, a, n);
  memcpy(m->mu_k, b, n);
  memcpy(m->rho_l_k, c, n);
}

#define NREC_BYTES(m) ((size_t)(m)->n_shots * (m)->nr * (m)->nt * sizeof(double))
void oracle_get_synthetic(const oracle_model *m, double *ux, double *uz) {
  memcpy(ux, m->rtf_ux, NREC_BYTES(m));
  memcpy(uz, m->rtf_uz, NREC_BYTES(m));
}
void oracle_set_synthetic(oracle_model *m, const double *ux, const double *uz) {
  memcpy(m->rtf_ux, ux, NREC_BYTES(m));
  memcpy(m->rtf_uz, uz, NREC_BYTES(m));
}
void oracle_get_observed(const oracle_model *m, double *ux, double *uz) {
  memcpy(ux, m->rtf_ux_true, NREC_BYTES(m));
  memcpy(uz, m->rtf_uz_true, NREC_BYTES(m));
}
void oracle_set_observed(oracle_model *m, const double *ux, const double *uz) {
  memcpy(m->rtf_ux_true, ux, NREC_BYTES(m));
  memcpy(m->rtf_uz_true, uz, NREC_BYTES(m));
}
void oracle_get_adjoint_sources(const oracle_model *m, double *ax, double *az) {
  memcpy(ax, m->a_ux, NREC_BYTES(m));
  memcpy(az, m->a_uz, NREC_BYTES(m));
}

/* src/fdModel.cpp:1358-1412: cell average, x-fastest parameter index */
void oracle_get_model_vector(const oracle_model *m, double *out) {
  int npp = m->nxf * m->nzf / (m->npx * m->npz);
  for (int ix = 0; ix < m->nxf / m->npx; ++ix) {
    for (int iz = 0; iz < m->nzf / m->npz; ++iz) {
      int ip = ix + iz * m->nxf / m->npx;
      int gix = ix * m->npx + m->nxib + m->np;
      int giz = iz * m->npz + m->nzib + m->np;
      out[ip] = 0;
      out[ip + npp] = 0;
      out[ip + 2 * npp] = 0;
      for (int sx = 0; sx < m->npx; ++sx) {
        for (int sz = 0; sz < m->npz; ++sz) {
          long i = I2(gix + sx, giz + sz);
          out[ip] += m->vp[i] / (m->npx * m->npz);
          out[ip + npp] += m->vs[i] / (m->npx * m->npz);
          out[ip + 2 * npp] += m->rho[i] / (m->npx * m->npz);
        }
      }
    }
  }
}

/* src/fdModel.cpp:1414-1457 */
void oracle_set_model_vector(oracle_model *m, const double *in) {
  int npp = m->nxf * m->nzf / (m->npx * m->npz);
  for (int ix = 0; ix < m->nxf / m->npx; ++ix) {
    for (int iz = 0; iz < m->nzf / m->npz; ++iz) {
      int ip = ix + iz * m->nxf / m->npx;
      int gix = ix * m->npx + m->nxib + m->np;
      int giz = iz * m->npz + m->nzib + m->np;
      for (int sx = 0; sx < m->npx; ++sx) {
        for (int sz = 0; sz < m->npz; ++sz) {
          long i = I2(gix + sx, giz + sz);
          m->vp[i] = in[ip];
          m->vs[i] = in[ip + npp];
          m->rho[i] = in[ip + 2 * npp];
        }
      }
    }
  }
  update_from_velocity(m);
}

/* src/fdModel.cpp:1459-1515: cell sum */
void oracle_get_gradient_vector(const oracle_model *m, double *out) {
  int npp = m->nxf * m->nzf / (m->npx * m->npz);
  for (int ix = 0; ix < m->nxf / m->npx; ++ix) {
    for (int iz = 0; iz < m->nzf / m->npz; ++iz) {
      int ip = ix + iz * m->nxf / m->npx;
      int gix = ix * m->npx + m->nxib + m->np;
      int giz = iz * m->npz + m->nzib + m->np;
      out[ip] = 0;
      out[ip + npp] = 0;
      out[ip + 2 * npp] = 0;
      for (int sx = 0; sx < m->npx; ++sx) {
        for (int sz = 0; sz < m->npz; ++sz) {
          long i = I2(gix + sx, giz + sz);
          out[ip] += m->vp_k[i];
          out[ip + npp] += m->vs_k[i];
          out[ip + 2 * npp] += m->rho_v_k[i];
        }
      }
    }
  }
}

void oracle_get_taper(const oracle_model *m, double *out) {
  memcpy(out, m->taper, (size_t)m->nx * m->nz * sizeof(double));
}
void oracle_get_stf(const oracle_model *m, double *out) {
  memcpy(out, m->stf, (size_t)m->n_sources * m->nt * sizeof(double));
}
void oracle_get_moment(const oracle_model *m, double *out) {
  memcpy(out, m->moment, (size_t)m->n_sources * 4 * sizeof(double));
}
void oracle_get_lame(const oracle_model *m, double *lm, double *la, double *mu,
                     double *b) {
  size_t n = (size_t)m->nx * m->nz * sizeof(double);
  memcpy(lm, m->lm, n);
  memcpy(la, m->la, n);
  memcpy(mu, m->mu, n);
  memcpy(b, m->b_vx, n);
}

void oracle_get_state(const oracle_model *m, double *vx, double *vz,
                      double *txx, double *tzz, double *txz) {
  size_t n = (size_t)m->nx * m->nz * sizeof(double);
  memcpy(vx, m->vx, n);
  memcpy(vz, m->vz, n);
  memcpy(txx, m->txx, n);
  memcpy(tzz, m->tzz, n);
  memcpy(txz, m->txz, n);
}

void oracle_set_state(oracle_model *m, const double *vx, const double *vz,
                      const double *txx, const double *tzz, const double *txz) {
  size_t n = (size_t)m->nx * m->nz * sizeof(double);
  memcpy(m->vx, vx, n);
  memcpy(m->vz, vz, n);
  memcpy(m->txx, txx, n);
  memcpy(m->tzz, tzz, n);
  memcpy(m->txz, txz, n);
}

int oracle_get_snapshot(const oracle_model *m, int i_shot, int i_snap,
                        double *vx, double *vz, double *txx, double *tzz,
                        double *txz) {
  if (!m->accu[0] || i_shot < 0 || i_shot >= m->n_shots || i_snap < 0 ||
      i_snap >= m->snapshots)
    return -1;
  long n = (long)m->nx * m->nz;
  long off = ((long)i_shot * m->snapshots + i_snap) * n;
  double *outs[5] = {vx, vz, txx, tzz, txz};
  for (int k = 0; k < 5; ++k)
    memcpy(outs[k], m->accu[k] + off, n * sizeof(double));
  return 0;
}
