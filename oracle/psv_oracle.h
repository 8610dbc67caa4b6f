/*
 * psv_oracle.h -- CPU restatement of psvWave's fdModel hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product path is compared against.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it; the product never does.
 *
 * Arrays crossing this interface use the reference layout: grids are
 * [ix][iz] with iz fastest (src/contiguous_arrays.cpp:43-50), data are
 * [shot][receiver][it], snapshots are [shot][snapshot][ix][iz].
 */
#ifndef PSV_ORACLE_H
#define PSV_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Explicit-argument configuration, the same 30 values as the reference's
 * variable constructor (src/fdModel.h:50-64). */
typedef struct {
  int nt, nx_inner, nz_inner, nx_inner_boundary, nz_inner_boundary;
  double dx, dz, dt;
  int np_boundary;
  double np_factor;
  double scalar_rho, scalar_vp, scalar_vs;
  int npx, npz;
  double peak_frequency, source_timeshift, delay_cycles_per_shot;
  int n_sources, n_shots;
  const int *ix_sources, *iz_sources;   /* n_sources, inner-grid indices */
  const double *moment_angles;          /* n_sources, degrees */
  const int *which_offsets;             /* n_shots + 1 (CSR) */
  const int *which_sources;             /* which_offsets[n_shots] */
  int nr;
  const int *ix_receivers, *iz_receivers; /* nr, inner-grid indices */
  int snapshot_interval;
} oracle_params;

typedef struct oracle_model oracle_model;

const char *oracle_last_error(void);
void oracle_set_threads(int n);

oracle_model *oracle_create(const oracle_params *p);
oracle_model *oracle_copy(const oracle_model *m);
void oracle_destroy(oracle_model *m);

/* nx, nz, nt, snapshots, free_parameters, n_shots, nr, n_sources */
void oracle_dims(const oracle_model *m, int *out8);

void oracle_set_parameter_fields(oracle_model *m, const double *vp,
                                 const double *vs, const double *rho);
void oracle_get_parameter_fields(const oracle_model *m, double *vp, double *vs,
                                 double *rho);

int oracle_forward(oracle_model *m, int i_shot, int store_fields);
int oracle_adjoint(oracle_model *m, int i_shot);
double oracle_l2_misfit(oracle_model *m);
void oracle_l2_adjoint_sources(oracle_model *m);
void oracle_reset_kernels(oracle_model *m);
void oracle_map_kernels(oracle_model *m);

void oracle_get_kernels(const oracle_model *m, double *vp_k, double *vs_k,
                        double *rho_k);
void oracle_get_lame_kernels(const oracle_model *m, double *lambda_k,
                             double *mu_k, double *rho_l_k);
void oracle_set_lame_kernels(oracle_model *m, const double *lambda_k,
                             const double *mu_k, const double *rho_l_k);
void oracle_get_synthetic(const oracle_model *m, double *ux, double *uz);
void oracle_set_synthetic(oracle_model *m, const double *ux, const double *uz);
void oracle_get_observed(const oracle_model *m, double *ux, double *uz);
void oracle_set_observed(oracle_model *m, const double *ux, const double *uz);
void oracle_get_adjoint_sources(const oracle_model *m, double *ax, double *az);

void oracle_get_model_vector(const oracle_model *m, double *out);
void oracle_set_model_vector(oracle_model *m, const double *in);
void oracle_get_gradient_vector(const oracle_model *m, double *out);

/* Setup products, for closed-form checks. */
void oracle_get_taper(const oracle_model *m, double *out);   /* [ix][iz] */
void oracle_get_stf(const oracle_model *m, double *out);     /* [src][it] */
void oracle_get_moment(const oracle_model *m, double *out);  /* [src][2][2] */
void oracle_get_lame(const oracle_model *m, double *lm, double *la,
                     double *mu, double *b);

/* Dynamic state access and single phases of one time step. */
void oracle_get_state(const oracle_model *m, double *vx, double *vz,
                      double *txx, double *tzz, double *txz);
void oracle_set_state(oracle_model *m, const double *vx, const double *vz,
                      const double *txx, const double *tzz, const double *txz);
void oracle_stress_phase(oracle_model *m, int adjoint);
void oracle_velocity_phase(oracle_model *m, int adjoint);
int oracle_set_snapshot(oracle_model *m, int i_shot, int i_snap, const double *vx,
                        const double *vz, const double *txx, const double *tzz,
                        const double *txz);
int oracle_correlate_step(oracle_model *m, int i_shot, int it);
int oracle_get_snapshot(const oracle_model *m, int i_shot, int i_snap,
                        double *vx, double *vz, double *txx, double *tzz,
                        double *txz);

#ifdef __cplusplus
}
#endif
#endif
