/*
 * psvwave.h -- C-ABI of the MI355X-native fdModel hot path
 * (libpsvwave_amd.so).
 *
 * Drop-in boundary for psvWave's fdModel: every entry point below replaces a
 * method of the reference's pybind11 class (src/psvWave.cpp:465-655, class
 * fdModelExtended at :85-458) or of the C++ class it wraps
 * (src/fdModel.h:35-318).  Plain pointers and sizes only.
 *
 * Conventions
 *  - Every function returns PSV_OK (0) or an error status; the message of
 *    the last error on the calling thread is psv_last_error().
 *  - PSV_EVALUE marks the conditions where the reference throws
 *    std::invalid_argument / std::length_error / py::value_error (surfaced
 *    in Python as ValueError); PSV_ERUNTIME marks HIP failures.
 *  - Host buffers are caller-owned and use the reference layout: grids
 *    (nx, nz) with iz fastest, data (n_shots, nr, nt), snapshots
 *    (n_shots, snapshots, nx, nz), vectors (free_parameters,).
 *  - psv_forward and psv_adjoint enqueue their time loops on the device and
 *    return (host-side argument errors are still reported at once); every
 *    other call on the model first waits for them, so results are complete
 *    when a call that returns them returns, and an error of an enqueued loop
 *    is reported by the next call.  psv_set_async(m, 0) makes them
 *    synchronous.  A model is not re-entrant; distinct models are
 *    independent.
 */
#ifndef PSVWAVE_H
#define PSVWAVE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSV_OK 0
#define PSV_EVALUE 1
#define PSV_ERUNTIME 2

typedef struct psv_model psv_model;

/* Configuration with the 30 values of the reference's explicit-argument
 * constructor (src/fdModel.h:50-64).  which_source_to_fire_in_which_shot is
 * given in CSR form: shot k fires which_sources[which_offsets[k] ..
 * which_offsets[k+1]). */
typedef struct {
  int nt, nx_inner, nz_inner, nx_inner_boundary, nz_inner_boundary;
  double dx, dz, dt;
  int np_boundary;
  double np_factor;
  double scalar_rho, scalar_vp, scalar_vs;
  int npx, npz;
  double peak_frequency, source_timeshift, delay_cycles_per_shot;
  int n_sources, n_shots;
  const int *ix_sources, *iz_sources;
  const double *moment_angles;
  const int *which_offsets, *which_sources;
  int nr;
  const int *ix_receivers, *iz_receivers;
  int snapshot_interval;
  const char *observed_data_folder, *stf_folder;
} psv_params;

/* Read-only attributes (src/psvWave.cpp:499-501, 557-588, 621-626). */
typedef struct {
  int nt, nx, nz, nx_inner, nz_inner, nx_inner_boundary, nz_inner_boundary;
  int np_boundary, nx_free_parameters, nz_free_parameters, free_parameters;
  int snapshot_interval, snapshots, n_shots, n_sources, nr, npx, npz;
  double dx, dz, dt, np_factor, peak_frequency, source_timeshift,
      delay_cycles_per_shot, scalar_rho, scalar_vp, scalar_vs, misfit;
} psv_info;

/* Message of the last failed call on this thread ("" if none). */
const char *psv_last_error(void);

/* Parsed and validated configuration, host only (no device work).  Covers
 * fdModel::parse_configuration_file / parse_parameters / initialize_arrays
 * (src/fdModel.cpp:390-523, 193-299) without allocating a model. */
typedef struct psv_config psv_config;
int psv_config_from_ini(const char *path, psv_config **out);
int psv_config_from_params(const psv_params *p, psv_config **out);
void psv_config_destroy(psv_config *c);
int psv_config_get_info(const psv_config *c, psv_info *out);
/* which_source_to_fire_in_which_shot (CSR), grid indices of sources and
 * receivers including np_boundary; any pointer may be NULL. */
int psv_config_get_lists(const psv_config *c, int *which_offsets,
                         int *which_sources, int *ix_src, int *iz_src,
                         int *ix_rec, int *iz_rec);
/* stf (n_sources, nt), moment (n_sources, 2, 2), taper (nx, nz); any
 * pointer may be NULL. */
int psv_config_get_setup(const psv_config *c, double *stf, double *moment,
                         double *taper);
int psv_create_from_config(const psv_config *c, psv_model **out);

/* fdModel(const char*) -- src/fdModel.cpp:16-27, psvWave.cpp:471. */
int psv_create_from_ini(const char *path, psv_model **out);
/* fdModel(nt, nx_inner, ...) -- src/fdModel.cpp:29-62. */
int psv_create(const psv_params *p, psv_model **out);
/* copy() -- src/psvWave.cpp:421-425, fdModel(const fdModel&) at
 * src/fdModel.cpp:64-97.  Deep copy of all device state. */
int psv_copy(const psv_model *m, psv_model **out);
/* ~fdModel() -- src/fdModel.cpp:99-137. */
void psv_destroy(psv_model *m);

/* def_readonly attributes -- src/psvWave.cpp:499-588, 621-626. */
int psv_get_info(const psv_model *m, psv_info *out);
/* which_source_to_fire_in_which_shot (src/psvWave.cpp:586-588), CSR form:
 * offsets has n_shots + 1 entries, sources n_sources. */
int psv_get_which(const psv_model *m, int *offsets, int *sources);
/* Source / receiver grid indices including the absorbing boundary
 * (ix_sources etc. after src/fdModel.cpp:197-212); used by get_sources /
 * get_receivers (src/psvWave.cpp:324-387). */
int psv_get_source_indices(const psv_model *m, int *ix, int *iz);
int psv_get_receiver_indices(const psv_model *m, int *ix, int *iz);
/* Output folders from the [output] section, copied into buf (size len). */
int psv_get_folders(const psv_model *m, char *observed, char *stf, int len);

/* forward_simulate -- src/fdModel.cpp:526-787, psvWave.cpp:112-158, 476-498.
 * omp_threads_override of the reference has no device meaning and is not
 * part of this call.  output_wavefields writes snapshots/vxNNNNN.txt and
 * vzNNNNN.txt every 10 steps as the reference does. */
int psv_forward(psv_model *m, int i_shot, int store_fields, int verbose,
                int output_wavefields);
/* adjoint_simulate -- src/fdModel.cpp:789-969, psvWave.cpp:160-205. */
int psv_adjoint(psv_model *m, int i_shot, int verbose);
/* Asynchronous per-shot calls (no reference counterpart: the reference's
 * calls return when their loop has finished).  With on = 1 (the default)
 * psv_forward and psv_adjoint return once their time loop is enqueued:
 * consecutive forwards run two at a time on two HIP streams, and
 * consecutive adjoints pair up, the second of a pair running beside the
 * first with its kernel terms deferred to a buffer (3 nx nz doubles per
 * snapshot) and added afterwards in its own order -- outputs are bitwise
 * those of synchronous calls.  Verbose calls, step timing, graph mode,
 * checkpointing and step variants 2, 5, 6, 7 and 14 stay synchronous.
 * psv_synchronize waits for the enqueued loops. */
int psv_set_async(psv_model *m, int on);
int psv_synchronize(psv_model *m);
/* calculate_l2_misfit -- src/fdModel.cpp:1128-1145.  Also stored as the
 * misfit attribute. */
int psv_calculate_l2_misfit(psv_model *m, double *misfit);
/* calculate_l2_adjoint_sources -- src/fdModel.cpp:1147-1162. */
int psv_calculate_l2_adjoint_sources(psv_model *m);
/* reset_kernels -- src/fdModel.cpp:1310-1324. */
int psv_reset_kernels(psv_model *m);
/* map_kernels_to_velocity -- src/fdModel.cpp:1164-1183. */
int psv_map_kernels_to_velocity(psv_model *m);
/* run_model -- src/fdModel.cpp:1291-1308 (C++-only in the reference). */
int psv_run_model(psv_model *m, int verbose, int simulate_adjoint);

/* get_parameter_fields / set_parameter_fields -- src/psvWave.cpp:248-297;
 * set also runs update_from_velocity (src/fdModel.cpp:1034-1051). */
int psv_get_parameter_fields(const psv_model *m, double *vp, double *vs,
                             double *rho);
int psv_set_parameter_fields(psv_model *m, const double *vp, const double *vs,
                             const double *rho);
/* get_kernels -- src/psvWave.cpp:299-306 (vp, vs, rho velocity kernels). */
int psv_get_kernels(const psv_model *m, double *vp_k, double *vs_k,
                    double *rho_k);
/* The Lame-basis kernels lambda_kernel, mu_kernel, density_l_kernel
 * (src/fdModel.h:224-226); public fields of the reference C++ class. */
int psv_get_lame_kernels(const psv_model *m, double *lambda_k, double *mu_k,
                         double *rho_l_k);
/* get/set_synthetic_data, get/set_observed_data --
 * src/psvWave.cpp:308-322, 389-419, 427-457. */
int psv_get_synthetic_data(const psv_model *m, double *ux, double *uz);
int psv_set_synthetic_data(psv_model *m, const double *ux, const double *uz);
int psv_get_observed_data(const psv_model *m, double *ux, double *uz);
int psv_set_observed_data(psv_model *m, const double *ux, const double *uz);
/* a_stf_ux / a_stf_uz (src/fdModel.h:246-247). */
int psv_get_adjoint_sources(const psv_model *m, double *ux, double *uz);
/* get_snapshots -- src/psvWave.cpp:90-97: five arrays of
 * (n_shots, snapshots, nx, nz). */
int psv_get_snapshots(const psv_model *m, double *vx, double *vz, double *txx,
                      double *tzz, double *txz);
/* One (shot, snapshot) slice of get_snapshots, five (nx, nz) arrays. */
int psv_get_snapshot(const psv_model *m, int i_shot, int i_snapshot,
                     double *vx, double *vz, double *txx, double *tzz,
                     double *txz);
/* get_model_vector / set_model_vector / get_gradient_vector --
 * src/fdModel.cpp:1358-1515.  n must equal free_parameters. */
int psv_get_model_vector(const psv_model *m, double *v, int64_t n);
int psv_set_model_vector(psv_model *m, const double *v, int64_t n);
int psv_get_gradient_vector(const psv_model *m, double *v, int64_t n);
/* load_vector -- src/fdModel.cpp:1517-1572 (reads 3*nxf*nzf values). */
int psv_load_vector(const psv_model *m, const char *path, int verbose,
                    double *v, int64_t n);
/* Setup products (src/fdModel.cpp:214-246, 263-298): stf (n_sources, nt),
 * moment (n_sources, 2, 2), taper (nx, nz). */
int psv_get_setup(const psv_model *m, double *stf, double *moment,
                  double *taper);

/* Text formats of write_receivers / write_sources / write_kernels
 * (src/fdModel.cpp:971-1032, 1326-1356) and load_receivers / load_model
 * (src/fdModel.cpp:1053-1126, 1185-1289). */
int psv_write_receivers(const psv_model *m, const char *prefix);
int psv_write_sources(const psv_model *m);
int psv_write_kernels(const psv_model *m);
int psv_load_receivers(psv_model *m, int verbose);
int psv_load_model(psv_model *m, const char *de_path, const char *vp_path,
                   const char *vs_path, int verbose);

/* Tuning and measurement (no reference counterpart).  variant: 0 fused
 * step kernel with 64x8 tiles, 1 fused 64x16, 2 split (two launches per
 * step), 3-5 the same with reciprocal-corrected divisions, 6 / 7 two steps
 * per launch (64x8 / 64x16 tiles), 8-13 lean one-step kernels (64 x 16,
 * 8, 12, 20, 24, 32 tiles), 14 the lean two-step kernel, 15-17 lean kernels whose blocks
 * advance 2, 3 or 4 tiles of a column strip in turn (the next tile's loads
 * in flight meanwhile), 18 the lean kernel on 64 x 4 tiles, 19 / 20 the
 * 64 x 8 / 64 x 12 lean kernels held to 6 / 5 waves per SIMD, 21 the 64 x 8
 * lean kernel with halo columns in the row transfers, 22 / 23 the 64 x 8 /
 * 64 x 4 lean kernels with separate velocity and stress LDS stages, 24 the
 * 64 x 8 lean kernel with halo columns and ring-point operands moved
 * straight into LDS (six waves per SIMD without spills), 25 the same
 * without the halo columns, 26
 * the 64 x 16 lean kernel with ring-point operands in LDS (four waves per
 * SIMD), 27 / 28 variants 24 / 26 with their loads capped to the grid (the
 * tiles on the right and bottom edges read no padding; 27 is the default
 * when dx, dz lie in [2^-40, 2^40]).
 * All variants give bitwise identical results.  xcd_remap: 1 keeps
 * neighbouring tiles on one XCD.
 * timing_stride k > 0 records the kernel start and end of every k-th step
 * launch; k < 0 brackets whole time loops with HIP events on the model's
 * stream (launches plus the gaps between them); psv_get_step_timing returns
 * the summed milliseconds and the number of time steps covered since the
 * last reset. */
int psv_set_step_variant(psv_model *m, int variant, int xcd_remap);
int psv_set_timing(psv_model *m, int timing_stride);
/* on = 1: capture each (shot, direction, store_fields) time loop once as a
 * HIP graph and replay it; with timing enabled, each replay is bracketed as
 * a whole (kernel time plus inter-kernel gaps). */
int psv_set_graph_mode(psv_model *m, int on);
/* Checkpointing (no reference counterpart; the reference stores every
 * snapshot, src/fdModel.cpp:555-572): keep only every `every`-th forward
 * snapshot (1: all, the default) and recompute the others during
 * adjoint_simulate, one segment of `every` snapshot intervals at a time.
 * Kernels are bit-identical to full storage; snapshot memory drops by
 * `every`, the adjoint costs one extra forward.  get_snapshots is then
 * unavailable.  Changing it discards stored snapshots. */
int psv_set_checkpointing(psv_model *m, int every);
int psv_get_step_timing(psv_model *m, double *total_ms, int64_t *timed,
                        int64_t *launches);
int psv_reset_step_timing(psv_model *m);
/* Device bytes held by the model. */
int psv_device_bytes(const psv_model *m, int64_t *bytes);
/* L2 misfit of device-resident traces without any allocation (no
 * counterpart in the reference; for callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same reduction as psv_calculate_l2_misfit, whose
 * result it reproduces bit for bit: *misfit (a device pointer) =
 * sum over i < n of 0.5 dt ((obs_ux[i] - syn_ux[i])^2 + (obs_uz[i] -
 * syn_uz[i])^2), the four arrays device buffers of n doubles each (a
 * model's traces are n = n_shots * nr * nt values in [shot][receiver][it]
 * order).  Enqueued on `stream` (a hipStream_t; NULL = the default stream),
 * not synchronised.  Its 512 compensated partial sums live in a fixed 8 KiB
 * device array of the library, one per GPU, held until the process ends;
 * every call overwrites them, so calls that overlap on one GPU race.
 * n outside [0, PSV_NOALLOC_MAX_TERMS] is rejected with PSV_EVALUE. */
#define PSV_NOALLOC_MAX_TERMS ((int64_t)1 << 40)
int psv_l2_misfit_noalloc(const double *syn_ux, const double *syn_uz,
                          const double *obs_ux, const double *obs_uz, int64_t n,
                          double dt, double *misfit, void *stream);
/* Streaming copy of `bytes` (a multiple of 16) device-to-device, `iters`
 * times; *ms is the mean time per copy (HIP events around all of them).
 * runtime_copy = 0: the library's plain copy kernel, the reference point for
 * byte counters; 1: hipMemcpyAsync (the runtime's own copy), the practical
 * streaming ceiling. */
int psv_calibrate_copy(int64_t bytes, int iters, int runtime_copy, double *ms);
/* No counterpart in the reference (its loops run on host threads).  HIP
 * streams the library holds open in this process and the runtime's
 * hardware-queue limit (GPU_MAX_HW_QUEUES, 4 by default): a model's two
 * lanes overlap only while they have a queue each, so more live streams
 * than queues (the default stream counts too) may run the lanes one after
 * the other -- same results, no overlap. */
int psv_stream_pressure(int *live_streams, int *hw_queues);

/* ---- Batched Monte-Carlo forward runs (no reference counterpart: the
 * reference runs one fdModel at a time, README.md:16-17 states the
 * Monte-Carlo purpose; copy() at src/psvWave.cpp:421-425 is its only way to
 * multiply models).  A batch holds n independent models of possibly
 * different sizes sharing nt and n_shots; psv_batch_forward advances every
 * model of one shot with one launch per time step over all their tiles.
 * Each member behaves as the model psv_create(&params[m]) would:
 * seismograms are bit-identical to running the members one by one. */
typedef struct psv_batch psv_batch;
int psv_batch_create(int n, const psv_params *params, psv_batch **out);
void psv_batch_destroy(psv_batch *b);
int psv_batch_size(const psv_batch *b, int *n);
int psv_batch_get_info(const psv_batch *b, int m, psv_info *out);
int psv_batch_set_parameter_fields(psv_batch *b, int m, const double *vp,
                                   const double *vs, const double *rho);
int psv_batch_set_observed_data(psv_batch *b, int m, const double *ux,
                                const double *uz);
int psv_batch_get_synthetic_data(const psv_batch *b, int m, double *ux,
                                 double *uz);
/* Forward simulation of shot i_shot for every model (no snapshots). */
int psv_batch_forward(psv_batch *b, int i_shot);
/* Forward simulation of every shot of every model (what psv_batch_forward
 * for shots 0 .. n_shots-1 does, bit for bit).  With 2 to 8 shots and a
 * lean step variant the shots advance together, one launch per time step
 * over every (tile, shot) of every model, shot s in the model's state buffer
 * set s (10 padded grids more per model and extra shot; shot by shot when
 * the device has no room for them and with the persistent loop).  In graph
 * mode that whole loop is captured once as one HIP graph and replayed. */
int psv_batch_forward_all(psv_batch *b);
/* *on = 1 when psv_batch_forward_all advances the shots together with the
 * current variant and modes (allocating the extra state sets if needed),
 * else 0. */
int psv_batch_lockstep(psv_batch *b, int *on);
/* L2 misfit of every model (n values), as calculate_l2_misfit. */
int psv_batch_misfits(psv_batch *b, double *out);
/* Step kernel for all members: 0, 1, 6, 7, 8-13 or 18-28 (see
 * psv_set_step_variant). */
int psv_batch_set_variant(psv_batch *b, int variant, int xcd_remap);
/* FWI gradient of every member against its observed data (what
 * psv_run_model(m, 0, 1) does for one model): kernels reset, then shot by
 * shot the forward storing snapshots, the adjoint sources and the adjoint
 * with kernel correlation -- one launch per time step over every member --
 * then misfits (psv_batch_misfits) and velocity kernels
 * (psv_batch_get_kernels, (nx, nz) each).  Needs one snapshot_interval for
 * all members, dividing nt, and step variant 0, 1, 8-13 or 18-28; each
 * member keeps one shot's snapshots.  Results equal each member's run_model bit for bit. */
int psv_batch_gradients(psv_batch *b);
int psv_batch_get_kernels(const psv_batch *b, int m, double *vp_k, double *vs_k,
                          double *rho_k);
int psv_batch_set_graph_mode(psv_batch *b, int on);
/* on = 1: each shot's forward time loop is one cooperative launch whose
 * resident blocks meet at a grid barrier between steps, instead of one
 * launch per step (step variant 9 only; the same results). */
int psv_batch_set_persistent(psv_batch *b, int on);
/* Times `steps` batched steps of variants 0, 1, 6, 7, 8-13, 18-20, 22-24,
 * 26-28 (ms_per_step[19]: ms per time step of one shot, -1 for a variant the batch
 * cannot use) in the loop psv_batch_forward_all runs with each, and keeps
 * the fastest (*best).  Overwrites the synthetic data. */
int psv_batch_autotune(psv_batch *b, int steps, int *best,
                       double *ms_per_step);
/* Summed wall time of psv_batch_forward time loops (HIP events) and the
 * number of grid points of all members. */
int psv_batch_timing(psv_batch *b, double *loop_ms, int64_t *points,
                     int reset);

/* Learned gradient preconditioner (no reference counterpart; the reference
 * has no preconditioner).  bfloat16 MFMA convolution on device buffers the
 * caller owns, enqueued on `stream` (a hipStream_t; NULL = the default
 * stream) and not synchronised.  Activations NHWC (one image) as bfloat16
 * bit patterns; weights [N][K] with K = (ky * ks + kx) * cin + ci; cin a
 * multiple of 8; in and weights 16-byte aligned; ks 1 or 3 ("same"
 * padding).  transposed = 1: 2x2 stride-2 transposed convolution (ks = 1,
 * N = 4 * cout, n = (2a + b) * cout + co writes pixel (2y + a, 2x + b) of
 * a (2h, 2w) output).  The
 * result (+ bias, then erf GELU if gelu = 1) goes to channel out_offset of
 * rows of out_stride channels, fp32 if out_f32 = 1 else bfloat16. */
int psv_conv2d_bf16(const void *in, const void *weights, const float *bias,
                    void *out, int out_f32, int h, int w, int cin, int cout,
                    int ks, int transposed, int gelu, int out_stride,
                    int out_offset, void *stream);
/* 2 x 2 max pooling of an NHWC bfloat16 image (h, w even; a batch of
 * images stacked along the rows pools as one image of batch * h rows). */
int psv_maxpool2_bf16(const void *in, void *out, int h, int w, int c,
                      void *stream);

/* Training kernels of the same network (bfloat16 operands, fp32
 * accumulation).  Images are stacked along the rows: `batch` images of
 * h x w pixels.  General convolution: the input has in_stride channels per
 * pixel, of which the first cin (a multiple of 8) are read; in and weights
 * are 16-byte aligned.
 *   mode 0: ks x ks convolution (ks 1 or 3, "same" padding per image),
 *           K = ks * ks * cin, output h x w;
 *   mode 1: 2x2 stride-2 transposed convolution of an h x w input
 *           (ks = 1, N = 4 * cout), output 2h x 2w;
 *   mode 2: 2x2 gather: the input has 2h x 2w pixels, K = 4 * cin and
 *           k = (2a + b) * cin + ci reads pixel (2y + a, 2x + b) -- the
 *           input gradient of mode 1 (weights [ci][(2a + b) * cout + co]).
 * bias may be NULL.  pre_out (mode 0, may be NULL) receives the bfloat16
 * pre-activation [pixels][cout]; gelu = 1 applies the erf GELU to it.  dz
 * (may be NULL) multiplies the result by gelu'(dz[pixel * dz_stride + co])
 * (the backward pass through a GELU whose pre-activation is dz). */
int psv_conv_bf16(const void *in, int in_stride, const void *weights, const float *bias,
                  void *out, int out_f32, int out_stride, int out_offset, void *pre_out,
                  const void *dz, int dz_stride, int batch, int h, int w, int cin, int cout,
                  int ks, int mode, int gelu, void *stream);
/* Weight gradient of a mode-0 convolution, dw[co][(ky * ks + kx) *
 * x_channels + ci] = sum over pixels of dy[p][co] x(p shifted by the tap)[ci]
 * (dy_gather2x2 = 0, dy at the h x w pixels), or of a mode-1 transposed
 * convolution, dw[(2a + b) * dy_channels + co][ci] (dy_gather2x2 = 1: dy at
 * 2h x 2w pixels, ks = 1).  dy_stride, x_channels and x_stride are
 * multiples of 8 (and dy_channels for the gather); dy and x are 16-byte
 * aligned.  db (may be NULL) receives the pixel sums of
 * dy's gathered columns: the bias gradient of a convolution ([co]); for the
 * 2x2 gather [(2a + b) * dy_channels + co], whose four positions add up to
 * the bias gradient.  fp32 results; the pixel sum is split over workspace
 * (up to workspace_floats floats; NULL = no split) and added in a fixed
 * order, so the result is deterministic. */
int psv_conv_wgrad_bf16(const void *dy, int dy_stride, int dy_gather2x2, int dy_channels,
                        const void *x, int x_stride, int x_channels, int ks, int batch, int h,
                        int w, float *dw, float *db, float *workspace,
                        int64_t workspace_floats, void *stream);
/* out[c] = sum over pixels of src[p * stride + offset + c] (bias gradients),
 * fixed-order fp32 sums. */
int psv_channel_sum_bf16(const void *src, int stride, int offset, int channels,
                         int64_t pixels, float *out, float *workspace,
                         int64_t workspace_floats, void *stream);
/* Gradient at a block output feeding a 2x2 max pooling and a skip
 * connection, through the block's GELU: out = (route(dpool) + dskip) *
 * gelu'(z), route sending each pooled gradient (fp32, may be NULL) to the
 * first maximum of its window of act in row-major order; dskip (bfloat16,
 * may be NULL) at channel skip_offset of rows of skip_stride channels. */
int psv_maxpool2_backward_bf16(const void *act, const void *z, const float *dpool,
                               const void *dskip, int skip_stride, int skip_offset, void *out,
                               int batch, int h, int w, int c, void *stream);

#ifdef __cplusplus
}
#endif
#endif
