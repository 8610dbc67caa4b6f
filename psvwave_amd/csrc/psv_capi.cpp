// psv_capi.cpp -- extern "C" entry points declared in include/psvwave.h.
// Each translates C++ exceptions into a status code plus a thread-local
// message: ValueError -> PSV_EVALUE (the reference's invalid_argument /
// length_error / value_error cases), everything else -> PSV_ERUNTIME.
#include "../../include/psvwave.h"

#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

#include "psv_batch.h"
#include "psv_model.h"

struct psv_model {
  psv::Model *impl;
};

namespace {

thread_local std::string g_error;

template <class F>
int guard(F &&f) {
  g_error.clear();
  try {
    f();
    return PSV_OK;
  } catch (const psv::ValueError &e) {
    g_error = e.what();
    return PSV_EVALUE;
  } catch (const std::bad_alloc &) {
    g_error = "host out of memory";
    return PSV_ERUNTIME;
  } catch (const std::exception &e) {
    g_error = e.what();
    return PSV_ERUNTIME;
  }
}

// The model, settled: every entry point but forward and adjoint first waits
// for the time loops those enqueue (and surfaces their errors).
psv::Model &M(psv_model *m) {
  if (!m || !m->impl) throw psv::ValueError("null model handle");
  m->impl->settle();
  return *m->impl;
}
const psv::Model &M(const psv_model *m) {
  if (!m || !m->impl) throw psv::ValueError("null model handle");
  m->impl->settle();
  return *m->impl;
}
// The model as it is (forward and adjoint enqueue behind pending work).
psv::Model &MA(psv_model *m) {
  if (!m || !m->impl) throw psv::ValueError("null model handle");
  return *m->impl;
}

void need(const void *p, const char *what) {
  if (!p) throw psv::ValueError(std::string("null pointer: ") + what);
}

psv_model *wrap(psv::Model *impl) {
  psv_model *h = new psv_model;
  h->impl = impl;
  return h;
}

}  // namespace

extern "C" {

const char *psv_last_error(void) { return g_error.c_str(); }

int psv_create_from_ini(const char *path, psv_model **out) {
  return guard([&] {
    need(path, "path");
    need(out, "out");
    psv::Config c = psv::parse_ini_file(path);
    *out = wrap(new psv::Model(c));
  });
}

struct psv_config {
  psv::Config cfg;
  psv::Derived d;
};

static psv::Config config_from_params(const psv_params *p) {
    need(p, "params");
    psv::Config c;
    c.nt = p->nt;
    c.nx_inner = p->nx_inner;
    c.nz_inner = p->nz_inner;
    c.nx_inner_boundary = p->nx_inner_boundary;
    c.nz_inner_boundary = p->nz_inner_boundary;
    c.dx = p->dx;
    c.dz = p->dz;
    c.dt = p->dt;
    c.np_boundary = p->np_boundary;
    c.np_factor = p->np_factor;
    c.scalar_rho = p->scalar_rho;
    c.scalar_vp = p->scalar_vp;
    c.scalar_vs = p->scalar_vs;
    c.npx = p->npx;
    c.npz = p->npz;
    c.peak_frequency = p->peak_frequency;
    c.source_timeshift = p->source_timeshift;
    c.delay_cycles_per_shot = p->delay_cycles_per_shot;
    c.n_sources = p->n_sources;
    c.n_shots = p->n_shots;
    if (c.n_sources < 0 || c.n_shots < 0 || p->nr < 0)
      throw psv::ValueError("negative counts");
    if (c.n_sources > 0) {
      need(p->ix_sources, "ix_sources");
      need(p->iz_sources, "iz_sources");
      need(p->moment_angles, "moment_angles");
      c.ix_sources.assign(p->ix_sources, p->ix_sources + c.n_sources);
      c.iz_sources.assign(p->iz_sources, p->iz_sources + c.n_sources);
      c.moment_angles.assign(p->moment_angles, p->moment_angles + c.n_sources);
    }
    need(p->which_offsets, "which_offsets");
    for (int s = 0; s < c.n_shots; ++s) {
      int b = p->which_offsets[s], e = p->which_offsets[s + 1];
      if (e < b) throw psv::ValueError("which_offsets must not decrease");
      if (e > b) need(p->which_sources, "which_sources");
      c.which_source_to_fire_in_which_shot.emplace_back(p->which_sources + b,
                                                        p->which_sources + e);
    }
    c.nr = p->nr;
    if (c.nr > 0) {
      need(p->ix_receivers, "ix_receivers");
      need(p->iz_receivers, "iz_receivers");
      c.ix_receivers.assign(p->ix_receivers, p->ix_receivers + c.nr);
      c.iz_receivers.assign(p->iz_receivers, p->iz_receivers + c.nr);
    }
    c.snapshot_interval = p->snapshot_interval;
    if (p->observed_data_folder) c.observed_data_folder = p->observed_data_folder;
    if (p->stf_folder) c.stf_folder = p->stf_folder;
    return c;
}

int psv_create(const psv_params *p, psv_model **out) {
  return guard([&] {
    need(out, "out");
    *out = wrap(new psv::Model(config_from_params(p)));
  });
}

int psv_config_from_ini(const char *path, psv_config **out) {
  return guard([&] {
    need(path, "path");
    need(out, "out");
    psv::Config c = psv::parse_ini_file(path);
    psv::Derived d = psv::derive(c);
    *out = new psv_config{c, d};
  });
}

int psv_config_from_params(const psv_params *p, psv_config **out) {
  return guard([&] {
    need(out, "out");
    psv::Config c = config_from_params(p);
    psv::Derived d = psv::derive(c);
    *out = new psv_config{c, d};
  });
}

void psv_config_destroy(psv_config *c) { delete c; }

static void fill_info(const psv::Config &c, const psv::Derived &d, double misfit,
                      psv_info *o) {
  o->nt = c.nt;
  o->nx = d.nx;
  o->nz = d.nz;
  o->nx_inner = c.nx_inner;
  o->nz_inner = c.nz_inner;
  o->nx_inner_boundary = c.nx_inner_boundary;
  o->nz_inner_boundary = c.nz_inner_boundary;
  o->np_boundary = c.np_boundary;
  o->nx_free_parameters = d.nx_free_parameters;
  o->nz_free_parameters = d.nz_free_parameters;
  o->free_parameters = d.free_parameters;
  o->snapshot_interval = c.snapshot_interval;
  o->snapshots = d.snapshots;
  o->n_shots = c.n_shots;
  o->n_sources = c.n_sources;
  o->nr = c.nr;
  o->npx = c.npx;
  o->npz = c.npz;
  o->dx = c.dx;
  o->dz = c.dz;
  o->dt = c.dt;
  o->np_factor = c.np_factor;
  o->peak_frequency = c.peak_frequency;
  o->source_timeshift = c.source_timeshift;
  o->delay_cycles_per_shot = c.delay_cycles_per_shot;
  o->scalar_rho = c.scalar_rho;
  o->scalar_vp = c.scalar_vp;
  o->scalar_vs = c.scalar_vs;
  o->misfit = misfit;
}

int psv_config_get_info(const psv_config *c, psv_info *out) {
  return guard([&] {
    need(c, "config");
    need(out, "out");
    fill_info(c->cfg, c->d, 0.0, out);
  });
}

int psv_config_get_lists(const psv_config *c, int *which_offsets,
                         int *which_sources, int *ix_src, int *iz_src,
                         int *ix_rec, int *iz_rec) {
  return guard([&] {
    need(c, "config");
    const auto &w = c->cfg.which_source_to_fire_in_which_shot;
    if (which_offsets) {
      which_offsets[0] = 0;
      int k = 0;
      for (size_t s = 0; s < w.size(); ++s) {
        for (int v : w[s])
          if (which_sources) which_sources[k++] = v;
          else ++k;
        which_offsets[s + 1] = k;
      }
    }
    auto put = [](int *dst, const std::vector<int> &v) {
      if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int));
    };
    put(ix_src, c->d.ix_src);
    put(iz_src, c->d.iz_src);
    put(ix_rec, c->d.ix_rec);
    put(iz_rec, c->d.iz_rec);
  });
}

int psv_config_get_setup(const psv_config *c, double *stf, double *moment,
                         double *taper) {
  return guard([&] {
    need(c, "config");
    const auto &d = c->d;
    if (stf && !d.stf.empty())
      std::memcpy(stf, d.stf.data(), d.stf.size() * sizeof(double));
    if (moment && !d.moment.empty())
      std::memcpy(moment, d.moment.data(), d.moment.size() * sizeof(double));
    if (taper)
      for (int ix = 0; ix < d.nx; ++ix)
        for (int iz = 0; iz < d.nz; ++iz)
          taper[(size_t)ix * d.nz + iz] = d.taper_lut[psv::taper_count(
              ix, iz, d.nx, d.nz, c->cfg.np_boundary)];
  });
}

int psv_create_from_config(const psv_config *c, psv_model **out) {
  return guard([&] {
    need(c, "config");
    need(out, "out");
    *out = wrap(new psv::Model(c->cfg));
  });
}

int psv_copy(const psv_model *m, psv_model **out) {
  return guard([&] {
    need(out, "out");
    *out = wrap(new psv::Model(M(m)));
  });
}

void psv_destroy(psv_model *m) {
  if (!m) return;
  delete m->impl;
  delete m;
}

int psv_get_info(const psv_model *m, psv_info *o) {
  return guard([&] {
    need(o, "out");
    fill_info(M(m).config(), M(m).derived(), M(m).misfit(), o);
  });
}

int psv_get_which(const psv_model *m, int *offsets, int *sources) {
  return guard([&] {
    need(offsets, "offsets");
    const auto &w = M(m).config().which_source_to_fire_in_which_shot;
    offsets[0] = 0;
    int k = 0;
    for (size_t s = 0; s < w.size(); ++s) {
      for (int v : w[s]) {
        need(sources, "sources");
        sources[k++] = v;
      }
      offsets[s + 1] = k;
    }
  });
}

int psv_get_source_indices(const psv_model *m, int *ix, int *iz) {
  return guard([&] {
    const auto &d = M(m).derived();
    if (d.ix_src.empty()) return;
    need(ix, "ix");
    need(iz, "iz");
    std::memcpy(ix, d.ix_src.data(), d.ix_src.size() * sizeof(int));
    std::memcpy(iz, d.iz_src.data(), d.iz_src.size() * sizeof(int));
  });
}

int psv_get_receiver_indices(const psv_model *m, int *ix, int *iz) {
  return guard([&] {
    const auto &d = M(m).derived();
    if (d.ix_rec.empty()) return;
    need(ix, "ix");
    need(iz, "iz");
    std::memcpy(ix, d.ix_rec.data(), d.ix_rec.size() * sizeof(int));
    std::memcpy(iz, d.iz_rec.data(), d.iz_rec.size() * sizeof(int));
  });
}

int psv_get_folders(const psv_model *m, char *observed, char *stf, int len) {
  return guard([&] {
    if (len <= 0) throw psv::ValueError("buffer length must be positive");
    const auto &c = M(m).config();
    if (observed) {
      std::strncpy(observed, c.observed_data_folder.c_str(), len - 1);
      observed[len - 1] = 0;
    }
    if (stf) {
      std::strncpy(stf, c.stf_folder.c_str(), len - 1);
      stf[len - 1] = 0;
    }
  });
}

int psv_forward(psv_model *m, int i_shot, int store_fields, int verbose,
                int output_wavefields) {
  return guard([&] {
    MA(m).forward(i_shot, store_fields != 0, verbose != 0,
                  output_wavefields != 0);
  });
}

int psv_adjoint(psv_model *m, int i_shot, int verbose) {
  return guard([&] { MA(m).adjoint(i_shot, verbose != 0); });
}

int psv_calculate_l2_misfit(psv_model *m, double *misfit) {
  return guard([&] {
    double v = M(m).calculate_l2_misfit();
    if (misfit) *misfit = v;
  });
}

int psv_l2_misfit_noalloc(const double *syn_ux, const double *syn_uz,
                          const double *obs_ux, const double *obs_uz, int64_t n,
                          double dt, double *misfit, void *stream) {
  return guard([&] {
    if (n < 0 || n > PSV_NOALLOC_MAX_TERMS)
      throw psv::ValueError("psv_l2_misfit_noalloc: n = " + std::to_string(n) +
                            " terms per component is outside [0, " +
                            std::to_string((long long)PSV_NOALLOC_MAX_TERMS) + "]");
    need(misfit, "misfit");
    if (n > 0) {
      need(syn_ux, "syn_ux");
      need(syn_uz, "syn_uz");
      need(obs_ux, "obs_ux");
      need(obs_uz, "obs_uz");
    }
    hipError_t e = psv::launch_misfit_noalloc(syn_ux, syn_uz, obs_ux, obs_uz, n, 0.5 * dt,
                                              misfit, static_cast<hipStream_t>(stream));
    if (e != hipSuccess)
      throw std::runtime_error(std::string("psv_l2_misfit_noalloc: ") + hipGetErrorString(e));
  });
}

int psv_calculate_l2_adjoint_sources(psv_model *m) {
  return guard([&] { M(m).calculate_l2_adjoint_sources(); });
}

int psv_reset_kernels(psv_model *m) {
  return guard([&] { M(m).reset_kernels(); });
}

int psv_map_kernels_to_velocity(psv_model *m) {
  return guard([&] { M(m).map_kernels_to_velocity(); });
}

int psv_run_model(psv_model *m, int verbose, int simulate_adjoint) {
  return guard([&] { M(m).run_model(verbose != 0, simulate_adjoint != 0); });
}

int psv_get_parameter_fields(const psv_model *m, double *vp, double *vs,
                             double *rho) {
  return guard([&] {
    need(vp, "vp");
    need(vs, "vs");
    need(rho, "rho");
    M(m).get_parameter_fields(vp, vs, rho);
  });
}

int psv_set_parameter_fields(psv_model *m, const double *vp, const double *vs,
                             const double *rho) {
  return guard([&] {
    need(vp, "vp");
    need(vs, "vs");
    need(rho, "rho");
    M(m).set_parameter_fields(vp, vs, rho);
  });
}

int psv_get_kernels(const psv_model *m, double *a, double *b, double *c) {
  return guard([&] {
    need(a, "vp_kernel");
    need(b, "vs_kernel");
    need(c, "rho_kernel");
    M(m).get_kernels(a, b, c);
  });
}

int psv_get_lame_kernels(const psv_model *m, double *a, double *b, double *c) {
  return guard([&] {
    need(a, "lambda_kernel");
    need(b, "mu_kernel");
    need(c, "density_l_kernel");
    M(m).get_lame_kernels(a, b, c);
  });
}

#define PSV_TRACES(get_name, set_name, which)                               \
  int get_name(const psv_model *m, double *ux, double *uz) {                \
    return guard([&] {                                                      \
      need(ux, "ux");                                                       \
      need(uz, "uz");                                                       \
      M(m).get_traces(which, ux, uz);                                       \
    });                                                                     \
  }                                                                         \
  int set_name(psv_model *m, const double *ux, const double *uz) {          \
    return guard([&] {                                                      \
      need(ux, "ux");                                                       \
      need(uz, "uz");                                                       \
      M(m).set_traces(which, ux, uz);                                       \
    });                                                                     \
  }
PSV_TRACES(psv_get_synthetic_data, psv_set_synthetic_data, 0)
PSV_TRACES(psv_get_observed_data, psv_set_observed_data, 1)
#undef PSV_TRACES

int psv_get_adjoint_sources(const psv_model *m, double *ux, double *uz) {
  return guard([&] {
    need(ux, "ux");
    need(uz, "uz");
    M(m).get_traces(2, ux, uz);
  });
}

int psv_get_snapshots(const psv_model *m, double *vx, double *vz, double *txx,
                      double *tzz, double *txz) {
  return guard([&] {
    double *out[5] = {vx, vz, txx, tzz, txz};
    for (auto p : out) need(p, "snapshot buffer");
    M(m).get_snapshots(out);
  });
}

int psv_get_snapshot(const psv_model *m, int i_shot, int i_snapshot,
                     double *vx, double *vz, double *txx, double *tzz,
                     double *txz) {
  return guard([&] {
    double *out[5] = {vx, vz, txx, tzz, txz};
    for (auto p : out) need(p, "snapshot buffer");
    M(m).get_snapshot(i_shot, i_snapshot, out);
  });
}

static void check_len(const psv_model *m, int64_t n) {
  if (n != M(m).derived().free_parameters)
    throw psv::ValueError("vector length " + std::to_string(n) +
                          " does not match free_parameters " +
                          std::to_string(M(m).derived().free_parameters));
}

int psv_get_model_vector(const psv_model *m, double *v, int64_t n) {
  return guard([&] {
    need(v, "vector");
    check_len(m, n);
    M(m).get_model_vector(v);
  });
}

int psv_set_model_vector(psv_model *m, const double *v, int64_t n) {
  return guard([&] {
    need(v, "vector");
    check_len(m, n);
    M(m).set_model_vector(v);
  });
}

int psv_get_gradient_vector(const psv_model *m, double *v, int64_t n) {
  return guard([&] {
    need(v, "vector");
    check_len(m, n);
    M(m).get_gradient_vector(v);
  });
}

int psv_load_vector(const psv_model *m, const char *path, int verbose,
                    double *v, int64_t n) {
  return guard([&] {
    need(path, "path");
    need(v, "vector");
    const auto &d = M(m).derived();
    int64_t want = (int64_t)3 * d.nx_free_parameters * d.nz_free_parameters;
    if (n != want)
      throw psv::ValueError("load_vector buffer must hold 3 * nx_free * nz_free "
                            "values");
    std::vector<double> r = M(m).load_vector(path, verbose != 0);
    std::memcpy(v, r.data(), r.size() * sizeof(double));
  });
}

int psv_get_setup(const psv_model *m, double *stf, double *moment,
                  double *taper) {
  return guard([&] { M(m).get_setup(stf, moment, taper); });
}

int psv_write_receivers(const psv_model *m, const char *prefix) {
  (void)prefix;
  return guard([&] { M(m).write_receivers(); });
}

int psv_write_sources(const psv_model *m) {
  return guard([&] { M(m).write_sources(); });
}

int psv_write_kernels(const psv_model *m) {
  return guard([&] { M(m).write_kernels(); });
}

int psv_load_receivers(psv_model *m, int verbose) {
  return guard([&] { M(m).load_receivers(verbose != 0); });
}

int psv_load_model(psv_model *m, const char *de_path, const char *vp_path,
                   const char *vs_path, int verbose) {
  return guard([&] {
    need(de_path, "de_path");
    need(vp_path, "vp_path");
    need(vs_path, "vs_path");
    M(m).load_model(de_path, vp_path, vs_path, verbose != 0);
  });
}

int psv_set_step_variant(psv_model *m, int variant, int xcd_remap) {
  return guard([&] { M(m).set_step_variant(variant, xcd_remap); });
}

int psv_set_graph_mode(psv_model *m, int on) {
  return guard([&] { M(m).set_graph_mode(on != 0); });
}

int psv_set_async(psv_model *m, int on) {
  return guard([&] { M(m).set_async(on != 0); });
}

int psv_synchronize(psv_model *m) {
  return guard([&] { M(m); });
}

int psv_set_checkpointing(psv_model *m, int every) {
  return guard([&] { M(m).set_checkpointing(every); });
}

int psv_set_timing(psv_model *m, int stride) {
  return guard([&] {
    M(m).set_timing(stride < 0 ? -1 : stride);
  });
}

int psv_get_step_timing(psv_model *m, double *ms, int64_t *timed,
                        int64_t *launches) {
  return guard([&] {
    need(ms, "ms");
    need(timed, "timed");
    need(launches, "launches");
    M(m).step_timing(ms, timed, launches);
  });
}

int psv_reset_step_timing(psv_model *m) {
  return guard([&] { M(m).reset_step_timing(); });
}

int psv_device_bytes(const psv_model *m, int64_t *bytes) {
  return guard([&] {
    need(bytes, "bytes");
    *bytes = M(m).device_bytes();
  });
}

int psv_stream_pressure(int *live_streams, int *hw_queues) {
  return guard([&] {
    need(live_streams, "live_streams");
    need(hw_queues, "hw_queues");
    *live_streams = psv::live_streams();
    *hw_queues = psv::hw_queue_limit();
  });
}

int psv_calibrate_copy(int64_t bytes, int iters, int runtime_copy, double *ms) {
  return guard([&] {
    need(ms, "ms");
    if (bytes <= 0 || bytes % 16 != 0 || iters < 1)
      throw psv::ValueError("bytes must be a positive multiple of 16, iters >= 1");
    auto chk = [](hipError_t e, const char *w) {
      if (e != hipSuccess) throw psv::RuntimeError(std::string(w) + ": " + hipGetErrorString(e));
    };
    double *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
      if (s) (void)hipStreamSynchronize(s);
      if (a) (void)hipFree(a);
      if (b) (void)hipFree(b);
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      if (s) (void)hipStreamDestroy(s);
    };
    try {
      long long n = bytes / 8;
      chk(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
      chk(hipMalloc((void **)&a, bytes), "hipMalloc");
      chk(hipMalloc((void **)&b, bytes), "hipMalloc");
      chk(psv::launch_fill(a, n, 1.0, s), "fill");
      auto copy = [&]() {
        return runtime_copy ? hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, s)
                            : psv::launch_copy(a, b, n, s);
      };
      chk(copy(), "copy");  // warm up
      chk(hipEventCreate(&e0), "hipEventCreate");
      chk(hipEventCreate(&e1), "hipEventCreate");
      chk(hipEventRecord(e0, s), "hipEventRecord");
      for (int i = 0; i < iters; ++i) chk(copy(), "copy");
      chk(hipEventRecord(e1, s), "hipEventRecord");
      chk(hipEventSynchronize(e1), "hipEventSynchronize");
      float t = 0.f;
      chk(hipEventElapsedTime(&t, e0, e1), "hipEventElapsedTime");
      *ms = t / iters;
    } catch (...) {
      cleanup();
      throw;
    }
    cleanup();
  });
}

struct psv_batch {
  psv::Batch *impl;
};

static psv::Batch &B(psv_batch *b) {
  if (!b || !b->impl) throw psv::ValueError("null batch handle");
  return *b->impl;
}
static const psv::Batch &B(const psv_batch *b) {
  if (!b || !b->impl) throw psv::ValueError("null batch handle");
  return *b->impl;
}

int psv_batch_create(int n, const psv_params *params, psv_batch **out) {
  return guard([&] {
    need(out, "out");
    if (n < 1) throw psv::ValueError("a batch needs at least one model");
    need(params, "params");
    std::vector<psv::Config> cfgs;
    for (int m = 0; m < n; ++m) {
      psv::Config c = config_from_params(&params[m]);
      psv::derive(c);  // validate before any device allocation
      cfgs.push_back(c);
    }
    psv_batch *h = new psv_batch;
    try {
      h->impl = new psv::Batch(cfgs);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  });
}

void psv_batch_destroy(psv_batch *b) {
  if (!b) return;
  delete b->impl;
  delete b;
}

int psv_batch_size(const psv_batch *b, int *n) {
  return guard([&] {
    need(n, "n");
    *n = B(b).size();
  });
}

int psv_batch_get_info(const psv_batch *b, int m, psv_info *out) {
  return guard([&] {
    need(out, "out");
    const psv::Model &mm = B(b).model(m);
    fill_info(mm.config(), mm.derived(), mm.misfit(), out);
  });
}

int psv_batch_set_parameter_fields(psv_batch *b, int m, const double *vp,
                                   const double *vs, const double *rho) {
  return guard([&] {
    need(vp, "vp");
    need(vs, "vs");
    need(rho, "rho");
    B(b).model(m).set_parameter_fields(vp, vs, rho);
  });
}

int psv_batch_set_observed_data(psv_batch *b, int m, const double *ux,
                                const double *uz) {
  return guard([&] {
    need(ux, "ux");
    need(uz, "uz");
    B(b).model(m).set_traces(1, ux, uz);
  });
}

int psv_batch_get_synthetic_data(const psv_batch *b, int m, double *ux,
                                 double *uz) {
  return guard([&] {
    need(ux, "ux");
    need(uz, "uz");
    B(b).model(m).get_traces(0, ux, uz);
  });
}

int psv_batch_forward(psv_batch *b, int i_shot) {
  return guard([&] { B(b).forward(i_shot); });
}

int psv_batch_forward_all(psv_batch *b) {
  return guard([&] { B(b).forward_all(); });
}

int psv_batch_lockstep(psv_batch *b, int *on) {
  return guard([&] {
    need(on, "on");
    *on = B(b).lockstep_available() ? 1 : 0;
  });
}

int psv_batch_misfits(psv_batch *b, double *out) {
  return guard([&] {
    need(out, "out");
    B(b).misfits(out);
  });
}

int psv_batch_set_persistent(psv_batch *b, int on) {
  return guard([&] { B(b).set_persistent(on != 0); });
}

int psv_batch_gradients(psv_batch *b) {
  return guard([&] { B(b).gradients(); });
}

int psv_batch_get_kernels(const psv_batch *b, int m, double *vp_k, double *vs_k,
                          double *rho_k) {
  return guard([&] {
    need(vp_k, "vp_k");
    need(vs_k, "vs_k");
    need(rho_k, "rho_k");
    B(b).model(m).get_kernels(vp_k, vs_k, rho_k);
  });
}

int psv_batch_set_variant(psv_batch *b, int variant, int xcd_remap) {
  return guard([&] { B(b).set_variant(variant, xcd_remap); });
}

int psv_batch_set_graph_mode(psv_batch *b, int on) {
  return guard([&] { B(b).set_graph_mode(on != 0); });
}

int psv_batch_autotune(psv_batch *b, int steps, int *best,
                       double *ms_per_step) {
  return guard([&] {
    need(best, "best");
    B(b).autotune(steps, best, ms_per_step);
  });
}

int psv_batch_timing(psv_batch *b, double *loop_ms, int64_t *points,
                     int reset) {
  return guard([&] {
    need(loop_ms, "loop_ms");
    need(points, "points");
    *loop_ms = B(b).loop_ms();
    *points = B(b).grid_points();
    if (reset) B(b).reset_timing();
  });
}

}  // extern "C"

int psv_conv2d_bf16(const void *in, const void *weights, const float *bias, void *out,
                    int out_f32, int h, int w, int cin, int cout, int ks, int transposed,
                    int gelu, int out_stride, int out_offset, void *stream) {
  return guard([&] {
    need(in, "in");
    need(weights, "weights");
    need(bias, "bias");
    need(out, "out");
    if (cin <= 0 || cin % 8 != 0) throw psv::ValueError("cin must be a positive multiple of 8");
    if (((uintptr_t)in | (uintptr_t)weights) & 15)
      throw psv::ValueError("in and weights must be 16-byte aligned");
    if (ks != 1 && ks != 3) throw psv::ValueError("ks must be 1 or 3");
    if (transposed && ks != 1) throw psv::ValueError("transposed convolutions use ks = 1");
    if (h <= 0 || w <= 0 || cout <= 0) throw psv::ValueError("empty convolution");
    if (out_offset < 0 || out_offset + cout > out_stride)
      throw psv::ValueError("output channels exceed out_stride");
    hipError_t e = psv::launch_conv_bf16(
        (const uint16_t *)in, (const uint16_t *)weights, bias, out, out_f32 != 0, h, w, cin,
        cout, ks, transposed != 0, gelu != 0, out_stride, out_offset, (hipStream_t)stream);
    if (e != hipSuccess) throw psv::RuntimeError(std::string("conv2d: ") + hipGetErrorString(e));
  });
}

int psv_maxpool2_bf16(const void *in, void *out, int h, int w, int c, void *stream) {
  return guard([&] {
    need(in, "in");
    need(out, "out");
    if (h <= 0 || w <= 0 || h % 2 || w % 2 || c <= 0)
      throw psv::ValueError("maxpool2 needs positive even h, w and c > 0");
    hipError_t e = psv::launch_maxpool2_bf16((const uint16_t *)in, (uint16_t *)out, h, w, c,
                                             (hipStream_t)stream);
    if (e != hipSuccess) throw psv::RuntimeError(std::string("maxpool2: ") + hipGetErrorString(e));
  });
}

int psv_conv_bf16(const void *in, int in_stride, const void *weights, const float *bias,
                  void *out, int out_f32, int out_stride, int out_offset, void *pre_out,
                  const void *dz, int dz_stride, int batch, int h, int w, int cin, int cout,
                  int ks, int mode, int gelu, void *stream) {
  return guard([&] {
    need(in, "in");
    need(weights, "weights");
    need(out, "out");
    if (mode < psv::kConv || mode > psv::kGather2x2)
      throw psv::ValueError("mode must be 0 (convolution), 1 (transposed) or 2 (2x2 gather)");
    if (cin <= 0 || cin % 8 != 0) throw psv::ValueError("cin must be a positive multiple of 8");
    if (in_stride < cin || in_stride % 8 != 0)
      throw psv::ValueError("in_stride must be a multiple of 8 and at least cin");
    if (((uintptr_t)in | (uintptr_t)weights) & 15)
      throw psv::ValueError("in and weights must be 16-byte aligned");
    if (mode == psv::kConv ? (ks != 1 && ks != 3) : ks != 1)
      throw psv::ValueError("ks must be 1 or 3 (1 for modes 1 and 2)");
    if (batch <= 0 || h <= 0 || w <= 0 || cout <= 0) throw psv::ValueError("empty convolution");
    if (out_offset < 0 || out_offset + cout > out_stride)
      throw psv::ValueError("output channels exceed out_stride");
    if (pre_out && mode != psv::kConv)
      throw psv::ValueError("pre-activations are kept for mode 0 only");
    if (dz && dz_stride < cout) throw psv::ValueError("dz_stride must be at least cout");
    psv::ConvGemm a{};
    a.in = (const uint16_t *)in;
    a.in_stride = in_stride;
    a.w = (const uint16_t *)weights;
    a.bias = bias;
    a.out16 = out_f32 ? nullptr : (uint16_t *)out;
    a.out32 = out_f32 ? (float *)out : nullptr;
    a.out_stride = out_stride;
    a.out_off = out_offset;
    a.pre16 = (uint16_t *)pre_out;
    a.dz = (const uint16_t *)dz;
    a.dz_stride = dz_stride;
    a.rows = batch * h;
    a.H = h;
    a.W = w;
    a.cin = cin;
    a.cout = cout;
    a.ks = ks;
    a.mode = mode;
    a.gelu = gelu ? 1 : 0;
    hipError_t e = psv::launch_conv_gemm(a, (hipStream_t)stream);
    if (e != hipSuccess) throw psv::RuntimeError(std::string("conv: ") + hipGetErrorString(e));
  });
}

int psv_conv_wgrad_bf16(const void *dy, int dy_stride, int dy_gather2x2, int dy_channels,
                        const void *x, int x_stride, int x_channels, int ks, int batch, int h,
                        int w, float *dw, float *db, float *workspace,
                        int64_t workspace_floats, void *stream) {
  return guard([&] {
    need(dy, "dy");
    need(x, "x");
    need(dw, "dw");
    if (dy_channels <= 0 || dy_stride < dy_channels || dy_stride % 8)
      throw psv::ValueError("dy_stride must be a multiple of 8 and at least dy_channels > 0");
    if (x_channels <= 0 || x_channels % 8 || x_stride < x_channels || x_stride % 8)
      throw psv::ValueError("x_channels and x_stride must be multiples of 8, x_stride >= "
                            "x_channels");
    if (dy_gather2x2 && dy_channels % 8)
      throw psv::ValueError("the 2x2 gather needs dy_channels a multiple of 8");
    if (((uintptr_t)dy | (uintptr_t)x) & 15)
      throw psv::ValueError("dy and x must be 16-byte aligned");
    if (ks != 1 && ks != 3) throw psv::ValueError("ks must be 1 or 3");
    if (dy_gather2x2 && ks != 1) throw psv::ValueError("the 2x2 gather uses ks = 1");
    if (batch <= 0 || h <= 0 || w <= 0) throw psv::ValueError("empty weight gradient");
    if (workspace_floats < 0) throw psv::ValueError("negative workspace size");
    psv::Wgrad a{};
    a.l = (const uint16_t *)dy;
    a.l_stride = dy_stride;
    a.l_s2d = dy_gather2x2 ? 1 : 0;
    a.lc = dy_channels;
    a.I = dy_gather2x2 ? 4 * dy_channels : dy_channels;
    a.r = (const uint16_t *)x;
    a.r_stride = x_stride;
    a.rc = x_channels;
    a.ks = ks;
    a.J = ks * ks * x_channels;
    a.rows = batch * h;
    a.H = h;
    a.W = w;
    hipError_t e = psv::launch_wgrad(a, dw, db, workspace, workspace ? workspace_floats : 0,
                                     (hipStream_t)stream);
    if (e != hipSuccess) throw psv::RuntimeError(std::string("wgrad: ") + hipGetErrorString(e));
  });
}

int psv_channel_sum_bf16(const void *src, int stride, int offset, int channels,
                         int64_t pixels, float *out, float *workspace,
                         int64_t workspace_floats, void *stream) {
  return guard([&] {
    need(src, "src");
    need(out, "out");
    if (channels <= 0 || pixels <= 0 || offset < 0 || offset + channels > stride)
      throw psv::ValueError("channel_sum needs 0 <= offset, offset + channels <= stride, "
                            "pixels > 0");
    hipError_t e = psv::launch_channel_sum((const uint16_t *)src, stride, offset, channels,
                                           pixels, out, workspace,
                                           workspace ? workspace_floats : 0,
                                           (hipStream_t)stream);
    if (e != hipSuccess)
      throw psv::RuntimeError(std::string("channel_sum: ") + hipGetErrorString(e));
  });
}

int psv_maxpool2_backward_bf16(const void *act, const void *z, const float *dpool,
                               const void *dskip, int skip_stride, int skip_offset, void *out,
                               int batch, int h, int w, int c, void *stream) {
  return guard([&] {
    need(z, "z");
    need(out, "out");
    if (dpool) need(act, "act");
    if (batch <= 0 || h <= 0 || w <= 0 || h % 2 || w % 2 || c <= 0)
      throw psv::ValueError("maxpool2_backward needs positive even h, w and c > 0");
    if (dskip && (skip_offset < 0 || skip_offset + c > skip_stride))
      throw psv::ValueError("skip channels exceed skip_stride");
    hipError_t e = psv::launch_pool_backward(
        (const uint16_t *)act, (const uint16_t *)z, dpool, (const uint16_t *)dskip,
        skip_stride, skip_offset, (uint16_t *)out, batch * h, w, c, (hipStream_t)stream);
    if (e != hipSuccess)
      throw psv::RuntimeError(std::string("maxpool2_backward: ") + hipGetErrorString(e));
  });
}
