// psv_model.hip -- host runtime of the MI355X fdModel.
//
// One Model owns every array of the reference class (src/fdModel.h:209-252)
// in HBM, plus the injection tables and ping-pong state buffers the fused
// step kernel needs.  All work is enqueued on the model's own stream; each
// public call synchronises before returning.
#include "psv_model.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>

namespace psv {

namespace {

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw RuntimeError(std::string(what) + ": " + hipGetErrorString(e));
}

long long round_up(long long v, long long m) { return (v + m - 1) / m * m; }

std::string zero_pad(int num, int pad) {
  std::ostringstream ss;
  ss << std::setw(pad) << std::setfill('0') << num;
  return ss.str();
}

std::atomic<int> g_live_streams{0};

}  // namespace

void create_stream(hipStream_t *s) {
  hip_check(hipStreamCreateWithFlags(s, hipStreamNonBlocking), "hipStreamCreate");
  ++g_live_streams;
}

void destroy_stream(hipStream_t s) {
  if (!s) return;
  (void)hipStreamDestroy(s);
  --g_live_streams;
}

int live_streams() { return g_live_streams.load(); }

int hw_queue_limit() {
  const char *v = std::getenv("GPU_MAX_HW_QUEUES");
  if (!v || !*v) return 4;
  char *end = nullptr;
  const long n = std::strtol(v, &end, 10);
  return (*end == '\0' && n > 0 && n < 1024) ? (int)n : 4;
}

Model::Model(const Config &c, hipStream_t shared_stream) : cfg_(c), d_(derive(c)) {
  if (shared_stream) {
    stream_ = shared_stream;
    owns_stream_ = false;
  }
  lean_ok_ = lean_eligible();
  if (lean_ok_) variant_ = kLeanTz8RingClamp;
  allocate();
  // homogeneous background then Lame fields (src/fdModel.cpp:248-261)
  hip_check(launch_fill(vp_, ngrid_, cfg_.scalar_vp, stream_), "fill vp");
  hip_check(launch_fill(vs_, ngrid_, cfg_.scalar_vs, stream_), "fill vs");
  hip_check(launch_fill(rho_, ngrid_, cfg_.scalar_rho, stream_), "fill rho");
  hip_check(launch_update_from_velocity(vp_, vs_, rho_, lm_, mu_, b_, d_.nx,
                                        d_.nz, pitch_, origin_, stream_),
            "update_from_velocity");
  sync();
}

Model::Model(const Model &o) : cfg_(o.cfg_), d_(o.d_) {
  lean_ok_ = o.lean_ok_;
  variant_ = o.variant_;
  xcd_remap_ = o.xcd_remap_;
  allocate();
  auto cp = [&](double *dst, const double *src, long long n) {
    if (n > 0)
      hip_check(hipMemcpyAsync(dst, src, n * sizeof(double),
                               hipMemcpyDeviceToDevice, stream_),
                "copy");
  };
  for (int k = 0; k < 2; ++k)
    for (int f = 0; f < 5; ++f) cp(state_[k][f], o.state_[k][f], npad_);
  cp(lm_, o.lm_, npad_);
  cp(mu_, o.mu_, npad_);
  cp(b_, o.b_, npad_);
  for (auto pr : {std::make_pair(vp_, o.vp_), std::make_pair(vs_, o.vs_),
                  std::make_pair(rho_, o.rho_), std::make_pair(lam_k_, o.lam_k_),
                  std::make_pair(mu_k_, o.mu_k_), std::make_pair(rho_l_, o.rho_l_),
                  std::make_pair(vp_k_, o.vp_k_), std::make_pair(vs_k_, o.vs_k_),
                  std::make_pair(rho_v_, o.rho_v_)})
    cp(pr.first, pr.second, ngrid_);
  for (int k = 0; k < 2; ++k) {
    cp(syn_[k], o.syn_[k], nrec_);
    cp(obs_[k], o.obs_[k], nrec_);
    cp(adj_[k], o.adj_[k], nrec_);
  }
  ckpt_every_ = o.ckpt_every_;
  async_ = o.async_;
  if (o.snap_[0]) {
    ensure_snapshots();
    long long n = (long long)cfg_.n_shots * stored_snapshots() * ngrid_;
    for (int f = 0; f < 5; ++f) cp(snap_[f], o.snap_[f], n);
  }
  misfit_ = o.misfit_;
  sync();
}

Model::~Model() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  if (lane_stream_) (void)hipStreamSynchronize(lane_stream_);
  clear_graphs();
  for (auto e : events_) (void)hipEventDestroy(e);
  for (auto e : lane_events_) (void)hipEventDestroy(e);
  for (auto e : ckpt_events_) (void)hipEventDestroy(e);
  for (auto e : job_events_) (void)hipEventDestroy(e);
  if (ev_join_) (void)hipEventDestroy(ev_join_);
  destroy_stream(lane_stream_);
  for (void *p : allocations_) (void)hipFree(p);
  if (staging_) (void)hipFree(staging_);
  if (owns_stream_) destroy_stream(stream_);
}

double *Model::dalloc(size_t n) {
  void *p = nullptr;
  size_t bytes = std::max<size_t>(n, 1) * sizeof(double);
  hip_check(hipMalloc(&p, bytes), "hipMalloc");
  allocations_.push_back(p);
  device_bytes_ += (int64_t)bytes;
  hip_check(launch_fill((double *)p, (long long)std::max<size_t>(n, 1), 0.0,
                        stream_),
            "zero-fill");
  return (double *)p;
}

void Model::sync() const {
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
}

void Model::allocate() {
  if (!stream_) create_stream(&stream_);
  ngrid_ = (long long)d_.nx * d_.nz;
  nrec_ = (long long)cfg_.n_shots * cfg_.nr * cfg_.nt;
  // Padded stencil layout, large enough for any tile height in use.
  ntx_ = (int)((d_.nx + kTileX - 1) / kTileX);
  pitch_ = (int)round_up(kHaloX + (long long)ntx_ * kTileX + 8, 64);
  // rows down to the last tile's bottom plus the kernels' halo, for every
  // tile height up to kMaxTileHeight
  rows_ = kHaloZ + (int)round_up(d_.nz + kMaxTileHeight, 16) + 8;
  origin_ = (long long)kHaloZ * pitch_ + kHaloX;
  npad_ = (long long)rows_ * pitch_;
  for (int k = 0; k < 2; ++k)
    for (int f = 0; f < 5; ++f) state_[k][f] = dalloc(npad_);
  lm_ = dalloc(npad_);
  mu_ = dalloc(npad_);
  b_ = dalloc(npad_);
  vp_ = dalloc(ngrid_);
  vs_ = dalloc(ngrid_);
  rho_ = dalloc(ngrid_);
  lam_k_ = dalloc(ngrid_);
  mu_k_ = dalloc(ngrid_);
  rho_l_ = dalloc(ngrid_);
  vp_k_ = dalloc(ngrid_);
  vs_k_ = dalloc(ngrid_);
  rho_v_ = dalloc(ngrid_);
  for (int k = 0; k < 2; ++k) {
    syn_[k] = dalloc(nrec_);
    obs_[k] = dalloc(nrec_);
    adj_[k] = dalloc(nrec_);
  }
  stf_ = dalloc(d_.stf.size());
  upload(stf_, d_.stf.data(), d_.stf.size());
  taper_lut_ = dalloc(d_.taper_lut.size());
  upload(taper_lut_, d_.taper_lut.data(), d_.taper_lut.size());
  {
    // nesting count along one axis: min(np, 1 + distance to the nearer edge)
    auto count1 = [&](int i, int n) {
      return std::min(cfg_.np_boundary, 1 + std::min(i, n - 1 - i));
    };
    // padded like the stencil arrays (entry kHaloX + x, kHaloZ + z), so the
    // lean kernels stage a tile's segment with its halo by plain transfers;
    // the margins hold 1 and are never used as factors
    std::vector<double> tx(pitch_, 1.0), tz(rows_, 1.0);
    for (int x = 0; x < d_.nx; ++x) tx[kHaloX + x] = taper_of_count(count1(x, d_.nx));
    for (int z = 0; z < d_.nz; ++z) tz[kHaloZ + z] = taper_of_count(count1(z, d_.nz));
    taper_xp_ = dalloc(tx.size());
    upload(taper_xp_, tx.data(), tx.size());
    taper_zp_ = dalloc(tz.size());
    upload(taper_zp_, tz.data(), tz.size());
    taper_x_ = taper_xp_ + kHaloX;
    taper_z_ = taper_zp_ + kHaloZ;
  }
  std::vector<long long> ridx(std::max(cfg_.nr, 1), 0);
  for (int r = 0; r < cfg_.nr; ++r)
    ridx[r] = origin_ + (long long)d_.iz_rec[r] * pitch_ + d_.ix_rec[r];
  rec_idx_ = (long long *)dalloc(ridx.size());
  sync();
  hip_check(hipMemcpy(rec_idx_, ridx.data(), ridx.size() * sizeof(long long),
                      hipMemcpyHostToDevice),
            "upload receivers");
  misfit_buf_ = dalloc(2 * kMisfitBlocks + 2);
  build_injection_tables();
  sync();
}

// Injection terms in the reference's order, bucketed by owning tile (stable,
// so terms hitting one point keep their order).  Forward: 12 moment-tensor
// terms per source (src/fdModel.cpp:681-751); adjoint: one vx and one vz
// term per receiver (src/fdModel.cpp:951-958).
void Model::build_injection_tables() {
  tz_ = step_tile_height(variant_);
  ntz_ = (d_.nz + tz_ - 1) / tz_;
  ntiles_ = ntx_ * ntz_;
  struct Term {
    int x, z, comp, src;
    double coef;
  };
  auto bucket = [&](const std::vector<Term> &terms, std::vector<int> &off,
                    std::vector<InjEntry> &out, int base) {
    std::vector<int> count(ntiles_ + 1, 0);
    for (const Term &t : terms) count[(t.z / tz_) * ntx_ + t.x / kTileX]++;
    off.assign(ntiles_ + 1, 0);
    for (int i = 0; i < ntiles_; ++i) off[i + 1] = off[i] + count[i];
    std::vector<int> fill(off.begin(), off.end() - 1);
    std::vector<InjEntry> local(terms.size());
    for (const Term &t : terms) {
      int tile = (t.z / tz_) * ntx_ + t.x / kTileX;
      InjEntry e;
      e.local = (t.z % tz_) * kTileX + (t.x % kTileX);
      e.comp = t.comp;
      e.src = t.src;
      e.pad = 0;
      e.coef = t.coef;
      local[fill[tile]++] = e;
    }
    for (auto &o : off) o += base;
    out.insert(out.end(), local.begin(), local.end());
  };

  std::vector<int> fwd_off, off;
  std::vector<InjEntry> fwd_entries;
  for (int s = 0; s < cfg_.n_shots; ++s) {
    std::vector<Term> terms;
    for (int src : cfg_.which_source_to_fire_in_which_shot[s]) {
      int x = d_.ix_src[src], z = d_.iz_src[src];
      const double *M = &d_.moment[(size_t)src * 4];
      double q01 = 0.25 * M[1], q10 = 0.25 * M[2];
      terms.push_back({x - 1, z, 0, src, -M[0]});
      terms.push_back({x, z, 0, src, M[0]});
      terms.push_back({x, z - 1, 1, src, -M[3]});
      terms.push_back({x, z, 1, src, M[3]});
      terms.push_back({x - 1, z + 1, 0, src, q01});
      terms.push_back({x, z + 1, 0, src, q01});
      terms.push_back({x - 1, z - 1, 0, src, -q01});
      terms.push_back({x, z - 1, 0, src, -q01});
      terms.push_back({x + 1, z - 1, 1, src, q10});
      terms.push_back({x + 1, z, 1, src, q10});
      terms.push_back({x - 1, z - 1, 1, src, -q10});
      terms.push_back({x - 1, z, 1, src, -q10});
    }
    bucket(terms, off, fwd_entries, (int)fwd_entries.size());
    fwd_off.insert(fwd_off.end(), off.begin(), off.end());
  }
  std::vector<Term> aterms;
  for (int r = 0; r < cfg_.nr; ++r) {
    aterms.push_back({d_.ix_rec[r], d_.iz_rec[r], 0, r, 0.0});
    aterms.push_back({d_.ix_rec[r], d_.iz_rec[r], 1, r, 0.0});
  }
  std::vector<int> adj_off;
  std::vector<InjEntry> adj_entries;
  bucket(aterms, adj_off, adj_entries, 0);

  // Halo tables: every term whose point lies in a tile's 4-point halo
  // region, local index over the (64 + 8) x (tz + 8) region, in term order.
  auto halo_bucket = [&](const std::vector<Term> &terms, std::vector<int> &off,
                         std::vector<InjEntry> &out, int base) {
    std::vector<std::vector<InjEntry>> per(ntiles_);
    const int w4 = kTileX + 8;
    for (const Term &t : terms) {
      int tx0 = std::max(0, (t.x - 67 + kTileX * 1024) / kTileX - 1024);
      int tx1 = std::min(ntx_ - 1, (t.x + 4) / kTileX);
      int tz0 = std::max(0, (t.z - tz_ - 3 + tz_ * 1024) / tz_ - 1024);
      int tz1 = std::min(ntz_ - 1, (t.z + 4) / tz_);
      for (int a = tz0; a <= tz1; ++a)
        for (int b = tx0; b <= tx1; ++b) {
          int x0 = b * kTileX, z0 = a * tz_;
          if (t.x < x0 - 4 || t.x >= x0 + kTileX + 4 || t.z < z0 - 4 ||
              t.z >= z0 + tz_ + 4)
            continue;
          InjEntry e;
          e.local = (t.z - (z0 - 4)) * w4 + (t.x - (x0 - 4));
          e.comp = t.comp;
          e.src = t.src;
          e.pad = 0;
          e.coef = t.coef;
          per[a * ntx_ + b].push_back(e);
        }
    }
    off.assign(ntiles_ + 1, base);
    for (int i = 0; i < ntiles_; ++i) {
      off[i + 1] = off[i] + (int)per[i].size();
      out.insert(out.end(), per[i].begin(), per[i].end());
    }
  };
  std::vector<int> fwd_halo_off;
  std::vector<InjEntry> fwd_halo;
  for (int s = 0; s < cfg_.n_shots; ++s) {
    std::vector<Term> terms;
    for (int src : cfg_.which_source_to_fire_in_which_shot[s]) {
      int x = d_.ix_src[src], z = d_.iz_src[src];
      const double *M = &d_.moment[(size_t)src * 4];
      double q01 = 0.25 * M[1], q10 = 0.25 * M[2];
      terms.push_back({x - 1, z, 0, src, -M[0]});
      terms.push_back({x, z, 0, src, M[0]});
      terms.push_back({x, z - 1, 1, src, -M[3]});
      terms.push_back({x, z, 1, src, M[3]});
      terms.push_back({x - 1, z + 1, 0, src, q01});
      terms.push_back({x, z + 1, 0, src, q01});
      terms.push_back({x - 1, z - 1, 0, src, -q01});
      terms.push_back({x, z - 1, 0, src, -q01});
      terms.push_back({x + 1, z - 1, 1, src, q10});
      terms.push_back({x + 1, z, 1, src, q10});
      terms.push_back({x - 1, z - 1, 1, src, -q10});
      terms.push_back({x - 1, z, 1, src, -q10});
    }
    halo_bucket(terms, off, fwd_halo, (int)fwd_halo.size());
    fwd_halo_off.insert(fwd_halo_off.end(), off.begin(), off.end());
  }
  std::vector<int> adj_halo_off;
  std::vector<InjEntry> adj_halo;
  halo_bucket(aterms, adj_halo_off, adj_halo, 0);

  // (Re)upload; these small tables are reallocated when the tiling changes.
  auto put = [&](void *&dst, const void *src, size_t bytes) {
    if (dst) {
      hip_check(hipStreamSynchronize(stream_), "sync");
      auto it = std::find(allocations_.begin(), allocations_.end(), dst);
      if (it != allocations_.end()) allocations_.erase(it);
      hip_check(hipFree(dst), "hipFree");
    }
    dst = nullptr;
    hip_check(hipMalloc(&dst, std::max<size_t>(bytes, 16)), "hipMalloc");
    allocations_.push_back(dst);
    if (bytes)
      hip_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "upload");
  };
  if (fwd_off.empty()) fwd_off.assign(ntiles_ + 1, 0);
  if (fwd_halo_off.empty()) fwd_halo_off.assign(ntiles_ + 1, 0);
  put((void *&)fwd_off_, fwd_off.data(), fwd_off.size() * sizeof(int));
  put((void *&)fwd_entries_, fwd_entries.data(),
      fwd_entries.size() * sizeof(InjEntry));
  put((void *&)adj_off_, adj_off.data(), adj_off.size() * sizeof(int));
  put((void *&)adj_entries_, adj_entries.data(),
      adj_entries.size() * sizeof(InjEntry));
  put((void *&)fwd_halo_off_, fwd_halo_off.data(),
      fwd_halo_off.size() * sizeof(int));
  put((void *&)fwd_halo_, fwd_halo.data(), fwd_halo.size() * sizeof(InjEntry));
  put((void *&)adj_halo_off_, adj_halo_off.data(),
      adj_halo_off.size() * sizeof(int));
  put((void *&)adj_halo_, adj_halo.data(), adj_halo.size() * sizeof(InjEntry));
}

// Taper of a point with nesting count c (taper_at in the kernels).
double Model::taper_of_count(int c) const {
  return c >= cfg_.np_boundary ? 1.0 : d_.taper_lut[c];
}

// The lean kernels divide with FastDiv and form the taper as the minimum of
// a column and a row factor; both are exact only for spacings in
// [2^-40, 2^40] and a taper nondecreasing in the nesting count.
bool Model::lean_eligible() const {
  if (!fast_division_ok(cfg_.dx) || !fast_division_ok(cfg_.dz)) return false;
  for (int c = 0; c < cfg_.np_boundary; ++c)
    if (!(taper_of_count(c) <= taper_of_count(c + 1))) return false;
  return true;
}

void Model::set_step_variant(int variant, int xcd_remap) {
  settle();
  if (variant < 0 || variant >= kNumVariants)
    throw ValueError("unknown step variant");
  if (step_uses_fastdiv(variant) && !lean_ok_)
    throw ValueError("step variants 8-20 need dx and dz in [2^-40, 2^40] "
                     "and a taper that grows towards the interior");
  variant_ = variant;
  xcd_remap_ = xcd_remap ? 1 : 0;
  clear_graphs();  // graphs hold pointers into the injection tables
  build_injection_tables();
}

void Model::clear_graphs() {
  for (auto &kv : graphs_) (void)hipGraphExecDestroy(kv.second);
  graphs_.clear();
}

void Model::set_graph_mode(bool on) {
  settle();
  if (!on) clear_graphs();
  graph_mode_ = on;
}

void Model::ensure_snapshots() {
  if (snap_[0]) return;
  clear_graphs();
  long long n = (long long)cfg_.n_shots * stored_snapshots() * ngrid_;
  for (int f = 0; f < 5; ++f) snap_[f] = dalloc(n);
}

void Model::upload(double *dst, const double *src, size_t n) const {
  sync();  // fills and kernels on the model stream may still touch dst
  if (n)
    hip_check(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyHostToDevice),
              "hipMemcpy H2D");
}

void Model::download(double *dst, const double *src, size_t n) const {
  sync();
  if (n)
    hip_check(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost),
              "hipMemcpy D2H");
}

double *Model::staging(size_t n) const {
  if (n > staging_n_) {
    sync();
    if (staging_) (void)hipFree(staging_);
    staging_ = nullptr;
    hip_check(hipMalloc((void **)&staging_, n * sizeof(double)),
              "hipMalloc staging");
    staging_n_ = n;
  }
  return staging_;
}

void Model::to_device_layout(double *dst, const double *host_ref) const {
  double *s = staging(ngrid_);
  upload(s, host_ref, ngrid_);
  hip_check(launch_transpose(s, dst, d_.nx, d_.nz, 1, 0, stream_), "transpose");
  sync();
}

void Model::to_host_layout(double *host_ref, const double *src,
                           int slices) const {
  double *s = staging((size_t)ngrid_ * slices);
  hip_check(launch_transpose(src, s, d_.nx, d_.nz, slices, 1, stream_),
            "transpose");
  download(host_ref, s, (size_t)ngrid_ * slices);
}

void Model::check_shot(int i_shot) const {
  if (i_shot < 0 || i_shot >= cfg_.n_shots)
    throw ValueError("i_shot " + std::to_string(i_shot) +
                     " is outside [0, n_shots)");
}

StepArgs Model::base_args() const {
  StepArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nx = d_.nx;
  a.nz = d_.nz;
  a.pitch = pitch_;
  a.ntx = ntx_;
  a.ntz = ntz_;
  a.tz = tz_;
  a.origin = origin_;
  a.lm = lm_;
  a.mu = mu_;
  a.b = b_;
  a.taper_lut = taper_lut_;
  a.np = cfg_.np_boundary;
  a.dt = cfg_.dt;
  a.dx = cfg_.dx;
  a.dz = cfg_.dz;
  a.dxdz = cfg_.dx * cfg_.dz;
  a.d4x = cfg_.dx * cfg_.dx * cfg_.dx * cfg_.dx;
  a.d4z = cfg_.dz * cfg_.dz * cfg_.dz * cfg_.dz;
  a.rdx = 1.0 / cfg_.dx;
  a.rdz = 1.0 / cfg_.dz;
  if (lean_ok_) {
    fast_division_constants(cfg_.dx, &a.zhx, &a.zlx);
    fast_division_constants(cfg_.dz, &a.zhz, &a.zlz);
  }
  a.taper_x = taper_x_;
  a.taper_z = taper_z_;
  a.nt = cfg_.nt;
  a.lam_k = lam_k_;
  a.mu_k = mu_k_;
  a.rho_k = rho_l_;
  a.kx0 = cfg_.np_boundary + cfg_.nx_inner_boundary;
  a.kx1 = cfg_.np_boundary + cfg_.nx_inner - cfg_.nx_inner_boundary;
  a.kz0 = cfg_.np_boundary + cfg_.nz_inner_boundary;
  a.kz1 = cfg_.np_boundary + cfg_.nz_inner - cfg_.nz_inner_boundary;
  a.wcorr = cfg_.snapshot_interval * cfg_.dt;
  a.stf = stf_;
  a.nr = cfg_.nr;
  a.rec_idx = rec_idx_;
  a.xcd_remap = xcd_remap_;
  a.ntiles = ntiles_;
  return a;
}

// Enqueues the launches of loop indices [k0, k1) (time step it = k
// forward, nt - 1 - k adjoint) on state buffer sets st[2][5], starting from
// set `parity`; returns the set holding the state afterwards.  One launch per
// step (two for the split variants) or per pair of steps (two-step variants;
// an unpaired step uses the matching one-step kernel).  `slot(it, ptrs)`
// decides whether snapshot step it is stored / correlated and where.  With
// `pairs`, every timing_stride-th launch records its kernel start and end;
// with `dump`, the reference's text snapshot is written every 10 steps.
int Model::enqueue_range(StepArgs &a, double *(*st)[5], int parity,
                         bool forward, int k0, int k1, bool snaps,
                         const SlotFn &slot,
                         std::vector<std::pair<hipEvent_t, hipEvent_t>> *pairs,
                         bool dump, hipStream_t s, double *defer) {
  hipStream_t ls = s ? s : stream_;
  const bool inplace = step_is_inplace(variant_);
  const bool two = step_is_two_step(variant_) && !dump;
  const int single = (two || dump) ? step_single_variant(variant_) : variant_;
  const int nt = cfg_.nt, si = cfg_.snapshot_interval;
  // timing events continue from event_cursor_, which the caller resets once
  // per timed loop: the segments of a checkpointed adjoint share one loop
  auto event = [&]() {
    if (event_cursor_ == events_.size()) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      events_.push_back(e);
    }
    return events_[event_cursor_++];
  };
  for (int k = k0; k < k1;) {
    const int pair = two && k + 1 < k1;
    const int cur = inplace ? 0 : parity;
    const int nxt = inplace ? 0 : 1 - parity;
    a.vx_i = st[cur][0];
    a.vz_i = st[cur][1];
    a.txx_i = st[cur][2];
    a.tzz_i = st[cur][3];
    a.txz_i = st[cur][4];
    a.vx_o = st[nxt][0];
    a.vz_o = st[nxt][1];
    a.txx_o = st[nxt][2];
    a.tzz_o = st[nxt][3];
    a.txz_o = st[nxt][4];
    const int it = forward ? k : nt - 1 - k;
    const int it1 = forward ? k + 1 : nt - 2 - k;
    a.it = it;
    a.do_snap = snaps && (it % si == 0) && slot(it, a.snap);
    a.it1 = it1;
    a.do_snap1 = pair && snaps && (it1 % si == 0) && slot(it1, a.snap1);
    a.kfresh = 0;
    if (defer && a.do_snap && !forward) {
      // deferred correlation (one-step kernels only): this snapshot slot's
      // terms go to their own slice instead of the kernels, written without
      // being read (as if the slice started at zero), so the slice needs no
      // clearing; the replay reads only the kernel region, which every
      // correlation step writes in full
      const long long off = (long long)(it / si) * 3 * ngrid_;
      a.rho_k = defer + off;
      a.lam_k = defer + off + ngrid_;
      a.mu_k = defer + off + 2 * ngrid_;
      a.kfresh = 1;
    }
    const int covered = pair ? 2 : 1;
    bool timed = pairs && timing_stride_ > 0 && (launches_ / covered) % timing_stride_ == 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timed) {
      e0 = event();
      e1 = event();
      pairs->emplace_back(e0, e1);
    }
    if (pair)
      hip_check(launch_two_steps(variant_, a, ls, e0, e1), "two-step kernel");
    else
      hip_check(launch_step(single, a, ls, e0, e1), "step kernel");
    launches_ += covered;
    if (timed) timed_ += covered;
    if (!inplace) parity = nxt;
    if (dump && it % 10 == 0) {
      // src/fdModel.cpp:753-776: state after step it, default stream format
      std::vector<double> hx(ngrid_), hz(ngrid_);
      double *sb = staging(ngrid_);
      hip_check(launch_unpad(st[nxt][0], sb, d_.nx, d_.nz, pitch_, origin_,
                             ls),
                "unpad");
      to_host_layout(hx.data(), sb);
      hip_check(launch_unpad(st[nxt][1], sb, d_.nx, d_.nz, pitch_, origin_,
                             ls),
                "unpad");
      to_host_layout(hz.data(), sb);
      std::ofstream fx("snapshots/vx" + zero_pad(it, 5) + ".txt");
      std::ofstream fz("snapshots/vz" + zero_pad(it, 5) + ".txt");
      for (int ix = 0; ix < d_.nx; ++ix) {
        for (int iz = 0; iz < d_.nz; ++iz) {
          fx << hx[(size_t)ix * d_.nz + iz] << " ";
          fz << hz[(size_t)ix * d_.nz + iz] << " ";
        }
        fx << std::endl;
        fz << std::endl;
      }
    }
    k += covered;
  }
  return parity;
}

// Snapshot slot of time step it in the stored snapshots: every snapshot,
// or with checkpointing every ckpt_every_-th one.
bool Model::stored_slot(int i_shot, int it, double **dst) const {
  const int slot = it / cfg_.snapshot_interval;
  if (slot % ckpt_every_ != 0) return false;
  long long off =
      ((long long)i_shot * stored_snapshots() + slot / ckpt_every_) * ngrid_;
  for (int f = 0; f < 5; ++f) dst[f] = snap_[f] + off;
  return true;
}

int Model::stored_snapshots() const {
  return (d_.snapshots + ckpt_every_ - 1) / ckpt_every_;
}

// One whole time loop from a zero state (src/fdModel.cpp:530-542, 792-805;
// zeroed with fill kernels, which are safe inside a captured graph).
void Model::enqueue_loop(StepArgs &a, int i_shot, bool forward, bool store,
                         std::vector<std::pair<hipEvent_t, hipEvent_t>> *pairs,
                         bool dump) {
  const bool inplace = step_is_inplace(variant_);
  for (int k = 0; k < (inplace ? 1 : 2); ++k)
    for (int f = 0; f < 5; ++f)
      hip_check(launch_fill(state_[k][f], npad_, 0.0, stream_), "zero state");
  a.rec_increments = (forward && step_is_two_step(variant_)) ? 1 : 0;
  a.rec_off = adj_off_;
  a.rec_entries = adj_entries_;
  SlotFn slot = [this, i_shot](int it, double **dst) {
    return stored_slot(i_shot, it, dst);
  };
  enqueue_range(a, state_, 0, forward, 0, cfg_.nt, store || !forward, slot,
                pairs, dump);
  if (a.rec_increments)
    hip_check(launch_receiver_scan(a.rtf_x, a.rtf_z, cfg_.nr, cfg_.nt, stream_),
              "receiver scan");
}

// ---- asynchronous per-shot calls ----

enum { kJobForward, kJobAdjoint, kJobDeferred, kJobReplay };

void Model::set_async(bool on) {
  settle();
  async_ = on;
}

// The asynchronous path covers one-step, ping-pong kernels without timing,
// graphs, checkpointing, verbose output or wavefield dumps.
bool Model::async_path(bool verbose, bool dump) const {
  return async_ && !verbose && !dump && timing_stride_ == 0 && !graph_mode_ &&
         ckpt_every_ == 1 && !step_is_two_step(variant_) && !step_is_inplace(variant_);
}

// Lane Y: its stream and state buffers.  Allocation zero-fills on the model
// stream, so the first use waits for it.
void Model::ensure_lane() {
  if (!lane_stream_) create_stream(&lane_stream_);
  if (!ev_join_) hip_check(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming),
                           "hipEventCreate");
  if ((int)fwd_job_.size() < cfg_.n_shots) fwd_job_.assign(cfg_.n_shots, -1);
  ensure_second_state();
}

int Model::push_job(int lane, int kind, int shot, bool store, int dep0, int dep1) {
  LaneJob j;
  j.kind = kind;
  j.shot = shot;
  j.store = store;
  j.dep[0] = dep0;
  j.dep[1] = dep1;
  j.lane = lane;
  j.k = j.parity = 0;
  j.started = j.done = false;
  const int id = (int)jobs_.size();
  jobs_.push_back(j);
  while (job_events_.size() < jobs_.size()) {
    hipEvent_t e;
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    job_events_.push_back(e);
  }
  lane_queue_[lane].push_back(id);
  return id;
}

// Enqueues the jobs' time steps, 64 at a time per lane in turn: one round
// (all = false) or until every job is enqueued.  The oldest unfinished job
// never waits on a younger one, so every round makes progress.
void Model::pump(bool all) {
  constexpr int kChunk = 64;
  const int nt = cfg_.nt;
  for (;;) {
    bool progress = false;
    for (int l = 0; l < 2; ++l) {
      if (lane_queue_[l].empty()) continue;
      LaneJob &jb = jobs_[lane_queue_[l].front()];
      const int id = lane_queue_[l].front();
      if ((jb.dep[0] >= 0 && !jobs_[jb.dep[0]].done) ||
          (jb.dep[1] >= 0 && !jobs_[jb.dep[1]].done))
        continue;
      hipStream_t s = l ? lane_stream_ : stream_;
      double *(*st)[5] = l ? lane_state_ : state_;
      if (!jb.started) {
        for (int d : jb.dep)
          if (d >= 0 && jobs_[d].lane != l)
            hip_check(hipStreamWaitEvent(s, job_events_[d], 0), "hipStreamWaitEvent");
        jb.started = true;
        if (jb.kind == kJobReplay) {
          // the deferred adjoint's kernel terms, in its own order
          hip_check(launch_replay_correlation(
                        rho_l_, lam_k_, mu_k_, defer_, d_.snapshots, ngrid_, d_.nx,
                        cfg_.np_boundary + cfg_.nx_inner_boundary,
                        cfg_.np_boundary + cfg_.nx_inner - cfg_.nx_inner_boundary,
                        cfg_.np_boundary + cfg_.nz_inner_boundary,
                        cfg_.np_boundary + cfg_.nz_inner - cfg_.nz_inner_boundary, s),
                    "replay correlation");
          jb.k = nt;
        } else {
          jb.a = jb.kind == kJobForward ? forward_args(jb.shot) : adjoint_args(jb.shot);
          jb.a.rec_increments = 0;
          jb.a.rec_off = adj_off_;
          jb.a.rec_entries = adj_entries_;
          for (int p = 0; p < 2; ++p)
            for (int f = 0; f < 5; ++f)
              hip_check(launch_fill(st[p][f], npad_, 0.0, s), "zero state");
        }
      }
      if (jb.k < nt) {
        const int shot = jb.shot;
        SlotFn slot = [this, shot](int it, double **dst) {
          return stored_slot(shot, it, dst);
        };
        const bool fwd = jb.kind == kJobForward;
        const int k1 = std::min(jb.k + kChunk, nt);
        jb.parity = enqueue_range(jb.a, st, jb.parity, fwd, jb.k, k1, jb.store || !fwd, slot,
                                  nullptr, false, s,
                                  jb.kind == kJobDeferred ? defer_ : nullptr);
        jb.k = k1;
      }
      progress = true;
      if (jb.k == nt) {
        hip_check(hipEventRecord(job_events_[id], s), "hipEventRecord");
        jb.done = true;
        lane_queue_[l].pop_front();
      }
    }
    if (!all || !progress) break;
  }
}

// Enqueues every pending job, adds pending deferred terms and waits for the
// device.  Errors of the enqueued work surface here.
void Model::settle() {
  if (jobs_.empty()) return;
  if (deferred_job_ >= 0) {
    push_job(0, kJobReplay, -1, false, deferred_job_, -1);
    deferred_job_ = -1;
  }
  pump(true);
  bool y = false;
  for (const LaneJob &j : jobs_) y = y || j.lane == 1;
  jobs_.clear();
  lane_queue_[0].clear();
  lane_queue_[1].clear();
  std::fill(fwd_job_.begin(), fwd_job_.end(), -1);
  fwd_calls_ = adj_calls_ = 0;
  last_replay_job_ = -1;
  if (y) {
    hip_check(hipEventRecord(ev_join_, lane_stream_), "hipEventRecord");
    hip_check(hipStreamWaitEvent(stream_, ev_join_, 0), "hipStreamWaitEvent");
  }
  sync();
}

// Adjoint with checkpointing: segments of ckpt_every_ snapshot intervals
// are processed last to first.  For each, the forward state is restored
// from the segment's checkpoint into a second buffer pair, the segment's
// snapshots are recomputed into one of two small buffers (on the second
// stream, one segment ahead of the adjoint), and the adjoint steps of the
// segment correlate with them.  The recomputed forward runs the same
// kernels on the same state, so the kernels equal those of full storage
// bit for bit.
void Model::enqueue_checkpointed_adjoint(
    StepArgs &aa, int i_shot,
    std::vector<std::pair<hipEvent_t, hipEvent_t>> *pairs) {
  const int nt = cfg_.nt, si = cfg_.snapshot_interval, k = ckpt_every_;
  if (!state2_[0][0]) {
    for (int p = 0; p < 2; ++p)
      for (int f = 0; f < 5; ++f) state2_[p][f] = dalloc(npad_);
    // all-zero CSR offsets for the finest tiling (64 x 8 tiles)
    const long long max_tiles = (long long)ntx_ * ((d_.nz + 7) / 8);
    scratch_off_ = (int *)dalloc((size_t)(max_tiles + 2) / 2 + 1);
  }
  if (seg_k_ < k) {  // two segment buffers of k snapshot slots
    for (int f = 0; f < 5; ++f) {
      if (seg_[f]) {
        auto it = std::find(allocations_.begin(), allocations_.end(), (void *)seg_[f]);
        if (it != allocations_.end()) allocations_.erase(it);
        hip_check(hipFree(seg_[f]), "hipFree");
        device_bytes_ -= (int64_t)(2LL * seg_k_ * ngrid_ * sizeof(double));
      }
      seg_[f] = dalloc((size_t)2 * k * ngrid_);
    }
    seg_k_ = k;
  }
  if (!lane_stream_) create_stream(&lane_stream_);
  while (ckpt_events_.size() < 6) {
    hipEvent_t e;
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    ckpt_events_.push_back(e);
  }
  hipEvent_t ev_start = ckpt_events_[0], ev_join = ckpt_events_[1];
  hipEvent_t ev_ready[2] = {ckpt_events_[2], ckpt_events_[3]};  // segment recomputed
  hipEvent_t ev_done[2] = {ckpt_events_[4], ckpt_events_[5]};   // segment correlated
  const bool inplace = step_is_inplace(variant_);
  for (int p = 0; p < (inplace ? 1 : 2); ++p)
    for (int f = 0; f < 5; ++f)
      hip_check(launch_fill(state_[p][f], npad_, 0.0, stream_), "zero state");
  hip_check(hipEventRecord(ev_start, stream_), "hipEventRecord");
  hip_check(hipStreamWaitEvent(lane_stream_, ev_start, 0), "hipStreamWaitEvent");
  StepArgs fa = base_args();
  fa.adjoint = 0;
  fa.dts = cfg_.dt;
  fa.tile_off = fwd_off_ + (size_t)i_shot * (ntiles_ + 1);
  fa.entries = fwd_entries_;
  fa.halo_off = fwd_halo_off_ + (size_t)i_shot * (ntiles_ + 1);
  fa.halo_entries = fwd_halo_;
  fa.nr = 0;  // the recomputation records no receivers
  fa.rec_off = scratch_off_;  // all-zero CSR: no entries
  fa.rec_entries = adj_entries_;
  // Segment c is recomputed on lane_stream_ into buffer c % 2 while the
  // adjoint of segment c + 1 runs on stream_ from the other buffer; the
  // adjoint steps stay in order on stream_, so the kernels are those of one
  // stream bit for bit.
  const int L = k * si;
  const int nseg = (nt + L - 1) / L;
  int parity = 0;
  for (int c = nseg - 1; c >= 0; --c) {
    const int t0 = c * L, t1 = std::min(t0 + L, nt), b = c & 1;
    SlotFn seg = [this, c, k, si, b](int it, double **dst) {
      long long off = ((long long)b * k + (it / si - c * k)) * ngrid_;
      for (int f = 0; f < 5; ++f) dst[f] = seg_[f] + off;
      return true;
    };
    if (c + 2 < nseg)  // buffer b was read by the adjoint of segment c + 2
      hip_check(hipStreamWaitEvent(lane_stream_, ev_done[b], 0), "hipStreamWaitEvent");
    double *ck[5];
    stored_slot(i_shot, t0, ck);
    for (int f = 0; f < 5; ++f)
      hip_check(launch_pad(ck[f], state2_[0][f], d_.nx, d_.nz, pitch_, origin_,
                           lane_stream_),
                "restore checkpoint");
    enqueue_range(fa, state2_, 0, true, t0, t1, true, seg, nullptr, false, lane_stream_);
    hip_check(hipEventRecord(ev_ready[b], lane_stream_), "hipEventRecord");
    hip_check(hipStreamWaitEvent(stream_, ev_ready[b], 0), "hipStreamWaitEvent");
    parity = enqueue_range(aa, state_, parity, false, nt - t1, nt - t0, true,
                           seg, pairs, false);
    hip_check(hipEventRecord(ev_done[b], stream_), "hipEventRecord");
  }
  hip_check(hipEventRecord(ev_join, lane_stream_), "hipEventRecord");
  hip_check(hipStreamWaitEvent(stream_, ev_join, 0), "hipStreamWaitEvent");
}

// The time loop.  Plain mode launches step by step; graph mode replays a
// captured graph of the whole loop (with timing, bracketing the replay).
void Model::run_steps(StepArgs &a, int i_shot, bool forward, bool store,
                      bool dump) {
  const int nt = cfg_.nt;
  if (graph_mode_ && !dump) {
    long long key = ((long long)i_shot * 2 + (forward ? 1 : 0)) * 2 + (store ? 1 : 0);
    auto it = graphs_.find(key);
    if (it == graphs_.end()) {
      sync();
      hipGraph_t graph = nullptr;
      hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture");
      int64_t saved = launches_;
      try {
        enqueue_loop(a, i_shot, forward, store, nullptr, false);
      } catch (...) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(stream_, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        throw;
      }
      launches_ = saved;
      hip_check(hipStreamEndCapture(stream_, &graph), "hipStreamEndCapture");
      hipGraphExec_t exec = nullptr;
      hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      hip_check(e, "hipGraphInstantiate");
      it = graphs_.emplace(key, exec).first;
    }
    bool timed = timing_stride_ != 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (timed) {
      while (events_.size() < 2) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        events_.push_back(e);
      }
      e0 = events_[0];
      e1 = events_[1];
      hip_check(hipEventRecord(e0, stream_), "hipEventRecord");
    }
    hip_check(hipGraphLaunch(it->second, stream_), "hipGraphLaunch");
    if (timed) hip_check(hipEventRecord(e1, stream_), "hipEventRecord");
    sync();
    launches_ += nt;
    if (timed) {
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      timed_ms_ += ms;  // includes the state fills and inter-kernel gaps
      timed_ += nt;
    }
    return;
  }
  if (timing_stride_ < 0) {  // bracket the whole loop
    timed_loop([&] { enqueue_loop(a, i_shot, forward, store, nullptr, dump); });
    return;
  }
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs;
  event_cursor_ = 0;
  enqueue_loop(a, i_shot, forward, store, &pairs, dump);
  sync();
  for (auto &p : pairs) {
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, p.first, p.second),
              "hipEventElapsedTime");
    timed_ms_ += ms;
  }
}

// Loop timing: HIP events recorded on the model stream before and after the
// enqueued launches; the bracket covers every launch of the loop and the gaps
// between them, and counts the time steps those launches advanced.
void Model::timed_loop(const std::function<void()> &enqueue) {
  while (events_.size() < 2) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    events_.push_back(e);
  }
  const int64_t before = launches_;
  hip_check(hipEventRecord(events_[0], stream_), "hipEventRecord");
  enqueue();
  hip_check(hipEventRecord(events_[1], stream_), "hipEventRecord");
  sync();
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, events_[0], events_[1]), "hipEventElapsedTime");
  timed_ms_ += ms;
  timed_ += launches_ - before;
}

// State buffer set `set`: 0 the model's, 1 the second lane's, 2.. extra.
double *const (*Model::state_set(int set) const)[5] {
  if (set < 0 || set >= kMaxStateSets) throw ValueError("state buffer set out of range");
  return set == 0 ? state_ : set == 1 ? lane_state_ : extra_state_[set - 2];
}

void Model::ensure_state_sets(int n) {
  if (n > kMaxStateSets) throw ValueError("too many state buffer sets");
  if (n > 1) ensure_second_state();
  bool added = false;
  for (int k = 2; k < n; ++k)
    if (!extra_state_[k - 2][0][0]) {
      for (int p = 0; p < 2; ++p)
        for (int f = 0; f < 5; ++f) extra_state_[k - 2][p][f] = dalloc(npad_);
      added = true;
    }
  if (added) sync();
}

StepArgs Model::forward_launch_args(int i_shot, int parity, int set) const {
  check_shot(i_shot);
  double *const(*state)[5] = state_set(set);
  if (!state[0][0]) throw RuntimeError("the state buffer set is not allocated");
  StepArgs a = base_args();
  a.adjoint = 0;
  a.dts = cfg_.dt;
  a.tile_off = fwd_off_ + (size_t)i_shot * (ntiles_ + 1);
  a.entries = fwd_entries_;
  a.halo_off = fwd_halo_off_ + (size_t)i_shot * (ntiles_ + 1);
  a.halo_entries = fwd_halo_;
  a.rtf_x = syn_[0] + (size_t)i_shot * cfg_.nr * cfg_.nt;
  a.rtf_z = syn_[1] + (size_t)i_shot * cfg_.nr * cfg_.nt;
  a.rec_increments = step_is_two_step(variant_) ? 1 : 0;
  a.rec_off = adj_off_;
  a.rec_entries = adj_entries_;
  const int cur = parity & 1, nxt = 1 - cur;
  a.vx_i = state[cur][0];
  a.vz_i = state[cur][1];
  a.txx_i = state[cur][2];
  a.tzz_i = state[cur][3];
  a.txz_i = state[cur][4];
  a.vx_o = state[nxt][0];
  a.vz_o = state[nxt][1];
  a.txx_o = state[nxt][2];
  a.tzz_o = state[nxt][3];
  a.txz_o = state[nxt][4];
  a.do_snap = 0;
  a.do_snap1 = 0;
  return a;
}

StepArgs Model::adjoint_launch_args(int i_shot, int parity) const {
  check_shot(i_shot);
  StepArgs a = adjoint_args(i_shot);
  a.rec_off = adj_off_;
  a.rec_entries = adj_entries_;
  const int cur = parity & 1, nxt = 1 - cur;
  a.vx_i = state_[cur][0];
  a.vz_i = state_[cur][1];
  a.txx_i = state_[cur][2];
  a.tzz_i = state_[cur][3];
  a.txz_i = state_[cur][4];
  a.vx_o = state_[nxt][0];
  a.vz_o = state_[nxt][1];
  a.txx_o = state_[nxt][2];
  a.tzz_o = state_[nxt][3];
  a.txz_o = state_[nxt][4];
  a.do_snap = 0;
  a.do_snap1 = 0;
  return a;
}

double *const *Model::shot_snapshot_store() {
  if (!shot_snap_[0]) {
    for (int f = 0; f < 5; ++f) shot_snap_[f] = dalloc((size_t)d_.snapshots * ngrid_);
    sync();
  }
  return shot_snap_;
}

void Model::enqueue_zero_state(hipStream_t s, int set) const {
  double *const(*state)[5] = state_set(set);
  for (int k = 0; k < 2; ++k)
    for (int f = 0; f < 5; ++f)
      hip_check(launch_fill(state[k][f], npad_, 0.0, s), "zero state");
}

void Model::state_arrays(int set, std::vector<double *> &out) const {
  double *const(*state)[5] = state_set(set);
  for (int k = 0; k < 2; ++k)
    for (int f = 0; f < 5; ++f) out.push_back(state[k][f]);
}

void Model::ensure_second_state() {
  if (!lane_state_[0][0]) {
    for (int p = 0; p < 2; ++p)
      for (int f = 0; f < 5; ++f) lane_state_[p][f] = dalloc(npad_);
    sync();
  }
}

void Model::enqueue_receiver_scan(int i_shot, hipStream_t s) const {
  hip_check(launch_receiver_scan(syn_[0] + (size_t)i_shot * cfg_.nr * cfg_.nt,
                                 syn_[1] + (size_t)i_shot * cfg_.nr * cfg_.nt,
                                 cfg_.nr, cfg_.nt, s),
            "receiver scan");
}

void Model::forward(int i_shot, bool store_fields, bool verbose,
                    bool output_wavefields) {
  check_shot(i_shot);
  if (store_fields && cfg_.nt % cfg_.snapshot_interval != 0)
    throw ValueError("Snapshot interval and size of accumulator don't match!"
                     " (nt must be a multiple of snapshot_interval)");
  if (async_path(verbose, output_wavefields)) {
    ensure_lane();
    // a pending adjoint may still read this shot's snapshots and a pending
    // forward of the same shot writes the same traces: wait for them; a new
    // snapshot store is zero-filled on the model stream before any lane
    // writes to it
    if (adj_calls_ > 0 || fwd_job_[i_shot] >= 0 || (store_fields && !snap_[0])) settle();
    if (store_fields && !snap_[0]) {
      ensure_snapshots();
      sync();
    }
    fwd_job_[i_shot] = push_job(fwd_calls_ & 1, kJobForward, i_shot, store_fields, -1, -1);
    ++fwd_calls_;
    pump(false);
    return;
  }
  settle();
  if (store_fields) ensure_snapshots();
  StepArgs a = forward_args(i_shot);
  auto t0 = std::chrono::steady_clock::now();
  run_steps(a, i_shot, true, store_fields, output_wavefields);
  if (verbose) {
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                   .count();
    std::cout << "Seconds elapsed for forward wave simulation: " << s
              << std::endl;
  }
}

void Model::adjoint(int i_shot, bool verbose) {
  check_shot(i_shot);
  if (cfg_.nt % cfg_.snapshot_interval != 0)
    throw ValueError("Snapshot interval and size of accumulator don't match!"
                     " (nt must be a multiple of snapshot_interval)");
  if (!snap_[0])
    throw ValueError("adjoint_simulate needs stored forward wavefields; run "
                     "forward_simulate with store_fields=True first");
  if (async_path(verbose, false)) {
    ensure_lane();
    if (!defer_ && !no_defer_room_) {
      // a new buffer is zero-filled on the model stream: allocate it on a
      // settled model and let the fill finish before a lane writes to it
      settle();
      no_defer_room_ = !ensure_defer_buffer();
      sync();
    }
    const int lane = defer_ ? (adj_calls_ & 1) : 0;
    const int fwd = fwd_job_[i_shot];  // this shot's forward, if pending
    if (lane == 0) {
      // the previous adjoint's deferred terms go into the kernels first
      if (deferred_job_ >= 0) {
        last_replay_job_ = push_job(0, kJobReplay, -1, false, deferred_job_, -1);
        deferred_job_ = -1;
      }
      push_job(0, kJobAdjoint, i_shot, false, fwd, -1);
    } else {
      // the deferred-term buffer is free once the last replay has read it
      deferred_job_ = push_job(1, kJobDeferred, i_shot, false, fwd, last_replay_job_);
    }
    ++adj_calls_;
    pump(false);
    return;
  }
  settle();
  StepArgs a = adjoint_args(i_shot);
  auto t0 = std::chrono::steady_clock::now();
  if (ckpt_every_ > 1 && timing_stride_ < 0) {
    timed_loop([&] { enqueue_checkpointed_adjoint(a, i_shot, nullptr); });
  } else if (ckpt_every_ > 1) {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pairs;
    event_cursor_ = 0;
    enqueue_checkpointed_adjoint(a, i_shot, &pairs);
    sync();
    for (auto &p : pairs) {
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, p.first, p.second),
                "hipEventElapsedTime");
      timed_ms_ += ms;
    }
  } else {
    run_steps(a, i_shot, false, false, false);
  }
  if (verbose) {
    double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0)
                   .count();
    std::cout << "Seconds elapsed for adjoint wave simulation: " << s
              << std::endl;
  }
}

double Model::calculate_l2_misfit() {
  settle();
  hip_check(launch_misfit(syn_[0], syn_[1], obs_[0], obs_[1], nrec_,
                          0.5 * cfg_.dt, misfit_buf_,
                          misfit_buf_ + 2 * kMisfitBlocks, stream_),
            "misfit");
  download(&misfit_, misfit_buf_ + 2 * kMisfitBlocks, 1);
  return misfit_;
}

void Model::calculate_l2_adjoint_sources() {
  settle();
  hip_check(launch_adjoint_sources(syn_[0], syn_[1], obs_[0], obs_[1], adj_[0],
                                   adj_[1], nrec_, stream_),
            "adjoint sources");
  sync();
}

void Model::reset_kernels() {
  settle();
  hip_check(launch_fill(lam_k_, ngrid_, 0.0, stream_), "reset");
  hip_check(launch_fill(mu_k_, ngrid_, 0.0, stream_), "reset");
  hip_check(launch_fill(rho_l_, ngrid_, 0.0, stream_), "reset");
  sync();
}

void Model::map_kernels_to_velocity() {
  settle();
  hip_check(launch_map_kernels(vp_, vs_, rho_, lam_k_, mu_k_, rho_l_, vp_k_,
                               vs_k_, rho_v_, ngrid_, stream_),
            "map kernels");
  sync();
}

// Arguments of the time loops of one shot (src/fdModel.cpp:526-551, 789-815).
StepArgs Model::forward_args(int i_shot) const {
  StepArgs a = base_args();
  a.adjoint = 0;
  a.dts = cfg_.dt;
  a.tile_off = fwd_off_ + (size_t)i_shot * (ntiles_ + 1);
  a.entries = fwd_entries_;
  a.halo_off = fwd_halo_off_ + (size_t)i_shot * (ntiles_ + 1);
  a.halo_entries = fwd_halo_;
  a.rtf_x = syn_[0] + (size_t)i_shot * cfg_.nr * cfg_.nt;
  a.rtf_z = syn_[1] + (size_t)i_shot * cfg_.nr * cfg_.nt;
  return a;
}

StepArgs Model::adjoint_args(int i_shot) const {
  StepArgs a = base_args();
  a.adjoint = 1;
  a.dts = -cfg_.dt;
  a.tile_off = adj_off_;
  a.entries = adj_entries_;
  a.halo_off = adj_halo_off_;
  a.halo_entries = adj_halo_;
  a.adj_x = adj_[0] + (size_t)i_shot * cfg_.nr * cfg_.nt;
  a.adj_z = adj_[1] + (size_t)i_shot * cfg_.nr * cfg_.nt;
  return a;
}

// src/fdModel.cpp:1291-1308: forward all shots (storing snapshots), misfit,
// and optionally adjoint sources, reset, adjoint all shots, map kernels.
void Model::run_model(bool verbose, bool simulate_adjoint) {
  if (cfg_.nt % cfg_.snapshot_interval != 0)
    throw ValueError("Snapshot interval and size of accumulator don't match!"
                     " (nt must be a multiple of snapshot_interval)");
  if (!verbose && ckpt_every_ == 1 && timing_stride_ <= 0 && !graph_mode_ &&
      cfg_.n_shots > 1 && !step_is_two_step(variant_) && !step_is_inplace(variant_)) {
    // the per-shot calls on two lanes (forwards two at a time, adjoints in
    // pairs with deferred kernel terms), bracketed by loop timing events
    settle();
    const bool timed = timing_stride_ < 0, was_async = async_;
    while (events_.size() < 2) {
      hipEvent_t e;
      hip_check(hipEventCreate(&e), "hipEventCreate");
      events_.push_back(e);
    }
    const int64_t before = launches_;
    if (timed) hip_check(hipEventRecord(events_[0], stream_), "hipEventRecord");
    timing_stride_ = 0;
    async_ = true;
    try {
      for (int s = 0; s < cfg_.n_shots; ++s) forward(s, true, false, false);
      calculate_l2_misfit();
      if (simulate_adjoint) {
        calculate_l2_adjoint_sources();
        reset_kernels();
        for (int s = 0; s < cfg_.n_shots; ++s) adjoint(s, false);
        settle();
        map_kernels_to_velocity();
      }
    } catch (...) {
      timing_stride_ = timed ? -1 : 0;
      async_ = was_async;
      throw;
    }
    timing_stride_ = timed ? -1 : 0;
    async_ = was_async;
    if (timed) {
      hip_check(hipEventRecord(events_[1], stream_), "hipEventRecord");
      sync();
      float ms = 0.f;
      hip_check(hipEventElapsedTime(&ms, events_[0], events_[1]), "hipEventElapsedTime");
      timed_ms_ += ms;
      timed_ += launches_ - before;
    }
    return;
  }
  if (!verbose && (ckpt_every_ == 1 || !simulate_adjoint) && timing_stride_ <= 0 &&
      cfg_.n_shots > 1) {
    run_model_two_lanes(simulate_adjoint);
    return;
  }
  for (int s = 0; s < cfg_.n_shots; ++s) forward(s, true, verbose, false);
  calculate_l2_misfit();
  if (simulate_adjoint) {
    calculate_l2_adjoint_sources();
    reset_kernels();
    for (int s = 0; s < cfg_.n_shots; ++s) adjoint(s, verbose);
    map_kernels_to_velocity();
  }
}

// Buffer for the deferred correlation of one adjoint: 3 arrays per snapshot
// slot.  Allocated once when the device has room for it (and a margin).
bool Model::ensure_defer_buffer() {
  const long long n = (long long)d_.snapshots * 3 * ngrid_;
  if (defer_ && defer_n_ == n) return true;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
  const double need = (double)n * sizeof(double) + (double)(2ull << 30);
  if ((double)free_b < need) return false;
  defer_ = dalloc((size_t)n);
  defer_n_ = n;
  return true;
}

// run_model on two HIP streams.  A single time loop leaves the GPU partly
// idle between launches (the completion-to-dispatch latency of dependent
// launches and the last blocks of each launch); two independent loops on
// two streams fill each other's gaps.  Lane X (stream_, state_) and lane Y
// (lane_stream_, lane_state_) run, for one-step kernels:
//   X: forward 0, adjoint 0, 1, ..., n-2, replay of adjoint n-1
//   Y: forward 1, 2, ..., n-1, adjoint n-1 (deferred)
// so every time loop has a partner.  The adjoints of shots 0 .. n-2
// accumulate the kernels in shot order; the last one runs beside them with
// its kernel pointers aimed at a zeroed buffer per snapshot slot, which
// therefore receives exactly the terms it would have applied, and the replay
// adds them to the kernels in the adjoint's own order after adjoint n-2
// (K - e == K + (0 - e) in IEEE arithmetic).  With two-step kernels (two
// correlation steps may share a launch) or without room for the buffer:
//   X: adjoint 0, ..., n-1 (each after its forward);  Y: forward 0, ..., n-1.
// Without the adjoint the lanes forward alternate shots.  Every output is
// bitwise that of the sequential calls.  The host enqueues the lanes in
// interleaved chunks of steps so that neither stream's queue runs dry.
// (With checkpointing only the forward-only driver uses two lanes: the
// checkpointed adjoint recomputes segments on the model stream.)  In graph
// mode each job -- one time loop with its prologue -- is captured once as a
// HIP graph per (lane, shot, kind) and replayed, so the host issues one
// launch per time loop.  Two lanes need two hardware queues: with more live
// streams in the process than the runtime's queues (GPU_MAX_HW_QUEUES, 4 by
// default) the lanes may share one and run one after the other -- the
// results are the same, the overlap is lost.
void Model::run_model_two_lanes(bool adj) {
  const int ns = cfg_.n_shots, nt = cfg_.nt;
  ensure_snapshots();
  if (!lane_stream_) create_stream(&lane_stream_);
  if (!lane_state_[0][0])
    for (int p = 0; p < 2; ++p)
      for (int f = 0; f < 5; ++f) lane_state_[p][f] = dalloc(npad_);
  while ((int)lane_events_.size() < ns + 3) {
    hipEvent_t e;
    hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    lane_events_.push_back(e);
  }
  while (events_.size() < 2) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    events_.push_back(e);
  }
  const bool timed = timing_stride_ < 0;
  const bool inplace = step_is_inplace(variant_);
  const bool two = step_is_two_step(variant_);
  const bool balanced = adj && !two && ensure_defer_buffer();
  const int64_t before = launches_;
  if (timed) hip_check(hipEventRecord(events_[0], stream_), "hipEventRecord");
  if (adj) {
    hip_check(launch_fill(lam_k_, ngrid_, 0.0, stream_), "reset");
    hip_check(launch_fill(mu_k_, ngrid_, 0.0, stream_), "reset");
    hip_check(launch_fill(rho_l_, ngrid_, 0.0, stream_), "reset");
  }
  // lane Y starts after everything enqueued so far on stream_
  hip_check(hipEventRecord(lane_events_[ns], stream_), "hipEventRecord");
  hip_check(hipStreamWaitEvent(lane_stream_, lane_events_[ns], 0), "hipStreamWaitEvent");

  enum Kind { kFwd, kAdj, kAdjDeferred, kReplay };
  struct Job {
    int shot;
    Kind kind;
  };
  struct Lane {
    hipStream_t s;
    double *(*st)[5];
    std::vector<Job> jobs;
    size_t j = 0;
    int k = 0, parity = 0;
    bool started = false;
    StepArgs a;
  };
  Lane lanes[2];  // 0: X, 1: Y
  lanes[0].s = stream_;
  lanes[0].st = state_;
  lanes[1].s = lane_stream_;
  lanes[1].st = lane_state_;
  std::vector<int> fwd_lane(ns, -1);  // lane of each shot's forward
  if (balanced) {
    lanes[0].jobs.push_back({0, kFwd});
    for (int sh = 1; sh < ns; ++sh) lanes[1].jobs.push_back({sh, kFwd});
    for (int sh = 0; sh + 1 < ns; ++sh) lanes[0].jobs.push_back({sh, kAdj});
    lanes[1].jobs.push_back({ns - 1, kAdjDeferred});
    lanes[0].jobs.push_back({ns - 1, kReplay});
  } else {
    for (int sh = 0; sh < ns; ++sh) {
      if (adj) {
        lanes[1].jobs.push_back({sh, kFwd});
        lanes[0].jobs.push_back({sh, kAdj});
      } else {
        lanes[1 - (sh & 1)].jobs.push_back({sh, kFwd});
      }
    }
  }
  for (int l = 0; l < 2; ++l)
    for (const Job &jb : lanes[l].jobs)
      if (jb.kind == kFwd) fwd_lane[jb.shot] = l;
  std::vector<char> fwd_enqueued(ns, 0);
  bool deferred_enqueued = false;
  const int chunk = 64;
  while (lanes[0].j < lanes[0].jobs.size() || lanes[1].j < lanes[1].jobs.size()) {
    for (int l = 1; l >= 0; --l) {
      Lane &ln = lanes[l];
      if (ln.j >= ln.jobs.size()) continue;
      const Job jb = ln.jobs[ln.j];
      const bool fwd = jb.kind == kFwd;
      if ((jb.kind == kAdj || jb.kind == kAdjDeferred) && !fwd_enqueued[jb.shot]) continue;
      if (jb.kind == kReplay && !deferred_enqueued) continue;
      if (jb.kind == kReplay) {
        // after adjoint n-2 (this lane) and the deferred adjoint (lane Y)
        hip_check(hipStreamWaitEvent(ln.s, lane_events_[ns + 2], 0), "hipStreamWaitEvent");
        hip_check(launch_replay_correlation(
                      rho_l_, lam_k_, mu_k_, defer_, d_.snapshots, ngrid_, d_.nx,
                      cfg_.np_boundary + cfg_.nx_inner_boundary,
                      cfg_.np_boundary + cfg_.nx_inner - cfg_.nx_inner_boundary,
                      cfg_.np_boundary + cfg_.nz_inner_boundary,
                      cfg_.np_boundary + cfg_.nz_inner - cfg_.nz_inner_boundary, ln.s),
                  "replay correlation");
        ++ln.j;
        continue;
      }
      const int shot = jb.shot;
      // everything a job enqueues before its time steps: the adjoint source
      // of its shot, the zeroed deferred-term buffer, the zeroed state
      auto prologue = [&]() {
        if (fwd) {
          ln.a = forward_args(shot);
        } else {
          const size_t off = (size_t)shot * cfg_.nr * nt;
          hip_check(launch_adjoint_sources(syn_[0] + off, syn_[1] + off, obs_[0] + off,
                                           obs_[1] + off, adj_[0] + off, adj_[1] + off,
                                           (long long)cfg_.nr * nt, ln.s),
                    "adjoint sources");
          ln.a = adjoint_args(shot);
        }
        ln.a.rec_increments = (fwd && two) ? 1 : 0;
        ln.a.rec_off = adj_off_;
        ln.a.rec_entries = adj_entries_;
        for (int p = 0; p < (inplace ? 1 : 2); ++p)
          for (int f = 0; f < 5; ++f)
            hip_check(launch_fill(ln.st[p][f], npad_, 0.0, ln.s), "zero state");
        ln.k = 0;
        ln.parity = 0;
      };
      SlotFn slot = [this, shot](int it, double **dst) {
        return stored_slot(shot, it, dst);
      };
      auto steps = [&](int k1) {
        ln.parity = enqueue_range(ln.a, ln.st, ln.parity, fwd, ln.k, k1, true, slot, nullptr,
                                  false, ln.s, jb.kind == kAdjDeferred ? defer_ : nullptr);
        ln.k = k1;
        if (ln.k == nt && ln.a.rec_increments)
          hip_check(launch_receiver_scan(ln.a.rtf_x, ln.a.rtf_z, cfg_.nr, nt, ln.s),
                    "receiver scan");
      };
      if (!ln.started && !fwd && fwd_lane[shot] != l)  // the shot's forward has run
        hip_check(hipStreamWaitEvent(ln.s, lane_events_[shot], 0), "hipStreamWaitEvent");
      if (graph_mode_) {
        // the whole job as one captured graph (kept per lane, shot and kind)
        const long long key = -1 - (((long long)l * ns + shot) * 4 + (long long)jb.kind);
        auto gi = graphs_.find(key);
        if (gi == graphs_.end()) {
          hipGraph_t graph = nullptr;
          hip_check(hipStreamBeginCapture(ln.s, hipStreamCaptureModeThreadLocal),
                    "hipStreamBeginCapture");
          const int64_t saved = launches_;
          try {
            prologue();
            steps(nt);
          } catch (...) {
            hipGraph_t g2 = nullptr;
            (void)hipStreamEndCapture(ln.s, &g2);
            if (g2) (void)hipGraphDestroy(g2);
            launches_ = saved;
            throw;
          }
          launches_ = saved;
          hip_check(hipStreamEndCapture(ln.s, &graph), "hipStreamEndCapture");
          hipGraphExec_t exec = nullptr;
          hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
          (void)hipGraphDestroy(graph);
          hip_check(e, "hipGraphInstantiate");
          gi = graphs_.emplace(key, exec).first;
        }
        hip_check(hipGraphLaunch(gi->second, ln.s), "hipGraphLaunch");
        launches_ += nt;
        ln.k = nt;
      } else {
        if (!ln.started) {
          prologue();
          ln.started = true;
        }
        steps(std::min(ln.k + chunk, nt));
      }
      if (ln.k == nt) {
        if (fwd) {
          hip_check(hipEventRecord(lane_events_[shot], ln.s), "hipEventRecord");
          fwd_enqueued[shot] = 1;
        }
        if (jb.kind == kAdjDeferred) {
          hip_check(hipEventRecord(lane_events_[ns + 2], ln.s), "hipEventRecord");
          deferred_enqueued = true;
        }
        ln.started = false;
        ++ln.j;
      }
    }
  }
  // join the lanes
  hip_check(hipEventRecord(lane_events_[ns + 1], lane_stream_), "hipEventRecord");
  hip_check(hipStreamWaitEvent(stream_, lane_events_[ns + 1], 0), "hipStreamWaitEvent");
  if (timed) hip_check(hipEventRecord(events_[1], stream_), "hipEventRecord");
  if (adj) map_kernels_to_velocity();
  calculate_l2_misfit();  // synchronises stream_
  if (timed) {
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, events_[0], events_[1]), "hipEventElapsedTime");
    timed_ms_ += ms;
    timed_ += launches_ - before;
  }
}

void Model::get_parameter_fields(double *vp, double *vs, double *rho) const {
  to_host_layout(vp, vp_);
  to_host_layout(vs, vs_);
  to_host_layout(rho, rho_);
}

void Model::set_parameter_fields(const double *vp, const double *vs,
                                 const double *rho) {
  settle();
  to_device_layout(vp_, vp);
  to_device_layout(vs_, vs);
  to_device_layout(rho_, rho);
  hip_check(launch_update_from_velocity(vp_, vs_, rho_, lm_, mu_, b_, d_.nx,
                                        d_.nz, pitch_, origin_, stream_),
            "update_from_velocity");
  sync();
}

void Model::get_kernels(double *a, double *b, double *c) const {
  to_host_layout(a, vp_k_);
  to_host_layout(b, vs_k_);
  to_host_layout(c, rho_v_);
}

void Model::get_lame_kernels(double *a, double *b, double *c) const {
  to_host_layout(a, lam_k_);
  to_host_layout(b, mu_k_);
  to_host_layout(c, rho_l_);
}

void Model::get_traces(int which, double *ux, double *uz) const {
  double *const *src = which == 0 ? syn_ : which == 1 ? obs_ : adj_;
  download(ux, src[0], nrec_);
  download(uz, src[1], nrec_);
}

void Model::set_traces(int which, const double *ux, const double *uz) {
  settle();
  double *const *dst = which == 0 ? syn_ : which == 1 ? obs_ : adj_;
  upload(dst[0], ux, nrec_);
  upload(dst[1], uz, nrec_);
}

void Model::set_checkpointing(int every) {
  settle();
  if (every < 1) throw ValueError("checkpoint interval must be >= 1");
  if (every == ckpt_every_) return;
  sync();
  clear_graphs();
  for (int f = 0; f < 5; ++f) {
    if (!snap_[f]) continue;
    auto it = std::find(allocations_.begin(), allocations_.end(), (void *)snap_[f]);
    if (it != allocations_.end()) allocations_.erase(it);
    hip_check(hipFree(snap_[f]), "hipFree");
    device_bytes_ -= (int64_t)((long long)cfg_.n_shots * stored_snapshots() *
                               ngrid_ * sizeof(double));
    snap_[f] = nullptr;
  }
  ckpt_every_ = every;
}

void Model::get_snapshot(int i_shot, int i_snap, double *const out[5]) const {
  check_shot(i_shot);
  if (ckpt_every_ > 1)
    throw ValueError("snapshots are not stored with checkpointing enabled");
  if (i_snap < 0 || i_snap >= d_.snapshots)
    throw ValueError("snapshot index out of range");
  if (!snap_[0]) {
    for (int f = 0; f < 5; ++f) std::memset(out[f], 0, ngrid_ * sizeof(double));
    return;
  }
  long long off = ((long long)i_shot * d_.snapshots + i_snap) * ngrid_;
  for (int f = 0; f < 5; ++f) to_host_layout(out[f], snap_[f] + off);
}

void Model::get_snapshots(double *const out[5]) const {
  if (ckpt_every_ > 1)
    throw ValueError("snapshots are not stored with checkpointing enabled");
  long long slices = (long long)cfg_.n_shots * d_.snapshots;
  if (!snap_[0]) {
    for (int f = 0; f < 5; ++f)
      std::memset(out[f], 0, slices * ngrid_ * sizeof(double));
    return;
  }
  // bounded staging: about 512 MiB per transfer
  long long chunk = std::max<long long>(1, (64ll << 20) / std::max<long long>(ngrid_, 1));
  for (int f = 0; f < 5; ++f)
    for (long long s0 = 0; s0 < slices; s0 += chunk) {
      int cnt = (int)std::min(chunk, slices - s0);
      to_host_layout(out[f] + s0 * ngrid_, snap_[f] + s0 * ngrid_, cnt);
    }
}

void Model::get_model_vector(double *v) const {
  double *s = staging(d_.free_parameters);
  hip_check(launch_model_vector(0, vp_, vs_, rho_, s, d_.nx,
                                d_.nx_free_parameters, d_.nz_free_parameters,
                                cfg_.npx, cfg_.npz,
                                cfg_.nx_inner_boundary + cfg_.np_boundary,
                                cfg_.nz_inner_boundary + cfg_.np_boundary,
                                stream_),
            "model vector");
  download(v, s, d_.free_parameters);
}

void Model::set_model_vector(const double *v) {
  settle();
  double *s = staging(d_.free_parameters);
  upload(s, v, d_.free_parameters);
  hip_check(launch_model_vector(1, vp_, vs_, rho_, s, d_.nx,
                                d_.nx_free_parameters, d_.nz_free_parameters,
                                cfg_.npx, cfg_.npz,
                                cfg_.nx_inner_boundary + cfg_.np_boundary,
                                cfg_.nz_inner_boundary + cfg_.np_boundary,
                                stream_),
            "model vector");
  hip_check(launch_update_from_velocity(vp_, vs_, rho_, lm_, mu_, b_, d_.nx,
                                        d_.nz, pitch_, origin_, stream_),
            "update_from_velocity");
  sync();
}

void Model::get_gradient_vector(double *v) const {
  double *s = staging(d_.free_parameters);
  hip_check(launch_gradient_vector(vp_k_, vs_k_, rho_v_, s, d_.nx,
                                   d_.nx_free_parameters,
                                   d_.nz_free_parameters, cfg_.npx, cfg_.npz,
                                   cfg_.nx_inner_boundary + cfg_.np_boundary,
                                   cfg_.nz_inner_boundary + cfg_.np_boundary,
                                   stream_),
            "gradient vector");
  download(v, s, d_.free_parameters);
}

void Model::get_setup(double *stf, double *moment, double *taper) const {
  if (stf) std::copy(d_.stf.begin(), d_.stf.end(), stf);
  if (moment) std::copy(d_.moment.begin(), d_.moment.end(), moment);
  if (taper)
    for (int ix = 0; ix < d_.nx; ++ix)
      for (int iz = 0; iz < d_.nz; ++iz)
        taper[(size_t)ix * d_.nz + iz] =
            d_.taper_lut[taper_count(ix, iz, d_.nx, d_.nz, cfg_.np_boundary)];
}

// src/fdModel.cpp:973-1006 (the prefix argument is not used there either)
void Model::write_receivers() const {
  std::vector<double> ux(nrec_), uz(nrec_);
  get_traces(0, ux.data(), uz.data());
  for (int s = 0; s < cfg_.n_shots; ++s) {
    std::ofstream fx(cfg_.observed_data_folder + "/rtf_ux" + std::to_string(s) + ".txt");
    std::ofstream fz(cfg_.observed_data_folder + "/rtf_uz" + std::to_string(s) + ".txt");
    fx.precision(std::numeric_limits<double>::digits10 + 10);
    fz.precision(std::numeric_limits<double>::digits10 + 10);
    for (int r = 0; r < cfg_.nr; ++r) {
      fx << std::endl;
      fz << std::endl;
      for (int it = 0; it < cfg_.nt; ++it) {
        size_t k = ((size_t)s * cfg_.nr + r) * cfg_.nt + it;
        fx << ux[k] << " ";
        fz << uz[k] << " ";
      }
    }
  }
}

// src/fdModel.cpp:1008-1032
void Model::write_sources() const {
  for (int s = 0; s < cfg_.n_shots; ++s) {
    std::ofstream f(cfg_.stf_folder + "/sources_shot_" + std::to_string(s) + ".txt");
    f.precision(std::numeric_limits<double>::digits10 + 10);
    for (int src : cfg_.which_source_to_fire_in_which_shot[s]) {
      f << std::endl;
      for (int it = 0; it < cfg_.nt; ++it)
        f << d_.stf[(size_t)src * cfg_.nt + it] << " ";
    }
  }
}

// src/fdModel.cpp:1326-1356
void Model::write_kernels() const {
  std::vector<double> a(ngrid_), b(ngrid_), c(ngrid_);
  get_kernels(a.data(), b.data(), c.data());
  std::ofstream fa("kernel_vp.txt"), fb("kernel_vs.txt"),
      fc("kernel_density.txt");
  for (int ix = 0; ix < d_.nx; ++ix) {
    for (int iz = 0; iz < d_.nz; ++iz) {
      size_t i = (size_t)ix * d_.nz + iz;
      fa << a[i] << " ";
      fb << b[i] << " ";
      fc << c[i] << " ";
    }
    fa << std::endl;
    fb << std::endl;
    fc << std::endl;
  }
}

// src/fdModel.cpp:1053-1126: stream extraction with the reference's
// good()-state checks.
void Model::load_receivers(bool verbose) {
  settle();
  std::vector<double> ux(nrec_), uz(nrec_);
  for (int s = 0; s < cfg_.n_shots; ++s) {
    std::string nx_ = cfg_.observed_data_folder + "/rtf_ux" + std::to_string(s) + ".txt";
    std::string nz_ = cfg_.observed_data_folder + "/rtf_uz" + std::to_string(s) + ".txt";
    std::ifstream fx(nx_), fz(nz_);
    if (verbose) {
      std::cout << "File for ux data at shot " << s << " is "
                << (fx.good() ? "good (exists at least)." : "ungood.") << std::endl;
      std::cout << "File for uz data at shot " << s << " is "
                << (fz.good() ? "good (exists at least)." : "ungood.") << std::endl;
    }
    if (!fx.good() || !fz.good()) throw ValueError("Not all data is present!");
    double px, pz;
    for (int r = 0; r < cfg_.nr; ++r)
      for (int it = 0; it < cfg_.nt; ++it) {
        fx >> px;
        fz >> pz;
        size_t k = ((size_t)s * cfg_.nr + r) * cfg_.nt + it;
        ux[k] = px;
        uz[k] = pz;
      }
    if (!fx.good() || !fz.good()) throw ValueError("Not enough data is present!");
    fx >> px;
    fz >> pz;
    if (fx.good() || fz.good()) throw ValueError("Too much data is present!");
  }
  set_traces(1, ux.data(), uz.data());
}

// src/fdModel.cpp:1185-1289
void Model::load_model(const std::string &de, const std::string &vp,
                       const std::string &vs, bool verbose) {
  std::ifstream fd(de), fp(vp), fs(vs);
  if (verbose) {
    std::cout << "Loading models." << std::endl;
    std::cout << "File: " << de << std::endl
              << "File for density is "
              << (fd.good() ? "good (exists at least)." : "ungood.") << std::endl;
    std::cout << "File: " << vp << std::endl
              << "File for P-wave velocity is "
              << (fp.good() ? "good (exists at least)." : "ungood.") << std::endl;
    std::cout << "File: " << vs << std::endl
              << "File for S-wave velocity is "
              << (fs.good() ? "good (exists at least)." : "ungood.") << std::endl;
  }
  if (!fd.good() || !fp.good() || !fs.good())
    throw ValueError("The files for target models don't seem to exist.");
  std::vector<double> hd(ngrid_), hp(ngrid_), hs(ngrid_);
  for (int ix = 0; ix < d_.nx; ++ix)
    for (int iz = 0; iz < d_.nz; ++iz) {
      size_t i = (size_t)ix * d_.nz + iz;
      fd >> hd[i];
      fp >> hp[i];
      fs >> hs[i];
      if (!fd.good() || !fp.good() || !fs.good())
        throw ValueError("Received bad state of one of the files at end of "
                         "reading. Does the data match the domain?");
    }
  double extra;
  fd >> extra;
  fp >> extra;
  fs >> extra;
  if (fd.good() || fp.good() || fs.good())
    throw ValueError("Too much data is present!");
  set_parameter_fields(hp.data(), hs.data(), hd.data());
}

// src/fdModel.cpp:1517-1572 (size 3 * nxf * nzf regardless of the basis)
std::vector<double> Model::load_vector(const std::string &path,
                                       bool verbose) const {
  size_t n = (size_t)3 * d_.nx_free_parameters * d_.nz_free_parameters;
  std::vector<double> m(n);
  std::ifstream f(path);
  if (verbose) {
    std::cout << "Loading model vector." << std::endl;
    std::cout << "File: " << path << std::endl;
    std::cout << "File is " << (f.good() ? "good (exists at least)." : "ungood.")
              << std::endl;
  }
  if (!f.good())
    throw ValueError("The files for target models don't seem to exist.");
  double p;
  for (size_t i = 0; i < n; ++i) {
    f >> p;
    m[i] = p;
  }
  if (!f.good()) throw ValueError("Not enough data is present!");
  f >> p;
  if (f.good()) throw ValueError("Too much data is present!");
  return m;
}

}  // namespace psv
