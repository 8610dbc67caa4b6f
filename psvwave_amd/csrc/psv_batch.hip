// psv_batch.hip -- batched forward runs over many models of uneven size.
//
// Each member is a full Model (its own HBM buffers and injection tables).
// The batch keeps, in device memory, the launch arguments of every
// (member, shot, buffer parity) and the list of (member, tile) pairs of all
// members, so one launch per time step advances every model: the grid is
// the union of all members' tiles, large enough to fill the GPU even when
// each model alone is small.  A shot's whole time loop can be captured as
// one HIP graph and replayed.
#include "psv_batch.h"

#include <algorithm>
#include <limits>

namespace psv {

namespace {
void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess)
    throw RuntimeError(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

Batch::Batch(const std::vector<Config> &cfgs) {
  if (cfgs.empty()) throw ValueError("a batch needs at least one model");
  nt_ = cfgs[0].nt;
  n_shots_ = cfgs[0].n_shots;
  for (const Config &c : cfgs)
    if (c.nt != nt_ || c.n_shots != n_shots_)
      throw ValueError("all models of a batch must share nt and n_shots");
  // the members run their own calls (parameters, adjoint sources, misfits)
  // on the batch's stream: one stream for the whole batch
  create_stream(&stream_);
  try {
    for (const Config &c : cfgs) models_.emplace_back(new Model(c, stream_));
  } catch (...) {
    models_.clear();
    destroy_stream(stream_);
    throw;
  }
  hip_check(hipEventCreate(&ev_[0]), "hipEventCreate");
  hip_check(hipEventCreate(&ev_[1]), "hipEventCreate");
  for (auto &m : models_) points_ += m->grid_points();
  bool lean = true;
  for (auto &m : models_) lean = lean && m->lean_ok();
  if (lean) variant_ = kLeanTz16RingClamp;
  build_tables();
}

Batch::~Batch() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  clear_graphs();
  if (descs_) (void)hipFree(descs_);
  if (ldescs_) (void)hipFree(ldescs_);
  if (tiles_) (void)hipFree(tiles_);
  for (auto p : zero_ptrs_)
    if (p) (void)hipFree(p);
  if (zero_lens_) (void)hipFree(zero_lens_);
  clear_gradient_tables();
  for (auto e : ev_)
    if (e) (void)hipEventDestroy(e);
  models_.clear();  // the members run on stream_
  destroy_stream(stream_);
}

Model &Batch::model(int m) {
  if (m < 0 || m >= size()) throw ValueError("model index out of range");
  return *models_[m];
}

const Model &Batch::model(int m) const {
  if (m < 0 || m >= size()) throw ValueError("model index out of range");
  return *models_[m];
}

void Batch::clear_graphs() {
  for (auto g : graphs_)
    if (g) (void)hipGraphExecDestroy(g);
  graphs_.clear();
  if (graph_all_) (void)hipGraphExecDestroy(graph_all_);
  graph_all_ = nullptr;
}

void Batch::set_persistent(bool on) {
  if (on && variant_ != kLeanTz8)
    throw ValueError("the persistent batch loop runs step variant 9 (lean 64 x 8)");
  persistent_ = on;
  clear_graphs();
}

void Batch::set_graph_mode(bool on) {
  if (!on) clear_graphs();
  graph_mode_ = on;
}

void Batch::clear_gradient_tables() {
  if (gfwd_) (void)hipFree(gfwd_);
  if (gadj_) (void)hipFree(gadj_);
  gfwd_ = gadj_ = nullptr;
}

void Batch::set_variant(int variant, int xcd_remap) {
  if (variant != kFusedTz8 && variant != kFusedTz16 && variant != kTwoStepTz8 &&
      variant != kTwoStepTz16 && !step_is_lean(variant))
    throw ValueError("batched runs use step variants 0, 1, 6, 7, 8-13 or 18-28");
  if (step_is_lean(variant))
    for (auto &m : models_)
      if (!m->lean_ok())
        throw ValueError("a member of the batch cannot use the lean step variants");
  variant_ = variant;
  xcd_remap_ = xcd_remap ? 1 : 0;
  if (variant != kLeanTz8) persistent_ = false;
  clear_graphs();
  clear_gradient_tables();  // descriptors hold injection-table pointers
  build_tables();
}

void Batch::build_tables() {
  std::vector<int2> tiles;
  std::vector<StepArgs> descs;
  for (size_t m = 0; m < models_.size(); ++m) {
    models_[m]->set_step_variant(variant_, xcd_remap_);
    for (int t = 0; t < models_[m]->tile_count(); ++t)
      tiles.push_back(make_int2((int)m, t));
    for (int s = 0; s < n_shots_; ++s)
      for (int p = 0; p < 2; ++p) descs.push_back(models_[m]->forward_launch_args(s, p));
  }
  if (tiles.size() > (size_t)std::numeric_limits<int>::max())
    throw ValueError("too many tiles in one batch");
  ntiles_ = (int)tiles.size();
  hip_check(hipStreamSynchronize(stream_), "sync");
  if (descs_) hip_check(hipFree(descs_), "hipFree");
  if (ldescs_) hip_check(hipFree(ldescs_), "hipFree");
  if (tiles_) hip_check(hipFree(tiles_), "hipFree");
  descs_ = ldescs_ = nullptr;
  tiles_ = nullptr;
  hip_check(hipMalloc((void **)&descs_, std::max<size_t>(1, descs.size()) * sizeof(StepArgs)),
            "hipMalloc");
  hip_check(hipMalloc((void **)&tiles_, std::max<size_t>(1, tiles.size()) * sizeof(int2)),
            "hipMalloc");
  hip_check(hipMemcpy(descs_, descs.data(), descs.size() * sizeof(StepArgs),
                      hipMemcpyHostToDevice),
            "upload");
  hip_check(hipMemcpy(tiles_, tiles.data(), tiles.size() * sizeof(int2),
                      hipMemcpyHostToDevice),
            "upload");
  ensure_zero_table(0);
}

// The device tables of every member's state set `set` (built before any
// stream capture: the allocation and upload may not be captured).
void Batch::ensure_zero_table(int set) {
  if (!zero_ptrs_[set]) {
    std::vector<double *> ptrs;
    std::vector<long long> lens;
    for (auto &m : models_) {
      m->state_arrays(set, ptrs);
      lens.resize(ptrs.size(), m->padded_points());
      zero_max_ = std::max(zero_max_, m->padded_points());
    }
    for (double *p : ptrs)
      if (!p) throw RuntimeError("a state buffer set is not allocated");
    hip_check(hipMalloc((void **)&zero_ptrs_[set], ptrs.size() * sizeof(double *)),
              "hipMalloc");
    hip_check(hipMemcpy(zero_ptrs_[set], ptrs.data(), ptrs.size() * sizeof(double *),
                        hipMemcpyHostToDevice),
              "upload");
    if (!zero_lens_) {
      hip_check(hipMalloc((void **)&zero_lens_, lens.size() * sizeof(long long)), "hipMalloc");
      hip_check(hipMemcpy(zero_lens_, lens.data(), lens.size() * sizeof(long long),
                          hipMemcpyHostToDevice),
                "upload");
    }
  }
}

void Batch::enqueue_zero_states(int set) {
  if (!zero_ptrs_[set]) throw RuntimeError("state zeroing table missing");
  const int n = 10 * size();
  const int bx = (int)std::min<long long>(64, (zero_max_ + 2047) / 2048);
  hip_check(launch_zero_many(zero_ptrs_[set], zero_lens_, n, std::max(1, bx), stream_),
            "zero states");
}

// Zero every member's state, then advance `nsteps` steps of shot i_shot
// with one launch per step (or per pair of steps), then form the receiver
// running sums when the two-step kernels recorded increments.
void Batch::enqueue_loop(int i_shot, int nsteps) {
  enqueue_zero_states(0);
  if (persistent_) {
    BatchLaunch b{descs_, tiles_, ntiles_, n_shots_, i_shot, 0, 0, 0, xcd_remap_, 0, 1};
    hip_check(launch_batch_persistent(b, 0, nsteps, stream_), "persistent batch loop");
    return;
  }
  const bool two = step_is_two_step(variant_);
  int launch = 0;
  for (int k = 0; k < nsteps;) {
    const bool pair = two && k + 1 < nsteps;
    BatchLaunch b{descs_, tiles_, ntiles_, n_shots_, i_shot, launch & 1,
                  k, k + 1, xcd_remap_, 0, 1};
    hip_check(launch_batch_step(variant_, pair, b, stream_), "batch step");
    k += pair ? 2 : 1;
    ++launch;
  }
  if (two && nsteps == nt_)
    for (auto &m : models_) m->enqueue_receiver_scan(i_shot, stream_);
}

void Batch::forward(int i_shot) {
  if (i_shot < 0 || i_shot >= n_shots_)
    throw ValueError("i_shot is outside [0, n_shots)");
  hip_check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
  if (graph_mode_) {
    if (graphs_.size() != (size_t)n_shots_) graphs_.assign(n_shots_, nullptr);
    if (!graphs_[i_shot]) {
      hip_check(hipStreamSynchronize(stream_), "sync");
      hipGraph_t graph = nullptr;
      hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture");
      try {
        enqueue_loop(i_shot, nt_);
      } catch (...) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(stream_, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        throw;
      }
      hip_check(hipStreamEndCapture(stream_, &graph), "hipStreamEndCapture");
      hipGraphExec_t exec = nullptr;
      hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      hip_check(e, "hipGraphInstantiate");
      graphs_[i_shot] = exec;
      hip_check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
    }
    hip_check(hipGraphLaunch(graphs_[i_shot], stream_), "hipGraphLaunch");
  } else {
    enqueue_loop(i_shot, nt_);
  }
  hip_check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "hipEventElapsedTime");
  loop_ms_ += ms;
}

bool Batch::lockstep_available() {
  if (n_shots_ < 2 || n_shots_ > Model::kMaxStateSets || !step_is_lean(variant_) ||
      persistent_ ||
      no_lockstep_room_ || (long long)ntiles_ * n_shots_ > std::numeric_limits<int>::max())
    return false;
  if (ldescs_) return true;
  try {
    for (auto &m : models_) m->ensure_state_sets(n_shots_);
  } catch (const RuntimeError &) {
    // no room: keep what was allocated and run shot by shot; clear the
    // failed allocation's error so that the next launch check does not
    // report it
    (void)hipGetLastError();
    no_lockstep_room_ = true;
    return false;
  }
  for (int set = 1; set < n_shots_; ++set) ensure_zero_table(set);
  std::vector<StepArgs> descs;
  for (auto &m : models_)
    for (int s = 0; s < n_shots_; ++s)
      for (int p = 0; p < 2; ++p) descs.push_back(m->forward_launch_args(s, p, s));
  hip_check(hipMalloc((void **)&ldescs_, descs.size() * sizeof(StepArgs)), "hipMalloc");
  hip_check(hipMemcpy(ldescs_, descs.data(), descs.size() * sizeof(StepArgs),
                      hipMemcpyHostToDevice),
            "upload");
  return true;
}

// Zero the state sets of every member (one per shot), then advance `nsteps`
// steps of every shot with one launch per step (lockstep_available must
// hold).
void Batch::enqueue_lockstep_loop(int nsteps) {
  for (int set = 0; set < n_shots_; ++set) enqueue_zero_states(set);
  for (int k = 0; k < nsteps; ++k) {
    BatchLaunch b{ldescs_, tiles_, ntiles_ * n_shots_, n_shots_, -1, k & 1,
                  k, k + 1, xcd_remap_, 0, 1};
    hip_check(launch_batch_step(variant_, false, b, stream_), "batch step");
  }
}

void Batch::forward_all() {
  if (!lockstep_available()) {
    for (int s = 0; s < n_shots_; ++s) forward(s);
    return;
  }
  hip_check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
  if (graph_mode_) {
    // the whole loop (state zeroing and nt launches) captured once and
    // replayed: one host launch per call; the tables it reads were built
    // by lockstep_available, outside the capture
    if (!graph_all_) {
      hip_check(hipStreamSynchronize(stream_), "sync");
      hipGraph_t graph = nullptr;
      hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
                "hipStreamBeginCapture");
      try {
        enqueue_lockstep_loop(nt_);
      } catch (...) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(stream_, &g2);
        if (g2) (void)hipGraphDestroy(g2);
        throw;
      }
      hip_check(hipStreamEndCapture(stream_, &graph), "hipStreamEndCapture");
      hipGraphExec_t exec = nullptr;
      hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      hip_check(e, "hipGraphInstantiate");
      graph_all_ = exec;
      hip_check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
    }
    hip_check(hipGraphLaunch(graph_all_, stream_), "hipGraphLaunch");
  } else {
    enqueue_lockstep_loop(nt_);
  }
  hip_check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "hipEventElapsedTime");
  loop_ms_ += ms;
}

// Time loop of shot i_shot for every member with the gradient descriptors:
// forward storing each member's snapshots of this shot, or adjoint (time
// steps nt-1 .. 0) correlating with them.
void Batch::enqueue_gradient_loop(int i_shot, bool forward) {
  enqueue_zero_states(0);
  const StepArgs *descs = forward ? gfwd_ : gadj_;
  for (int k = 0; k < nt_; ++k) {
    const int it = forward ? k : nt_ - 1 - k;
    BatchLaunch b{descs, tiles_, ntiles_, n_shots_, i_shot, k & 1, it, it, xcd_remap_,
                  1, si_};
    hip_check(launch_batch_step(variant_, false, b, stream_, nullptr, nullptr,
                                !forward && it % si_ == 0),
              "batch gradient step");
  }
}

void Batch::gradients() {
  if (!batch_gradient_variant_ok(variant_))
    throw ValueError("batched gradients need step variant 0, 1, 8-13 or 18-28");
  si_ = models_[0]->config().snapshot_interval;
  for (auto &m : models_)
    if (m->config().snapshot_interval != si_)
      throw ValueError("batched gradients need one snapshot_interval for all members");
  if (nt_ % si_ != 0)
    throw ValueError("Snapshot interval and size of accumulator don't match! "
                     "(nt must be a multiple of snapshot_interval)");
  if (!gfwd_) {
    std::vector<StepArgs> f, a;
    for (auto &m : models_) {
      double *const *store = m->shot_snapshot_store();
      for (int s = 0; s < n_shots_; ++s)
        for (int p = 0; p < 2; ++p) {
          f.push_back(m->forward_launch_args(s, p));
          a.push_back(m->adjoint_launch_args(s, p));
          for (int q = 0; q < 5; ++q) f.back().snap[q] = a.back().snap[q] = store[q];
        }
    }
    hip_check(hipMalloc((void **)&gfwd_, f.size() * sizeof(StepArgs)), "hipMalloc");
    hip_check(hipMalloc((void **)&gadj_, a.size() * sizeof(StepArgs)), "hipMalloc");
    hip_check(hipMemcpy(gfwd_, f.data(), f.size() * sizeof(StepArgs), hipMemcpyHostToDevice),
              "upload");
    hip_check(hipMemcpy(gadj_, a.data(), a.size() * sizeof(StepArgs), hipMemcpyHostToDevice),
              "upload");
  }
  for (auto &m : models_) m->reset_kernels();
  hip_check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
  for (int s = 0; s < n_shots_; ++s) {
    // forward s, its adjoint sources (each member's own stream), adjoint s:
    // the kernels accumulate shot by shot, as in the reference's order
    enqueue_gradient_loop(s, true);
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    for (auto &m : models_) m->calculate_l2_adjoint_sources();
    enqueue_gradient_loop(s, false);
  }
  hip_check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
  hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "hipEventElapsedTime");
  loop_ms_ += ms;
  for (auto &m : models_) {
    m->calculate_l2_misfit();
    m->map_kernels_to_velocity();
  }
}

void Batch::misfits(double *out) {
  for (size_t m = 0; m < models_.size(); ++m)
    out[m] = models_[m]->calculate_l2_misfit();
}

double Batch::autotune(int steps, int *best, double *ms_per_variant) {
  if (steps < 2) throw ValueError("autotune needs at least 2 steps");
  const int candidates[kBatchCandidates] = {
      kFusedTz8, kFusedTz16, kTwoStepTz8, kTwoStepTz16, kLeanTz16,
      kLeanTz8,  kLeanTz12,  kLeanTz20,   kLeanTz24,    kLeanTz32,
      kLeanTz4,  kLeanTz8W6, kLeanTz12W5, kLeanTz8Sep, kLeanTz4Sep, kLeanTz8Ring,
      kLeanTz16Ring, kLeanTz8RingClamp, kLeanTz16RingClamp};
  const int keep_remap = xcd_remap_;
  double best_ms = std::numeric_limits<double>::infinity();
  int best_v = variant_;
  bool lean = true;
  for (auto &m : models_) lean = lean && m->lean_ok();
  for (int k = 0; k < kBatchCandidates; ++k) {
    if (ms_per_variant) ms_per_variant[k] = -1.0;
    if (step_is_lean(candidates[k]) && !lean) continue;
    set_variant(candidates[k], keep_remap);
    // the loop forward_all runs with this variant: the shots in lockstep
    // where it can, else shot 0 alone; timed per step of one shot
    const bool lock = lockstep_available();
    const int n = std::min(steps, nt_);
    auto run = [&]() {
      if (lock)
        enqueue_lockstep_loop(n);
      else
        enqueue_loop(0, n);
    };
    run();  // warm up
    hip_check(hipEventRecord(ev_[0], stream_), "hipEventRecord");
    run();
    hip_check(hipEventRecord(ev_[1], stream_), "hipEventRecord");
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, ev_[0], ev_[1]), "hipEventElapsedTime");
    double per = ms / (n * (lock ? n_shots_ : 1));
    if (ms_per_variant) ms_per_variant[k] = per;
    if (per < best_ms) {
      best_ms = per;
      best_v = candidates[k];
    }
  }
  set_variant(best_v, keep_remap);
  if (best) *best = best_v;
  return best_ms;
}

}  // namespace psv
