// psv_batch.h -- many independent fdModels advanced together: one launch per
// time step covers every tile of every model (Monte-Carlo forward runs over
// models of uneven size).
#pragma once

#include <memory>
#include <vector>

#include "psv_model.h"

namespace psv {

// Variants autotune() times, in this order: 0, 1, 6, 7, 8, 9, 10, 11, 12, 13,
// 18, 19, 20, 22, 23, 24, 26, 27, 28.
constexpr int kBatchCandidates = 19;

class Batch {
 public:
  // All members must share nt and n_shots.
  explicit Batch(const std::vector<Config> &cfgs);
  ~Batch();
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  int size() const { return (int)models_.size(); }
  Model &model(int m);
  const Model &model(int m) const;

  // Step kernel for every member: 0 / 1 one-step 64x8 / 64x16 tiles,
  // 6 / 7 two-step 64x8 / 64x16 tiles, 8-13 lean one-step (64 x 16, 8,
  // 12, 20, 24, 32 tiles).
  void set_variant(int variant, int xcd_remap);
  int variant() const { return variant_; }
  void set_graph_mode(bool on);
  // One cooperative launch per shot's whole time loop (variant 9 only): the
  // blocks stay resident and meet at a grid barrier between steps, instead
  // of one launch per step.
  void set_persistent(bool on);
  // Forward simulation of shot i_shot for every member (no snapshots).
  void forward(int i_shot);
  // Forward simulation of every shot of every member.  With 2 to
  // Model::kMaxStateSets shots and a lean step variant (no persistent loop)
  // the shots advance together: one launch per time step over every (tile,
  // shot) of every member, shot s in the member's state buffer set s, so
  // the reads of a tile's materials meet in L2; in graph mode the whole loop
  // is one captured HIP graph.  Otherwise, or when the device has no room
  // for the extra state sets, shot by shot.
  // Seismograms are those of forward(0), forward(1), ... bit for bit.
  void forward_all();
  bool lockstep_available();
  // Per-member L2 misfit over all shots (calculate_l2_misfit of each).
  void misfits(double *out);
  // FWI gradient of every member against its observed data (the reference's
  // run_model, src/fdModel.cpp:1291-1308): kernels reset, then shot by shot
  // the forward storing snapshots, the adjoint sources and the adjoint with
  // kernel correlation, each one launch per time step over every member's
  // tiles; then misfits and the kernel mapping.  Each member keeps one
  // shot's snapshots (5 nx nz doubles per snapshot).  Kernels and misfits
  // equal the members' own run_model bit for bit.
  void gradients();
  // Times `steps` batched steps of each candidate variant (ms_per_variant
  // gets kBatchCandidates entries, -1 for a variant the batch cannot use),
  // keeps the fastest; returns its time per step of one shot in ms.  Each
  // variant is timed in the loop forward_all runs with it: all shots in
  // lockstep where that applies, else shot 0 alone.
  double autotune(int steps, int *best, double *ms_per_variant);
  long long grid_points() const { return points_; }
  double loop_ms() const { return loop_ms_; }
  void reset_timing() { loop_ms_ = 0.0; }

 private:
  void build_tables();
  void enqueue_loop(int i_shot, int nsteps);
  void enqueue_lockstep_loop(int nsteps);
  // every member's state buffer set `set` zeroed in one launch
  void ensure_zero_table(int set);
  void enqueue_zero_states(int set);
  double **zero_ptrs_[Model::kMaxStateSets] = {};
  long long *zero_lens_ = nullptr;
  long long zero_max_ = 0;
  void clear_graphs();
  void clear_gradient_tables();
  void enqueue_gradient_loop(int i_shot, bool forward);

  std::vector<std::unique_ptr<Model>> models_;
  int nt_ = 0, n_shots_ = 0, variant_ = kFusedTz8, xcd_remap_ = 1;
  hipStream_t stream_ = nullptr;
  StepArgs *descs_ = nullptr;  // [model][shot][parity]
  StepArgs *ldescs_ = nullptr; // the same, shot s in state set s (forward_all)
  bool no_lockstep_room_ = false;
  int2 *tiles_ = nullptr;      // (model, tile) pairs
  int ntiles_ = 0;
  long long points_ = 0;
  bool graph_mode_ = false;
  bool persistent_ = false;
  std::vector<hipGraphExec_t> graphs_;  // per shot
  hipGraphExec_t graph_all_ = nullptr;  // forward_all's loop of all shots
  hipEvent_t ev_[2] = {nullptr, nullptr};
  double loop_ms_ = 0.0;
  // gradient descriptors [model][shot][parity] with snapshot-store bases
  StepArgs *gfwd_ = nullptr, *gadj_ = nullptr;
  int si_ = 1;
};

}  // namespace psv
