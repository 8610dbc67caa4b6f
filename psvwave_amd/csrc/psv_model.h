// psv_model.h -- device-resident fdModel: owns the HBM state of one model and
// drives the time loops.  The C-ABI (psv_capi.hip) is a thin layer over it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <string>
#include <vector>

#include "psv_config.h"
#include "psv_kernels.h"

namespace psv {

struct RuntimeError : public std::runtime_error {
  explicit RuntimeError(const std::string &m) : std::runtime_error(m) {}
};

// HIP streams the library holds open in this process (model, lane and
// batch streams).  Two time loops overlap only on two hardware queues: the
// HIP runtime creates at most GPU_MAX_HW_QUEUES (4 by default) per device
// and shares them among the streams beyond that.
void create_stream(hipStream_t *s);  // hipStreamNonBlocking; throws on failure
void destroy_stream(hipStream_t s);
int live_streams();
int hw_queue_limit();  // GPU_MAX_HW_QUEUES, or 4 when unset or invalid

class Model {
 public:
  // shared_stream: run on that stream (a batch's) instead of one of its
  // own; the caller keeps it alive until the model is destroyed
  explicit Model(const Config &c, hipStream_t shared_stream = nullptr);
  Model(const Model &other);  // deep copy of device state
  ~Model();
  Model &operator=(const Model &) = delete;

  const Config &config() const { return cfg_; }
  const Derived &derived() const { return d_; }
  double misfit() const { return misfit_; }

  // hot path (src/fdModel.cpp:526-969, 1128-1183, 1310-1324)
  void forward(int i_shot, bool store_fields, bool verbose,
               bool output_wavefields);
  void adjoint(int i_shot, bool verbose);
  double calculate_l2_misfit();
  void calculate_l2_adjoint_sources();
  void reset_kernels();
  void map_kernels_to_velocity();
  void run_model(bool verbose, bool simulate_adjoint);

  // Asynchronous per-shot calls (on by default): forward() and adjoint()
  // enqueue their time loops and return.  Forwards alternate between two
  // lanes (HIP streams with their own state buffers); adjoints pair up --
  // the first of a pair updates the kernels, the second runs beside it on
  // the other lane with its kernel terms deferred to a buffer and added
  // afterwards in its own order -- so every result is bitwise that of the
  // synchronous calls.  settle() waits for the enqueued loops (and adds
  // pending deferred terms); every other method must be called on a settled
  // model -- the C-ABI settles before each of them.  Verbose calls, timing,
  // graph mode, checkpointing and the two-step and in-place variants run
  // synchronously.
  void set_async(bool on);
  bool async() const { return async_; }
  void settle();

  // parameters and data, host buffers in the reference layout
  void get_parameter_fields(double *vp, double *vs, double *rho) const;
  void set_parameter_fields(const double *vp, const double *vs,
                            const double *rho);
  void get_kernels(double *vp_k, double *vs_k, double *rho_k) const;
  void get_lame_kernels(double *lam_k, double *mu_k, double *rho_l) const;
  void get_traces(int which, double *ux, double *uz) const;  // 0 syn 1 obs 2 adj
  void set_traces(int which, const double *ux, const double *uz);
  void get_snapshot(int i_shot, int i_snap, double *const out[5]) const;
  void get_snapshots(double *const out[5]) const;
  void get_model_vector(double *v) const;
  void set_model_vector(const double *v);
  void get_gradient_vector(double *v) const;
  void get_setup(double *stf, double *moment, double *taper) const;

  // text formats (src/fdModel.cpp:971-1126, 1185-1289, 1326-1356, 1517-1572)
  void write_receivers() const;
  void write_sources() const;
  void write_kernels() const;
  void load_receivers(bool verbose);
  void load_model(const std::string &de, const std::string &vp,
                  const std::string &vs, bool verbose);
  std::vector<double> load_vector(const std::string &path, bool verbose) const;

  // tuning and measurement
  void set_step_variant(int variant, int xcd_remap);
  // stride k > 0: HIP start/stop timestamps of every k-th step launch;
  // k < 0: events bracketing whole time loops (launches and the gaps
  // between them); 0: off.
  void set_timing(int stride) { timing_stride_ = stride; }
  // Capture each (direction, shot, store) time loop once as a HIP graph and
  // replay it; step timing then brackets whole loops.
  void set_graph_mode(bool on);
  // Keep every `every`-th forward snapshot only (1: all) and recompute the
  // others during the adjoint, one segment at a time.
  void set_checkpointing(int every);
  int checkpointing() const { return ckpt_every_; }
  void step_timing(double *ms, int64_t *timed, int64_t *launches) const {
    *ms = timed_ms_;
    *timed = timed_;
    *launches = launches_;
  }
  void reset_step_timing() { timed_ms_ = 0, timed_ = 0, launches_ = 0; }
  int64_t device_bytes() const { return device_bytes_; }

  // Pieces of a forward time loop for batched launches across models
  // (psv_batch.hip): the arguments of a launch that reads state buffer set
  // `parity`, zeroing the state, and the receiver scan of two-step loops.
  // set s > 0: another state buffer set (ensure_state_sets), so that up to
  // kMaxStateSets shots can advance side by side.
  StepArgs forward_launch_args(int i_shot, int parity, int set = 0) const;
  // Batched gradients: the arguments of an adjoint launch reading state set
  // `parity` (adjoint sources of shot i_shot, kernel correlation), and a
  // snapshot store for one shot at a time ([slot][iz][ix] per field) whose
  // bases both launch kinds get in StepArgs::snap.
  StepArgs adjoint_launch_args(int i_shot, int parity) const;
  double *const *shot_snapshot_store();
  void enqueue_zero_state(hipStream_t s, int set = 0) const;
  // The ten padded arrays of a state buffer set (2 parities x 5 fields) and
  // their length, for batched zeroing.
  void state_arrays(int set, std::vector<double *> &out) const;
  long long padded_points() const { return npad_; }
  void ensure_second_state();
  // state buffer sets 0 .. n-1 allocated (n <= kMaxStateSets); set 1 is the
  // second lane's
  void ensure_state_sets(int n);
  static constexpr int kMaxStateSets = 8;
  void enqueue_receiver_scan(int i_shot, hipStream_t s) const;
  int tile_count() const { return ntiles_; }
  bool lean_ok() const { return lean_ok_; }
  int step_variant() const { return variant_; }
  long long grid_points() const { return ngrid_; }

 private:
  void allocate();
  void build_injection_tables();
  void ensure_snapshots();
  double *dalloc(size_t n);
  void upload(double *dst, const double *src, size_t n) const;
  void download(double *dst, const double *src, size_t n) const;
  double *staging(size_t n) const;
  void to_device_layout(double *dst_compact, const double *host_ref) const;
  void to_host_layout(double *host_ref, const double *src_compact,
                      int slices = 1) const;
  StepArgs base_args() const;
  void run_steps(StepArgs &a, int i_shot, bool forward, bool store,
                 bool dump);
  void sync() const;
  void check_shot(int i_shot) const;

  Config cfg_;
  Derived d_;
  int variant_ = kFusedTz8, xcd_remap_ = 1;
  int tz_ = 8, ntx_ = 0, ntz_ = 0, ntiles_ = 0;
  int pitch_ = 0, rows_ = 0;
  long long origin_ = 0, npad_ = 0, ngrid_ = 0, nrec_ = 0;
  hipStream_t stream_ = nullptr;
  bool owns_stream_ = true;
  std::vector<void *> allocations_;
  int64_t device_bytes_ = 0;

  double *state_[2][5] = {};  // ping-pong: vx, vz, txx, tzz, txz (padded)
  double *lm_ = nullptr, *mu_ = nullptr, *b_ = nullptr;  // padded
  double *vp_ = nullptr, *vs_ = nullptr, *rho_ = nullptr;  // compact
  double *lam_k_ = nullptr, *mu_k_ = nullptr, *rho_l_ = nullptr;
  double *vp_k_ = nullptr, *vs_k_ = nullptr, *rho_v_ = nullptr;
  double *stf_ = nullptr, *taper_lut_ = nullptr;
  double *taper_x_ = nullptr, *taper_z_ = nullptr;  // per column / per row
  double *taper_xp_ = nullptr, *taper_zp_ = nullptr; // their padded storage
  bool lean_ok_ = false;
  bool lean_eligible() const;
  double taper_of_count(int c) const;
  double *syn_[2] = {}, *obs_[2] = {}, *adj_[2] = {};
  double *snap_[5] = {};
  double *shot_snap_[5] = {};  // one shot's snapshots (batched gradients)
  long long *rec_idx_ = nullptr;
  int *fwd_off_ = nullptr, *adj_off_ = nullptr;
  InjEntry *fwd_entries_ = nullptr, *adj_entries_ = nullptr;
  int *fwd_halo_off_ = nullptr, *adj_halo_off_ = nullptr;
  InjEntry *fwd_halo_ = nullptr, *adj_halo_ = nullptr;
  double *misfit_buf_ = nullptr;  // partials then the result
  mutable double *staging_ = nullptr;
  mutable size_t staging_n_ = 0;
  double misfit_ = 0.0;

  int timing_stride_ = 0;
  double timed_ms_ = 0.0;
  int64_t timed_ = 0, launches_ = 0;
  std::vector<hipEvent_t> events_;
  size_t event_cursor_ = 0;  // next free entry of events_ in a timed loop
  bool graph_mode_ = false;
  std::map<long long, hipGraphExec_t> graphs_;
  void clear_graphs();
  void timed_loop(const std::function<void()> &enqueue);
  using SlotFn = std::function<bool(int, double **)>;
  int enqueue_range(StepArgs &a, double *(*st)[5], int parity, bool forward,
                    int k0, int k1, bool snaps, const SlotFn &slot,
                    std::vector<std::pair<hipEvent_t, hipEvent_t>> *pairs,
                    bool dump, hipStream_t s = nullptr, double *defer = nullptr);
  StepArgs forward_args(int i_shot) const;
  StepArgs adjoint_args(int i_shot) const;
  // run_model with two lanes: forward time loops on lane_stream_ (state in
  // lane_state_), adjoint time loops in shot order on stream_; see
  // psv_model.hip.
  void run_model_two_lanes(bool simulate_adjoint);
  hipStream_t lane_stream_ = nullptr;
  double *lane_state_[2][5] = {};
  double *extra_state_[kMaxStateSets - 2][2][5] = {};  // sets 2 ..
  double *const (*state_set(int set) const)[5];
  std::vector<hipEvent_t> lane_events_;
  double *defer_ = nullptr;  // deferred correlation terms [slot][3][ngrid]
  long long defer_n_ = 0;
  bool ensure_defer_buffer();
  void enqueue_loop(StepArgs &a, int i_shot, bool forward, bool store,
                    std::vector<std::pair<hipEvent_t, hipEvent_t>> *pairs,
                    bool dump);
  void enqueue_checkpointed_adjoint(
      StepArgs &aa, int i_shot,
      std::vector<std::pair<hipEvent_t, hipEvent_t>> *pairs);
  bool stored_slot(int i_shot, int it, double **dst) const;
  int stored_snapshots() const;
  int ckpt_every_ = 1;
  double *state2_[2][5] = {};  // forward recomputation buffers (padded)
  double *seg_[5] = {};        // snapshots of two recomputed segments
  int seg_k_ = 0;              // slots per segment buffer in seg_
  std::vector<hipEvent_t> ckpt_events_;
  int *scratch_off_ = nullptr; // all-zero CSR offsets

  // asynchronous per-shot calls (see set_async): each call adds a job --
  // one time loop on one lane, or the replay of deferred kernel terms --
  // and pump() enqueues the jobs' time steps in chunks, alternating between
  // the lanes so that neither stream's queue runs dry while the host is
  // blocked on the other one.  A job waits for its dependencies (jobs added
  // before it) through their end events.
  struct LaneJob {
    int kind;    // kJobForward, kJobAdjoint, kJobDeferred, kJobReplay
    int shot;
    bool store;
    int dep[2];  // ids of jobs that must end first, -1 for none
    int lane, k, parity;
    bool started, done;
    StepArgs a;
  };
  bool async_ = true;
  std::vector<LaneJob> jobs_;           // this batch of calls, by id
  std::deque<int> lane_queue_[2];       // unfinished job ids per lane, in order
  std::vector<hipEvent_t> job_events_;  // end event of each job
  std::vector<int> fwd_job_;            // per shot: its pending forward job, or -1
  int fwd_calls_ = 0, adj_calls_ = 0;   // since the last settle
  int deferred_job_ = -1;               // deferred adjoint not yet replayed
  int last_replay_job_ = -1;            // its deferred-term buffer is free after it
  bool no_defer_room_ = false;          // the device had no room for defer_
  hipEvent_t ev_join_ = nullptr;
  bool async_path(bool verbose, bool dump) const;
  void ensure_lane();
  int push_job(int lane, int kind, int shot, bool store, int dep0, int dep1);
  void pump(bool all);
};

}  // namespace psv
