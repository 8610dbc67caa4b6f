"""Benchmark: fwd+adj grid-point updates per second on one MI355X.

Workload: the layered-faulted single model -- 1700 x 350 inner grid plus a
50-point absorbing taper (1800 x 450 = 810,000 points), dx = dz = 5 m,
dt = 0.3 ms, nt = 6000, FP64, 8 shots, 300 receivers, snapshots every 10
steps.  One step = one FWI gradient: forward of all 8 shots with snapshot
storage, L2 misfit, adjoint sources, kernel reset, adjoint of all 8 shots
with kernel correlation, kernel mapping -- 7.776e10 point-updates.

Prints one JSON line.  ``value`` is point-updates per second of the whole
step with all inputs resident in HBM.  ``roofline`` prices the step kernel
(the dominant kernel) at its algorithmic bytes per launch over its average
duration when it runs alone (HIP start/end events of every launch of one
shot's forward + adjoint on one stream, after the timed region);
``roofline.device_aggregate`` prices the timed region instead: HIP events
around every run_model (two time loops at a time on two streams), divided
by the time steps launched.  ``roofline.traffic`` is the bytes per launch
from the committed rocprofv3 counter summary for the kernel.
``cpu_baseline`` times the CPU oracle (OpenMP over the CPUs this process may
use, threads bound to cores) on a bounded sample of the same grid.
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("fwd+adj Gpt-updates/s and % HBM peak; FWI time-to-misfit, "
          "learned vs plain")
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published


def _progress(msg):
    """One progress line on stderr (the JSON result goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (profiles/<kernel>_hbm_traffic.md, line `bytes_per_launch = N`),
    or None when no summary for this kernel exists."""
    name = kernel.replace("<", "").replace(">", "").replace(",", "_").replace(" ", "_")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"{name}_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.strip().startswith("bytes_per_launch ="):
                    return float(line.split("=", 1)[1])
    except OSError:
        return None
    return None


def _args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=2)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--variant", type=int, default=27,
                   help="step kernel: 0 fused 64x8, 1 fused 64x16, 2 split; "
                        "3-5 the same with reciprocal-multiply division; "
                        "6, 7 two steps per launch (64x8, 64x16); 8-13 lean "
                        "one-step kernels with 64 x 16, 8, 12, 20, 24, 32 "
                        "tiles; 14 lean two "
                        "steps per launch (64x8); 15-17 lean with 2, 3, 4 "
                        "tiles of a strip per block; 18 lean 64x4; 19, 20 "
                        "lean 64x8 / 64x12 at 6 / 5 waves per SIMD; 21 lean "
                        "64x8 with halo columns by DMA; 22, 23 lean 64x8 / "
                        "64x4 with a separate stress stage; 24 lean 64x8 "
                        "with halo columns and ring "
                        "operands by DMA, 6 waves per SIMD; 25 the same "
                        "without the halo columns; 26 lean 64x16 with ring "
                        "operands by DMA, 4 waves per SIMD; 27 (the library "
                        "default), 28 variants 24, 26 with loads capped to "
                        "the grid")
    p.add_argument("--no-xcd-remap", action="store_true")
    p.add_argument("--sequential", action="store_true",
                   help="drive the shots call by call (forward_simulate per "
                        "shot, misfit, adjoint sources, adjoint_simulate per "
                        "shot) instead of run_model, which runs forward and "
                        "adjoint loops of different shots side by side on two "
                        "HIP streams (bitwise the same results)")
    p.add_argument("--graph", action="store_true",
                   help="replay each shot's time loop as a captured HIP graph")
    p.add_argument("--timing-stride", type=int, default=-1,
                   help="-1: HIP events around every time loop (average per "
                        "time step, gaps included); k > 0: kernel start/end "
                        "of every k-th step launch")
    p.add_argument("--nt", type=int, default=6000)
    p.add_argument("--shots", type=int, default=8)
    p.add_argument("--cpu-baseline-nt", type=int, default=4000,
                   help="time steps of the CPU sample (fwd + adj, 1 shot)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                   help="first run a plain copy kernel over BYTES (the "
                        "reference point for byte counters) and report it")
    p.add_argument("--only-calibrate", action="store_true")
    p.add_argument("--workload", choices=["single", "mc", "c4", "c5", "unet"],
                   default="single",
                   help="single: the layered-faulted model (default); mc: the "
                        "512-model Monte-Carlo forward batch; c4: gradient "
                        "through torch.autograd with checkpointing vs full "
                        "snapshot storage, checked by finite differences; c5: "
                        "FWI time-to-misfit with the learned preconditioner; "
                        "unet: the preconditioner network's training step and "
                        "inference on the library's MFMA kernels vs torch")
    p.add_argument("--c4-directions", type=int, default=20)
    p.add_argument("--c5-pairs", type=int, default=8192)
    p.add_argument("--c5-epochs", type=int, default=16)
    p.add_argument("--c5-iterations", type=int, default=30)
    p.add_argument("--c5-along-path", type=float, default=0.5,
                   help="fraction of the training pairs taken part of the way "
                        "from the background to the true model")
    p.add_argument("--c5-seeds", type=int, default=4,
                   help="networks trained (seeds 0, 1, ...); each one's "
                        "time-to-misfit is reported, with the median")
    p.add_argument("--c5-holdout", type=int, default=2,
                   help="held-out models the FWI runs on")
    p.add_argument("--c5-widths", default="32,48,64,128",
                   help="U-Net channel widths of the four levels")
    p.add_argument("--c5-no-mirror", action="store_true",
                   help="use the network's prediction alone instead of its "
                        "mean with the mirrored prediction of the mirrored "
                        "gradient")
    p.add_argument("--c4-every", type=int, default=8)
    p.add_argument("--mc-models", type=int, default=512)
    p.add_argument("--persistent", action="store_true",
                   help="mc: run each shot's forward loop as one cooperative "
                        "launch with a grid barrier between steps (variant 9)")
    return p.parse_args()


def calibrate_copy(nbytes, iters=5, runtime_copy=False):
    """Mean time of a device copy of nbytes (the library's plain copy kernel,
    or hipMemcpyAsync with runtime_copy); returns (ms, GB/s of read + write
    traffic)."""
    import ctypes as C
    from psvwave_amd import _lib
    lib = _lib.load()
    ms = C.c_double()
    _lib.check(lib.psv_calibrate_copy(int(nbytes), int(iters), int(bool(runtime_copy)),
                                      C.byref(ms)))
    return ms.value, 2.0 * nbytes / (ms.value * 1e-3) / 1e9


def stream_ceiling(launch_bytes, achieved_gbs):
    """The device's copy rate -- read + write bytes over time, HIP events
    around 20 copies, the better of the library's plain copy kernel and
    hipMemcpyAsync device-to-device (the runtime's copy is faster on sizes
    the Infinity Cache holds, the kernel on HBM-sized ones) -- with as many
    bytes as one step launch moves and with 2 GB: the practical ceiling of
    a streaming kernel beside the published peak."""
    out = {}
    for label, total in (("launch_bytes", launch_bytes), ("2GB", 2e9)):
        nb = int(total / 2) // 16 * 16
        kernel = calibrate_copy(nb, iters=20)[1]
        runtime = calibrate_copy(nb, iters=20, runtime_copy=True)[1]
        out[f"copy_GBs_at_{label}"] = round(max(kernel, runtime), 1)
        out[f"copy_kernel_runtime_GBs_at_{label}"] = [round(kernel, 1), round(runtime, 1)]
    out["frac_of_copy_at_launch_bytes"] = round(achieved_gbs / out["copy_GBs_at_launch_bytes"], 4)
    return out


def _host_cpu():
    """CPU model, sockets and physical cores of the host (lscpu), and the
    logical CPUs this process may run on."""
    import subprocess
    info = {"model": None, "sockets": None, "physical_cores": None,
            "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {k.strip(): v.strip() for k, _, v in
              (line.partition(":") for line in out.splitlines())}
        info["model"] = kv.get("Model name")
        sockets = int(kv.get("Socket(s)", "0") or 0)
        per = int(kv.get("Core(s) per socket", "0") or 0)
        info["sockets"] = sockets or None
        info["physical_cores"] = sockets * per or None
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity_cpus"] = os.cpu_count()
    return info


def cpu_baseline(nt_sample):
    """CPU oracle on a bounded sample: one shot, forward + adjoint, the
    benchmark grid, nt_sample steps.  The oracle runs one OpenMP thread per
    CPU this process may use (its affinity mask), or OMP_NUM_THREADS if that
    is set lower (a shared host's per-job CPU share), with threads bound
    close, one place per core; the library reads the binding at load, so it
    is set before the oracle is loaded."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle.pyoracle import OracleModel
    from psvwave_amd.synthetic import layered_faulted_config, layered_faulted_model
    host = _host_cpu()
    threads = host["affinity_cpus"] or 1
    env_cap = os.environ.get("OMP_NUM_THREADS", "")
    if env_cap.isdigit() and int(env_cap) > 0:
        threads = min(threads, int(env_cap))
    OracleModel.set_threads(threads)
    cfg = layered_faulted_config(nt=nt_sample, n_shots=1)
    true = layered_faulted_model(cfg, faulted=False)
    start = layered_faulted_model(cfg, faulted=True)
    o = OracleModel(cfg)
    o.set_parameter_fields(*true)
    o.forward_simulate(0, store_fields=False)
    o.set_observed_data(*o.get_synthetic_data())
    o.set_parameter_fields(*start)
    t0 = time.perf_counter()
    o.forward_simulate(0)
    o.calculate_l2_misfit()
    o.calculate_l2_adjoint_sources()
    o.reset_kernels()
    o.adjoint_simulate(0)
    o.map_kernels_to_velocity()
    secs = time.perf_counter() - t0
    points = 2.0 * o.nx * o.nz * nt_sample
    return {"value": points / secs / 1e9, "unit": "Gpt-updates/s",
            "cores": threads, "kind": "port",
            "sample": (f"1 shot fwd+adj, {o.nx}x{o.nz} grid, nt={nt_sample} "
                       f"({points:.3e} point-updates, {secs:.2f} s, "
                       f"OpenMP {threads} threads, OMP_PROC_BIND="
                       f"{os.environ.get('OMP_PROC_BIND')}, OMP_PLACES="
                       f"{os.environ.get('OMP_PLACES')})"),
            "host": host, "omp_num_threads_env": env_cap or None}


def kernel_alone_us(model):
    """Average duration in microseconds of a step launch on its own: kernel
    start/end events of every launch of one shot's forward (with snapshots)
    and adjoint (with correlation), called shot by shot on one stream, so
    the launches sample ordinary, snapshot and correlation steps in the
    benchmark's proportions."""
    model.set_timing(1)
    model.reset_step_timing()
    model.forward_simulate(0)
    model.adjoint_simulate(0)
    ms, timed, _ = model.step_timing()
    model.set_timing(0)
    return ms / max(timed, 1) * 1e3


def monte_carlo(a):
    """Forward-only throughput of the Monte-Carlo batch: 512 Gaussian
    random-field models of uneven size (87M grid points in total), 2 shots,
    nt = 3000; one step = forward of both shots for every model + misfits."""
    import psvwave_amd
    from psvwave_amd.synthetic import monte_carlo_configs, monte_carlo_model
    cfgs = monte_carlo_configs(n_models=a.mc_models)
    t0 = time.perf_counter()
    batch = psvwave_amd.Batch(cfgs)
    rng = np.random.default_rng(3)
    # observed data: the homogeneous background of each member (setup)
    for m, c in enumerate(cfgs):
        i = batch.infos[m]
        shape = (i["nx"], i["nz"])
        med = c["medium"]
        batch.set_parameter_fields(m, np.full(shape, med["scalar_vp"]),
                                   np.full(shape, med["scalar_vs"]),
                                   np.full(shape, med["scalar_rho"]))
    if a.persistent:
        batch.set_variant(9)
        batch.set_persistent(True)
        best, tuned = 9, None
    else:
        best, tuned = batch.autotune(steps=20)
    if a.graph:
        batch.set_graph_mode(True)
    batch.forward_all()
    for m in range(len(cfgs)):
        batch.set_observed_data(m, *batch.get_synthetic_data(m))
        batch.set_parameter_fields(m, *monte_carlo_model(cfgs[m], rng))
    setup = time.perf_counter() - t0

    def step():
        batch.forward_all()
        return batch.misfits()

    _progress("setup done")
    for k in range(a.warmup):
        step()
        _progress(f"warmup step {k + 1}/{a.warmup}")
    batch.timing(reset=True)
    t0 = time.perf_counter()
    for k in range(a.steps):
        mis = step()
        _progress(f"timed step {k + 1}/{a.steps}")
    secs = time.perf_counter() - t0
    loop_ms, points = batch.timing()
    nt = batch.nt
    upd = 2.0 * points * nt  # two shots
    value = upd * a.steps / secs / 1e9
    loop_s_per_step = loop_ms / 1e3 / a.steps
    # Algorithmic bytes per point-update: state read + written (80 B) and
    # the materials lm, mu, b (24 B) -- read once per time step for both
    # shots when forward_all runs them in lockstep (92 B per update).
    lock = batch.lockstep()
    bytes_per_update = 92.0 if lock else 104.0
    achieved = bytes_per_update * upd / loop_s_per_step / 1e9
    launch_bytes = bytes_per_update * points * (2 if lock else 1)
    return {
        "metric": METRIC + " (Monte-Carlo forward batch)",
        "value": round(value, 3), "unit": "Gpt-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(secs / a.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian random fields, observed data from "
                "each member's homogeneous background)",
        "config": {"workload": "Monte-Carlo forward batch", "models": len(cfgs),
                   "grid_points": points, "nt": nt, "shots": 2, "receivers": 32,
                   "point_updates_per_step": upd, "variant": best,
                   "autotune_ms_per_step": tuned, "hip_graph": a.graph,
                   "persistent": a.persistent,
                   "shots_in_lockstep": lock,
                   "setup_s": round(setup, 1)},
        "roofline": {"bound": "hbm", "kernel": f"batch step (variant {best})",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "algorithmic_bytes_per_update": bytes_per_update,
                     "algorithmic_bytes_per_launch": launch_bytes,
                     "note": ("time = whole forward loops (HIP events), gaps "
                              "included; " + ("each launch advances both shots "
                              "(materials read once for the two)" if lock else
                              "each launch advances one shot")),
                     "stream_ceiling": stream_ceiling(launch_bytes, achieved)},
        "misfit_mean": float(np.mean(mis)),
    }


def gradient_check(a):
    """C4: the FWI gradient of a 680 x 380 Gaussian-random-field model (4
    shots, 64 receivers, nt = 4000) through psvwave_amd.autograd, once with
    every snapshot stored (16.5 GB) and once with checkpointing (every k-th
    snapshot kept, the rest recomputed); the two must agree bit for bit.
    The gradient is then compared with central differences of the misfit
    along smooth random directions supported on the kernel region."""
    import torch
    import psvwave_amd
    from psvwave_amd.autograd import misfit
    from psvwave_amd.synthetic import c4_config, c4_models, gaussian_random_field
    cfg = c4_config()
    true, start = c4_models(cfg)
    res = {}
    grads = {}
    obs = None
    for every in (1, a.c4_every):
        # release the previous model first: its streams would share the
        # runtime's hardware queues with this one's two lanes
        m = f = params = None
        gc.collect()
        m = psvwave_amd.fdModel(cfg)
        m.set_checkpointing(every)
        if obs is None:
            m.set_parameter_fields(*true)
            for s in range(m.n_shots):
                m.forward_simulate(s, store_fields=False)
            obs = m.get_synthetic_data()
        m.set_observed_data(*obs)
        times = []
        for rep in range(2):  # the first call allocates snapshot and lane buffers
            params = [torch.tensor(f, dtype=torch.float64, requires_grad=True) for f in start]
            torch.cuda.synchronize() if torch.cuda.is_available() else None
            t0 = time.perf_counter()
            f = misfit(*params, m)
            f.backward()
            times.append(time.perf_counter() - t0)
        secs = times[1]
        grads[every] = [p.grad.numpy().copy() for p in params]
        res[every] = {"seconds": round(secs, 3), "first_call_seconds": round(times[0], 3),
                      "device_bytes": m.device_bytes(), "misfit": float(f.detach())}
        _progress(f"gradient with checkpoint interval {every}: {secs:.2f} s "
                  f"(first call {times[0]:.2f} s)")
        if every == a.c4_every:
            model = m
    bitwise = all(np.array_equal(x, y) for x, y in zip(grads[1], grads[a.c4_every]))
    # central differences along smooth random directions (each parameter
    # perturbed by a smooth field of up to 1 % of its mean)
    rng = np.random.default_rng(5)
    shape = start[0].shape

    def plain_misfit(fields):
        model.set_parameter_fields(*fields)
        for s in range(model.n_shots):
            model.forward_simulate(s, store_fields=False)
        model.calculate_l2_misfit()
        return model.misfit

    # the reference gradient is accumulated on the kernel region only
    # (src/fdModel.cpp:817-861), so directions live there
    dd_ = cfg["domain"]
    npb = cfg["boundary"]["np_boundary"]
    mask = np.zeros(shape)
    mask[npb + dd_["nx_inner_boundary"]:npb + dd_["nx_inner"] - dd_["nx_inner_boundary"],
         npb + dd_["nz_inner_boundary"]:npb + dd_["nz_inner"] - dd_["nz_inner_boundary"]] = 1.0
    errors = []
    for k in range(a.c4_directions):
        d = [mask * gaussian_random_field(shape, 20, rng) * rng.choice([-1.0, 1.0]) * 0.01
             * f.mean() for f in start]
        h = 0.5
        fp = plain_misfit([f + h * dd for f, dd in zip(start, d)])
        fm = plain_misfit([f - h * dd for f, dd in zip(start, d)])
        fd = (fp - fm) / (2 * h)
        g = float(sum(np.sum(gg * dd) for gg, dd in zip(grads[a.c4_every], d)))
        errors.append(abs(g - fd) / abs(fd))
        _progress(f"direction {k + 1}/{a.c4_directions}: gradient {g:.6e} "
                  f"finite difference {fd:.6e}")
    full, ck = res[1], res[a.c4_every]
    n = shape[0] * shape[1] * cfg["domain"]["nt"] * cfg["sources"]["n_shots"]
    return {
        "metric": "C4 FWI gradient through torch.autograd with checkpointing",
        "value": round(3.0 * n / ck["seconds"] / 1e9, 3),
        "unit": "Gpt-updates/s (forward + recompute + adjoint of the checkpointed gradient)",
        "n_gpus": 1, "higher_is_better": True, "dtype": "f64", "data": "synthetic (seeded GRF)",
        "config": {"workload": "C4 gradient check", "grid": f"{shape[0]}x{shape[1]}",
                   "nt": cfg["domain"]["nt"], "shots": cfg["sources"]["n_shots"],
                   "receivers": cfg["receivers"]["nr"], "checkpoint_interval": a.c4_every},
        "full_storage": full, "checkpointed": ck,
        "device_bytes_ratio": round(ck["device_bytes"] / full["device_bytes"], 4),
        "bitwise_equal_gradients": bitwise,
        "fd_relative_errors": [round(e, 5) for e in errors],
        "fd_max_relative_error": round(max(errors), 5) if errors else None,
        "fd_median_relative_error": round(float(np.median(errors)), 5) if errors else None,
    }


def learned_fwi(a):
    """C5: FWI time-to-misfit with the learned preconditioner against plain
    steepest descent on the same GPU.  Training pairs (velocity kernels at a
    background or part of the way to the true model, remaining relative
    perturbation) come from the batched device solver; --c5-seeds U-Nets
    are trained on the library's bfloat16 MFMA kernels (HipUNetTrainer), one
    per seed; on each held-out model the plain run sets the target misfit
    (its final one) and each network's learned run (its direction tried
    first, the plain gradient as fallback, the same line search) is timed to
    that misfit.  `value` is the median over seeds and held-out models of
    learned time / plain time (lower is better); runs that never reach the
    target count as infinite."""
    import torch
    import psvwave_amd
    from psvwave_amd import precond
    widths = tuple(int(c) for c in a.c5_widths.split(","))
    t0 = time.perf_counter()
    grads, perts = precond.gradient_pairs(a.c5_pairs, seed=7, log=_progress,
                                          along_path=a.c5_along_path)
    t_pairs = time.perf_counter() - t0
    _progress(f"{a.c5_pairs} training pairs in {t_pairs:.1f} s")
    unets, hist, t_train = [], [], []
    for seed in range(a.c5_seeds):
        t0 = time.perf_counter()
        torch.manual_seed(seed)
        net = precond.make_unet(widths)
        hist.append(precond.train_hip(net, grads, perts, epochs=a.c5_epochs, log=_progress))
        unets.append(precond.HipUNet(net.eval()))
        t_train.append(time.perf_counter() - t0)
        _progress(f"network {seed}: trained in {t_train[-1]:.1f} s")
    params = sum(p.numel() for p in net.parameters())
    del grads, perts
    model = psvwave_amd.fdModel(precond.c5_config())
    runs = []
    for h in range(a.c5_holdout):
        bg, true = precond.c5_models(np.random.default_rng(12345 + h))
        model.set_parameter_fields(*true)
        for s in range(model.n_shots):
            model.forward_simulate(s, store_fields=False)
        model.set_observed_data(*model.get_synthetic_data())
        h_plain, _ = precond.fwi(model, bg, iterations=a.c5_iterations,
                                 log=lambda m: _progress(f"model {h} plain " + m))
        target = h_plain[-1][1]
        t_plain = precond.time_to_misfit(h_plain, target)
        for seed, unet in enumerate(unets):
            h_learn, _ = precond.fwi(
                model, bg, iterations=a.c5_iterations,
                direction=lambda g, f, u=unet: precond.learned_direction_hip(
                    u, g, f, mirror=not a.c5_no_mirror),
                target=target, log=lambda m: _progress(f"model {h} seed {seed} " + m))
            t_learn = precond.time_to_misfit(h_learn, target)
            runs.append({
                "holdout": h, "seed": seed,
                "ratio": round(t_learn / t_plain, 4) if t_learn is not None else None,
                "plain_seconds": round(h_plain[-1][0], 2),
                "learned_seconds_to_target": round(t_learn, 2) if t_learn is not None else None,
                "learned_iterations": len(h_learn) - 1,
                "fallbacks": int(sum(1 for x in h_learn if x[2])),
                "initial_misfit": h_plain[0][1], "target_misfit": target,
                "plain_history": [[round(t, 3), f] for t, f, _ in h_plain],
                "learned_history": [[round(t, 3), f, bool(fb)] for t, f, fb in h_learn]})
            _progress(f"model {h} seed {seed}: ratio {runs[-1]['ratio']}")
    ratios = [r["ratio"] if r["ratio"] is not None else float("inf") for r in runs]
    med = float(np.median(ratios))
    return {
        "metric": "FWI time-to-misfit, learned vs plain",
        "value": round(med, 4) if np.isfinite(med) else None,
        "unit": "learned time / plain time to the plain run's final misfit "
                "(median over seeds and held-out models)",
        "n_gpus": 1, "higher_is_better": False,
        "dtype": "f64 solver, bf16 network (hand-written MFMA kernels)",
        "data": "synthetic (seeded GRF background + perturbation)",
        "config": {"workload": "C5 learned preconditioner", "grid": "316x188", "nt": 2500,
                   "shots": 6, "receivers": 48, "training_pairs": a.c5_pairs,
                   "pairs_along_path": a.c5_along_path, "mirror_average": not a.c5_no_mirror,
                   "widths": list(widths), "parameters": params, "seeds": a.c5_seeds,
                   "holdout_models": a.c5_holdout, "epochs": a.c5_epochs,
                   "fwi_iterations": a.c5_iterations},
        "ratio_min": min(ratios), "ratio_max": max(ratios) if np.isfinite(max(ratios)) else None,
        "runs_not_reaching_target": int(sum(1 for r in ratios if not np.isfinite(r))),
        "runs_at_or_below_half": int(sum(1 for r in ratios if r <= 0.5)),
        "pairs_seconds": round(t_pairs, 1), "train_seconds": [round(t, 1) for t in t_train],
        "train_loss_final": [round(h[-1], 5) for h in hist],
        "runs": runs,
    }


def _unet_flops(h, w, batch, widths):
    """Multiply-add FLOPs of one forward pass of the U-Net with channel
    widths `widths` on batch images of h x w (padded) pixels:
    2 * pixels * K * N per convolution."""
    c = list(widths)
    f, hh, ww = 0, h, w
    ins = [3, c[0], c[1]]
    for lvl in range(3):
        f += 2 * hh * ww * 9 * (ins[lvl] * c[lvl] + c[lvl] * c[lvl])
        hh, ww = hh // 2, ww // 2
    f += 2 * hh * ww * 9 * (c[2] * c[3] + c[3] * c[3])
    for lvl in range(3):
        cu = c[2 - lvl]
        f += 2 * hh * ww * 4 * c[3 - lvl] * cu  # 2x2 stride-2 transposed convolution
        hh, ww = 2 * hh, 2 * ww
        f += 2 * hh * ww * 9 * (2 * cu * cu + cu * cu)
    f += 2 * hh * ww * c[0] * 3
    return batch * f


def unet_kernels(a):
    """The preconditioner network on the library's bfloat16 MFMA kernels:
    one AdamW training step on a batch of 16 gradients of the 316 x 188 C5
    grid (forward, input gradients, weight gradients), and the inference of
    one gradient, against torch's bfloat16 autocast (MIOpen) on the same
    data.  FLOPs count 2 * pixels * K * N per convolution on the padded
    320 x 192 images; a training step is 3 x the forward (forward, input
    gradient, weight gradient)."""
    import copy
    import torch
    from psvwave_amd import precond
    dev = torch.device("cuda", 0)
    torch.manual_seed(0)
    net = precond.make_unet().to(dev)
    net_t = copy.deepcopy(net)
    tr = precond.HipUNetTrainer(net)
    rng = np.random.default_rng(0)
    nx, nz, batch = 316, 188, 16
    x = torch.tensor(precond.normalise_gradient(
        rng.standard_normal((batch, 3, nx, nz)).astype(np.float32)), device=dev)
    y = torch.tensor(0.02 * rng.standard_normal((batch, 3, nx, nz)), dtype=torch.float32,
                     device=dev)
    mask = torch.tensor(precond.kernel_mask((nx, nz)), dtype=torch.float32, device=dev)
    opt = torch.optim.AdamW(net_t.parameters(), lr=2e-3, weight_decay=1e-4)

    def torch_step():
        with torch.autocast(device_type="cuda", dtype=torch.bfloat16):
            pred = net_t(x)
        loss = (((pred.float() - y) * mask) ** 2).mean() / (mask.mean() * 1e-3)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()

    unet = precond.HipUNet(net.eval())
    x1 = x[0].cpu().numpy()
    x1d = x[:1]

    def torch_infer():
        with torch.no_grad(), torch.autocast(device_type="cuda", dtype=torch.bfloat16):
            net_t(x1d)

    def timed(f, n):
        for _ in range(a.warmup + 1):
            f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            f()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / n * 1e3

    n = max(a.steps, 10)
    res = {"hip_train_ms": timed(lambda: tr.step(x, y, mask), n),
           "torch_train_ms": timed(torch_step, n),
           "hip_infer_ms": timed(lambda: unet(x1), n),
           "torch_infer_ms": timed(torch_infer, n)}
    _progress(" ".join(f"{k} {v:.3f}" for k, v in res.items()))
    widths = [net.enc[0][0].out_channels, net.enc[1][0].out_channels,
              net.enc[2][0].out_channels, net.mid[0].out_channels]
    flops = 3 * _unet_flops(320, 192, batch, widths)
    tf = flops / (res["hip_train_ms"] * 1e-3) / 1e12
    return {
        "metric": "U-Net training step time (bf16 MFMA kernels)",
        "value": round(res["hip_train_ms"], 3), "unit": "ms per training step",
        "higher_is_better": False, "n_gpus": 1, "steps": n, "warmup": a.warmup,
        "dtype": "bf16 operands, fp32 accumulation", "data": "synthetic (random gradients)",
        "config": {"workload": "C5 preconditioner network", "batch": batch,
                   "grid": f"{nx}x{nz} padded to 320x192", "widths": widths,
                   "parameters": sum(p.numel() for p in net.parameters())},
        "torch_autocast_train_ms": round(res["torch_train_ms"], 3),
        "speedup_vs_torch_train": round(res["torch_train_ms"] / res["hip_train_ms"], 3),
        "hip_infer_ms": round(res["hip_infer_ms"], 3),
        "torch_autocast_infer_ms": round(res["torch_infer_ms"], 3),
        "roofline": {"bound": "mfma", "achieved": round(tf, 2), "peak": 2500.0,
                     "unit": "TFLOP/s", "frac": round(tf / 2500.0, 4),
                     "note": "whole training step (all kernels, host work "
                             "included) against dense bf16 MFMA peak"},
    }


def main():
    a = _args()
    if a.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; "
                          "multi-GPU runs (shots sharded across GPUs) are not "
                          "part of this build", "n_gpus": a.gpus}))
        sys.exit(2)
    copy = None
    if a.calibrate_copy:
        ms, gbs = calibrate_copy(a.calibrate_copy)
        copy = {"bytes": a.calibrate_copy, "ms": round(ms, 4),
                "read_write_GBs": round(gbs, 1)}
        print(json.dumps({"copy_calibration": copy}), file=sys.stderr, flush=True)
        if a.only_calibrate:
            return
    if a.workload == "mc":
        print(json.dumps(monte_carlo(a)), flush=True)
        return
    if a.workload == "c4":
        print(json.dumps(gradient_check(a)), flush=True)
        return
    if a.workload == "unet":
        print(json.dumps(unet_kernels(a)), flush=True)
        return
    if a.workload == "c5":
        print(json.dumps(learned_fwi(a)), flush=True)
        return
    import psvwave_amd
    from psvwave_amd.synthetic import (algorithmic_bytes_per_shot,
                                       layered_faulted_config,
                                       layered_faulted_model)

    cfg = layered_faulted_config(nt=a.nt, n_shots=a.shots)
    model = psvwave_amd.fdModel(cfg)
    model.set_step_variant(a.variant, not a.no_xcd_remap)
    if a.graph:
        model.set_graph_mode(True)
    true = layered_faulted_model(cfg, faulted=False)
    start = layered_faulted_model(cfg, faulted=True)
    # observed data from the unfaulted model (setup, not timed)
    model.set_parameter_fields(*true)
    for s in range(model.n_shots):
        model.forward_simulate(s, store_fields=False)
    model.set_observed_data(*model.get_synthetic_data())
    model.set_parameter_fields(*start)

    def step():
        if not a.sequential:
            # src/fdModel.cpp:1291-1308 (forward all shots, misfit, adjoint
            # sources, reset, adjoint all shots, map kernels)
            model.run_model(False, True)
            return
        for s in range(model.n_shots):
            model.forward_simulate(s)
        model.calculate_l2_misfit()
        model.calculate_l2_adjoint_sources()
        model.reset_kernels()
        for s in range(model.n_shots):
            model.adjoint_simulate(s)
        model.map_kernels_to_velocity()

    _progress("setup done")
    for k in range(a.warmup):
        step()
        _progress(f"warmup step {k + 1}/{a.warmup}")
    # the per-shot calls run asynchronously only without step timing, so
    # the call-by-call driver is timed by the wall clock alone
    events = not a.sequential or a.graph or a.timing_stride > 0
    model.set_timing(a.timing_stride if events else 0)
    model.reset_step_timing()
    t0 = time.perf_counter()
    for k in range(a.steps):
        step()  # the step's last call waits for the device
        _progress(f"timed step {k + 1}/{a.steps}")
    secs = time.perf_counter() - t0
    if events:
        kern_ms, timed, launches = model.step_timing()
    else:
        launches = timed = 2 * model.n_shots * model.nt * a.steps
        kern_ms = secs * 1e3
    model.set_timing(0)
    alone_us = kernel_alone_us(model)

    nx, nz, nt = model.nx, model.nz, model.nt
    points_per_step = 2.0 * model.n_shots * nx * nz * nt
    value = points_per_step * a.steps / secs / 1e9
    d = cfg["domain"]
    kpts = (d["nx_inner"] - 2 * d["nx_inner_boundary"]) * \
        (d["nz_inner"] - 2 * d["nz_inner_boundary"])
    bytes_per_launch = algorithmic_bytes_per_shot(
        nx, nz, nt, cfg["inversion"]["snapshot_interval"], kpts) / (2.0 * nt)
    launches_per_kernel = {2: 2, 5: 2, 6: 0.5, 7: 0.5, 14: 0.5}.get(a.variant, 1)
    avg_launch_s = kern_ms / 1e3 / max(timed, 1)
    device_achieved = bytes_per_launch / avg_launch_s / 1e9
    achieved = bytes_per_launch / (alone_us * 1e-6) / 1e9
    kernel_name = {0: "step_fused<8,ieee>", 1: "step_fused<16,ieee>",
                   2: "step_split_stress<8,ieee>+step_split_velocity<8,ieee>",
                   3: "step_fused<8,recip>", 4: "step_fused<16,recip>",
                   5: "step_split_stress<8,recip>+step_split_velocity<8,recip>",
                   6: "step_two<8,256>", 7: "step_two<16,512>",
                   8: "step_lean<16>", 9: "step_lean<8>", 10: "step_lean<12>",
                   11: "step_lean<20>", 12: "step_lean<24>", 13: "step_lean<32>",
                   14: "step_lean2<8>", 15: "step_lean_pipe<8,2>",
                   16: "step_lean_pipe<8,3>", 17: "step_lean_pipe<8,4>",
                   18: "step_lean<4>", 19: "step_lean<8,6 waves>",
                   20: "step_lean<12,5 waves>", 21: "step_lean<8,halo DMA>",
                   22: "step_lean<8,separate stages>", 23: "step_lean<4,separate stages>",
                   24: "step_lean<8,halo+ring DMA>", 25: "step_lean<8,ring DMA>",
                   26: "step_lean<16,ring DMA>",
                   27: "step_lean<8,halo+ring DMA,capped>",
                   28: "step_lean<16,ring DMA,capped>",
                   }[a.variant]
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gpt-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(secs / a.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded layered-faulted model, observed data from "
                "the unfaulted model)",
        "config": {
            "workload": "layered-faulted single model, fwd+adj FWI gradient",
            "grid": f"{d['nx_inner']}x{d['nz_inner']} inner + "
                    f"{cfg['boundary']['np_boundary']} taper = {nx}x{nz}",
            "nt": nt, "shots": model.n_shots, "receivers": model.nr,
            "snapshot_interval": cfg["inversion"]["snapshot_interval"],
            "point_updates_per_step": points_per_step,
            "step_kernel": kernel_name,
            "xcd_remap": not a.no_xcd_remap,
            "hip_graph": a.graph,
            "driver": ("call by call, one stream" if a.graph or a.timing_stride > 0 else
                       "call by call (forward_simulate per shot, misfit, adjoint "
                       "sources, adjoint_simulate per shot): asynchronous calls, "
                       "forwards two at a time, adjoints in pairs with deferred "
                       "kernel terms" if a.sequential else
                       "run_model on two HIP streams: forwards of shots 1.. "
                       "beside the adjoints of shots 0.., the last adjoint's "
                       "kernel terms deferred and replayed in order (bitwise "
                       "the sequential results)"),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kernel_name,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": _measured_traffic(kernel_name),
            "traffic_note": ("rocprofv3 FETCH_SIZE + WRITE_SIZE per launch, each "
                             "scaled by the same counter on a plain copy kernel: "
                             "bytes between L2 and the memory fabric, Infinity-"
                             "Cache hits included"),
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "kernel_us": round(alone_us, 3),
            "timing": ("kernel alone: HIP start/end events of every step launch "
                       "(hipExtLaunchKernel) of one shot's forward + adjoint on "
                       "the model's stream, after the timed region"),
            "device_aggregate": {
                "achieved": round(device_achieved, 1),
                "frac": round(device_achieved / HBM_PEAK_GBS, 4),
                "wall_us_per_launch": round(avg_launch_s * 1e6, 3),
                "timed_launches": timed,
                "launches": launches,
                "launches_per_time_step": launches_per_kernel,
                "timing": (f"kernel start/end of every {a.timing_stride}th launch"
                           if a.timing_stride > 0 else
                           "HIP events around each time loop, per time step, gaps "
                           "included" if a.graph else
                           "wall clock of the timed steps per time step launched"
                           if a.sequential else
                           "HIP events around each run_model (both streams), per "
                           "time step of every shot: wall time per launch while "
                           "two launches overlap"),
                "note": ("one stream" if a.graph or a.timing_stride > 0 else
                         "two time loops at a time share the GPU; their two "
                         "state sets and materials (2 x 84 MB) fit the 256 MB "
                         "Infinity Cache, so part of this rate is served by it"),
            },
        },
    }
    if copy:
        out["roofline"]["copy_calibration"] = copy
    out["roofline"]["stream_ceiling"] = stream_ceiling(bytes_per_launch, achieved)
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_baseline_nt)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
