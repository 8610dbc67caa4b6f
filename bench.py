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
time per launch: HIP events on the model's stream bracket every forward and
adjoint time loop (the step launches and the gaps between them), divided by
the time steps launched.  ``roofline.traffic`` is the HBM bytes per launch
from the committed rocprofv3 counter summary for the kernel.
``cpu_baseline`` times the CPU oracle (OpenMP, all host cores up to 16) on a
bounded sample of the same grid.
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
    name = kernel.replace("<", "").replace(">", "").replace(",", "_")
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
    p.add_argument("--variant", type=int, default=9,
                   help="step kernel: 0 fused 64x8, 1 fused 64x16, 2 split; "
                        "3-5 the same with reciprocal-multiply division; "
                        "6, 7 two steps per launch (64x8, 64x16); 8-13 lean "
                        "one-step kernels with 64 x 16, 8, 12, 20, 24, 32 "
                        "tiles (9 is the library default); 14 lean two "
                        "steps per launch (64x8)")
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
    p.add_argument("--c5-pairs", type=int, default=2048)
    p.add_argument("--c5-epochs", type=int, default=20)
    p.add_argument("--c5-iterations", type=int, default=30)
    p.add_argument("--c5-along-path", type=float, default=0.5,
                   help="fraction of the training pairs taken part of the way "
                        "from the background to the true model")
    p.add_argument("--c5-networks", type=int, default=2,
                   help="networks trained (seeds 0, 1, ...) whose directions "
                        "are averaged")
    p.add_argument("--c5-no-mirror", action="store_true",
                   help="use the network's prediction alone instead of its "
                        "mean with the mirrored prediction of the mirrored "
                        "gradient")
    p.add_argument("--c4-every", type=int, default=8)
    p.add_argument("--mc-models", type=int, default=512)
    return p.parse_args()


def calibrate_copy(nbytes, iters=5):
    """Mean time of a plain device copy of nbytes; returns (ms, GB/s of
    read + write traffic)."""
    import ctypes as C
    from psvwave_amd import _lib
    lib = _lib.load()
    ms = C.c_double()
    _lib.check(lib.psv_calibrate_copy(int(nbytes), int(iters), C.byref(ms)))
    return ms.value, 2.0 * nbytes / (ms.value * 1e-3) / 1e9


def cpu_baseline(nt_sample):
    """CPU oracle on a bounded sample: one shot, forward + adjoint, the
    benchmark grid, nt_sample steps."""
    from oracle.pyoracle import OracleModel
    from psvwave_amd.synthetic import layered_faulted_config, layered_faulted_model
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(16, cores))
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
                       f"OpenMP {threads} threads)")}


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
    achieved = 104.0 * upd / loop_s_per_step / 1e9
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
                   "setup_s": round(setup, 1)},
        "roofline": {"bound": "hbm", "kernel": f"batch step (variant {best})",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None,
                     "note": "time = whole forward loops (HIP events), gaps included"},
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
    steepest descent on the same GPU.  Training pairs (gradient at a smooth
    background, true perturbation) come from the device solver; the U-Net is
    trained in bfloat16 autocast; both FWI runs start from the background of
    a held-out model and use the same line search.  `value` is the learned
    run's time to reach the misfit the plain run ends at, divided by the
    plain run's time (lower is better).  Training and inference run on the
    library's bfloat16 MFMA convolution kernels (HipUNetTrainer, HipUNet);
    the learned direction is the mean over --c5-networks trainings."""
    import torch
    import psvwave_amd
    from psvwave_amd import precond
    t0 = time.perf_counter()
    grads, perts = precond.gradient_pairs(a.c5_pairs, seed=7, workers=8, log=_progress,
                                          along_path=a.c5_along_path)
    t_pairs = time.perf_counter() - t0
    _progress(f"{a.c5_pairs} training pairs in {t_pairs:.1f} s")
    unets, hist = [], []
    t0 = time.perf_counter()
    for seed in range(a.c5_networks):
        torch.manual_seed(seed)
        net = precond.make_unet()
        hist.append(precond.train_hip(net, grads, perts, epochs=a.c5_epochs, log=_progress))
        unets.append(precond.HipUNet(net.eval()))
    t_train = time.perf_counter() - t0
    bg, true = precond.c5_models(np.random.default_rng(12345))
    model = psvwave_amd.fdModel(precond.c5_config())
    model.set_parameter_fields(*true)
    for s in range(model.n_shots):
        model.forward_simulate(s, store_fields=False)
    model.set_observed_data(*model.get_synthetic_data())
    h_plain, _ = precond.fwi(model, bg, iterations=a.c5_iterations,
                             log=lambda m: _progress("plain " + m))
    target = h_plain[-1][1]
    h_learn, _ = precond.fwi(model, bg, iterations=a.c5_iterations,
                             direction=lambda g, f: precond.learned_direction_ensemble(
                                 unets, g, f, mirror=not a.c5_no_mirror),
                             target=target, log=lambda m: _progress("learned " + m))
    t_plain = precond.time_to_misfit(h_plain, target)
    t_learn = precond.time_to_misfit(h_learn, target)
    return {
        "metric": "FWI time-to-misfit, learned vs plain",
        "value": round(t_learn / t_plain, 4) if t_learn is not None else None,
        "unit": "learned time / plain time to the plain run's final misfit",
        "n_gpus": 1, "higher_is_better": False,
        "dtype": "f64 solver, bf16 network (hand-written MFMA kernels)",
        "data": "synthetic (seeded GRF background + perturbation)",
        "config": {"workload": "C5 learned preconditioner", "grid": "316x188", "nt": 2500,
                   "shots": 6, "receivers": 48, "training_pairs": a.c5_pairs,
                   "pairs_along_path": a.c5_along_path, "mirror_average": not a.c5_no_mirror,
                   "networks": a.c5_networks,
                   "epochs": a.c5_epochs, "fwi_iterations": a.c5_iterations},
        "pairs_seconds": round(t_pairs, 1), "train_seconds": round(t_train, 1),
        "train_loss": [[round(x, 5) for x in h] for h in hist],
        "plain": {"seconds": round(h_plain[-1][0], 2), "initial_misfit": h_plain[0][1],
                  "final_misfit": h_plain[-1][1], "iterations": len(h_plain) - 1},
        "learned": {"seconds_to_target": round(t_learn, 2) if t_learn is not None else None,
                    "final_misfit": h_learn[-1][1], "iterations": len(h_learn) - 1,
                    "fallbacks": int(sum(1 for h in h_learn if h[2]))},
    }


def _unet_flops(h, w, batch, base=32):
    """Multiply-add FLOPs of one forward pass of the U-Net on batch images
    of h x w (padded) pixels: 2 * pixels * K * N per convolution."""
    c = [base, 2 * base, 4 * base, 8 * base]
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
    flops = 3 * _unet_flops(320, 192, batch)
    tf = flops / (res["hip_train_ms"] * 1e-3) / 1e12
    return {
        "metric": "U-Net training step time (bf16 MFMA kernels)",
        "value": round(res["hip_train_ms"], 3), "unit": "ms per training step",
        "higher_is_better": False, "n_gpus": 1, "steps": n, "warmup": a.warmup,
        "dtype": "bf16 operands, fp32 accumulation", "data": "synthetic (random gradients)",
        "config": {"workload": "C5 preconditioner network", "batch": batch,
                   "grid": f"{nx}x{nz} padded to 320x192", "parameters": 1.9e6},
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
    model.set_timing(a.timing_stride)
    model.reset_step_timing()
    t0 = time.perf_counter()
    for k in range(a.steps):
        step()  # every call synchronises the model stream before returning
        _progress(f"timed step {k + 1}/{a.steps}")
    secs = time.perf_counter() - t0
    kern_ms, timed, launches = model.step_timing()
    model.set_timing(0)

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
    achieved = bytes_per_launch / avg_launch_s / 1e9
    kernel_name = {0: "step_fused<8,ieee>", 1: "step_fused<16,ieee>",
                   2: "step_split_stress<8,ieee>+step_split_velocity<8,ieee>",
                   3: "step_fused<8,recip>", 4: "step_fused<16,recip>",
                   5: "step_split_stress<8,recip>+step_split_velocity<8,recip>",
                   6: "step_two<8,256>", 7: "step_two<16,512>",
                   8: "step_lean<16>", 9: "step_lean<8>", 10: "step_lean<12>",
                   11: "step_lean<20>", 12: "step_lean<24>", 13: "step_lean<32>",
                   14: "step_lean2<8>",
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
            "driver": ("call by call, one stream" if a.sequential or a.graph or
                       a.timing_stride > 0 else
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
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "avg_launch_us": round(avg_launch_s * 1e6, 3),
            "timed_launches": timed,
            "launches": launches,
            "launches_per_time_step": launches_per_kernel,
            "timing": (f"kernel start/end of every {a.timing_stride}th launch"
                       if a.timing_stride > 0 else
                       "HIP events around each time loop, per time step, gaps "
                       "included" if a.sequential or a.graph else
                       "HIP events around each run_model (both streams), per "
                       "time step of every shot: wall time per launch while "
                       "two launches overlap"),
        },
    }
    if copy:
        out["roofline"]["copy_calibration"] = copy
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(a.cpu_baseline_nt)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
