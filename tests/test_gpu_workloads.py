"""Oracle parity on the workloads the benchmarks run, at their own grid
shapes and through the drivers they time.

* The small CPU-runnable workload exactly as stated (200 x 100 inner + 50
  taper, nt = 4000, two shots, 19 receivers at depth 10, +10 % vp/vs
  inclusion): per-shot calls and run_model.
* The layered-faulted benchmark model on its full 1800 x 450 grid (water
  layer with vs = 0, nz_inner_boundary = 20, 300 receivers) at a reduced nt,
  through run_model with three shots -- the two HIP streams and the replay
  of the last adjoint's deferred kernel terms run at the benchmark's tiling
  (29 x 57 tiles of 64 x 8) -- and through the per-shot calls.
* The Monte-Carlo batch from its generator (uneven members, 30-point taper,
  Gaussian random fields): every member against a single fdModel, three
  against the oracle.

Seismograms, Lame and velocity kernels and gradient vectors must equal the
oracle's bit for bit; the misfit (a compensated tree sum on the device, a
serial sum in the oracle) to 1e-12 relative.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _oracle_run(pyoracle, cfg, true, start):
    """Observed data from `true`, then forward, misfit, adjoint sources,
    reset, adjoint and kernel mapping at `start`, shot by shot."""
    o = pyoracle.OracleModel(cfg)
    o.set_parameter_fields(*true)
    for s in range(o.n_shots):
        o.forward_simulate(s, store_fields=False)
    obs = o.get_synthetic_data()
    o.set_observed_data(*obs)
    o.set_parameter_fields(*start)
    for s in range(o.n_shots):
        o.forward_simulate(s)
    o.calculate_l2_misfit()
    o.calculate_l2_adjoint_sources()
    o.reset_kernels()
    for s in range(o.n_shots):
        o.adjoint_simulate(s)
    o.map_kernels_to_velocity()
    return obs, _outputs(o)


def _outputs(m):
    return {"misfit": m.misfit, "syn": m.get_synthetic_data(),
            "lame": m.get_lame_kernels(), "kernels": m.get_kernels(),
            "gradient": m.get_gradient_vector()}


def _device_model(cfg, true, start, obs_oracle):
    import psvwave_amd
    m = psvwave_amd.fdModel(cfg)
    m.set_parameter_fields(*true)
    for s in range(m.n_shots):
        m.forward_simulate(s, store_fields=False)
    obs = m.get_synthetic_data()
    for a, b in zip(obs, obs_oracle):
        assert np.array_equal(a, b), "observed data differ from the oracle"
    m.set_observed_data(*obs)
    m.set_parameter_fields(*start)
    return m


def _per_shot(m):
    for s in range(m.n_shots):
        m.forward_simulate(s)
    m.calculate_l2_misfit()
    m.calculate_l2_adjoint_sources()
    m.reset_kernels()
    for s in range(m.n_shots):
        m.adjoint_simulate(s)
    m.map_kernels_to_velocity()


def _assert_matches(got, want, tag):
    assert abs(got["misfit"] - want["misfit"]) <= 1e-12 * abs(want["misfit"]), tag
    assert want["misfit"] > 0, tag
    for key in ("syn", "lame", "kernels"):
        for k, (a, b) in enumerate(zip(got[key], want[key])):
            assert np.array_equal(a, b), (tag, key, k, rel_l2(a, b))
    assert np.array_equal(got["gradient"], want["gradient"]), tag
    # the comparison is not vacuous: the kernels carry signal
    assert all(np.abs(k).max() > 0 for k in want["lame"]), tag


def test_small_workload_matches_oracle(oracle_lib):
    from psvwave_amd.synthetic import c1_config, c1_models
    cfg = c1_config()
    true, start = c1_models(cfg)
    obs, want = _oracle_run(oracle_lib, cfg, true, start)
    m = _device_model(cfg, true, start, obs)
    _per_shot(m)
    _assert_matches(_outputs(m), want, "per-shot calls")
    m.run_model(False, True)
    _assert_matches(_outputs(m), want, "run_model")


def test_benchmark_grid_matches_oracle(oracle_lib):
    """The layered-faulted model on the benchmark's full grid, nt = 1000
    (the 10 Hz wavelet peaks at step 467), three shots."""
    from psvwave_amd.synthetic import layered_faulted_config, layered_faulted_model
    cfg = layered_faulted_config(nt=1000, n_shots=3)
    true = layered_faulted_model(cfg, faulted=False)
    start = layered_faulted_model(cfg, faulted=True)
    obs, want = _oracle_run(oracle_lib, cfg, true, start)
    m = _device_model(cfg, true, start, obs)
    assert (m.nx, m.nz) == (1800, 450)
    m.run_model(False, True)
    _assert_matches(_outputs(m), want, "run_model")
    _per_shot(m)
    _assert_matches(_outputs(m), want, "per-shot calls")


def test_monte_carlo_members_match_single_models_and_oracle(oracle_lib):
    import psvwave_amd
    from psvwave_amd.synthetic import monte_carlo_configs, monte_carlo_model
    nt = 600
    cfgs = monte_carlo_configs(n_models=24, nt=nt)
    rng = np.random.default_rng(3)
    batch = psvwave_amd.Batch(cfgs)
    models, singles = [], []
    for m, cfg in enumerate(cfgs):
        fields = monte_carlo_model(cfg, rng)
        med = cfg["medium"]
        shape = fields[0].shape
        background = (np.full(shape, med["scalar_vp"]), np.full(shape, med["scalar_vs"]),
                      np.full(shape, med["scalar_rho"]))
        single = psvwave_amd.fdModel(cfg)
        single.set_parameter_fields(*background)
        for s in range(2):
            single.forward_simulate(s, store_fields=False)
        obs = single.get_synthetic_data()
        single.set_observed_data(*obs)
        single.set_parameter_fields(*fields)
        batch.set_observed_data(m, *obs)
        batch.set_parameter_fields(m, *fields)
        models.append((fields, obs))
        singles.append(single)
    batch.forward_all()
    misfits = batch.misfits()
    for m, single in enumerate(singles):
        for s in range(2):
            single.forward_simulate(s, store_fields=False)
        single.calculate_l2_misfit()
        for a, b in zip(batch.get_synthetic_data(m), single.get_synthetic_data()):
            assert np.array_equal(a, b), m
        assert misfits[m] == single.misfit, m
    for m in (0, 7, 23):
        fields, obs = models[m]
        o = oracle_lib.OracleModel(cfgs[m])
        o.set_parameter_fields(*fields)
        o.set_observed_data(*obs)
        for s in range(2):
            o.forward_simulate(s, store_fields=False)
        o.calculate_l2_misfit()
        for a, b in zip(batch.get_synthetic_data(m), o.get_synthetic_data()):
            assert np.array_equal(a, b), m
        assert abs(misfits[m] - o.misfit) <= 1e-12 * o.misfit, m
        assert o.misfit > 0
