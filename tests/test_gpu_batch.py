"""Batched forward runs: every member must equal the same model run alone
(bit for bit), for each batched step kernel and in graph mode."""
import numpy as np
import pytest

from conftest import small_config

pytestmark = pytest.mark.gpu


def _members():
    sizes = [(90, 50, 12), (37, 13, 0), (130, 21, 7), (200, 64, 10), (64, 16, 4)]
    cfgs = []
    for k, (nxi, nzi, npb) in enumerate(sizes):
        c = small_config(**{"domain.nt": 97, "domain.nx_inner": nxi,
                            "domain.nz_inner": nzi, "boundary.np_boundary": npb,
                            "domain.nx_inner_boundary": 1,
                            "domain.nz_inner_boundary": 1})
        c["sources"].update(n_sources=2, n_shots=2, moment_angles=[20 + 30 * k, 200],
                            ix_sources=[max(1, nxi // 3), max(1, nxi // 2)],
                            iz_sources=[max(1, nzi // 3), max(1, nzi // 2)],
                            which_source_to_fire_in_which_shot=[[0], [1]])
        c["receivers"].update(nr=3, ix_receivers=[0, nxi // 2, nxi - 1],
                              iz_receivers=[nzi - 1, nzi // 2, 0])
        cfgs.append(c)
    return cfgs


def _fields(nx, nz, seed):
    rng = np.random.default_rng(seed)
    return (2000 + 200 * rng.random((nx, nz)), 800 + 100 * rng.random((nx, nz)),
            1500 + 100 * rng.random((nx, nz)))


@pytest.mark.parametrize("variant,graph", [(0, False), (1, False), (6, False),
                                           (7, False), (6, True), (8, False),
                                           (9, False), (8, True), (10, False),
                                           (11, False), (12, False), (13, False),
                                           (18, False), (19, False), (20, False),
                                           (21, False), (22, False), (23, False),
                                           (24, False), (25, False), (26, False),
                                           (27, False), (28, False)])
def test_batch_members_equal_single_models(oracle_lib, variant, graph, persistent=False):
    import psvwave_amd
    cfgs = _members()
    batch = psvwave_amd.Batch(cfgs)
    batch.set_variant(variant)
    if graph:
        batch.set_graph_mode(True)
    if persistent:
        batch.set_persistent(True)
    singles = []
    for m, cfg in enumerate(cfgs):
        info = batch.infos[m]
        f = _fields(info["nx"], info["nz"], m)
        batch.set_parameter_fields(m, *f)
        single = psvwave_amd.fdModel(cfg)
        single.set_parameter_fields(*f)
        obs = tuple(np.full((2, 3, 97), 1e-9 * (m + 1)) for _ in range(2))
        batch.set_observed_data(m, *obs)
        single.set_observed_data(*obs)
        singles.append(single)
    for rep in range(2 if graph else 1):
        batch.forward_all()
        misfits = batch.misfits()
        for m, single in enumerate(singles):
            for s in range(2):
                single.forward_simulate(s, store_fields=False)
            single.calculate_l2_misfit()
            for a, b in zip(batch.get_synthetic_data(m), single.get_synthetic_data()):
                assert np.array_equal(a, b), (variant, m, rep)
            assert misfits[m] == single.misfit
    # one member against the oracle
    o = oracle_lib.OracleModel(cfgs[3])
    o.set_parameter_fields(*_fields(o.nx, o.nz, 3))
    for s in range(2):
        o.forward_simulate(s, store_fields=False)
    for a, b in zip(batch.get_synthetic_data(3), o.get_synthetic_data()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("variant", [9, 10, 18, 22, 24, 27, 28])
def test_forward_all_lockstep_equals_shot_by_shot(variant):
    """forward_all advances both shots in the same launches (second shot in a
    second state buffer set); the seismograms equal forward(0), forward(1)
    of the same batch bit for bit, on repeated calls too."""
    import psvwave_amd
    cfgs = _members()
    batch = psvwave_amd.Batch(cfgs)
    batch.set_variant(variant)
    for m in range(len(cfgs)):
        info = batch.infos[m]
        batch.set_parameter_fields(m, *_fields(info["nx"], info["nz"], 50 + m))
    print(f"[lockstep] variant {variant}", flush=True)
    batch.forward(0)
    batch.forward(1)
    want = [batch.get_synthetic_data(m) for m in range(len(cfgs))]
    assert batch.lockstep()
    for rep in range(2):
        batch.forward_all()
        for m in range(len(cfgs)):
            for a, b in zip(batch.get_synthetic_data(m), want[m]):
                assert np.array_equal(a, b), (variant, m, rep)
    batch.set_variant(0)  # fused steps: shot by shot
    assert not batch.lockstep()
    # graph mode: the loop of both shots captured once, then replayed with
    # host work and synchronisation between the replays
    batch.set_variant(variant)
    batch.set_graph_mode(True)
    assert batch.lockstep()
    for rep in range(3):
        batch.forward_all()
        for m in range(len(cfgs)):
            for a, b in zip(batch.get_synthetic_data(m), want[m]):
                assert np.array_equal(a, b), (variant, m, "graph", rep)
        info = batch.infos[rep]
        batch.set_parameter_fields(rep, *_fields(info["nx"], info["nz"], 50 + rep))


def test_persistent_batch_loop_equals_single_models(oracle_lib):
    """One cooperative launch per shot's loop, a grid barrier between
    steps: the same results as per-step launches and as single models."""
    test_batch_members_equal_single_models(oracle_lib, 9, False, persistent=True)
    import psvwave_amd
    batch = psvwave_amd.Batch(_members()[:2])
    batch.set_variant(10)
    with pytest.raises(ValueError, match="variant 9"):
        batch.set_persistent(True)


def test_batch_autotune_and_validation():
    import psvwave_amd
    cfgs = _members()
    batch = psvwave_amd.Batch(cfgs[:2])
    best, times = batch.autotune(steps=6)
    tuned = {0, 1, 6, 7, 8, 9, 10, 11, 12, 13, 18, 19, 20, 22, 23, 24, 26, 27, 28}
    assert best in tuned
    assert set(times) == tuned
    assert all(t > 0 for t in times.values())
    bad = small_config(**{"domain.nt": 50})
    with pytest.raises(ValueError, match="share nt"):
        psvwave_amd.Batch([cfgs[0], bad])
    with pytest.raises(ValueError, match="variants"):
        batch.set_variant(2)


@pytest.mark.parametrize("variant", [9, 0, 18, 22, 23, 8, 13, 24, 26, 27, 28])
def test_batch_gradients_equal_single_models(oracle_lib, variant):
    """Batched FWI gradients (forward with snapshots, adjoint with kernel
    correlation, shot by shot over all members) equal each member's own
    run_model bit for bit, and one member equals the oracle."""
    import psvwave_amd
    cfgs = _members()
    for c in cfgs:
        c["domain"]["nt"] = 100
    batch = psvwave_amd.Batch(cfgs)
    batch.set_variant(variant)
    singles, starts = [], []
    for m, cfg in enumerate(cfgs):
        info = batch.infos[m]
        true = _fields(info["nx"], info["nz"], m)
        start = _fields(info["nx"], info["nz"], 100 + m)
        batch.set_parameter_fields(m, *true)
        starts.append(start)
    batch.forward_all()
    for m, cfg in enumerate(cfgs):
        obs = batch.get_synthetic_data(m)
        batch.set_observed_data(m, *obs)
        batch.set_parameter_fields(m, *starts[m])
        single = psvwave_amd.fdModel(cfg)
        single.set_observed_data(*obs)
        single.set_parameter_fields(*starts[m])
        singles.append((single, obs))
    for rep in range(2):
        misfits = batch.gradients()
        for m, (single, _) in enumerate(singles):
            if rep == 0:
                single.run_model(False, True)
            assert misfits[m] == single.misfit, (variant, m)
            for a, b in zip(batch.get_kernels(m), single.get_kernels()):
                assert np.array_equal(a, b), (variant, m, rep)
            assert any(np.abs(k).max() > 0 for k in single.get_kernels())
    o = oracle_lib.OracleModel(cfgs[3])
    o.set_parameter_fields(*starts[3])
    o.set_observed_data(*singles[3][1])
    for s in range(2):
        o.forward_simulate(s)
    o.calculate_l2_misfit()
    o.calculate_l2_adjoint_sources()
    o.reset_kernels()
    for s in range(2):
        o.adjoint_simulate(s)
    o.map_kernels_to_velocity()
    for a, b in zip(batch.get_kernels(3), o.get_kernels()):
        assert np.array_equal(a, b)
    assert abs(misfits[3] - o.misfit) <= 1e-12 * o.misfit


def _members3():
    """Three members with three shots each (one source per shot)."""
    cfgs = []
    for k, (nxi, nzi, npb) in enumerate([(90, 50, 12), (130, 21, 7), (64, 16, 4)]):
        c = small_config(**{"domain.nt": 97, "domain.nx_inner": nxi,
                            "domain.nz_inner": nzi, "boundary.np_boundary": npb,
                            "domain.nx_inner_boundary": 1,
                            "domain.nz_inner_boundary": 1})
        c["sources"].update(n_sources=3, n_shots=3, moment_angles=[20 + 30 * k, 200, 110],
                            ix_sources=[max(1, nxi // 3), max(1, nxi // 2),
                                        max(1, 2 * nxi // 3)],
                            iz_sources=[max(1, nzi // 3), max(1, nzi // 2),
                                        max(1, nzi // 4)],
                            which_source_to_fire_in_which_shot=[[0], [1], [2]])
        c["receivers"].update(nr=3, ix_receivers=[0, nxi // 2, nxi - 1],
                              iz_receivers=[nzi - 1, nzi // 2, 0])
        cfgs.append(c)
    return cfgs


@pytest.mark.parametrize("variant", [27, 28, 9])
def test_forward_all_lockstep_three_shots(oracle_lib, variant):
    """Three shots advance in the same launches, shot s in state buffer set
    s: the seismograms equal forward(0), forward(1), forward(2) bit for bit,
    and one member equals the oracle."""
    import psvwave_amd
    cfgs = _members3()
    batch = psvwave_amd.Batch(cfgs)
    batch.set_variant(variant)
    fields = []
    for m in range(len(cfgs)):
        info = batch.infos[m]
        fields.append(_fields(info["nx"], info["nz"], 70 + m))
        batch.set_parameter_fields(m, *fields[m])
    print(f"[lockstep3] variant {variant}", flush=True)
    for s in range(3):
        batch.forward(s)
    want = [batch.get_synthetic_data(m) for m in range(len(cfgs))]
    assert batch.lockstep()
    for rep in range(2):
        batch.forward_all()
        for m in range(len(cfgs)):
            for a, b in zip(batch.get_synthetic_data(m), want[m]):
                assert np.array_equal(a, b), (variant, m, rep)
    o = oracle_lib.OracleModel(cfgs[1])
    o.set_parameter_fields(*fields[1])
    for s in range(3):
        o.forward_simulate(s, store_fields=False)
    for a, b in zip(batch.get_synthetic_data(1), o.get_synthetic_data()):
        assert np.array_equal(a, b)
