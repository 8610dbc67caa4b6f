"""run_model (src/fdModel.cpp:1291-1308) runs forward and adjoint loops of
different shots on two HIP streams; with one-step kernels the last adjoint
defers its kernel terms to a buffer that is replayed in order.  Every output
must be bitwise that of the sequential calls, for every step-kernel family,
with and without the adjoint, for two and four shots, and the fallbacks
(checkpointing, graph mode, per-launch timing) must agree too."""
import numpy as np
import pytest

from conftest import small_config

pytestmark = pytest.mark.gpu


def _config(n_shots=4):
    c = small_config(**{"domain.nt": 200})
    c["sources"].update(n_sources=n_shots, n_shots=n_shots,
                        moment_angles=[90, 45, 180, 10][:n_shots],
                        ix_sources=[20, 51, 70, 33][:n_shots],
                        iz_sources=[10, 7, 12, 30][:n_shots],
                        which_source_to_fire_in_which_shot=[[k] for k in range(n_shots)])
    return c


def _fields(nx, nz, seed):
    rng = np.random.default_rng(seed)
    return (2000 + 200 * rng.random((nx, nz)), 800 + 100 * rng.random((nx, nz)),
            1500 + 100 * rng.random((nx, nz)))


def _prepared(cfg, variant):
    import psvwave_amd
    m = psvwave_amd.fdModel(cfg)
    m.set_step_variant(variant)
    m.set_parameter_fields(*_fields(m.nx, m.nz, 1))
    for s in range(m.n_shots):
        m.forward_simulate(s, store_fields=False)
    m.set_observed_data(*m.get_synthetic_data())
    m.set_parameter_fields(*_fields(m.nx, m.nz, 2))
    return m


def _sequential(m, adjoint):
    for s in range(m.n_shots):
        m.forward_simulate(s, store_fields=True)
    m.calculate_l2_misfit()
    if adjoint:
        m.calculate_l2_adjoint_sources()
        m.reset_kernels()
        for s in range(m.n_shots):
            m.adjoint_simulate(s)
        m.map_kernels_to_velocity()


def _outputs(m, adjoint):
    out = [m.misfit, *m.get_synthetic_data(), *m.get_snapshots()]
    if adjoint:
        out += [*m.get_kernels(), m.get_gradient_vector()]
    return out


@pytest.mark.parametrize("variant", [27, 9, 0, 2, 6, 14])
@pytest.mark.parametrize("adjoint", [True, False])
@pytest.mark.parametrize("n_shots", [4, 2])
def test_two_lane_run_model_is_bitwise_sequential(variant, adjoint, n_shots):
    cfg = _config(n_shots)
    a, b = _prepared(cfg, variant), _prepared(cfg, variant)
    a.run_model(False, adjoint)
    _sequential(b, adjoint)
    for k, (x, y) in enumerate(zip(_outputs(a, adjoint), _outputs(b, adjoint))):
        assert np.array_equal(x, y), (variant, adjoint, k)
    # a second call reuses the lane's buffers and events
    a.run_model(False, adjoint)
    for k, (x, y) in enumerate(zip(_outputs(a, adjoint), _outputs(b, adjoint))):
        assert np.array_equal(x, y), (variant, adjoint, "again", k)


@pytest.mark.parametrize("variant", [27, 0])
def test_deferred_terms_ignore_earlier_calls(variant):
    """The deferred adjoint writes its per-slot kernel terms without reading
    its buffer, which is not cleared between calls: run_model at new
    parameters, after a run that left other terms in the buffer, equals the
    sequential calls at the new parameters."""
    cfg = _config(4)
    a, b = _prepared(cfg, variant), _prepared(cfg, variant)
    a.run_model(False, True)
    new = _fields(a.nx, a.nz, 3)
    a.set_parameter_fields(*new)
    b.set_parameter_fields(*new)
    a.run_model(False, True)
    _sequential(b, True)
    for k, (x, y) in enumerate(zip(_outputs(a, True), _outputs(b, True))):
        assert np.array_equal(x, y), (variant, k)


def test_run_model_timing_and_fallbacks():
    cfg = _config(3)
    ref = _prepared(cfg, 9)
    _sequential(ref, True)
    want = _outputs(ref, True)
    # loop timing brackets both lanes and counts every step of every shot
    m = _prepared(cfg, 9)
    m.set_timing(-1)
    m.reset_step_timing()
    m.run_model(False, True)
    ms, timed, launches = m.step_timing()
    assert timed == 2 * 3 * cfg["domain"]["nt"] and ms > 0
    for x, y in zip(_outputs(m, True), want):
        assert np.array_equal(x, y)
    # per-launch timing and checkpointing take the sequential path; graph
    # mode keeps the two lanes with one captured graph per time loop
    for setup in (lambda mm: mm.set_timing(7), lambda mm: mm.set_graph_mode(True),
                  lambda mm: mm.set_checkpointing(4)):
        mm = _prepared(cfg, 9)
        setup(mm)
        mm.run_model(False, True)
        got = [mm.misfit, *mm.get_synthetic_data(), *mm.get_kernels()]
        assert all(np.array_equal(x, y) for x, y in
                   zip(got, [want[0], want[1], want[2], *want[-4:-1]]))


def test_forward_lanes_with_checkpointing():
    """The forward-only driver keeps two lanes under checkpointing; the
    checkpointed adjoint afterwards must give the sequential kernels."""
    cfg = _config()
    a, b = _prepared(cfg, 9), _prepared(cfg, 9)
    for m in (a, b):
        m.set_checkpointing(4)
    a.run_model(False, False)
    for s in range(b.n_shots):
        b.forward_simulate(s, store_fields=True)
    b.calculate_l2_misfit()
    assert a.misfit == b.misfit
    for x, y in zip(a.get_synthetic_data(), b.get_synthetic_data()):
        assert np.array_equal(x, y)
    for m in (a, b):
        m.calculate_l2_adjoint_sources()
        m.reset_kernels()
        for s in range(m.n_shots):
            m.adjoint_simulate(s)
        m.map_kernels_to_velocity()
    for x, y in zip(a.get_kernels(), b.get_kernels()):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("variant,adjoint", [(9, True), (9, False), (6, True), (0, True)])
def test_two_lane_graphs_are_bitwise_sequential(variant, adjoint):
    """Graph mode: each time loop of the two-lane run_model is captured once
    and replayed; the second call replays the cached graphs."""
    cfg = _config(3)
    a, b = _prepared(cfg, variant), _prepared(cfg, variant)
    a.set_graph_mode(True)
    _sequential(b, adjoint)
    want = _outputs(b, adjoint)
    for rep in range(2):
        a.run_model(False, adjoint)
        for k, (x, y) in enumerate(zip(_outputs(a, adjoint), want)):
            assert np.array_equal(x, y), (variant, adjoint, rep, k)
