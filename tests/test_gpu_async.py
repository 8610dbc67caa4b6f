"""Asynchronous per-shot calls (forward_simulate / adjoint_simulate return
once their time loop is enqueued; forwards run two at a time, adjoints in
pairs with deferred kernel terms) give bitwise the results of synchronous
calls, for the reference's call pattern and for irregular ones."""
import numpy as np
import pytest

from conftest import small_config

pytestmark = pytest.mark.gpu


def _fields(m, seed):
    rng = np.random.default_rng(seed)
    vp, vs, rho = m.get_parameter_fields()
    ix = np.arange(m.nx)[:, None]
    iz = np.arange(m.nz)[None, :]
    blob = ((ix - m.nx // 2 - 7) ** 2 + (iz - m.nz // 2) ** 2) < 64
    return (vp * (1 + 0.06 * blob), vs * (1 - 0.04 * blob),
            rho * (1 + 0.02 * rng.random(vp.shape)))


def _model(cfg, asynchronous, variant=None):
    import psvwave_amd
    m = psvwave_amd.fdModel(cfg)
    if variant is not None:
        m.set_step_variant(variant, True)
    m.set_async(asynchronous)
    true = _fields(m, 1)
    start = m.get_parameter_fields()
    m.set_parameter_fields(*true)
    for s in range(m.n_shots):
        m.forward_simulate(s, store_fields=False)
    m.set_observed_data(*m.get_synthetic_data())
    m.set_parameter_fields(*start)
    return m


def _outputs(m):
    return [m.misfit, *m.get_synthetic_data(), *m.get_lame_kernels(), *m.get_kernels(),
            m.get_gradient_vector(), *m.get_snapshots()]


def _gradient(m, adjoint_order=None):
    for s in range(m.n_shots):
        m.forward_simulate(s)
    m.calculate_l2_misfit()
    m.calculate_l2_adjoint_sources()
    m.reset_kernels()
    for s in adjoint_order or range(m.n_shots):
        m.adjoint_simulate(s)
    m.map_kernels_to_velocity()
    return _outputs(m)


def _assert_same(a, b, tag):
    assert len(a) == len(b)
    for k, (x, y) in enumerate(zip(a, b)):
        assert np.array_equal(np.asarray(x), np.asarray(y)), (tag, k)


def _shots(n):
    cfg = small_config(**{"domain.nt": 200})
    cfg["sources"].update(n_sources=n, n_shots=n, moment_angles=[30 + 25 * k for k in range(n)],
                          ix_sources=[10 + 17 * k for k in range(n)], iz_sources=[9] * n,
                          which_source_to_fire_in_which_shot=[[k] for k in range(n)])
    return cfg


@pytest.mark.parametrize("n_shots", [1, 2, 3, 4])
def test_reference_call_pattern_is_bitwise_synchronous(n_shots):
    cfg = _shots(n_shots)
    print(f"[async] {n_shots} shots", flush=True)
    want = _gradient(_model(cfg, False))
    got = _gradient(_model(cfg, True))
    _assert_same(got, want, n_shots)


def test_adjoints_in_any_order_and_repeated():
    cfg = _shots(4)
    order = [2, 0, 3, 3, 1]
    want = _gradient(_model(cfg, False), order)
    got = _gradient(_model(cfg, True), order)
    _assert_same(got, want, "order")


def test_interleaved_and_repeated_calls():
    cfg = _shots(3)
    outs = []
    for asynchronous in (False, True):
        m = _model(cfg, asynchronous)
        m.forward_simulate(1)
        m.forward_simulate(1)          # the same shot again
        m.forward_simulate(0)
        m.forward_simulate(2, store_fields=False)
        m.forward_simulate(2)
        m.calculate_l2_misfit()
        m.calculate_l2_adjoint_sources()
        m.reset_kernels()
        m.adjoint_simulate(0)
        m.forward_simulate(0)          # a forward while an adjoint is pending
        m.adjoint_simulate(1)
        m.adjoint_simulate(2)
        k1 = m.get_lame_kernels()      # observation between adjoints
        m.adjoint_simulate(0)
        m.map_kernels_to_velocity()
        outs.append([*k1, *_outputs(m)])
    _assert_same(outs[1], outs[0], "interleaved")


def test_synchronous_variants_and_copies():
    """Two-step and in-place variants take the synchronous path; copies of a
    model with pending work see the settled state."""
    cfg = _shots(2)
    want = _gradient(_model(cfg, False))
    for variant in (2, 14, 9, 15):
        got = _gradient(_model(cfg, True, variant))
        _assert_same(got, want, variant)
    m = _model(cfg, True)
    for s in range(m.n_shots):
        m.forward_simulate(s)
    c = m.copy()
    _assert_same(c.get_synthetic_data(), m.get_synthetic_data(), "copy")


def test_warning_when_lanes_may_share_a_queue():
    """With more live library streams (the default stream included) than
    the runtime's hardware queues, a model's two lanes may share a queue:
    the library warns once per process, and results stay those of the
    synchronous calls."""
    import gc
    import psvwave_amd
    import psvwave_amd.model as pm
    gc.collect()
    cfg = _shots(2)
    want = _gradient(_model(cfg, False))
    pm._queue_warning_issued = False
    models = []
    with pytest.warns(RuntimeWarning, match="hardware"):
        while not pm._queue_warning_issued:
            assert len(models) < 64
            m = psvwave_amd.fdModel(cfg)
            m.forward_simulate(0, store_fields=False)  # opens the model's second lane
            models.append(m)
    got = _gradient(_model(cfg, True))
    _assert_same(got, want, "shared queues")
    del models
    gc.collect()


def test_stream_count_follows_models_and_batches():
    """The library's live-stream count rises with each model (its own
    stream, and its second lane once used) and with each batch (one stream
    for all members), and falls back when they are released."""
    import ctypes as C
    import gc
    import psvwave_amd
    from psvwave_amd import _lib
    lib = _lib.load()

    def live():
        n, q = C.c_int(), C.c_int()
        _lib.check(lib.psv_stream_pressure(C.byref(n), C.byref(q)))
        return n.value

    gc.collect()
    base = live()
    cfg = _shots(2)
    m = psvwave_amd.fdModel(cfg)
    assert live() == base + 1
    m.forward_simulate(0, store_fields=False)  # opens the second lane
    m.synchronize()
    assert live() == base + 2
    b = psvwave_amd.Batch([cfg] * 3)
    assert live() == base + 3
    del m, b
    gc.collect()
    assert live() == base
