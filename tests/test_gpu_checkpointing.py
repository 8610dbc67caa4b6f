"""Checkpointed adjoint (every k-th snapshot kept, the rest recomputed per
segment) and the autograd wrapper."""
import numpy as np
import pytest

from conftest import small_config

pytestmark = pytest.mark.gpu


def _run(model, fields, obs):
    model.set_parameter_fields(*fields)
    for s in range(model.n_shots):
        model.forward_simulate(s)
    model.set_observed_data(*obs)
    model.calculate_l2_misfit()
    model.calculate_l2_adjoint_sources()
    model.reset_kernels()
    for s in range(model.n_shots):
        model.adjoint_simulate(s)
    model.map_kernels_to_velocity()
    return model.get_lame_kernels() + model.get_kernels() + (model.get_gradient_vector(),)


@pytest.mark.parametrize("variant,every,nt", [(0, 4, 300), (6, 8, 295), (2, 3, 300),
                                              (7, 5, 300), (8, 4, 300),
                                              (14, 5, 295)])
def test_checkpointed_adjoint_is_bitwise_full_storage(oracle_lib, variant, every, nt):
    import psvwave_amd
    cfg = small_config(**{"domain.nt": nt})
    rng = np.random.default_rng(5)
    full = psvwave_amd.fdModel(cfg)
    full.set_step_variant(variant)
    ck = psvwave_amd.fdModel(cfg)
    ck.set_step_variant(variant)
    ck.set_checkpointing(every)
    shape = (full.nx, full.nz)
    fields = (2000 + 100 * rng.random(shape), 800 + 50 * rng.random(shape),
              1500 + 80 * rng.random(shape))
    obs = tuple(1e-10 * rng.standard_normal((full.n_shots, full.nr, nt)) for _ in range(2))
    a = _run(full, fields, obs)
    b = _run(ck, fields, obs)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert ck.device_bytes() < full.device_bytes()
    with pytest.raises(ValueError, match="checkpointing"):
        ck.get_snapshots()
    o = oracle_lib.OracleModel(cfg)
    c = _run(o, fields, obs)
    for x, y in zip(b, c):
        assert np.array_equal(x, y)


def test_autograd_matches_kernels_and_finite_differences():
    """Same perturbations as the oracle's finite-difference pin; the
    reference gradient agrees with central differences to a few percent."""
    torch = pytest.importorskip("torch")
    import psvwave_amd
    from psvwave_amd.autograd import misfit
    cfg = small_config(**{"domain.nt": 600})
    m = psvwave_amd.fdModel(cfg)
    m.set_checkpointing(4)
    vp0, vs0, rho0 = m.get_parameter_fields()
    nx, nz = vp0.shape
    ix = np.arange(nx)[:, None]
    iz = np.arange(nz)[None, :]

    def bump(cx, cz, r):
        return np.exp(-((ix - cx) ** 2 + (iz - cz) ** 2) / (2 * r * r))

    true = (vp0 * (1 + 0.05 * bump(50, 40, 5)), vs0 * (1 - 0.05 * bump(60, 35, 5)),
            rho0 * (1 + 0.05 * bump(70, 45, 5)))
    m.set_parameter_fields(*true)
    for s in range(m.n_shots):
        m.forward_simulate(s, store_fields=False)
    m.set_observed_data(*m.get_synthetic_data())
    start = [torch.tensor(a, dtype=torch.float64, requires_grad=True)
             for a in (vp0, vs0, rho0)]
    f = misfit(*start, m)
    f.backward()
    kernels = m.get_kernels()
    for t, k in zip(start, kernels):
        assert np.array_equal(t.grad.numpy(), k)
    d = torch.from_numpy(bump(55, 40, 6))
    for p, scale in enumerate((2000.0, 800.0, 1500.0)):
        eps = 1e-3 * scale
        fp = [t.detach().clone() for t in start]
        fm = [t.detach().clone() for t in start]
        fp[p] += eps * d
        fm[p] -= eps * d
        fd = (float(misfit(*fp, m)) - float(misfit(*fm, m))) / (2 * eps)
        g = float((start[p].grad * d).sum())
        assert g == pytest.approx(fd, rel=0.10), (p, g, fd)


def test_autograd_backward_after_another_forward():
    """A second misfit() on the same model between forward and backward (a
    line-search trial) must not change the first graph's gradient; changed
    observed data make the backward raise."""
    torch = pytest.importorskip("torch")
    import psvwave_amd
    from psvwave_amd.autograd import misfit
    cfg = small_config(**{"domain.nt": 300})
    m = psvwave_amd.fdModel(cfg)
    vp0, vs0, rho0 = m.get_parameter_fields()
    rng = np.random.default_rng(8)
    m.set_parameter_fields(vp0 * (1 + 0.02 * rng.random(vp0.shape)), vs0, rho0)
    for s in range(m.n_shots):
        m.forward_simulate(s, store_fields=False)
    m.set_observed_data(*m.get_synthetic_data())
    a = [torch.tensor(x, requires_grad=True) for x in (vp0, vs0, rho0)]
    fa = misfit(*a, m)
    fa.backward()
    want = [t.grad.clone() for t in a]
    b = [torch.tensor(x, requires_grad=True) for x in (vp0, vs0, rho0)]
    fb = misfit(*b, m)
    c = [torch.tensor(x * 1.01) for x in (vp0, vs0, rho0)]
    fc = misfit(*c, m)  # the model now holds the state at c
    assert float(fc) != float(fb.detach())
    fb.backward()
    for g, w in zip(b, want):
        assert torch.equal(g.grad, w)
    d = [torch.tensor(x, requires_grad=True) for x in (vp0, vs0, rho0)]
    fd = misfit(*d, m)
    m.set_observed_data(*(o * 1.5 for o in m.get_observed_data()))
    with pytest.raises(RuntimeError, match="observed data"):
        fd.backward()
