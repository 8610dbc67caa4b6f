"""The allocation-free misfit entry point (psv_l2_misfit_noalloc) against the
model's own psv_calculate_l2_misfit, through the C-ABI: bit for bit, on two
calls in a row, the second with a smaller input (the fixed partial-sum
array keeps the first call's contents, which the second must overwrite)."""
import ctypes as C

import numpy as np
import pytest

from conftest import small_config

pytestmark = pytest.mark.gpu


def _model_and_traces(cfg):
    """A model whose synthetic and observed traces differ; returns the model's
    own misfit (psv_calculate_l2_misfit) and the four trace arrays."""
    import psvwave_amd
    m = psvwave_amd.fdModel(cfg)
    vp, vs, rho = m.get_parameter_fields()
    ix = np.arange(m.nx)[:, None]
    iz = np.arange(m.nz)[None, :]
    blob = 1.0 + 0.08 * (((ix - m.nx // 2) ** 2 + (iz - m.nz // 2) ** 2) < 49)
    m.set_parameter_fields(vp * blob, vs, rho * blob)
    for s in range(m.n_shots):
        m.forward_simulate(s, store_fields=False)
    m.set_observed_data(*m.get_synthetic_data())
    m.set_parameter_fields(vp, vs, rho)
    for s in range(m.n_shots):
        m.forward_simulate(s, store_fields=False)
    m.calculate_l2_misfit()
    syn = m.get_synthetic_data()
    obs = m.get_observed_data()
    return m, syn, obs


def _noalloc(lib, syn, obs, dt):
    import torch
    dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (*syn, *obs)]
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    n = dev[0].numel()
    assert all(t.numel() == n for t in dev)
    torch.cuda.synchronize()
    st = lib.psv_l2_misfit_noalloc(*(C.c_void_p(t.data_ptr()) for t in dev), n, dt,
                                   C.c_void_p(out.data_ptr()), None)
    assert st == 0, lib.psv_last_error()
    torch.cuda.synchronize()
    return out.item()


def test_noalloc_misfit_matches_model_misfit_twice():
    from psvwave_amd import _lib
    lib = _lib.load()
    big = small_config()
    small = small_config(**{"domain.nt": 120, "receivers.nr": 3,
                            "receivers.ix_receivers": [5, 40, 80],
                            "receivers.iz_receivers": [40, 3, 44]})
    results = []
    for cfg in (big, small):
        m, syn, obs = _model_and_traces(cfg)
        assert syn[0].size == m.n_shots * m.nr * m.nt
        got = _noalloc(lib, syn, obs, m.dt)
        assert m.misfit > 0
        assert got == m.misfit, (got, m.misfit)
        results.append(got)
    assert results[1] < results[0]


def test_noalloc_misfit_rejects_bad_sizes():
    from psvwave_amd import _lib
    lib = _lib.load()
    out = C.c_double(0.0)
    for n in (-1, (1 << 40) + 1):
        st = lib.psv_l2_misfit_noalloc(None, None, None, None, n, 1e-3,
                                       C.cast(C.pointer(out), C.c_void_p), None)
        assert st == _lib.PSV_EVALUE
        assert b"outside" in lib.psv_last_error()


def test_noalloc_misfit_of_nothing_is_zero():
    import torch
    from psvwave_amd import _lib
    lib = _lib.load()
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    st = lib.psv_l2_misfit_noalloc(None, None, None, None, 0, 1e-3,
                                   C.c_void_p(out.data_ptr()), None)
    assert st == 0, lib.psv_last_error()
    torch.cuda.synchronize()
    assert out.item() == 0.0
