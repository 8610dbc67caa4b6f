"""Parity of the HIP path with the CPU oracle (run on an MI355X).

The bar is the one the fdModel path is held to: relative L2 error <= 1e-12
on seismograms, misfit, kernels and gradient.  Built with contraction off and
IEEE divisions, the device path is expected to be bit-identical for
everything but the misfit (a compensated parallel sum instead of a serial
one), and the tests also assert that where it holds.
"""
import json
import os

import numpy as np
import pytest

from conftest import (DEFAULT_INI, GOLDEN, default_config, notebook2_models,
                      rel_l2, small_config)

pytestmark = pytest.mark.gpu

TOL = 1e-12  # relative L2, fixed by the parity requirement


def _product(cfg_or_path, variant=None, xcd_remap=True):
    import psvwave_amd
    m = psvwave_amd.fdModel(cfg_or_path)
    if variant is not None:
        m.set_step_variant(variant, xcd_remap)
    return m


def _chain(model, true_fields, start_fields, adjoint=True):
    """obs = forward(true); then forward(start), misfit, adjoint sources,
    reset, adjoint, map -- the canonical order of test_descent.py:37-83."""
    model.set_parameter_fields(*true_fields)
    for s in range(model.n_shots):
        model.forward_simulate(s)
    ux, uz = model.get_synthetic_data()
    model.set_observed_data(ux, uz)
    model.set_parameter_fields(*start_fields)
    for s in range(model.n_shots):
        model.forward_simulate(s)
    model.calculate_l2_misfit()
    out = {"syn": model.get_synthetic_data(), "obs": (ux, uz),
           "misfit": model.misfit}
    if adjoint:
        model.calculate_l2_adjoint_sources()
        model.reset_kernels()
        for s in range(model.n_shots):
            model.adjoint_simulate(s)
        model.map_kernels_to_velocity()
        out["lame"] = model.get_lame_kernels()
        out["kernels"] = model.get_kernels()
        out["gradient"] = model.get_gradient_vector()
    return out


def _inclusion(nx, nz, cx, cz, r, scale):
    ix = np.arange(nx)[:, None]
    iz = np.arange(nz)[None, :]
    return 1.0 + scale * (((ix - cx) ** 2 + (iz - cz) ** 2) < r * r)


def _compare(p, o, bitwise=True):
    for key in ("syn", "obs"):
        for a, b in zip(p[key], o[key]):
            assert rel_l2(a, b) <= TOL, key
            if bitwise:
                assert np.array_equal(a, b), f"{key} not bit-identical"
    assert abs(p["misfit"] - o["misfit"]) <= TOL * abs(o["misfit"])
    if "lame" in o:
        for key in ("lame", "kernels"):
            for a, b in zip(p[key], o[key]):
                assert rel_l2(a, b) <= TOL, key
                if bitwise:
                    assert np.array_equal(a, b), f"{key} not bit-identical"
        assert rel_l2(p["gradient"], o["gradient"]) <= TOL
        if bitwise:
            assert np.array_equal(p["gradient"], o["gradient"])


def _oracle_chain(oracle_lib, cfg, true_fields, start_fields, adjoint=True):
    o = oracle_lib.OracleModel(cfg)
    o.set_parameter_fields(*true_fields)
    for s in range(o.n_shots):
        o.forward_simulate(s)
    ux, uz = o.get_synthetic_data()
    o.set_observed_data(ux, uz)
    o.set_parameter_fields(*start_fields)
    for s in range(o.n_shots):
        o.forward_simulate(s)
    out = {"syn": o.get_synthetic_data(), "obs": (ux, uz),
           "misfit": o.calculate_l2_misfit()}
    if adjoint:
        o.calculate_l2_adjoint_sources()
        o.reset_kernels()
        for s in range(o.n_shots):
            o.adjoint_simulate(s)
        o.map_kernels_to_velocity()
        out["lame"] = o.get_lame_kernels()
        out["kernels"] = o.get_kernels()
        out["gradient"] = o.get_gradient_vector()
    return out, o


def _small_models(cfg):
    np_ = cfg["boundary"]["np_boundary"]
    nx = cfg["domain"]["nx_inner"] + 2 * np_
    nz = cfg["domain"]["nz_inner"] + 2 * np_
    start = (np.full((nx, nz), 2000.0), np.full((nx, nz), 800.0),
             np.full((nx, nz), 1500.0))
    bump = _inclusion(nx, nz, nx * 0.55, nz * 0.6, 8, 0.08)
    true = (start[0] * bump, start[1] * _inclusion(nx, nz, nx * 0.4, nz * 0.5, 6, -0.1),
            start[2] * _inclusion(nx, nz, nx * 0.6, nz * 0.45, 5, 0.05))
    return true, start


@pytest.mark.parametrize("variant", list(range(29)))
def test_small_chain_matches_oracle(oracle_lib, variant):
    cfg = small_config()
    true, start = _small_models(cfg)
    print(f"[parity] small chain, variant {variant}", flush=True)
    p = _chain(_product(cfg, variant), true, start)
    o, _ = _oracle_chain(oracle_lib, cfg, true, start)
    _compare(p, o)


@pytest.mark.parametrize("variant", [6, 7, 14])
def test_two_step_odd_nt_matches_oracle(oracle_lib, variant):
    """Two-step launches with an odd number of steps (the last step runs on
    the one-step kernel) and snapshot steps at odd time indices."""
    cfg = small_config(**{"domain.nt": 151, "inversion.snapshot_interval": 151})
    true, start = _small_models(cfg)
    p = _chain(_product(cfg, variant), true, start)
    o, _ = _oracle_chain(oracle_lib, cfg, true, start)
    _compare(p, o)
    cfg = small_config(**{"domain.nt": 147, "inversion.snapshot_interval": 7})
    p = _chain(_product(cfg, variant), true, start)
    o, _ = _oracle_chain(oracle_lib, cfg, true, start)
    _compare(p, o)


def test_notebook2_known_answers_on_gpu():
    """The whole chain on the device reproduces the reference's recorded
    Notebook-2 outputs (misfit 1920401.82, |g|_inf 77.1590)."""
    with open(os.path.join(GOLDEN, "notebook2_known_answers.json")) as f:
        ka = json.load(f)
    m = _product(DEFAULT_INI)
    true, start = notebook2_models(m.nx, m.nz, m.np_boundary, m.dx, m.dz)
    print("[parity] notebook-2 chain on the device", flush=True)
    out = _chain(m, true, start)
    assert f"{out['misfit']:.2f}" == "1920401.82"
    assert float(f"{np.abs(out['gradient']).max():.5e}") == pytest.approx(
        ka["lbfgsb_iterate0"]["proj_g_inf"], abs=0)


def test_notebook2_chain_matches_oracle(oracle_lib):
    cfg = default_config()
    m = _product(DEFAULT_INI)
    true, start = notebook2_models(m.nx, m.nz, m.np_boundary, m.dx, m.dz)
    print("[parity] notebook-2 chain vs oracle", flush=True)
    p = _chain(m, true, start)
    o, _ = _oracle_chain(oracle_lib, cfg, true, start)
    _compare(p, o)


def test_snapshots_match_oracle(oracle_lib):
    cfg = small_config()
    true, start = _small_models(cfg)
    m = _product(cfg)
    m.set_parameter_fields(*true)
    m.forward_simulate(1)
    o = oracle_lib.OracleModel(cfg)
    o.set_parameter_fields(*true)
    o.forward_simulate(1)
    snaps = m.get_snapshots()
    for k in (0, 7, m.snapshots - 1):
        ref = o.get_snapshot(1, k)
        got = m.get_snapshot(1, k)
        for a, b, c in zip(got, ref, snaps):
            assert np.array_equal(a, b)
            assert np.array_equal(c[1, k], b)
