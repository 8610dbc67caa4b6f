"""Pins of the CPU oracle against what fixes the reference's results
independently of any retyping of its formulas.

* Notebook 2's recorded outputs (the reference's only absolute numbers):
  whole forward + misfit + adjoint + kernel + gradient chain.
* Closed forms: taper by rectangle counting, Ricker wavelet, moment tensor
  with the truncated pi, receiver integration, injection terms.
* Exactness of the 4th-order staggered differences on quartic polynomials:
  each of the five updates, forward and adjoint, with distinct lm / la / mu
  and dx != dz, so a wrong coefficient, offset, sign or spacing fails.
* Physics: the gradient matches central finite differences of the misfit
  per parameter block, the velocity-kernel map is the chain rule of
  lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2, and one correlation step is
  the textbook strain (and velocity) correlation, strains solved from the
  stresses by Hooke's law.
"""
import json
import math
import os

import numpy as np
import pytest

from conftest import (GOLDEN, default_config, notebook2_models, small_config)


def _tiny(nx_inner=16, nz_inner=12, np_boundary=0, nt=1, **extra):
    cfg = small_config()
    cfg["domain"].update(nt=nt, nx_inner=nx_inner, nz_inner=nz_inner,
                         nx_inner_boundary=1, nz_inner_boundary=1,
                         dx=1.25, dz=0.8, dt=0.01)
    cfg["boundary"].update(np_boundary=np_boundary, np_factor=0.05)
    cfg["sources"].update(n_sources=1, n_shots=1, moment_angles=[30.0],
                          ix_sources=[nx_inner // 2], iz_sources=[nz_inner // 2],
                          which_source_to_fire_in_which_shot=[[0]],
                          delay_cycles_per_shot=0)
    cfg["receivers"].update(nr=1, ix_receivers=[1], iz_receivers=[1])
    cfg["inversion"]["snapshot_interval"] = 1
    for k, v in extra.items():
        sec, key = k.split(".")
        cfg[sec][key] = v
    return cfg


def _grid_model(oracle_lib, cfg, rho=2.0, vp=3.0, vs=1.0):
    o = oracle_lib.OracleModel(cfg)
    shape = (o.nx, o.nz)
    o.set_parameter_fields(np.full(shape, vp), np.full(shape, vs),
                           np.full(shape, rho))
    return o


# ----------------------------------------------------------------- known answers
@pytest.mark.slow
def test_notebook2_known_answers(oracle_lib):
    with open(os.path.join(GOLDEN, "notebook2_known_answers.json")) as f:
        ka = json.load(f)
    o = oracle_lib.OracleModel(default_config())
    true, start = notebook2_models(o.nx, o.nz, 25, 1.249, 1.249)
    o.set_parameter_fields(*true)
    o.forward_simulate(0)
    o.set_observed_data(*o.get_synthetic_data())
    o.set_parameter_fields(*start)
    o.forward_simulate(0)
    misfit = o.calculate_l2_misfit()
    assert f"{misfit:.2f}" == ka["misfit_at_start"]["printed"].split(": ")[1]
    o.calculate_l2_adjoint_sources()
    o.reset_kernels()
    o.adjoint_simulate(0)
    o.map_kernels_to_velocity()
    g = o.get_gradient_vector()
    # L-BFGS-B prints |proj g| (= max |g| without bounds) with 6 digits
    assert float(f"{np.abs(g).max():.5e}") == ka["lbfgsb_iterate0"]["proj_g_inf"]
    assert float(f"{misfit:.5e}") == ka["lbfgsb_iterate0"]["f"]


# ------------------------------------------------------------------ closed forms
def test_taper_matches_rectangle_counting(oracle_lib):
    cfg = _tiny(nx_inner=9, nz_inner=7, np_boundary=6)
    o = oracle_lib.OracleModel(cfg)
    nx, nz, npb, f = o.nx, o.nz, 6, 0.05
    count = np.zeros((nx, nz))
    for d in range(npb):  # the d-th rectangle [d, n-d) in both axes
        count[d:nx - d, d:nz - d] += 1
    expect = np.exp(-(f * (npb - count)) ** 2)
    np.testing.assert_allclose(o.get_taper(), expect, rtol=1e-15, atol=0)
    assert np.all(o.get_taper()[npb:nx - npb, npb:nz - npb] == 1.0)


def test_ricker_wavelets_and_shot_delays(oracle_lib):
    cfg = small_config(**{"domain.nt": 400})
    o = oracle_lib.OracleModel(cfg)
    stf = o.get_stf()
    f, dt = 60.0, 0.00025
    delay = cfg["sources"]["delay_cycles_per_shot"]
    t = np.arange(400) * dt
    # shot 0 fires sources [0, 2]: positions 0 and 1; shot 1 fires [1]
    for src, pos in ((0, 0), (2, 1), (1, 0)):
        tau = t - 1.4 / f - pos * delay / f
        ricker = (1 - 2 * (np.pi * f * tau) ** 2) * np.exp(-(np.pi * f * tau) ** 2)
        np.testing.assert_allclose(stf[src], ricker, rtol=1e-12, atol=1e-14)
    # source_timeshift is parsed but does not move the wavelet
    cfg2 = small_config(**{"domain.nt": 400, "sources.source_timeshift": 0.3})
    assert np.array_equal(oracle_lib.OracleModel(cfg2).get_stf(), stf)


def test_moment_tensor_uses_truncated_pi(oracle_lib):
    cfg = small_config()
    M = oracle_lib.OracleModel(cfg).get_moment()
    for s, ang in enumerate(cfg["sources"]["moment_angles"]):
        a = ang * 3.14159265 / 180.0
        expect = np.array([[math.cos(a), -math.sin(a)],
                           [-math.sin(a), -math.cos(a)]]) * 1e15
        np.testing.assert_allclose(M[s], expect, rtol=1e-15)
    # 90 degrees is not exactly vertical with the truncated constant
    assert M[0, 0, 0] == pytest.approx(1.79489674e6, rel=1e-8)


def test_lame_parameters(oracle_lib):
    o = _grid_model(oracle_lib, _tiny(), rho=2.0, vp=3.0, vs=1.0)
    lm, la, mu, b = o.get_lame()
    assert np.all(lm == 18.0) and np.all(mu == 2.0) and np.all(la == 14.0)
    assert np.all(b == 0.5)


# ------------------------------------------------------------ stencil exactness
def _quartic(c):
    return lambda x: c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3 + c[4] * x ** 4


def _dquartic(c):
    return lambda x: c[1] + 2 * c[2] * x + 3 * c[3] * x ** 2 + 4 * c[4] * x ** 3


P = [0.3, -1.1, 0.07, 0.004, -0.0002]
Q = [-0.5, 0.9, -0.05, 0.003, 0.0001]


@pytest.mark.parametrize("adjoint", [False, True])
def test_stress_update_is_exact_on_quartics(oracle_lib, adjoint):
    cfg = _tiny()
    o = _grid_model(oracle_lib, cfg)
    lm, la, mu, dt, dx, dz = 18.0, 14.0, 2.0, 0.01, 1.25, 0.8
    sgn = -1.0 if adjoint else 1.0
    nx, nz = o.nx, o.nz
    X = np.arange(nx, dtype=float)[:, None] * np.ones((1, nz))
    Z = np.ones((nx, 1)) * np.arange(nz, dtype=float)[None, :]
    inner = np.zeros((nx, nz), bool)
    inner[2:nx - 2, 2:nz - 2] = True
    zero = np.zeros((nx, nz))
    # normal stresses: vx varies along x, vz along z
    o.set_state(_quartic(P)(X), _quartic(Q)(Z), zero, zero, zero)
    o.stress_phase(adjoint)
    _, _, txx, tzz, txz = o.get_state()
    dP = _dquartic(P)(X + 0.5) / dx  # vx staggered half a cell forward in x
    dQ = _dquartic(Q)(Z - 0.5) / dz  # vz half a cell back in z
    np.testing.assert_allclose(txx[inner], (sgn * dt * (lm * dP + la * dQ))[inner],
                               rtol=1e-10)
    np.testing.assert_allclose(tzz[inner], (sgn * dt * (la * dP + lm * dQ))[inner],
                               rtol=1e-10)
    assert np.all(txz == 0) and np.all(txx[~inner] == 0)
    # shear stress: vx varies along z, vz along x
    o.set_state(_quartic(P)(Z), _quartic(Q)(X), zero, zero, zero)
    o.stress_phase(adjoint)
    _, _, txx, tzz, txz = o.get_state()
    expect = sgn * dt * mu * (_dquartic(P)(Z + 0.5) / dz + _dquartic(Q)(X - 0.5) / dx)
    np.testing.assert_allclose(txz[inner], expect[inner], rtol=1e-10)
    assert np.all(txx == 0) and np.all(tzz == 0)


@pytest.mark.parametrize("adjoint", [False, True])
def test_velocity_update_is_exact_on_quartics(oracle_lib, adjoint):
    cfg = _tiny()
    o = _grid_model(oracle_lib, cfg)
    b, dt, dx, dz = 0.5, 0.01, 1.25, 0.8
    sgn = -1.0 if adjoint else 1.0
    nx, nz = o.nx, o.nz
    X = np.arange(nx, dtype=float)[:, None] * np.ones((1, nz))
    Z = np.ones((nx, 1)) * np.arange(nz, dtype=float)[None, :]
    inner = np.zeros((nx, nz), bool)
    inner[2:nx - 2, 2:nz - 2] = True
    zero = np.zeros((nx, nz))
    # vx from txx(x) and txz(z); vz stays zero
    o.set_state(zero, zero, _quartic(P)(X), zero, _quartic(Q)(Z))
    o.velocity_phase(adjoint)
    vx, vz, *_ = o.get_state()
    expect = sgn * b * dt * (_dquartic(P)(X - 0.5) / dx + _dquartic(Q)(Z - 0.5) / dz)
    np.testing.assert_allclose(vx[inner], expect[inner], rtol=1e-10)
    assert np.all(vz == 0)
    # vz from txz(x) and tzz(z); vx stays zero
    o.set_state(zero, zero, zero, _quartic(Q)(Z), _quartic(P)(X))
    o.velocity_phase(adjoint)
    vx, vz, *_ = o.get_state()
    expect = sgn * b * dt * (_dquartic(P)(X + 0.5) / dx + _dquartic(Q)(Z + 0.5) / dz)
    np.testing.assert_allclose(vz[inner], expect[inner], rtol=1e-10)
    assert np.all(vx == 0)


def test_taper_multiplies_updated_values_and_rim_is_frozen(oracle_lib):
    cfg = _tiny(np_boundary=4)
    o = _grid_model(oracle_lib, cfg)
    nx, nz = o.nx, o.nz
    c = np.full((nx, nz), 3.0)
    zero = np.zeros((nx, nz))
    o.set_state(zero, zero, c, c, c)
    o.stress_phase(False)
    _, _, txx, tzz, txz = o.get_state()
    taper = o.get_taper()
    inner = np.zeros((nx, nz), bool)
    inner[2:nx - 2, 2:nz - 2] = True
    for s in (txx, tzz, txz):
        assert np.array_equal(s[inner], (taper * 3.0)[inner])
        assert np.all(s[~inner] == 3.0)


# ------------------------------------------------- injection and receiver terms
def test_moment_tensor_injection_terms(oracle_lib):
    cfg = _tiny(nt=2)
    cfg["receivers"].update(nr=2, ix_receivers=[8, 7], iz_receivers=[6, 7])
    o = _grid_model(oracle_lib, cfg, rho=2.0)
    o.forward_simulate(0, store_fields=False)
    # after step 1 the injected field has propagated; rerun with nt=1
    cfg1 = _tiny(nt=1)
    o1 = _grid_model(oracle_lib, cfg1, rho=2.0)
    o1.forward_simulate(0, store_fields=False)
    vx, vz, *_ = o1.get_state()
    M = o1.get_moment()[0]
    S = o1.get_stf()[0, 0]
    dt, dx, dz, b = 0.01, 1.25, 0.8, 0.5
    x, z = 8, 6
    tx = lambda k: k * S * dt * b / dx ** 4
    tz = lambda k: k * S * dt * b / dz ** 4
    ex = np.zeros_like(vx)
    ez = np.zeros_like(vz)
    ex[x - 1, z] -= tx(M[0, 0]); ex[x, z] += tx(M[0, 0])
    ez[x, z - 1] -= tz(M[1, 1]); ez[x, z] += tz(M[1, 1])
    q = 0.25 * M[0, 1]
    ex[x - 1, z + 1] += tx(q); ex[x, z + 1] += tx(q)
    ex[x - 1, z - 1] -= tx(q); ex[x, z - 1] -= tx(q)
    q = 0.25 * M[1, 0]
    ez[x + 1, z - 1] += tz(q); ez[x + 1, z] += tz(q)
    ez[x - 1, z - 1] -= tz(q); ez[x - 1, z] -= tz(q)
    np.testing.assert_allclose(vx, ex, rtol=1e-14, atol=0)
    np.testing.assert_allclose(vz, ez, rtol=1e-14, atol=0)
    # receivers integrate dt * v / (dx dz) from the state at the start of
    # each step: zero at it = 0, the injected value at it = 1
    ux, uz = o.get_synthetic_data()
    assert np.all(ux[:, :, 0] == 0) and np.all(uz[:, :, 0] == 0)
    np.testing.assert_allclose(ux[0, 0, 1], dt * ex[8, 6] / (dx * dz), rtol=1e-15)
    np.testing.assert_allclose(uz[0, 0, 1], dt * ez[8, 6] / (dx * dz), rtol=1e-15)
    np.testing.assert_allclose(ux[0, 1, 1], dt * ex[7, 7] / (dx * dz), rtol=1e-15)


def test_adjoint_source_injection(oracle_lib):
    cfg = _tiny(nt=1)
    cfg["receivers"].update(nr=3, ix_receivers=[5, 5, 9], iz_receivers=[4, 4, 8])
    o = _grid_model(oracle_lib, cfg, rho=2.0)
    o.forward_simulate(0)  # snapshots (zero state at it = 0)
    syn = o.get_synthetic_data()
    assert all(np.all(s == 0) for s in syn)
    a = np.array([1.5, -0.25, 2.0])
    obs = -a.reshape(1, 3, 1)
    o.set_observed_data(obs, 2 * obs)
    o.calculate_l2_adjoint_sources()
    ax, az = o.get_adjoint_sources()
    assert np.array_equal(ax.ravel(), a) and np.array_equal(az.ravel(), 2 * a)
    o.reset_kernels()
    o.adjoint_simulate(0)
    vx, vz, *_ = o.get_state()
    dt, dx, dz, b = 0.01, 1.25, 0.8, 0.5
    k = dt * b / (dx * dz)
    np.testing.assert_allclose(vx[5, 4], k * (a[0] + a[1]), rtol=1e-15)
    np.testing.assert_allclose(vz[9, 8], k * 2 * a[2], rtol=1e-15)
    assert np.count_nonzero(vx) == 2 and np.count_nonzero(vz) == 2


# ------------------------------------------------------------ misfit, vectors
def test_misfit_is_half_dt_sum_of_squares(oracle_lib):
    cfg = small_config(**{"domain.nt": 50})
    o = oracle_lib.OracleModel(cfg)
    rng = np.random.default_rng(0)
    shape = (o.n_shots, o.nr, o.nt)
    s = [rng.normal(size=shape) for _ in range(4)]
    o.set_synthetic_data(s[0], s[1])
    o.set_observed_data(s[2], s[3])
    exact = math.fsum((0.5 * 0.00025 * ((s[2] - s[0]) ** 2 + (s[3] - s[1]) ** 2)).ravel()
                      .tolist())
    assert o.calculate_l2_misfit() == pytest.approx(exact, rel=1e-13)
    o.calculate_l2_adjoint_sources()
    ax, az = o.get_adjoint_sources()
    assert np.array_equal(ax, s[0] - s[2]) and np.array_equal(az, s[1] - s[3])


def test_model_vector_cell_average_and_gradient_cell_sum(oracle_lib):
    cfg = small_config(**{"basis.npx": 4, "basis.npz": 2})
    o = oracle_lib.OracleModel(cfg)
    nx, nz = o.nx, o.nz
    X = np.arange(nx, dtype=float)[:, None] * np.ones((1, nz))
    Z = np.ones((nx, 1)) * np.arange(nz, dtype=float)[None, :]
    o.set_parameter_fields(1000 + X + 100 * Z, 500 + 2 * X, 1500 + Z)
    m = o.get_model_vector()
    npb, nxib, nzib = 12, 5, 6
    nxf, nzf = 90 - 2 * nxib, 50 - 2 * nzib
    ncx, ncz = nxf // 4, nzf // 2
    n = ncx * ncz
    assert m.size == 3 * n == o.free_parameters
    for iz in (0, 3, ncz - 1):
        for ix in (0, 5, ncx - 1):
            p = ix + iz * ncx  # x-fastest parameter index
            gx = npb + nxib + 4 * ix + 1.5  # mean of the 4 x-indices
            gz = npb + nzib + 2 * iz + 0.5
            assert m[p] == pytest.approx(1000 + gx + 100 * gz, rel=1e-14)
            assert m[p + n] == pytest.approx(500 + 2 * gx, rel=1e-14)
            assert m[p + 2 * n] == pytest.approx(1500 + gz, rel=1e-14)
    # set_model_vector fills whole cells, leaving the rest of the grid
    m2 = np.arange(3 * n, dtype=float) + 1.0
    o.set_model_vector(m2)
    vp, vs, rho = o.get_parameter_fields()
    assert vp[npb + nxib + 4 * 3 + 2, npb + nzib + 2 * 1 + 1] == m2[3 + 1 * ncx]
    assert vp[0, 0] == 1000.0
    np.testing.assert_array_equal(o.get_model_vector(), m2)


def test_velocity_kernel_map_is_the_chain_rule(oracle_lib):
    """vp/vs/rho kernels from Lame kernels by the chain rule of
    lambda = rho (vp^2 - 2 vs^2), mu = rho vs^2 (numerical Jacobian)."""
    cfg = _tiny(nt=1)
    o = oracle_lib.OracleModel(cfg)
    rng = np.random.default_rng(1)
    shape = (o.nx, o.nz)
    vp = 2000 + 100 * rng.random(shape)
    vs = 800 + 50 * rng.random(shape)
    rho = 1500 + 80 * rng.random(shape)
    lam_k = rng.normal(size=shape)
    mu_k = rng.normal(size=shape)
    rho_k = rng.normal(size=shape)

    def lame(vp, vs, rho):
        return rho * (vp ** 2 - 2 * vs ** 2), rho * vs ** 2, rho

    h = 1e-3
    jac = []
    for k, base in enumerate((vp, vs, rho)):
        args_p = [vp, vs, rho]
        args_m = [vp, vs, rho]
        args_p[k] = base + h
        args_m[k] = base - h
        lp, mp, rp = lame(*args_p)
        lmn, mmn, rmn = lame(*args_m)
        jac.append(((lp - lmn) / (2 * h), (mp - mmn) / (2 * h), (rp - rmn) / (2 * h)))
    expect = [jac[k][0] * lam_k + jac[k][1] * mu_k + jac[k][2] * rho_k for k in range(3)]
    got = _map_with_oracle(oracle_lib, cfg, (vp, vs, rho), (lam_k, mu_k, rho_k))
    for g, e in zip(got, expect):
        np.testing.assert_allclose(g, e, rtol=1e-6, atol=1e-9 * np.abs(e).max())


def test_correlation_is_the_strain_correlation(oracle_lib):
    """One correlation step of the adjoint (forward snapshot x adjoint state)
    against the textbook sensitivity densities: strains from stresses by
    solving the plane-strain Hooke system [[l + 2m, l], [l, l + 2m]] e = s
    (numpy), e_xz = s_xz / (2 m); then
        K_lambda += w tr(e) tr(e'),   K_mu += w 2 e : e',
        K_rho    -= w v . v'
    with w = snapshot_interval * dt, on the kernel region only.  This pins the
    lambda, mu and rho correlations point by point, and with the chain-rule
    pin of the velocity map the vp and rho gradient blocks."""
    cfg = _tiny(nx_inner=16, nz_inner=12, nt=3, **{"inversion.snapshot_interval": 3,
                                                    "domain.nx_inner_boundary": 2,
                                                    "domain.nz_inner_boundary": 3})
    o = oracle_lib.OracleModel(cfg)
    rng = np.random.default_rng(11)
    shape = (o.nx, o.nz)
    vp = 2.5 + rng.random(shape)
    vs = 0.8 + 0.4 * rng.random(shape)
    rho = 1.5 + rng.random(shape)
    o.set_parameter_fields(vp, vs, rho)
    fwd = [rng.normal(size=shape) for _ in range(5)]
    adj = [rng.normal(size=shape) for _ in range(5)]
    o.set_snapshot(0, 0, *fwd)
    o.set_state(*adj)
    o.reset_kernels()
    o.correlate_step(0, 0)
    lam_k, mu_k, rho_k = o.get_lame_kernels()

    lam = rho * (vp ** 2 - 2 * vs ** 2)
    mu = rho * vs ** 2
    hooke = np.empty(shape + (2, 2))
    hooke[..., 0, 0] = hooke[..., 1, 1] = lam + 2 * mu
    hooke[..., 0, 1] = hooke[..., 1, 0] = lam

    def strains(txx, tzz, txz):
        e = np.linalg.solve(hooke, np.stack([txx, tzz], axis=-1)[..., None])[..., 0]
        return e[..., 0], e[..., 1], txz / (2 * mu)

    exx, ezz, exz = strains(*fwd[2:])
    axx, azz, axz = strains(*adj[2:])
    w = 3 * cfg["domain"]["dt"]
    region = np.zeros(shape, bool)
    d = cfg["domain"]
    region[d["nx_inner_boundary"]:d["nx_inner"] - d["nx_inner_boundary"],
           d["nz_inner_boundary"]:d["nz_inner"] - d["nz_inner_boundary"]] = True
    expect = {
        "lambda": w * (exx + ezz) * (axx + azz),
        "mu": w * 2 * (exx * axx + ezz * azz + 2 * exz * axz),
        "rho": -w * (fwd[0] * adj[0] + fwd[1] * adj[1]),
    }
    for name, got in (("lambda", lam_k), ("mu", mu_k), ("rho", rho_k)):
        e = np.where(region, expect[name], 0.0)
        np.testing.assert_allclose(got, e, rtol=1e-9, atol=1e-13 * np.abs(e).max(),
                                   err_msg=name)
        assert np.all(got[~region] == 0), name


def _map_with_oracle(oracle_lib, cfg, fields, kernels):
    """The oracle's map_kernels_to_velocity applied to given Lame kernels."""
    o = oracle_lib.OracleModel(cfg)
    o.set_parameter_fields(*fields)
    o.set_lame_kernels(*kernels)
    o.map_kernels_to_velocity()
    return o.get_kernels()


# ------------------------------------------------------------ gradient physics
@pytest.mark.slow
def test_gradient_matches_central_finite_differences(oracle_lib):
    cfg = small_config(**{"domain.nt": 600})
    o = oracle_lib.OracleModel(cfg)
    vp, vs, rho = o.get_parameter_fields()
    nx, nz = vp.shape
    ix = np.arange(nx)[:, None]
    iz = np.arange(nz)[None, :]

    def bump(cx, cz, r):
        return np.exp(-((ix - cx) ** 2 + (iz - cz) ** 2) / (2 * r * r))

    true = (vp * (1 + 0.05 * bump(50, 40, 5)), vs * (1 - 0.05 * bump(60, 35, 5)),
            rho * (1 + 0.05 * bump(70, 45, 5)))
    o.set_parameter_fields(*true)
    for s in range(o.n_shots):
        o.forward_simulate(s)
    o.set_observed_data(*o.get_synthetic_data())

    def misfit(fields):
        o.set_parameter_fields(*fields)
        for s in range(o.n_shots):
            o.forward_simulate(s)
        return o.calculate_l2_misfit()

    start = (vp, vs, rho)
    misfit(start)
    o.calculate_l2_adjoint_sources()
    o.reset_kernels()
    for s in range(o.n_shots):
        o.adjoint_simulate(s)
    o.map_kernels_to_velocity()
    kernels = o.get_kernels()
    d = bump(55, 40, 6)
    for p, scale in enumerate((2000.0, 800.0, 1500.0)):
        eps = 1e-3 * scale
        fp = [f.copy() for f in start]
        fm = [f.copy() for f in start]
        fp[p] = fp[p] + eps * d
        fm[p] = fm[p] - eps * d
        fd = (misfit(fp) - misfit(fm)) / (2 * eps)
        g = float(np.sum(kernels[p] * d))
        # the reference gradient is not the exact discrete adjoint: it agrees
        # with finite differences to a few percent
        assert g == pytest.approx(fd, rel=0.10), (p, g, fd)


def test_oracle_is_thread_count_independent(oracle_lib):
    cfg = small_config(**{"domain.nt": 120})
    out = []
    for threads in (1, 3):
        oracle_lib.OracleModel.set_threads(threads)
        o = oracle_lib.OracleModel(cfg)
        for s in range(o.n_shots):
            o.forward_simulate(s)
        out.append(o.get_synthetic_data())
    oracle_lib.OracleModel.set_threads(min(16, os.cpu_count() or 1))
    for a, b in zip(*out):
        assert np.array_equal(a, b)
