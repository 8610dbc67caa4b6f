"""Edge cases of the HIP path, each checked against the oracle (bitwise) or
against the reference's documented behaviour."""
import os

import numpy as np
import pytest

from conftest import DEFAULT_INI, small_config

pytestmark = pytest.mark.gpu


def _product(cfg, variant=0, remap=True):
    import psvwave_amd
    m = psvwave_amd.fdModel(cfg)
    m.set_step_variant(variant, remap)
    return m


def _random_fields(nx, nz, seed):
    rng = np.random.default_rng(seed)
    vp = 2000 + 150 * rng.random((nx, nz))
    vs = 800 + 80 * rng.random((nx, nz))
    rho = 1500 + 100 * rng.random((nx, nz))
    return vp, vs, rho


def _run_both(oracle_lib, cfg, variant=0, remap=True, seed=0):
    """forward all shots on random media, misfit against perturbed data,
    adjoint all shots; returns (product results, oracle results)."""
    res = []
    for which in ("product", "oracle"):
        m = _product(cfg, variant, remap) if which == "product" else \
            oracle_lib.OracleModel(cfg)
        f = _random_fields(m.nx, m.nz, seed)
        m.set_parameter_fields(*f)
        for s in range(m.n_shots):
            m.forward_simulate(s)
        ux, uz = m.get_synthetic_data()
        m.set_observed_data(ux * 0.9 + 1e-9, uz * 1.1)
        m.calculate_l2_misfit()
        m.calculate_l2_adjoint_sources()
        m.reset_kernels()
        for s in range(m.n_shots):
            m.adjoint_simulate(s)
        m.map_kernels_to_velocity()
        res.append({"syn": (ux, uz), "lame": m.get_lame_kernels(),
                    "kernels": m.get_kernels(), "misfit": m.misfit,
                    "grad": m.get_gradient_vector()})
    return res


def _assert_same(p, o):
    for a, b in zip(p["syn"], o["syn"]):
        assert np.array_equal(a, b)
    for key in ("lame", "kernels"):
        for a, b in zip(p[key], o[key]):
            assert np.array_equal(a, b), key
    assert np.array_equal(p["grad"], o["grad"])
    assert p["misfit"] == pytest.approx(o["misfit"], rel=1e-12)


def test_sources_and_receivers_on_tile_boundaries(oracle_lib):
    # np = 12: inner ix 52 is grid column 64 (first column of tile 1), inner
    # iz 4 is grid row 16 (first row of a 64x8 and of a 64x16 tile)
    cfg = small_config(**{"domain.nt": 200})
    cfg["sources"].update(ix_sources=[52, 52, 51], iz_sources=[4, 12, 4],
                          moment_angles=[30, 60, 90])
    cfg["receivers"].update(ix_receivers=[52, 51, 52, 52, 63, 0, 89],
                            iz_receivers=[4, 4, 3, 4, 11, 0, 49])
    for variant in range(29):
        p, o = _run_both(oracle_lib, cfg, variant)
        _assert_same(p, o)


def test_coincident_sources_and_receivers_keep_reference_order(oracle_lib):
    cfg = small_config(**{"domain.nt": 160})
    # two sources at one point in one shot, three receivers at one point
    cfg["sources"].update(ix_sources=[30, 30, 31], iz_sources=[20, 20, 20],
                          moment_angles=[10, 200, 77],
                          which_source_to_fire_in_which_shot=[[1, 0, 2], []],
                          n_shots=2)
    cfg["receivers"].update(nr=5, ix_receivers=[40, 40, 40, 3, 40],
                            iz_receivers=[30, 30, 30, 4, 31])
    for variant in (0, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25,
                    26, 27, 28):
        p, o = _run_both(oracle_lib, cfg, variant)
        _assert_same(p, o)


@pytest.mark.parametrize("nx_inner,nz_inner,np_b", [(37, 13, 0), (3, 3, 2), (130, 21, 7)])
def test_odd_grid_sizes(oracle_lib, nx_inner, nz_inner, np_b):
    cfg = small_config(**{"domain.nt": 90, "domain.nx_inner": nx_inner,
                          "domain.nz_inner": nz_inner, "boundary.np_boundary": np_b,
                          "domain.nx_inner_boundary": 1,
                          "domain.nz_inner_boundary": 1})
    cfg["sources"].update(n_sources=1, n_shots=1, ix_sources=[nx_inner // 2],
                          iz_sources=[nz_inner // 2], moment_angles=[33],
                          which_source_to_fire_in_which_shot=[[0]])
    cfg["receivers"].update(nr=2, ix_receivers=[0, nx_inner - 1],
                            iz_receivers=[nz_inner - 1, 0])
    if np_b == 0:  # a source needs one point of margin around it
        cfg["sources"]["ix_sources"] = [max(1, nx_inner // 2)]
    for variant in (0, 2, 3, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23,
                    24, 25, 26, 27, 28):
        p, o = _run_both(oracle_lib, cfg, variant)
        _assert_same(p, o)


def test_wavefront_tail_values_match_bitwise(oracle_lib):
    """Ahead of the physical front the stencil leaves values down to the
    subnormal range (|v| < 2^-960 at hundreds of points of this run), where
    the fast division must hand over to the IEEE one."""
    cfg = small_config(**{"domain.nt": 160, "domain.nx_inner": 400,
                          "domain.nz_inner": 24, "domain.nx_inner_boundary": 1,
                          "domain.nz_inner_boundary": 1, "boundary.np_boundary": 4})
    cfg["sources"].update(n_sources=1, n_shots=1, ix_sources=[10], iz_sources=[12],
                          moment_angles=[40], which_source_to_fire_in_which_shot=[[0]])
    cfg["receivers"].update(nr=2, ix_receivers=[100, 300], iz_receivers=[12, 12])
    p, o = _run_both(oracle_lib, cfg, 8)
    far = np.abs(o["syn"][0][0, 1])  # the far receiver's trace
    assert far[far > 0].min() < 2.0 ** -960
    _assert_same(p, o)
    for variant in (0, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26,
                    27, 28):
        _assert_same(_run_product_only(cfg, variant, True), o)


def test_variants_and_xcd_mapping_agree_bitwise(oracle_lib):
    cfg = small_config(**{"domain.nt": 150, "domain.nx_inner": 300,
                          "domain.nz_inner": 70})
    cfg["sources"]["ix_sources"] = [20, 151, 270]
    cfg["receivers"]["ix_receivers"] = [5, 17, 130, 151, 152, 275, 288]
    p, o = _run_both(oracle_lib, cfg, 0, True)
    _assert_same(p, o)
    for variant in range(29):
        for remap in (True, False):
            _assert_same(_run_product_only(cfg, variant, remap), o)


def _run_product_only(cfg, variant, remap):
    m = _product(cfg, variant, remap)
    f = _random_fields(m.nx, m.nz, 0)
    m.set_parameter_fields(*f)
    for s in range(m.n_shots):
        m.forward_simulate(s)
    ux, uz = m.get_synthetic_data()
    m.set_observed_data(ux * 0.9 + 1e-9, uz * 1.1)
    m.calculate_l2_misfit()
    m.calculate_l2_adjoint_sources()
    m.reset_kernels()
    for s in range(m.n_shots):
        m.adjoint_simulate(s)
    m.map_kernels_to_velocity()
    return {"syn": (ux, uz), "lame": m.get_lame_kernels(), "kernels": m.get_kernels(),
            "misfit": m.misfit, "grad": m.get_gradient_vector()}


def test_basis_functions_and_kernel_accumulation(oracle_lib):
    cfg = small_config(**{"domain.nt": 100, "basis.npx": 4, "basis.npz": 2})
    m = _product(cfg, 6)
    o = oracle_lib.OracleModel(cfg)
    for mod in (m, o):
        mod.set_parameter_fields(*_random_fields(mod.nx, mod.nz, 3))
        v = mod.get_model_vector()
        mod.set_model_vector(v * 1.01)
        for s in range(mod.n_shots):
            mod.forward_simulate(s)
        ux, uz = mod.get_synthetic_data()
        mod.set_observed_data(uz, ux)
        mod.calculate_l2_adjoint_sources()
        mod.reset_kernels()
        # kernels accumulate across adjoint runs until reset_kernels
        mod.adjoint_simulate(0)
        mod.adjoint_simulate(0)
        mod.adjoint_simulate(1)
        mod.map_kernels_to_velocity()
    assert np.array_equal(m.get_model_vector(), o.get_model_vector())
    for a, b in zip(m.get_parameter_fields(), o.get_parameter_fields()):
        assert np.array_equal(a, b)
    for a, b in zip(m.get_lame_kernels(), o.get_lame_kernels()):
        assert np.array_equal(a, b)
    assert np.array_equal(m.get_gradient_vector(), o.get_gradient_vector())


def test_store_fields_false_and_error_paths():
    cfg = small_config(**{"domain.nt": 40})
    m = _product(cfg)
    with pytest.raises(ValueError, match="i_shot"):
        m.forward_simulate(2)
    with pytest.raises(ValueError, match="i_shot"):
        m.forward_simulate(-1)
    with pytest.raises(ValueError, match="stored forward"):
        m.adjoint_simulate(0)  # no snapshots were ever stored
    m.forward_simulate(0, store_fields=False)
    assert np.any(m.get_synthetic_data()[0] != 0)
    with pytest.raises(ValueError, match="right shape"):
        m.set_parameter_fields(np.zeros((3, 3)), np.zeros((3, 3)), np.zeros((3, 3)))
    with pytest.raises(ValueError, match="right shape"):
        m.set_observed_data(np.zeros((1, 2, 3)), np.zeros((1, 2, 3)))
    with pytest.raises(ValueError, match="free_parameters"):
        m.set_model_vector(np.zeros(5))
    # nt not a multiple of snapshot_interval: storing would overrun the
    # snapshot buffer in the reference; here it is an error
    bad = _product(small_config(**{"domain.nt": 43}))
    bad.forward_simulate(0, store_fields=False)
    with pytest.raises(ValueError, match="Snapshot interval"):
        bad.forward_simulate(0, store_fields=True)


def test_geometry_getters_follow_the_reference():
    import psvwave_amd
    m = psvwave_amd.fdModel(DEFAULT_INI)
    assert m.get_extent() == (-25, 1.249 * 225, -25, 1.249 * 125)
    assert m.get_extent(False) == (0, 1.249 * 200, 0, 1.249 * 100)
    IX, IZ = m.get_coordinates(True)
    assert IX.shape == (250, 150) and IX[0, 0] == -25 * 1.249 and IZ[3, 7] == (7 - 25) * 1.249
    ix, iz = m.get_coordinates(False)
    assert ix[5, 6] == 5 and iz[5, 6] == 6
    x, z = m.get_sources()
    np.testing.assert_allclose(x, np.array([25, 75, 125, 175]) * 1.249)
    x, z = m.get_sources(False, True)
    assert list(x) == [50, 100, 150, 200] and list(z) == [35] * 4
    x, z = m.get_receivers(False, False)
    assert list(x) == list(range(10, 200, 10)) and list(z) == [90] * 19
    assert m.which_source_to_fire_in_which_shot == [[0, 1, 2, 3]]
    assert (m.nt, m.nx, m.nz, m.snapshots, m.free_parameters) == (8000, 250, 150, 800, 32400)


def test_text_formats_round_trip(tmp_path):
    cfg = small_config(**{"domain.nt": 30})
    cfg["output"] = {"observed_data_folder": str(tmp_path), "stf_folder": str(tmp_path)}
    m = _product(cfg)
    for s in range(m.n_shots):
        m.forward_simulate(s)
    m.write_receivers()
    m.write_sources()
    lines = open(tmp_path / "rtf_ux0.txt").read().split("\n")
    assert lines[0] == "" and len(lines) == m.nr + 1  # a newline before each receiver
    m2 = _product(cfg)
    m2.load_receivers()
    for a, b in zip(m2.get_observed_data(), m.get_synthetic_data()):
        assert np.array_equal(a, b)  # 25 significant digits round-trip doubles
    stf, _, _ = m.get_setup()
    txt = open(tmp_path / "sources_shot_0.txt").read().split("\n")
    assert len(txt) == 3 and float(txt[1].split()[5]) == stf[0, 5]
    (tmp_path / "bad.txt").write_text("1 2 3")
    with pytest.raises(ValueError, match="Not enough data"):
        m.load_vector(str(tmp_path / "bad.txt"), False)
    n = 3 * m.nx_free_parameters * m.nz_free_parameters
    (tmp_path / "v.txt").write_text(" ".join(str(float(i)) for i in range(n)) + "\n")
    v = m.load_vector(str(tmp_path / "v.txt"), False)
    assert np.array_equal(v, np.arange(n, dtype=float))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        (tmp_path / "snapshots").mkdir()
        m.forward_simulate(0, store_fields=False, output_wavefields=True)
        files = sorted(os.listdir(tmp_path / "snapshots"))
        assert files[:2] == ["vx00000.txt", "vx00010.txt"] and len(files) == 6
        rows = open(tmp_path / "snapshots" / "vz00010.txt").read().strip("\n").split("\n")
        assert len(rows) == m.nx and len(rows[0].split()) == m.nz
        m.write_kernels()
        assert os.path.exists(tmp_path / "kernel_vp.txt")
    finally:
        os.chdir(cwd)


def test_graph_replays_match_eager_mode(oracle_lib):
    """Each time loop captured once and replayed: every replay, with eager
    work (parameter changes, misfit, reset) between replays, must equal the
    eager result bit for bit."""
    cfg = small_config(**{"domain.nt": 125})  # odd: two-step tail launch
    eager = _product(cfg, 0)
    graph = _product(cfg, 6)
    graph.set_graph_mode(True)
    for rep in range(3):
        fields = _random_fields(eager.nx, eager.nz, rep)
        out = []
        for m in (eager, graph):
            m.set_parameter_fields(*fields)
            for s in range(m.n_shots):
                m.forward_simulate(s)
            ux, uz = m.get_synthetic_data()
            m.set_observed_data(ux * 0.5, uz * 2.0)
            m.calculate_l2_misfit()
            m.calculate_l2_adjoint_sources()
            m.reset_kernels()
            for s in range(m.n_shots):
                m.adjoint_simulate(s)
            m.map_kernels_to_velocity()
            out.append((m.get_synthetic_data(), m.get_lame_kernels(), m.misfit))
        (se, ke, me), (sg, kg, mg) = out
        for a, b in zip(se + ke, sg + kg):
            assert np.array_equal(a, b), f"replay {rep} differs"
        assert me == mg


def test_step_timing_modes_leave_results_unchanged(oracle_lib):
    """Kernel-timestamp timing (stride > 0), loop bracketing (stride < 0) and
    no timing give the same seismograms; the counters cover the time steps."""
    import psvwave_amd
    cfg = small_config(**{"domain.nt": 120})
    out = []
    for stride in (0, 7, -1):
        m = psvwave_amd.fdModel(cfg)
        m.set_timing(stride)
        m.reset_step_timing()
        for s in range(m.n_shots):
            m.forward_simulate(s, store_fields=False)
        ms, timed, launches = m.step_timing()
        assert launches == m.n_shots * 120
        if stride == 0:
            assert timed == 0
        elif stride < 0:
            assert timed == launches and ms > 0
        else:
            assert 0 < timed < launches and ms > 0
        out.append(m.get_synthetic_data())
    for a, b in zip(out[1:], out[:-1]):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
