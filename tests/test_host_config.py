"""Host-side logic of the product library, without a GPU: INI reading with
the reference's quirks, validation errors, setup arrays, and the C-ABI
symbol table.
"""
import os
import re

import numpy as np
import pytest

from conftest import DEFAULT_INI, ROOT, default_config, small_config

psvwave_amd = pytest.importorskip("psvwave_amd")


def _write(tmp_path, text, name="c.ini"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "psvwave.h")).read()
    names = set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(psv_\w+)\s*\(", header,
                           re.M))
    assert len(names) > 40
    from psvwave_amd import _lib
    lib = _lib.load()  # the package's loader (torch's HIP runtime first)
    assert os.path.samefile(_lib.LIB_PATH, os.path.join(ROOT, "psvwave_amd",
                                                        "libpsvwave_amd.so"))
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert names <= set(_lib._SIGNATURES), names - set(_lib._SIGNATURES)


def test_stream_pressure_reads_the_queue_limit(monkeypatch):
    """psv_stream_pressure makes no HIP call: with no model alive it reports
    no library streams, and the hardware-queue limit from GPU_MAX_HW_QUEUES
    (4 when unset or not a positive number)."""
    import ctypes as C
    from psvwave_amd import _lib
    lib = _lib.load()
    live, queues = C.c_int(-1), C.c_int(-1)

    def read():
        _lib.check(lib.psv_stream_pressure(C.byref(live), C.byref(queues)))
        return live.value, queues.value

    monkeypatch.delenv("GPU_MAX_HW_QUEUES", raising=False)
    assert read() == (0, 4)
    monkeypatch.setenv("GPU_MAX_HW_QUEUES", "7")
    assert read() == (0, 7)
    for bad in ("0", "-3", "x", "5x", ""):
        monkeypatch.setenv("GPU_MAX_HW_QUEUES", bad)
        assert read() == (0, 4), bad
    with pytest.raises(ValueError):
        _lib.check(lib.psv_stream_pressure(None, C.byref(queues)))


def test_default_ini_parses_like_the_explicit_values():
    """Restates tests/test_constructor_comparison.cpp at the configuration
    level: the INI file and the explicit values give the same model."""
    a = psvwave_amd.Config(DEFAULT_INI)
    b = psvwave_amd.Config(default_config())
    assert a.info == b.info
    assert a.lists()[0] == b.lists()[0] == [[0, 1, 2, 3]]
    for x, y in zip(a.setup(), b.setup()):
        assert np.array_equal(x, y)
    i = a.info
    assert (i["nx"], i["nz"], i["snapshots"], i["free_parameters"]) == (250, 150, 800, 32400)
    assert i["dx"] == 1.249 and i["dt"] == 0.00025


def test_setup_arrays_match_the_oracle(oracle_lib):
    for cfg in (default_config(), small_config(**{"basis.npx": 2})):
        c = psvwave_amd.Config(cfg)
        o = oracle_lib.OracleModel(cfg)
        stf, mom, taper = c.setup()
        assert np.array_equal(stf, o.get_stf())
        assert np.array_equal(mom, o.get_moment())
        assert np.array_equal(taper, o.get_taper())


def test_write_dictionary_round_trip(tmp_path):
    d = psvwave_amd.get_dictionary()
    path = str(tmp_path / "auto.conf")
    psvwave_amd.write_dictionary(path, d)
    c = psvwave_amd.Config(path)
    which, (ixs, izs), (ixr, izr) = c.lists()
    assert which == [[0, 1]]
    assert list(ixs) == [175, 375] and list(izs) == [75, 75]
    assert list(ixr) == [75, 175, 275, 375]
    assert c.info["nt"] == 2000 and c.info["nx"] == 450
    assert c.info == psvwave_amd.Config(d).info


INI = """[domain]
nt = {nt};   trailing junk after a semicolon with no space
nx_inner = 40 ; inline comment
nz_inner = 30
nx_inner_boundary = 2
nz_inner_boundary = 3
dx = 1.5e0junk
dz = 2.
dt = 0.001
[boundary]; comment after a section header
np_boundary = 5
np_factor = 0.015
[MEDIUM]
Scalar_Rho = 1500
scalar_vp = 2000
scalar_vs = 800
[basis]
npx = 1
npz = 1
[sources]
peak_frequency = 50
source_timeshift = 0
delay_cycles_per_shot = 3 // over f
n_sources = 3
n_shots = 2
moment_angles = {{ 90 , 45.5, 180 }} ;
ix_sources = {{10, 20.9, 30}}
iz_sources = {{5,5,5}}
which_source_to_fire_in_which_shot = {which}
[receivers]
nr = 2
ix_receivers = {{1, 2}}
iz_receivers = {{3, 4}}
[inversion]
snapshot_interval = 10
[output]
observed_data_folder = ./obs dir
stf_folder = .
"""


def test_ini_quirks(tmp_path):
    c = psvwave_amd.Config(_write(tmp_path, INI.format(nt="0x64", which="{{0, 2}, {1}};")))
    i = c.info
    assert i["nt"] == 100  # strtol base 0
    assert i["nx_inner"] == 40 and i["dx"] == 1.5 and i["dz"] == 2.0
    assert i["scalar_rho"] == 1500  # keys are case-insensitive
    assert i["delay_cycles_per_shot"] == 3
    which, (ixs, _), (ixr, izr) = c.lists()
    assert which == [[0, 2], [1]]
    assert list(ixs) == [15, 25, 35]  # 20.9 truncates to 20, +np
    assert list(ixr) == [6, 7] and list(izr) == [8, 9]
    # nested list without the closing "};" (as configparser writes it)
    c2 = psvwave_amd.Config(_write(tmp_path, INI.format(nt=100, which="{{0,2},{1}}")))
    assert c2.lists()[0] == [[0, 2], [1]]


def test_octal_prefix_and_missing_values(tmp_path):
    c = psvwave_amd.Config(_write(tmp_path, INI.format(nt="0100", which="{{0,2},{1}}")))
    assert c.info["nt"] == 64  # base-0 strtol reads a leading 0 as octal
    with pytest.raises(ValueError, match="Missing configuration value"):
        psvwave_amd.Config(_write(tmp_path, INI.replace("nr = 2\n", "").format(
            nt=100, which="{{0,2},{1}}")))


def test_validation_errors_match_the_reference(tmp_path):
    with pytest.raises(ValueError, match=r"Can't load .ini file."):
        psvwave_amd.Config(str(tmp_path / "nope.ini"))
    bad = small_config()
    bad["sources"]["ix_sources"] = [1, 2]
    with pytest.raises(ValueError, match="Dimension mismatch between n_sources"):
        psvwave_amd.Config(bad)
    bad = small_config()
    bad["sources"]["n_shots"] = 3
    with pytest.raises(ValueError, match="Mismatch between n_shots"):
        psvwave_amd.Config(bad)
    bad = small_config()
    bad["sources"]["which_source_to_fire_in_which_shot"] = [[0], [1]]
    with pytest.raises(ValueError, match="Mismatch between n_sources and sources.which"):
        psvwave_amd.Config(bad)
    bad = small_config()
    bad["receivers"]["nr"] = 3
    with pytest.raises(ValueError, match="Mismatch between nr"):
        psvwave_amd.Config(bad)
    # guards where the reference has undefined behaviour
    bad = small_config(**{"basis.npx": 7})
    with pytest.raises(ValueError, match="basis"):
        psvwave_amd.Config(bad)
    bad = small_config()
    bad["sources"]["ix_sources"] = [20, 51, 200]
    with pytest.raises(ValueError, match="outside the grid"):
        psvwave_amd.Config(bad)
    bad = small_config()
    bad["receivers"]["iz_receivers"] = [40, 40, 41, 3, 40, 40, 400]
    with pytest.raises(ValueError, match="outside the grid"):
        psvwave_amd.Config(bad)
    bad = small_config(**{"inversion.snapshot_interval": 0})
    with pytest.raises(ValueError, match="snapshot_interval"):
        psvwave_amd.Config(bad)


def test_list_parsing_edge_cases(tmp_path):
    # "{}" reads as one zero, as atof("") does in the reference
    text = INI.format(nt=100, which="{{0,2},{1}}").replace(
        "nr = 2\nix_receivers = {1, 2}\niz_receivers = {3, 4}",
        "nr = 1\nix_receivers = {}\niz_receivers = { }")
    c = psvwave_amd.Config(_write(tmp_path, text))
    _, _, (ixr, izr) = c.lists()
    assert list(ixr) == [5] and list(izr) == [5]
    with pytest.raises(ValueError, match="closing"):
        psvwave_amd.Config(_write(tmp_path, INI.format(nt=100, which="{{0,2},{1}}")
                                  .replace("iz_sources = {5,5,5}", "iz_sources = 5,5,5")))


def test_reference_module_name():
    """``import psvWave`` (the reference's module, psvWave/__init__.py:6)
    gives this build's class and INI helpers."""
    import psvWave
    import psvwave_amd
    assert psvWave.fdModel is psvwave_amd.fdModel
    assert psvWave.get_dictionary() == psvwave_amd.get_dictionary()
    assert psvWave.write_dictionary is psvwave_amd.write_dictionary
    assert psvWave.from_dictionary is psvwave_amd.from_dictionary
