import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")
DEFAULT_INI = os.path.join(GOLDEN, "default_testing_configuration.ini")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


def default_config():
    """The values of default_testing_configuration.ini as a dict, as the
    reference's explicit-argument test writes them
    (tests/test_variable_constructor.cpp:29-60)."""
    return {
        "domain": dict(nt=8000, nx_inner=200, nz_inner=100, nx_inner_boundary=10,
                       nz_inner_boundary=20, dx=1.249, dz=1.249, dt=0.00025),
        "boundary": dict(np_boundary=25, np_factor=0.015),
        "medium": dict(scalar_rho=1500.0, scalar_vp=2000.0, scalar_vs=800.0),
        "sources": dict(peak_frequency=50.0, n_sources=4, n_shots=1,
                        source_timeshift=0.005, delay_cycles_per_shot=24,
                        moment_angles=[90, 180, 90, 180],
                        ix_sources=[25, 75, 125, 175], iz_sources=[10, 10, 10, 10],
                        which_source_to_fire_in_which_shot=[[0, 1, 2, 3]]),
        "receivers": dict(nr=19, ix_receivers=list(range(10, 200, 10)),
                          iz_receivers=[90] * 19),
        "inversion": dict(snapshot_interval=10),
        "basis": dict(npx=1, npz=1),
        "output": dict(observed_data_folder=".", stf_folder="."),
    }


def small_config(**over):
    """A small two-shot configuration for fast tests; keyword overrides are
    'section.key' -> value."""
    cfg = {
        "domain": dict(nt=300, nx_inner=90, nz_inner=50, nx_inner_boundary=5,
                       nz_inner_boundary=6, dx=1.25, dz=1.1, dt=0.00025),
        "boundary": dict(np_boundary=12, np_factor=0.02),
        "medium": dict(scalar_rho=1500.0, scalar_vp=2000.0, scalar_vs=800.0),
        "sources": dict(peak_frequency=60.0, n_sources=3, n_shots=2,
                        source_timeshift=0.0, delay_cycles_per_shot=6,
                        moment_angles=[90, 45, 180],
                        ix_sources=[20, 51, 70], iz_sources=[10, 7, 12],
                        which_source_to_fire_in_which_shot=[[0, 2], [1]]),
        "receivers": dict(nr=7, ix_receivers=[5, 17, 30, 51, 52, 75, 88],
                          iz_receivers=[40, 40, 41, 3, 40, 40, 44]),
        "inversion": dict(snapshot_interval=5),
        "basis": dict(npx=1, npz=1),
        "output": dict(observed_data_folder=".", stf_folder="."),
    }
    for k, v in over.items():
        sec, key = k.split(".")
        cfg[sec][key] = v
    return cfg


def grid_coordinates(nx, nz, np_boundary, dx, dz):
    """get_coordinates(True) of the reference (src/psvWave.cpp:207-246)."""
    ix = np.arange(nx)[:, None] * np.ones((1, nz))
    iz = np.ones((nx, 1)) * np.arange(nz)[None, :]
    return (ix - np_boundary) * dx, (iz - np_boundary) * dz


def notebook2_models(nx, nz, np_boundary, dx, dz):
    """Target and start models of Notebook 2, cell 3."""
    IX, IZ = grid_coordinates(nx, nz, np_boundary, dx, dz)
    bg = np.array([2000, 800, 1500])
    vp0 = np.ones((nx, nz)) * bg[0]
    vs0 = np.ones((nx, nz)) * bg[1]
    rho0 = np.ones((nx, nz)) * bg[2]
    zm = (IZ.max() + IZ.min()) / 2
    xm = (IX.max() + IX.min()) / 2
    c = ((IX - xm) ** 2 + (IZ - zm) ** 2) ** 0.5 < 15
    vs = vs0 * (1 - 0.1 * c)
    vp = vp0 * (1 - 0.1 * c)
    xm = (IX.max() + IX.min()) / 6
    c = ((IX - xm) ** 2 + (IZ - zm) ** 2) ** 0.5 < 15
    vs = vs * (1 + 0.1 * c)
    rho = rho0 * (1 + 0.1 * c)
    xm = (IX.max() + IX.min()) * 5 / 6
    c = ((IX - xm) ** 2 + (IZ - zm) ** 2) ** 0.5 < 15
    vp = vp * (1 + 0.1 * c)
    rho = rho * (1 + 0.1 * c)
    return (vp, vs, rho), (vp0, vs0, rho0)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    num = np.linalg.norm((a - b).ravel())
    return num / den if den > 0 else num


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    pyoracle.OracleModel.set_threads(min(16, os.cpu_count() or 1))
    return pyoracle
