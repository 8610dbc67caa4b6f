"""Seeded synthetic workloads for benchmarks and large-grid tests.

``layered_faulted_config`` / ``layered_faulted_model`` build the single-model
benchmark workload: a 1700 x 350 inner grid (plus a 50-point taper, dx = dz
= 5 m, dt = 0.3 ms, nt = 6000, snapshot every 10 steps), a water layer over
40 layers with vp in 1500-5500 m/s, vs = vp / 1.73, rho = 310 vp^0.25, cut by
six normal faults; 8 shots of one 10 Hz Ricker source each and 300
receivers.  Everything is drawn from ``numpy.random.default_rng(seed)``.
"""
import math

import numpy as np

WATER_ROWS = 20


def layered_faulted_config(nt=6000, n_shots=8, nr=300, nx_inner=1700,
                           nz_inner=350, np_boundary=50, snapshot_interval=10):
    """Configuration dict (the layout of psvwave_amd.get_dictionary)."""
    ix_src = [100 + 214 * k for k in range(n_shots)]
    ix_rec = [int(math.floor(k * nx_inner / nr)) for k in range(nr)]
    return {
        "domain": dict(nt=nt, nx_inner=nx_inner, nz_inner=nz_inner,
                       nx_inner_boundary=0, nz_inner_boundary=WATER_ROWS,
                       dx=5.0, dz=5.0, dt=3e-4),
        "boundary": dict(np_boundary=np_boundary, np_factor=0.015),
        "medium": dict(scalar_rho=2000.0, scalar_vp=2500.0, scalar_vs=1445.0),
        "sources": dict(peak_frequency=10.0, n_sources=n_shots, n_shots=n_shots,
                        source_timeshift=0.0, delay_cycles_per_shot=0,
                        moment_angles=[45.0] * n_shots, ix_sources=ix_src,
                        iz_sources=[25] * n_shots,
                        which_source_to_fire_in_which_shot=[[k] for k in range(n_shots)]),
        "receivers": dict(nr=nr, ix_receivers=ix_rec, iz_receivers=[WATER_ROWS] * nr),
        "inversion": dict(snapshot_interval=snapshot_interval),
        "basis": dict(npx=1, npz=1),
        "output": dict(observed_data_folder=".", stf_folder="."),
    }


def layered_faulted_model(cfg, seed=2, faulted=True, n_layers=40, n_faults=6):
    """(vp, vs, rho) on the full (nx, nz) grid, reference layout.

    Inner rows 0..19 are water (vp 1500, vs 0).  Below, n_layers layers with
    sorted random interfaces and sorted vp; normal faults dipping 60 degrees
    shift their hanging wall down by 5-25 rows (the seafloor stays flat).
    The absorbing zone repeats the edge values.
    """
    d = cfg["domain"]
    nxi, nzi = d["nx_inner"], d["nz_inner"]
    npb = cfg["boundary"]["np_boundary"]
    rng = np.random.default_rng(seed)
    interfaces = np.sort(rng.choice(np.arange(WATER_ROWS + 1, nzi), n_layers - 1,
                                    replace=False))
    vps = np.sort(rng.uniform(1500.0, 5500.0, n_layers))
    layer_of_row = np.searchsorted(interfaces, np.arange(nzi), side="right")
    vp_col = vps[layer_of_row]
    vp = np.tile(vp_col, (nxi, 1))  # (nx_inner, nz_inner)
    faults = [(rng.uniform(100, 1600), int(rng.integers(5, 26)))
              for _ in range(n_faults)]
    if faulted:
        z = np.arange(nzi)
        x = np.arange(nxi)[:, None]
        for x0, throw in faults:
            xf = x0 + z / math.tan(math.radians(60.0))  # fault trace per row
            hanging = x > xf[None, :]
            src_rows = np.maximum(WATER_ROWS, z - throw)
            shifted = vp[:, src_rows]
            vp = np.where(hanging & (z[None, :] >= WATER_ROWS), shifted, vp)
    vp[:, :WATER_ROWS] = 1500.0
    vs = vp / 1.73
    vs[:, :WATER_ROWS] = 0.0
    rho = 310.0 * vp ** 0.25
    pad = ((npb, npb), (npb, npb))
    return tuple(np.ascontiguousarray(np.pad(a, pad, mode="edge"))
                 for a in (vp, vs, rho))


def algorithmic_bytes_per_shot(nx, nz, nt, si, kernel_points):
    """Algorithmic HBM bytes of one shot's forward + adjoint: 104 B per
    point-update (read vx, vz, txx, tzz, txz, lm, mu, b; write the five
    fields), 40 B per grid point per forward snapshot, 88 B per kernel-region
    point per adjoint correlation (read five snapshot values, read and write
    three kernels)."""
    n = nx * nz
    return n * nt * 208 + (nt // si) * (40 * n + 88 * kernel_points)


def gaussian_random_field(shape, corr_len, rng):
    """Zero-mean, unit-range smooth random field: white noise on a grid
    coarsened by corr_len / 4, Gaussian-smoothed (sigma = 4 coarse cells),
    interpolated back, then scaled to [0, 1]."""
    from scipy import ndimage
    f = max(1, int(corr_len // 4))
    coarse = (shape[0] // f + 3, shape[1] // f + 3)
    g = ndimage.gaussian_filter(rng.standard_normal(coarse), sigma=corr_len / f / 4.0,
                                mode="wrap")
    full = ndimage.zoom(g, f, order=1)[: shape[0], : shape[1]]
    lo, hi = full.min(), full.max()
    return (full - lo) / (hi - lo) if hi > lo else np.zeros(shape)


def monte_carlo_configs(n_models=512, seed=3, nt=3000, np_boundary=30, nr=32):
    """Configurations of the Monte-Carlo batch: nx_inner ~ U{200..800},
    nz_inner ~ U{100..400}, dx = dz = 5 m, dt = 0.5 ms, 10 Hz, 2 shots of
    one source each (at a quarter and three quarters of the width, depth 10),
    nr receivers evenly along the surface at depth 5."""
    rng = np.random.default_rng(seed)
    sizes = [(int(rng.integers(200, 801)), int(rng.integers(100, 401)))
             for _ in range(n_models)]
    cfgs = []
    for nxi, nzi in sizes:
        cfgs.append({
            "domain": dict(nt=nt, nx_inner=nxi, nz_inner=nzi, nx_inner_boundary=0,
                           nz_inner_boundary=0, dx=5.0, dz=5.0, dt=5e-4),
            "boundary": dict(np_boundary=np_boundary, np_factor=0.015),
            "medium": dict(scalar_rho=1800.0, scalar_vp=2400.0, scalar_vs=1200.0),
            "sources": dict(peak_frequency=10.0, n_sources=2, n_shots=2,
                            source_timeshift=0.0, delay_cycles_per_shot=0,
                            moment_angles=[45.0, 135.0],
                            ix_sources=[nxi // 4, (3 * nxi) // 4], iz_sources=[10, 10],
                            which_source_to_fire_in_which_shot=[[0], [1]]),
            "receivers": dict(nr=nr,
                              ix_receivers=[int((k + 0.5) * nxi / nr) for k in range(nr)],
                              iz_receivers=[5] * nr),
            "inversion": dict(snapshot_interval=10),
            "basis": dict(npx=1, npz=1),
            "output": dict(observed_data_folder=".", stf_folder="."),
        })
    return cfgs


def monte_carlo_model(cfg, rng, corr_len=20):
    """(vp, vs, rho) on the full grid: Gaussian random fields with
    correlation length corr_len points, vp in 1800-3000 m/s, vs/vp in
    0.45-0.55, rho in 1400-2200 kg/m3."""
    d, npb = cfg["domain"], cfg["boundary"]["np_boundary"]
    shape = (d["nx_inner"] + 2 * npb, d["nz_inner"] + 2 * npb)
    vp = 1800.0 + 1200.0 * gaussian_random_field(shape, corr_len, rng)
    vs = vp * (0.45 + 0.10 * gaussian_random_field(shape, corr_len, rng))
    rho = 1400.0 + 800.0 * gaussian_random_field(shape, corr_len, rng)
    return vp, vs, rho


def c4_config(nt=4000, n_shots=4, nr=64):
    """Configuration of the gradient-check workload: 600 x 300 inner grid
    plus a 40-point taper (680 x 380), dx = dz = 5 m, dt = 0.5 ms, 10 Hz,
    n_shots shots of one source each spread along the top at depth 10, nr
    receivers at depth 5, snapshot every 10 steps (with nt = 4000 and four
    shots the full snapshot store is 16.5 GB)."""
    nxi, nzi, npb = 600, 300, 40
    return {
        "domain": dict(nt=nt, nx_inner=nxi, nz_inner=nzi, nx_inner_boundary=10,
                       nz_inner_boundary=10, dx=5.0, dz=5.0, dt=5e-4),
        "boundary": dict(np_boundary=npb, np_factor=0.015),
        "medium": dict(scalar_rho=1800.0, scalar_vp=2400.0, scalar_vs=1200.0),
        "sources": dict(peak_frequency=10.0, n_sources=n_shots, n_shots=n_shots,
                        source_timeshift=0.0, delay_cycles_per_shot=0,
                        moment_angles=[45.0 + 30.0 * k for k in range(n_shots)],
                        ix_sources=[int((k + 0.5) * nxi / n_shots) for k in range(n_shots)],
                        iz_sources=[10] * n_shots,
                        which_source_to_fire_in_which_shot=[[k] for k in range(n_shots)]),
        "receivers": dict(nr=nr, ix_receivers=[int((k + 0.5) * nxi / nr) for k in range(nr)],
                          iz_receivers=[5] * nr),
        "inversion": dict(snapshot_interval=10),
        "basis": dict(npx=1, npz=1),
        "output": dict(observed_data_folder=".", stf_folder="."),
    }


def c4_models(cfg, seed=4):
    """(true, start) models of the gradient-check workload: the start model
    is a smooth Gaussian random field (vp 2200-2800 m/s, vs/vp 0.45-0.55,
    rho 1600-2000 kg/m3, correlation length 40 points); the true model adds
    a +-5 % perturbation with correlation length 12 points."""
    rng = np.random.default_rng(seed)
    d, npb = cfg["domain"], cfg["boundary"]["np_boundary"]
    shape = (d["nx_inner"] + 2 * npb, d["nz_inner"] + 2 * npb)
    vp = 2200.0 + 600.0 * gaussian_random_field(shape, 40, rng)
    vs = vp * (0.45 + 0.10 * gaussian_random_field(shape, 40, rng))
    rho = 1600.0 + 400.0 * gaussian_random_field(shape, 40, rng)
    start = (vp, vs, rho)
    true = tuple(f * (1.0 + 0.1 * (gaussian_random_field(shape, 12, rng) - 0.5))
                 for f in start)
    return true, start


def c1_config(nt=4000):
    """Configuration of the small CPU-runnable workload: 200 x 100 inner
    grid plus a 50-point taper (300 x 200), dx = dz = 1.25 m, dt = 0.25 ms,
    nt = 4000, snapshot every 10 steps; two shots of one 50 Hz source each
    (dip angles 90 and 180 degrees at inner (50, 90) and (150, 90)), 19
    receivers at inner depth 10.  np_factor, the inner boundaries (10, 20),
    the delay and the unused time shift are the reference test
    configuration's values (tests/test_configurations/
    default_testing_configuration.ini)."""
    return {
        "domain": dict(nt=nt, nx_inner=200, nz_inner=100, nx_inner_boundary=10,
                       nz_inner_boundary=20, dx=1.25, dz=1.25, dt=2.5e-4),
        "boundary": dict(np_boundary=50, np_factor=0.015),
        "medium": dict(scalar_rho=1500.0, scalar_vp=2000.0, scalar_vs=800.0),
        "sources": dict(peak_frequency=50.0, n_sources=2, n_shots=2,
                        source_timeshift=0.005, delay_cycles_per_shot=24,
                        moment_angles=[90.0, 180.0], ix_sources=[50, 150],
                        iz_sources=[90, 90],
                        which_source_to_fire_in_which_shot=[[0], [1]]),
        "receivers": dict(nr=19, ix_receivers=list(range(10, 200, 10)),
                          iz_receivers=[10] * 19),
        "inversion": dict(snapshot_interval=10),
        "basis": dict(npx=1, npz=1),
        "output": dict(observed_data_folder=".", stf_folder="."),
    }


def c1_models(cfg):
    """(true, start) models of the small workload on the full (nx, nz) grid:
    the start model is homogeneous (vp 2000, vs 800, rho 1500); the true
    model has vp and vs 10 % higher inside the circle of radius 15 grid
    points about the grid centre, (ix - 149.5)^2 + (iz - 99.5)^2 < 15^2."""
    d, npb = cfg["domain"], cfg["boundary"]["np_boundary"]
    nx, nz = d["nx_inner"] + 2 * npb, d["nz_inner"] + 2 * npb
    med = cfg["medium"]
    start = (np.full((nx, nz), med["scalar_vp"]), np.full((nx, nz), med["scalar_vs"]),
             np.full((nx, nz), med["scalar_rho"]))
    ix = np.arange(nx, dtype=np.float64)[:, None]
    iz = np.arange(nz, dtype=np.float64)[None, :]
    circle = ((ix - (nx - 1) / 2) ** 2 + (iz - (nz - 1) / 2) ** 2) < 15.0 ** 2
    true = (start[0] * (1 + 0.1 * circle), start[1] * (1 + 0.1 * circle), start[2].copy())
    return true, start
